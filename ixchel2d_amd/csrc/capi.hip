// Host side of the C-ABI (include/ixchel2d.h): device state, buffer
// rotation, phase launches, the SIMPLE control loop and the fused time step.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/ixchel2d.h"
#include "simple_device.hpp"
#include "adjoint_device.hpp"
#include "sweep_launch.hpp"

using namespace ixh;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(IXH_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                \
  } while (0)

#define CHK(expr)           \
  do {                      \
    int rc_ = (expr);       \
    if (rc_ != 0) return rc_; \
  } while (0)

// a kernel launch on the state's stream; with timing events set, the
// kernel's own start and end are recorded into them
#define IXH_LAUNCH(s, kern, grid, block, ...)                                                      \
  do {                                                                                             \
    if ((s)->t_start)                                                                              \
      hipExtLaunchKernelGGL(kern, grid, block, 0, (s)->launch, (s)->t_start, (s)->t_stop, 0,     \
                            __VA_ARGS__);                                                          \
    else                                                                                           \
      hipLaunchKernelGGL(kern, grid, block, 0, (s)->launch, __VA_ARGS__);                          \
  } while (0)

enum KernelId {
  K_U_Y = 0, K_U_X, K_V_Y, K_V_X, K_P_Y, K_P_X, K_CORR, K_T_Y, K_T_X, K_RES, K_FIN,
  K_L1, K_L2, K_L3, K_L4, K_L5, K_L6,  // the 6 launches of the fused step
  K_NUM
};
const char *kKernelNames[K_NUM] = {"sweep_u_y", "sweep_u_x", "sweep_v_y", "sweep_v_x", "sweep_p_y",
                                   "sweep_p_x", "correct",   "sweep_t_y", "sweep_t_x", "residual_u",
                                   "finalize",  "multi_v_y+residual", "multi_v_x+fold", "step_sweep_p_y",
                                   "step_p_x+correct", "multi_t_y+u_y", "multi_t_x+u_x"};
constexpr int kParts = 6;  // launches of one fused step

constexpr int kNbuf = 4;  // rotation buffers per solved field

}  // namespace

struct ixh_state {
  int mi = 0, nj = 0, P = 0, Q = 0;  // row pitch P >= mi+1, transposed pitch Q >= nj+1
  size_t fsize = 0;  // doubles per field buffer (either layout)
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipStream_t launch = nullptr;  // the stream kernels go to (the state's stream)
  int u_hold = -1;  // u buffer still read by a residual running after the next u sweeps
  int pv_alt = 0;   // which of the two v norm partial slots the v x-sweep writes
  ResArgs res_next{};  // fused steps: the residual of the last built step
  ixh_params prm{};
  // FP32-mixed: u, v, T, p' (all copies) and a_u, a_v are stored as float in
  // the same buffers (same element pitch); p, b_o, Res_u stay FP64
  bool f32 = false;
  // FP32-refined: every field stays FP64; the sweeps assemble in FP32 and
  // solve for FP32 increments against an FP64 residual
  bool delta = false;
  // FP32-stochastic: FP32-mixed storage (f32 set), the increment sweeps of
  // FP32-refined, and every new u, v, T stored with stochastic rounding
  bool sr = false;
  // geometry (one allocation, 1-based arrays)
  double *d_geo = nullptr;
  Geo geo{};
  std::vector<double> h_dx[6];  // host copies of dxp, dyp, dxu, dyu, dxv, dyv (1-based)
  std::vector<double> h_fx[4], h_fy[4];  // fexp, fexu, xu, xp / feyp, feyv, yv, yp
  double *d_fast = nullptr;     // FastGeo tables
  FastGeo fast{};
  float *d_f32 = nullptr;       // FP32 copies of the geometry and FastGeo tables
  GeoF geof{};
  FastGeoF fastf{};
  // solved fields with rotation buffers
  double *ub[kNbuf]{}, *vb[kNbuf]{}, *tb[kNbuf]{}, *pcb[kNbuf]{};
  // transposed twins (y contiguous), rotated together with the row buffers
  double *ubT[kNbuf]{}, *vbT[kNbuf]{}, *tbT[kNbuf]{}, *pcbT[kNbuf]{};
  int u_cur = 0, u_ant = 0, v_cur = 0, v_ant = 0, t_cur = 0, t_ant = 0, pc_cur = 0;
  int v_snap = -1, t_snap = -1, pc_snap = -1;  // -1: snapshot is zero
  bool pc_zero = true;
  double *p = nullptr, *bo = nullptr, *au = nullptr, *av = nullptr, *resu = nullptr;
  double *pT = nullptr, *auT = nullptr, *avT = nullptr;
  double *src_c[3]{}, *src_p[3]{}, *src_cT[3]{}, *src_pT[3]{};
  bool has_src[3]{};
  int8_t *bct[3][4]{};
  double *bcv[3][4]{};
  int bclen[3][4]{};
  bool bc_set[3][4]{};
  double *part = nullptr;  // 5 partial arrays
  int plen = 0;
  double *d_out = nullptr;   // 8 doubles: last folded norms
  double *d_hist = nullptr;  // 8 doubles per recorded iteration
  int hist_cap = 0;
  int *d_slot = nullptr;
  // energy adjoint: checkpoints (kCkArrays fields each) of the recorded run,
  // the u, v tape of one segment and the adjoint work arrays
  int adj_steps = 0, adj_every = 0;
  std::vector<double *> adj_ck;
  std::vector<double *> adj_tape;
  double *adj_w[7]{};  // lambda, lambda', mu, adjoint of T_ant, the sweep's row coefficients (3)
  double *adj_g[4]{};  // adjoints of the boundary values of T, sides a..d
  std::vector<double *> adj2_tape;  // coupled adjoint: 16 fields per step of a segment
  std::vector<double *> adj2_w;     // coupled adjoint: work arrays
  // profiling
  bool profiling = false;
  double kms[K_NUM]{};
  int kcount[K_NUM]{};
  // fused-step launch timing: kernel start/end events of every launch
  // (hipExtLaunchKernelGGL), a ring of kRing steps harvested as it wraps
  hipEvent_t t_start = nullptr, t_stop = nullptr;  // events of the launch being issued
  std::vector<hipEvent_t> lev;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // pinned host memory for downloads: one copy of a whole pitched field
  void *h_stage = nullptr;
  size_t h_stage_bytes = 0;
};

namespace {

double *pc_of(ixh_state *s, int idx) { return idx < 0 ? nullptr : s->pcb[idx]; }

int pick(int excl0, int excl1, int excl2 = -1) {
  for (int k = 0; k < kNbuf; ++k)
    if (k != excl0 && k != excl1 && k != excl2) return k;
  return -1;
}

int dmalloc(double **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, n * sizeof(double) + 256);
  if (e != hipSuccess) return fail(IXH_ERR_NOMEM, "hipMalloc(%zu doubles): %s", n, hipGetErrorString(e));
  e = hipMemset(*p, 0, n * sizeof(double) + 256);
  if (e != hipSuccess) return fail(IXH_ERR_HIP, "hipMemset: %s", hipGetErrorString(e));
  return 0;
}

void field_extent(const ixh_state *s, int f, int *m, int *n) {
  switch (f) {
    case IXH_U: case IXH_U_ANT: case IXH_AU: case IXH_RESU: *m = s->mi; *n = s->nj + 1; break;
    case IXH_V: case IXH_V_ANT: case IXH_AV: *m = s->mi + 1; *n = s->nj; break;
    default: *m = s->mi + 1; *n = s->nj + 1; break;
  }
}

int side_len(const ixh_state *s, int var, int side) {
  const bool ac = side == IXH_SIDE_A || side == IXH_SIDE_C;
  if (var == IXH_VAR_U) return ac ? s->nj + 1 : s->mi;
  if (var == IXH_VAR_V) return ac ? s->nj : s->mi + 1;
  return ac ? s->nj + 1 : s->mi + 1;
}

int check_params(const ixh_params *p) {
  if (!p) return fail(IXH_ERR_ARG, "params is NULL");
  if (!(p->dt > 0) || !(p->rel_pres > 0) || !(p->rel_vel > 0) || !(p->rel_ener > 0))
    return fail(IXH_ERR_ARG, "dt and relaxation factors must be positive");
  if (!(p->gamma_momen > 0) || !(p->gamma_energ > 0))
    return fail(IXH_ERR_ARG, "diffusivities must be positive");
  if (p->solver != IXH_SOLVER_PCR && p->solver != IXH_SOLVER_THOMAS && p->solver != IXH_SOLVER_PCR_EXACT)
    return fail(IXH_ERR_ARG, "unknown solver %d", p->solver);
  if (p->precision != IXH_FP64 && p->precision != IXH_FP32_MIXED && p->precision != IXH_FP32_REFINED &&
      p->precision != IXH_FP32_STOCHASTIC)
    return fail(IXH_ERR_ARG, "unknown precision %d", p->precision);
  if (p->precision == IXH_FP32_MIXED && p->solver != IXH_SOLVER_PCR)
    return fail(IXH_ERR_UNSUPP, "the FP32-mixed path uses the PCR solver only");
  if (p->precision == IXH_FP32_REFINED && p->solver != IXH_SOLVER_PCR)
    return fail(IXH_ERR_UNSUPP, "the FP32-refined path uses the PCR solver only");
  if (p->precision == IXH_FP32_STOCHASTIC && p->solver != IXH_SOLVER_PCR)
    return fail(IXH_ERR_UNSUPP, "the FP32-stochastic path uses the PCR solver only");
  return 0;
}

Phys phys_of(const ixh_state *s) {
  const ixh_params &p = s->prm;
  return Phys{p.dt, p.rel_pres, p.rel_vel, p.rel_ener, p.gamma_momen, p.gamma_energ, p.ri_x, p.ri_y};
}

// ---- kernel launch helpers (with optional event timing)
struct Timed {
  ixh_state *s;
  int id;
  Timed(ixh_state *s_, int id_) : s(s_), id(id_) {
    if (s->profiling) (void)hipEventRecord(s->ev0, s->stream);
  }
  int done() {
    if (!s->profiling) return 0;
    HIPCHK(hipEventRecord(s->ev1, s->stream));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->kms[id] += ms;
    s->kcount[id] += 1;
    return 0;
  }
};

LaunchCtx ctx_of(const ixh_state *s) { return LaunchCtx{s->launch, s->t_start, s->t_stop}; }

// the sweep kernels live in sweeps_*.hip, one translation unit per path
template <int EQ, int DIR, int SOLVER, typename R>
int launch_sweep_t(ixh_state *s, const SweepArgs &A, int n) {
  HIPCHK((launch_sweep_k<EQ, DIR, SOLVER, R>(ctx_of(s), A, n)));
  return 0;
}

template <int SOLVER, typename R>
int launch_pcorr_t(ixh_state *s, const SweepArgs &A, const CorrLine &K, int n) {
  HIPCHK((launch_pcorr_k<SOLVER, R>(ctx_of(s), A, K, n)));
  return 0;
}

// the p' x-sweep with the correction of its lines
int launch_pcorr(ixh_state *s, const SweepArgs &A, const CorrLine &K, int n) {
  if (s->sr) return launch_pcorr_t<SOLVE_DELTA_SR, float>(s, A, K, n);
  if (s->f32) return launch_pcorr_t<SOLVE_PCR, float>(s, A, K, n);
  if (s->delta) return launch_pcorr_t<SOLVE_DELTA, float>(s, A, K, n);
  if (s->prm.solver == IXH_SOLVER_THOMAS) return launch_pcorr_t<SOLVE_THOMAS, double>(s, A, K, n);
  if (s->prm.solver == IXH_SOLVER_PCR_EXACT) return launch_pcorr_t<SOLVE_PCR_EXACT, double>(s, A, K, n);
  return launch_pcorr_t<SOLVE_PCR, double>(s, A, K, n);
}

// the p' y-sweep of a fused step (p' zero at entry, row layout out)
template <int SOLVER, typename R>
int launch_sweep_pz_t(ixh_state *s, const SweepArgs &A, int n) {
  HIPCHK((launch_sweep_pz_k<SOLVER, R>(ctx_of(s), A, n)));
  return 0;
}

int launch_sweep_pz(ixh_state *s, const SweepArgs &A, int n) {
  if (!A.in_zero) return fail(IXH_ERR_ARG, "the zero-input p' y-sweep needs p' = 0 at entry");
  if (s->sr) return launch_sweep_pz_t<SOLVE_DELTA_SR, float>(s, A, n);
  if (s->f32) return launch_sweep_pz_t<SOLVE_PCR, float>(s, A, n);
  if (s->delta) return launch_sweep_pz_t<SOLVE_DELTA, float>(s, A, n);
  if (s->prm.solver == IXH_SOLVER_THOMAS) return launch_sweep_pz_t<SOLVE_THOMAS, double>(s, A, n);
  if (s->prm.solver == IXH_SOLVER_PCR_EXACT) return launch_sweep_pz_t<SOLVE_PCR_EXACT, double>(s, A, n);
  return launch_sweep_pz_t<SOLVE_PCR, double>(s, A, n);
}

template <int EQ, int DIR>
int launch_sweep(ixh_state *s, const SweepArgs &A, int n) {
  if (s->sr) return launch_sweep_t<EQ, DIR, SOLVE_DELTA_SR, float>(s, A, n);
  if (s->f32) return launch_sweep_t<EQ, DIR, SOLVE_PCR, float>(s, A, n);
  if (s->delta) return launch_sweep_t<EQ, DIR, SOLVE_DELTA, float>(s, A, n);
  if (s->prm.solver == IXH_SOLVER_THOMAS) return launch_sweep_t<EQ, DIR, SOLVE_THOMAS, double>(s, A, n);
  if (s->prm.solver == IXH_SOLVER_PCR_EXACT) return launch_sweep_t<EQ, DIR, SOLVE_PCR_EXACT, double>(s, A, n);
  return launch_sweep_t<EQ, DIR, SOLVE_PCR, double>(s, A, n);
}

// FastGeo tables for the current diffusivities, dt and relaxation factors
int build_fast(ixh_state *s) {
  const int mi = s->mi, nj = s->nj, gx = mi + 2, gy = nj + 2;
  const std::vector<double> &dxp = s->h_dx[0], &dyp = s->h_dx[1], &dxu = s->h_dx[2], &dyv = s->h_dx[5];
  const double gm = s->prm.gamma_momen, gt = s->prm.gamma_energ;
  // per equation: gx x-records (west, east) then gy y-records (south,
  // north), each {conductance, Peclet factor} of its low and high face
  const size_t per = 4 * (size_t)gx + 4 * (size_t)gy;
  std::vector<double> h(3 * per, 0.0);
  auto q = [](double a, double b) { return b != 0.0 ? a / b : 0.0; };
  auto rec = [&](double *r, double G, double dlo, double dhi) {
    r[0] = q(G, dlo); r[1] = q(dlo, G);
    r[2] = q(G, dhi); r[3] = q(dhi, G);
  };
  for (int e = 0; e < 3; ++e) {
    double *fx = &h[e * per], *fy = fx + 4 * (size_t)gx;
    const double G = e == 2 ? gt : gm;
    for (int i = 2; i <= mi; ++i) {
      // west/east distances: u between u-nodes (dxp), v and T between p-nodes (dxu)
      const double dw = e == 0 ? dxp[i] : dxu[i - 1];
      const double de = e == 0 ? (i + 1 <= mi ? dxp[i + 1] : 0.0) : dxu[i];
      rec(fx + 4 * (size_t)i, G, dw, de);
    }
    for (int j = 2; j <= nj; ++j) {
      // south/north distances: u and T between p-rows (dyv), v between v-rows (dyp)
      const double ds = e == 1 ? dyp[j] : dyv[j - 1];
      const double dn = e == 1 ? (j + 1 <= nj ? dyp[j + 1] : 0.0) : dyv[j];
      rec(fy + 4 * (size_t)j, G, ds, dn);
    }
  }
  if (!s->d_fast) CHK(dmalloc(&s->d_fast, 3 * per));
  HIPCHK(hipMemcpy(s->d_fast, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  static_assert(sizeof(FaceT<double>) == 4 * sizeof(double) && sizeof(FaceT<float>) == 4 * sizeof(float),
                "face records are four packed values");
  for (int e = 0; e < 3; ++e) {
    const double *b = s->d_fast + e * per;
    s->fast.fx[e] = reinterpret_cast<const FaceT<double> *>(b);
    s->fast.fy[e] = reinterpret_cast<const FaceT<double> *>(b + 4 * (size_t)gx);
  }
  s->fast.inv_dt = 1.0 / s->prm.dt;
  s->fast.inv_rel_vel = 1.0 / s->prm.rel_vel;
  s->fast.inv_rel_ener = 1.0 / s->prm.rel_ener;
  // FP32 copies: the 1-D geometry arrays (Geo order) followed by the tables,
  // which start on a 16-byte boundary (one record per access)
  const size_t ngeo = (7 * (size_t)gx + 7 * (size_t)gy + 3) & ~(size_t)3;
  std::vector<float> f(ngeo + h.size());
  {
    const std::vector<double> *xs[] = {&s->h_dx[0], &s->h_dx[2], &s->h_dx[4], &s->h_fx[0], &s->h_fx[1],
                                       &s->h_fx[2], &s->h_fx[3]};
    const std::vector<double> *ys[] = {&s->h_dx[1], &s->h_dx[3], &s->h_dx[5], &s->h_fy[0], &s->h_fy[1],
                                       &s->h_fy[2], &s->h_fy[3]};
    size_t off = 0;
    for (auto *v : xs) { for (int i = 0; i < gx; ++i) f[off + i] = (float)(*v)[i]; off += gx; }
    for (auto *v : ys) { for (int j = 0; j < gy; ++j) f[off + j] = (float)(*v)[j]; off += gy; }
    for (size_t k = 0; k < h.size(); ++k) f[ngeo + k] = (float)h[k];
  }
  if (!s->d_f32) {
    hipError_t e = hipMalloc((void **)&s->d_f32, f.size() * sizeof(float) + 256);
    if (e != hipSuccess) return fail(IXH_ERR_NOMEM, "hipMalloc(fp32 tables): %s", hipGetErrorString(e));
  }
  HIPCHK(hipMemcpy(s->d_f32, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
  const float *d = s->d_f32;
  s->geof.dxp = d; s->geof.dxu = d + gx; s->geof.dxv = d + 2 * gx; s->geof.fexp = d + 3 * gx;
  s->geof.fexu = d + 4 * gx; s->geof.xu = d + 5 * gx; s->geof.xp = d + 6 * gx;
  d += 7 * (size_t)gx;
  s->geof.dyp = d; s->geof.dyu = d + gy; s->geof.dyv = d + 2 * gy; s->geof.feyp = d + 3 * gy;
  s->geof.feyv = d + 4 * gy; s->geof.yv = d + 5 * gy; s->geof.yp = d + 6 * gy;
  d = s->d_f32 + ngeo;
  for (int e = 0; e < 3; ++e) {
    const float *b = d + e * per;
    s->fastf.fx[e] = reinterpret_cast<const FaceT<float> *>(b);
    s->fastf.fy[e] = reinterpret_cast<const FaceT<float> *>(b + 4 * (size_t)gx);
  }
  s->fastf.inv_dt = (float)s->fast.inv_dt;
  s->fastf.inv_rel_vel = (float)s->fast.inv_rel_vel;
  s->fastf.inv_rel_ener = (float)s->fast.inv_rel_ener;
  return 0;
}

// arguments of a sweep: the y-sweeps read the transposed copies
SweepArgs base_args(ixh_state *s, int dir) {
  const bool tr = dir == IXH_DIR_Y;
  SweepArgs A;
  memset(&A, 0, sizeof A);
  A.mi = s->mi;
  A.nj = s->nj;
  A.P = tr ? s->Q : s->P;
  A.P2 = tr ? s->P : s->Q;
  A.g = s->geo;
  A.ph = phys_of(s);
  A.fg = s->fast;
  A.gf = s->geof;
  A.ff = s->fastf;
  A.u = (tr ? s->ubT : s->ub)[s->u_cur];
  A.v = (tr ? s->vbT : s->vb)[s->v_cur];
  A.t = (tr ? s->tbT : s->tb)[s->t_cur];
  A.p = tr ? s->pT : s->p;
  A.uant = (tr ? s->ubT : s->ub)[s->u_ant];
  A.vant = (tr ? s->vbT : s->vb)[s->v_ant];
  A.tant = (tr ? s->tbT : s->tb)[s->t_ant];
  A.pc = (tr ? s->pcbT : s->pcb)[s->pc_cur];
  A.au = tr ? s->auT : s->au;
  A.av = tr ? s->avT : s->av;
  return A;
}

// output buffers k of a field: the sweep's own layout and the other one
void set_out(SweepArgs &A, int dir, double *const *rowb, double *const *trb, int k) {
  A.out = dir == IXH_DIR_Y ? trb[k] : rowb[k];
  A.out2 = dir == IXH_DIR_Y ? rowb[k] : trb[k];
}

void set_bc(ixh_state *s, SweepArgs &A, int var, int dir) {
  for (int sd = 0; sd < 4; ++sd) {
    A.bc[sd].type = s->bct[var][sd];
    A.bc[sd].val = s->bcv[var][sd];
  }
  const bool tr = dir == IXH_DIR_Y;
  A.sc = s->has_src[var] ? (tr ? s->src_cT : s->src_c)[var] : nullptr;
  A.sp = s->has_src[var] ? (tr ? s->src_pT : s->src_p)[var] : nullptr;
}

int check_bcs(ixh_state *s, int var) {
  for (int sd = 0; sd < 4; ++sd)
    if (!s->bc_set[var][sd])
      return fail(IXH_ERR_ARG, "boundary conditions of variable %d side %d are not set", var, sd);
  return 0;
}

double *part_of(ixh_state *s, int k) { return s->part + (size_t)k * s->plen; }

// ---- phases.  prep_* build the kernel arguments of one phase and advance
// the buffer rotation (shared by the single-cavity and the batch paths);
// the launchers run them on one cavity.
int prep_u(ixh_state *s, int dir, SweepArgs &A) {
  CHK(check_bcs(s, IXH_VAR_U));
  A = base_args(s, dir);
  set_bc(s, A, IXH_VAR_U, dir);
  const int out = pick(s->u_cur, s->u_ant, s->u_hold);
  set_out(A, dir, s->ub, s->ubT, out);
  A.acoef = dir == IXH_DIR_Y ? s->auT : s->au;
  A.acoef2 = dir == IXH_DIR_Y ? s->au : s->auT;
  A.line0 = 2;
  A.nlines = dir == IXH_DIR_Y ? s->mi - 2 : s->nj - 1;
  s->u_cur = out;
  return 0;
}

int prep_v(ixh_state *s, int dir, SweepArgs &A) {
  CHK(check_bcs(s, IXH_VAR_V));
  A = base_args(s, dir);
  set_bc(s, A, IXH_VAR_V, dir);
  A.acoef = dir == IXH_DIR_Y ? s->avT : s->av;
  A.acoef2 = dir == IXH_DIR_Y ? s->av : s->avT;
  A.line0 = 2;
  int out;
  if (dir == IXH_DIR_Y) {
    s->v_snap = s->v_cur;  // fv = v (main.f90:559-564)
    out = pick(s->v_cur, s->v_ant);
    A.nlines = s->mi - 1;
  } else {
    out = pick(s->v_cur, s->v_ant, s->v_snap);
    A.snap = s->v_snap >= 0 ? s->vb[s->v_snap] : s->vb[s->v_cur];
    A.part0 = part_of(s, s->pv_alt ? 5 : 0);
    A.nlines = s->nj - 2;
  }
  set_out(A, dir, s->vb, s->vbT, out);
  s->v_cur = out;
  return 0;
}

int prep_p(ixh_state *s, int dir, SweepArgs &A) {
  A = base_args(s, dir);
  A.in_zero = s->pc_zero ? 1 : 0;
  A.line0 = 2;
  if (dir == IXH_DIR_Y) s->pc_snap = s->pc_zero ? -1 : s->pc_cur;  // fcorr = p' (main.f90:762-767)
  const int out = pick(s->pc_cur, s->pc_snap);
  set_out(A, dir, s->pcb, s->pcbT, out);
  if (dir == IXH_DIR_Y) {
    A.nlines = s->mi - 1;
  } else {
    A.snap = pc_of(s, s->pc_snap);
    A.bo = s->bo;
    A.part0 = part_of(s, 1);
    A.part1 = part_of(s, 2);
    A.nlines = s->nj - 1;
  }
  s->pc_cur = out;
  s->pc_zero = false;
  return 0;
}

int prep_t(ixh_state *s, int dir, SweepArgs &A) {
  CHK(check_bcs(s, IXH_VAR_T));
  A = base_args(s, dir);
  set_bc(s, A, IXH_VAR_T, dir);
  A.line0 = 2;
  int out;
  if (dir == IXH_DIR_Y) {
    s->t_snap = s->t_cur;  // ftemp = T (main.f90:974-979)
    out = pick(s->t_cur, s->t_ant);
    A.nlines = s->mi - 1;
  } else {
    out = pick(s->t_cur, s->t_ant, s->t_snap);
    A.snap = s->t_snap >= 0 ? s->tb[s->t_snap] : s->tb[s->t_cur];
    A.part0 = part_of(s, 3);
    A.nlines = s->nj - 1;
  }
  set_out(A, dir, s->tb, s->tbT, out);
  s->t_cur = out;
  return 0;
}

// line length of a sweep (sets the LDS size of the kernel)
int line_len(const ixh_state *s, int eq, int dir) {
  if (dir == IXH_DIR_X) return eq == EQ_U ? s->mi : s->mi + 1;
  return eq == EQ_V ? s->nj : s->nj + 1;
}

int sweep_u(ixh_state *s, int dir) {
  SweepArgs A;
  CHK(prep_u(s, dir, A));
  Timed tm(s, dir == IXH_DIR_Y ? K_U_Y : K_U_X);
  if (dir == IXH_DIR_Y)
    CHK((launch_sweep<EQ_U, DIR_Y>(s, A, line_len(s, EQ_U, dir))));
  else
    CHK((launch_sweep<EQ_U, DIR_X>(s, A, line_len(s, EQ_U, dir))));
  return tm.done();
}

int sweep_v(ixh_state *s, int dir) {
  SweepArgs A;
  CHK(prep_v(s, dir, A));
  Timed tm(s, dir == IXH_DIR_Y ? K_V_Y : K_V_X);
  if (dir == IXH_DIR_Y)
    CHK((launch_sweep<EQ_V, DIR_Y>(s, A, line_len(s, EQ_V, dir))));
  else
    CHK((launch_sweep<EQ_V, DIR_X>(s, A, line_len(s, EQ_V, dir))));
  return tm.done();
}

int sweep_p(ixh_state *s, int dir) {
  SweepArgs A;
  CHK(prep_p(s, dir, A));
  Timed tm(s, dir == IXH_DIR_Y ? K_P_Y : K_P_X);
  if (dir == IXH_DIR_Y)
    CHK((launch_sweep<EQ_P, DIR_Y>(s, A, line_len(s, EQ_P, dir))));
  else
    CHK((launch_sweep<EQ_P, DIR_X>(s, A, line_len(s, EQ_P, dir))));
  return tm.done();
}

int sweep_t(ixh_state *s, int dir) {
  SweepArgs A;
  CHK(prep_t(s, dir, A));
  Timed tm(s, dir == IXH_DIR_Y ? K_T_Y : K_T_X);
  if (dir == IXH_DIR_Y)
    CHK((launch_sweep<EQ_T, DIR_Y>(s, A, line_len(s, EQ_T, dir))));
  else
    CHK((launch_sweep<EQ_T, DIR_X>(s, A, line_len(s, EQ_T, dir))));
  return tm.done();
}

// move field f (IXH_U/V/T or their _ANT) to a buffer of its own when the
// current and lagged copies share one
int unalias(ixh_state *s, int f) {
  double **b = nullptr, **bT = nullptr;
  int *cur = nullptr, *ant = nullptr;
  if (f == IXH_U || f == IXH_U_ANT) { b = s->ub; bT = s->ubT; cur = &s->u_cur; ant = &s->u_ant; }
  else if (f == IXH_V || f == IXH_V_ANT) { b = s->vb; bT = s->vbT; cur = &s->v_cur; ant = &s->v_ant; }
  else if (f == IXH_T || f == IXH_T_ANT) { b = s->tb; bT = s->tbT; cur = &s->t_cur; ant = &s->t_ant; }
  else return 0;
  if (*cur != *ant) return 0;
  const int k = pick(*cur, -1);
  HIPCHK(hipMemcpyAsync(b[k], b[*cur], s->fsize * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(bT[k], bT[*cur], s->fsize * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  if (f == IXH_U || f == IXH_V || f == IXH_T) *cur = k; else *ant = k;
  return 0;
}

// arguments of the correction; false when p' is zero (nothing changes)
int prep_correct(ixh_state *s, CorrArgs &A, bool &active) {
  memset(&A, 0, sizeof A);
  active = !s->pc_zero;
  if (!active) return 0;
  // the correction updates u, v in place; give them their own buffers when
  // they still share one with the lagged copies
  CHK(unalias(s, IXH_U));
  CHK(unalias(s, IXH_V));
  A.mi = s->mi;
  A.nj = s->nj;
  A.P = s->P;
  A.Q = s->Q;
  A.g = s->geo;
  A.p = s->p;
  A.pT = s->pT;
  A.pc = s->pcb[s->pc_cur];
  A.au = s->au;
  A.av = s->av;
  A.u = s->ub[s->u_cur];
  A.v = s->vb[s->v_cur];
  A.uT = s->ubT[s->u_cur];
  A.vT = s->vbT[s->v_cur];
  A.sr = s->sr ? 1 : 0;
  return 0;
}

// the correction as the p' x-sweep does it: corrected u and v into free
// buffers (the sweep's other workgroups still read the old ones while it
// runs), p in place (only its own line's workgroup reads p)
void prep_correct_line(ixh_state *s, CorrLine &K) {
  const int nu = pick(s->u_cur, s->u_ant, s->u_hold);
  const int nv = pick(s->v_cur, s->v_ant, s->v_snap);
  K.u = s->ub[nu];
  K.uT = s->ubT[nu];
  K.v = s->vb[nv];
  K.vT = s->vbT[nv];
  K.p = s->p;
  K.pT = s->pT;
  s->u_cur = nu;
  s->v_cur = nv;
}

int corr_tiles_x(const ixh_state *s) { return (s->mi + 1 + kTile - 1) / kTile; }
int corr_tiles_y(const ixh_state *s) { return (s->nj + 1 + kTile - 1) / kTile; }

int do_correct(ixh_state *s) {
  CorrArgs A;
  bool active;
  CHK(prep_correct(s, A, active));
  if (!active) return 0;
  dim3 block(kBlock), grid(corr_tiles_x(s), corr_tiles_y(s));
  Timed tm(s, K_CORR);
  if (s->f32)
    hipLaunchKernelGGL(k_correct<float>, grid, block, 0, s->launch, A);
  else
    hipLaunchKernelGGL(k_correct<double>, grid, block, 0, s->launch, A);
  HIPCHK(hipGetLastError());
  return tm.done();
}

void prep_residual(ixh_state *s, ResArgs &A) {
  memset(&A, 0, sizeof A);
  A.mi = s->mi;
  A.nj = s->nj;
  A.P = s->P;
  A.g = s->geo;
  A.ph = phys_of(s);
  A.u = s->ub[s->u_cur];
  A.v = s->vb[s->v_cur];
  A.t = s->tb[s->t_cur];
  A.p = s->p;
  A.uant = s->ub[s->u_ant];
  A.sc = s->has_src[0] ? s->src_c[0] : nullptr;
  A.sp = s->has_src[0] ? s->src_p[0] : nullptr;
  A.resu = s->resu;
  A.part = part_of(s, 4);
}

int do_residual(ixh_state *s) {
  ResArgs A;
  prep_residual(s, A);
  Timed tm(s, K_RES);
  if (s->f32)
    hipLaunchKernelGGL(k_residual_u<float>, dim3(s->nj - 1), dim3(kBlock), 0, s->launch, A);
  else
    hipLaunchKernelGGL(k_residual_u<double>, dim3(s->nj - 1), dim3(kBlock), 0, s->launch, A);
  HIPCHK(hipGetLastError());
  return tm.done();
}

void prep_finalize(ixh_state *s, FinArgs &F, int which, double *out, int *slot) {
  F.nj = s->nj;
  F.pv = part_of(s, s->pv_alt ? 5 : 0);
  F.pp = part_of(s, 1);
  F.pb = part_of(s, 2);
  F.pt = part_of(s, 3);
  F.pr = part_of(s, 4);
  F.out = out;
  F.slot = slot;
  F.which = which;
}

int do_finalize(ixh_state *s, int which, double *out, int *slot) {
  FinArgs F;
  prep_finalize(s, F, which, out, slot);
  Timed tm(s, K_FIN);
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, s->launch, F);
  HIPCHK(hipGetLastError());
  return tm.done();
}

// fold the selected norms and copy them to the host
int read_norms(ixh_state *s, int which, double out[8]) {
  CHK(do_finalize(s, which, s->d_out, nullptr));
  HIPCHK(hipMemcpyAsync(out, s->d_out, 8 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

// One single-pass time step in upstream's order, without host
// synchronisation; the norms land in d_hist[slot] through the device counter.
int fast_step(ixh_state *s) {
  CHK(sweep_u(s, IXH_DIR_Y));
  CHK(sweep_u(s, IXH_DIR_X));
  CHK(sweep_v(s, IXH_DIR_Y));
  CHK(sweep_v(s, IXH_DIR_X));
  s->pc_zero = true;
  CHK(sweep_p(s, IXH_DIR_Y));
  CHK(sweep_p(s, IXH_DIR_X));
  CHK(do_correct(s));
  CHK(sweep_t(s, IXH_DIR_Y));
  CHK(sweep_t(s, IXH_DIR_X));
  CHK(do_residual(s));
  CHK(do_finalize(s, 31, s->d_hist, s->d_slot));
  s->u_ant = s->u_cur;
  s->v_ant = s->v_cur;
  s->t_ant = s->t_cur;
  return 0;
}

// ---- fused step: 6 launches on the state's stream per time step
// k_multi with sweeps (EQ0, DIR0) and (EQ1, DIR1) at a common EPT
template <int EQ0, int DIR0, int EQ1, int DIR1, int SOLVER, typename R>
int launch_multi_t(ixh_state *s, const MultiArgs &M, int nwg, int ept, int maxn) {
  HIPCHK((launch_multi_k<EQ0, DIR0, EQ1, DIR1, SOLVER, R>(ctx_of(s), M, nwg, ept, maxn)));
  return 0;
}

template <int EQ0, int DIR0, int EQ1, int DIR1>
int launch_multi(ixh_state *s, const MultiArgs &M, int nwg, int ept, int maxn) {
  if (s->sr) return launch_multi_t<EQ0, DIR0, EQ1, DIR1, SOLVE_DELTA_SR, float>(s, M, nwg, ept, maxn);
  if (s->f32) return launch_multi_t<EQ0, DIR0, EQ1, DIR1, SOLVE_PCR, float>(s, M, nwg, ept, maxn);
  if (s->delta) return launch_multi_t<EQ0, DIR0, EQ1, DIR1, SOLVE_DELTA, float>(s, M, nwg, ept, maxn);
  if (s->prm.solver == IXH_SOLVER_THOMAS)
    return launch_multi_t<EQ0, DIR0, EQ1, DIR1, SOLVE_THOMAS, double>(s, M, nwg, ept, maxn);
  if (s->prm.solver == IXH_SOLVER_PCR_EXACT)
    return launch_multi_t<EQ0, DIR0, EQ1, DIR1, SOLVE_PCR_EXACT, double>(s, M, nwg, ept, maxn);
  return launch_multi_t<EQ0, DIR0, EQ1, DIR1, SOLVE_PCR, double>(s, M, nwg, ept, maxn);
}

int round8(int x) { return (x + 7) & ~7; }

// lay out the parts of a fused launch; returns the workgroup count
int multi_layout(MultiArgs &M, int n0, int n1, int nres, int nfin) {
  M.cnt[0] = n0; M.cnt[1] = n1; M.cnt[2] = nres; M.cnt[3] = nfin;
  M.off[0] = 0;
  M.off[1] = round8(n0);
  M.off[2] = M.off[1] + round8(n1);
  M.off[3] = M.off[2] + round8(nres);
  return M.off[3] + nfin;
}

// The 6 launches of one fused time step k (the u equation must not read T):
//   1: v_y(k) [+ residual of step k-1]   2: v_x(k) [+ fold of step k-1]
//   3: p'_y(k)   4: p'_x(k) + correction(k)   5: T_y(k) + u_y(k+1)
//   6: T_x(k) + u_x(k+1)
// The u sweeps of the first step run before the first step, the residual
// and the fold of the last step after it.  u_ant of step k+1 is u of step
// k; the u buffer the residual of step k reads as its u_ant (u of step k-1)
// is held out of the rotation while the u sweeps of step k+1 run, and
// nothing writes u, v or p between those sweeps and the residual.  The fold
// of step k runs beside v_x of step k+1, so the v norm partials alternate
// between two slots.
struct StepLaunches {
  MultiArgs m1, m2, m5, m6;
  SweepArgs py, px;
  CorrLine corr;
  int nwg1, nwg2, nwg5, nwg6, ept1, ept2, ept5, ept6;
  bool last;
};

// arguments of the launches of one step (advances the rotation as the
// step does); the residual of step k is kept in the state for step k+1
int build_step(ixh_state *s, bool first, bool last, StepLaunches &L) {
  memset(&L, 0, sizeof L);
  L.last = last;
  CHK(prep_v(s, IXH_DIR_Y, L.m1.s[0]));
  if (!first) L.m1.r = s->res_next;
  L.nwg1 = multi_layout(L.m1, L.m1.s[0].nlines, 0, first ? 0 : s->nj - 1, 0);
  L.ept1 = ept_for(line_len(s, EQ_V, IXH_DIR_Y));
  if (!first) prep_finalize(s, L.m2.f, 31, s->d_hist, s->d_slot);  // step k-1's v partials
  s->pv_alt ^= 1;
  CHK(prep_v(s, IXH_DIR_X, L.m2.s[0]));
  L.nwg2 = multi_layout(L.m2, L.m2.s[0].nlines, 0, 0, first ? 0 : 1);
  L.ept2 = ept_for(line_len(s, EQ_V, IXH_DIR_X));
  s->pc_zero = true;
  CHK(prep_p(s, IXH_DIR_Y, L.py));
  CHK(prep_p(s, IXH_DIR_X, L.px));
  prep_correct_line(s, L.corr);
  // the T sweeps and the residual of step k read u of step k: build their
  // arguments before the u sweeps of step k+1 advance the rotation
  CHK(prep_t(s, IXH_DIR_Y, L.m5.s[0]));
  CHK(prep_t(s, IXH_DIR_X, L.m6.s[0]));
  prep_residual(s, s->res_next);
  const int held = s->u_ant;
  s->u_ant = s->u_cur;
  L.ept5 = ept_for(line_len(s, EQ_T, IXH_DIR_Y));
  L.ept6 = ept_for(line_len(s, EQ_T, IXH_DIR_X));
  int nu5 = 0, nu6 = 0;
  if (!last) {
    s->u_hold = held;
    CHK(prep_u(s, IXH_DIR_Y, L.m5.s[1]));
    CHK(prep_u(s, IXH_DIR_X, L.m6.s[1]));
    s->u_hold = -1;
    nu5 = L.m5.s[1].nlines;
    nu6 = L.m6.s[1].nlines;
    L.ept5 = std::max(L.ept5, ept_for(line_len(s, EQ_U, IXH_DIR_Y)));
    L.ept6 = std::max(L.ept6, ept_for(line_len(s, EQ_U, IXH_DIR_X)));
  }
  L.nwg5 = multi_layout(L.m5, L.m5.s[0].nlines, nu5, 0, 0);
  L.nwg6 = multi_layout(L.m6, L.m6.s[0].nlines, nu6, 0, 0);
  s->v_ant = s->v_cur;
  s->t_ant = s->t_cur;
  return 0;
}

// launch part k (1..kParts) of a built step
int launch_part(ixh_state *s, const StepLaunches &L, int k) {
  switch (k) {
    case 1: return launch_multi<EQ_V, DIR_Y, -1, DIR_Y>(s, L.m1, L.nwg1, L.ept1, line_len(s, EQ_V, IXH_DIR_Y));
    case 2: return launch_multi<EQ_V, DIR_X, -1, DIR_X>(s, L.m2, L.nwg2, L.ept2, line_len(s, EQ_V, IXH_DIR_X));
    case 3: return launch_sweep_pz(s, L.py, line_len(s, EQ_P, IXH_DIR_Y));
    case 4: return launch_pcorr(s, L.px, L.corr, line_len(s, EQ_P, IXH_DIR_X));
    case 5:
      if (L.last) return launch_multi<EQ_T, DIR_Y, -1, DIR_Y>(s, L.m5, L.nwg5, L.ept5, s->nj + 1);
      return launch_multi<EQ_T, DIR_Y, EQ_U, DIR_Y>(s, L.m5, L.nwg5, L.ept5, s->nj + 1);
    case 6:
      if (L.last) return launch_multi<EQ_T, DIR_X, -1, DIR_X>(s, L.m6, L.nwg6, L.ept6, s->mi + 1);
      return launch_multi<EQ_T, DIR_X, EQ_U, DIR_X>(s, L.m6, L.nwg6, L.ept6, s->mi + 1);
  }
  return fail(IXH_ERR_ARG, "launch part %d", k);
}

constexpr int kRing = 32;  // steps of launch events in flight

// add the durations of the launches of ring slot `slot`
int harvest_slot(ixh_state *s, int slot) {
  hipEvent_t *e = &s->lev[(size_t)slot * 2 * kParts];
  HIPCHK(hipEventSynchronize(e[2 * kParts - 1]));
  for (int k = 0; k < kParts; ++k) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e[2 * k], e[2 * k + 1]));
    s->kms[K_L1 + k] += ms;
    s->kcount[K_L1 + k] += 1;
  }
  return 0;
}

// step k of a run; with profiling on, every launch records its kernel's
// start and end into the ring slot of step k
int fused_step(ixh_state *s, bool first, bool last, int k) {
  StepLaunches L;
  CHK(build_step(s, first, last, L));
  const bool timed = s->profiling;
  const int slot = k % kRing;
  if (timed) {
    if (s->lev.empty()) {
      s->lev.resize((size_t)kRing * 2 * kParts);
      for (hipEvent_t &e : s->lev) HIPCHK(hipEventCreate(&e));
    }
    if (k >= kRing) CHK(harvest_slot(s, slot));
  }
  for (int p = 1; p <= kParts; ++p) {
    if (timed) {
      s->t_start = s->lev[(size_t)slot * 2 * kParts + 2 * (p - 1)];
      s->t_stop = s->lev[(size_t)slot * 2 * kParts + 2 * (p - 1) + 1];
    }
    const int rc = launch_part(s, L, p);
    s->t_start = s->t_stop = nullptr;
    CHK(rc);
  }
  if (last) {
    // the residual and the fold of the last step
    if (s->f32)
      hipLaunchKernelGGL(k_residual_u<float>, dim3(s->nj - 1), dim3(kBlock), 0, s->launch, s->res_next);
    else
      hipLaunchKernelGGL(k_residual_u<double>, dim3(s->nj - 1), dim3(kBlock), 0, s->launch, s->res_next);
    HIPCHK(hipGetLastError());
    CHK(do_finalize(s, 31, s->d_hist, s->d_slot));
  }
  return 0;
}

// nsteps fused steps (the u equation must not read T): the u sweeps of the
// first step before it, the residual and the fold of the last after it; the
// norms of each step go to the history slot d_slot points at
int run_fused(ixh_state *s, int nsteps) {
  if (nsteps <= 0) return 0;
  CHK(sweep_u(s, IXH_DIR_Y));
  CHK(sweep_u(s, IXH_DIR_X));
  for (int k = 0; k < nsteps; ++k) CHK(fused_step(s, k == 0, k == nsteps - 1, k));
  if (s->profiling)
    for (int k = std::max(0, nsteps - kRing); k < nsteps; ++k) CHK(harvest_slot(s, k % kRing));
  return 0;
}

int ensure_hist(ixh_state *s, int n) {
  if (n <= s->hist_cap) return 0;
  if (s->d_hist) HIPCHK(hipFree(s->d_hist));
  s->d_hist = nullptr;
  CHK(dmalloc(&s->d_hist, (size_t)8 * n));
  s->hist_cap = n;
  return 0;
}

}  // namespace

// values per array of ixh_tridiagonal_noalloc's scratch
constexpr size_t kTriScratchPerArray = (size_t)1 << 18;
// Device scratch of ixh_tridiagonal_noalloc: the four arrays of the systems
// of one call, 8 MiB in all, held by every GPU the library runs on until the
// process ends.  In host code `scratch` names a host-side stand-in, not the
// device array: its device address is looked up once per device with
// hipGetSymbolAddress (with that device current) and only that address is
// used.  Calls on one GPU share its array, so calls that overlap on one GPU
// race; each call copies in every value it reads, so nothing needs clearing.
static __device__ __align__(256) double scratch[4 * kTriScratchPerArray];

namespace {

// ---- batched tridiagonal solves of host arrays (ixh_tridiagonal*)
int check_tridiag(const double *a, const double *b, const double *c, const double *r, int n, int nsys, int solver) {
  if (!a || !b || !c || !r || n < 1 || nsys < 0) return fail(IXH_ERR_ARG, "bad tridiagonal arguments");
  if (n > kNmax) return fail(IXH_ERR_UNSUPP, "systems longer than %d are not built", kNmax);
  if (solver != IXH_SOLVER_PCR && solver != IXH_SOLVER_THOMAS) return fail(IXH_ERR_ARG, "unknown solver");
  return 0;
}

// nsys systems of n rows through device memory d (4 n nsys doubles): copy
// a, b, c, r in, solve on the null stream, copy r and b back
int tridiag_on(double *d, const double *a, double *b, const double *c, double *r, int n, int nsys, int solver) {
  const size_t tot = (size_t)n * nsys;
  double *da = d, *db = d + tot, *dc = d + 2 * tot, *dr = d + 3 * tot;
  hipError_t e = hipSuccess;
  if ((e = hipMemcpy(da, a, tot * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(db, b, tot * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dc, c, tot * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dr, r, tot * 8, hipMemcpyHostToDevice)) != hipSuccess)
    return fail(IXH_ERR_HIP, "tridiagonal upload: %s", hipGetErrorString(e));
  const int ept = ept_for(n);
#define IXH_TRI(S, E) hipLaunchKernelGGL((k_tridiag<S, E>), dim3(nsys), dim3(kBlock), 0, 0, da, db, dc, dr, n)
  if (solver == IXH_SOLVER_THOMAS) {
    switch (ept) {
      case 1: IXH_TRI(SOLVE_THOMAS, 1); break; case 2: IXH_TRI(SOLVE_THOMAS, 2); break;
      case 3: IXH_TRI(SOLVE_THOMAS, 3); break; case 4: IXH_TRI(SOLVE_THOMAS, 4); break;
      case 6: IXH_TRI(SOLVE_THOMAS, 6); break; default: IXH_TRI(SOLVE_THOMAS, 8); break;
    }
  } else {
    switch (ept) {
      case 1: IXH_TRI(SOLVE_PCR, 1); break; case 2: IXH_TRI(SOLVE_PCR, 2); break;
      case 3: IXH_TRI(SOLVE_PCR, 3); break; case 4: IXH_TRI(SOLVE_PCR, 4); break;
      case 6: IXH_TRI(SOLVE_PCR, 6); break; default: IXH_TRI(SOLVE_PCR, 8); break;
    }
  }
#undef IXH_TRI
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(r, dr, tot * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(b, db, tot * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return fail(IXH_ERR_HIP, "tridiagonal: %s", hipGetErrorString(e));
  return 0;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int tri_scratch(double **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(IXH_ERR_UNSUPP, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

// =================================================================== C-ABI
extern "C" {

const char *ixh_last_error(void) { return g_err.c_str(); }
const char *ixh_version(void) { return "ixchel2d-amd 0.3 (gfx950; FP64, FP32-mixed, FP32-refined, FP32-stochastic)"; }
const char *ixh_kernel_name(int32_t id) { return (id >= 0 && id < K_NUM) ? kKernelNames[id] : ""; }

int ixh_create(const ixh_grid *grid, const ixh_params *params, ixh_state **out) {
  if (!grid || !out) return fail(IXH_ERR_ARG, "grid/out is NULL");
  *out = nullptr;
  CHK(check_params(params));
  const int mi = grid->mi, nj = grid->nj;
  if (mi < 4 || nj < 4) return fail(IXH_ERR_ARG, "mesh needs mi, nj >= 4 (got %d, %d)", mi, nj);
  if (mi + 1 > kNmax || nj + 1 > kNmax)
    return fail(IXH_ERR_UNSUPP, "lines longer than %d nodes are not built (mi=%d nj=%d)", kNmax, mi, nj);
  if (!grid->xu || !grid->yv || !grid->xp || !grid->yp) return fail(IXH_ERR_ARG, "grid coordinates are NULL");
  ixh_state *s = new ixh_state();
  s->mi = mi;
  s->nj = nj;
  s->P = ((mi + 1) + 15) / 16 * 16;  // rows start on 128-byte boundaries
  s->Q = ((nj + 1) + 15) / 16 * 16;  // so do the columns of the transposed copies
  s->fsize = std::max((size_t)s->P * (nj + 1), (size_t)s->Q * (mi + 1));
  s->prm = *params;
  s->f32 = params->precision == IXH_FP32_MIXED || params->precision == IXH_FP32_STOCHASTIC;
  s->sr = params->precision == IXH_FP32_STOCHASTIC;
  s->delta = params->precision == IXH_FP32_REFINED;
  int rc = 0;
  auto bail = [&](int code) {
    ixh_destroy(s);
    return code;
  };
  if (hipStreamCreate(&s->stream) != hipSuccess) return bail(fail(IXH_ERR_HIP, "hipStreamCreate failed"));
  s->own_stream = true;
  if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return bail(fail(IXH_ERR_HIP, "hipEventCreate failed"));
  s->launch = s->stream;
  // geometry on the host (mod_malla.f90:167-217), 1-based
  const int gx = mi + 2, gy = nj + 2;
  std::vector<double> xu(gx, 0.0), yv(gy, 0.0), xp(gx, 0.0), yp(gy, 0.0);
  for (int i = 1; i <= mi; ++i) xu[i] = grid->xu[i - 1];
  for (int j = 1; j <= nj; ++j) yv[j] = grid->yv[j - 1];
  for (int i = 1; i <= mi + 1; ++i) xp[i] = grid->xp[i - 1];
  for (int j = 1; j <= nj + 1; ++j) yp[j] = grid->yp[j - 1];
  std::vector<double> dxp(gx, 0.0), dyp(gy, 0.0), dxu(gx, 0.0), dyu(gy, 0.0), dxv(gx, 0.0), dyv(gy, 0.0);
  std::vector<double> fexp(gx, 0.0), feyp(gy, 0.0), fexu(gx, 0.0), feyv(gy, 0.0);
  for (int i = 2; i <= mi; ++i) dxp[i] = xu[i] - xu[i - 1];
  for (int j = 2; j <= nj; ++j) dyp[j] = yv[j] - yv[j - 1];
  for (int i = 1; i <= mi; ++i) dxu[i] = xp[i + 1] - xp[i];
  for (int j = 2; j <= nj; ++j) dyu[j] = yv[j] - yv[j - 1];
  for (int i = 2; i <= mi; ++i) dxv[i] = xu[i] - xu[i - 1];
  for (int j = 1; j <= nj; ++j) dyv[j] = yp[j + 1] - yp[j];
  for (int i = 1; i <= mi; ++i) fexp[i] = (xp[i + 1] - xu[i]) / (xp[i + 1] - xp[i]);
  for (int j = 1; j <= nj; ++j) feyp[j] = (yp[j + 1] - yv[j]) / (yp[j + 1] - yp[j]);
  for (int i = 1; i <= mi - 1; ++i) fexu[i] = (xu[i + 1] - xp[i + 1]) / (xu[i + 1] - xu[i]);
  for (int j = 1; j <= nj - 1; ++j) feyv[j] = (yv[j + 1] - yp[j + 1]) / (yv[j + 1] - yv[j]);
  s->h_dx[0] = dxp; s->h_dx[1] = dyp; s->h_dx[2] = dxu; s->h_dx[3] = dyu; s->h_dx[4] = dxv; s->h_dx[5] = dyv;
  s->h_fx[0] = fexp; s->h_fx[1] = fexu; s->h_fx[2] = xu; s->h_fx[3] = xp;
  s->h_fy[0] = feyp; s->h_fy[1] = feyv; s->h_fy[2] = yv; s->h_fy[3] = yp;
  const std::vector<double> *xs[] = {&dxp, &dxu, &dxv, &fexp, &fexu, &xu, &xp};
  const std::vector<double> *ys[] = {&dyp, &dyu, &dyv, &feyp, &feyv, &yv, &yp};
  const size_t ngeo = 7 * (size_t)gx + 7 * (size_t)gy;
  std::vector<double> host(ngeo);
  size_t off = 0;
  for (auto *v : xs) { memcpy(&host[off], v->data(), gx * sizeof(double)); off += gx; }
  for (auto *v : ys) { memcpy(&host[off], v->data(), gy * sizeof(double)); off += gy; }
  if ((rc = dmalloc(&s->d_geo, ngeo))) return bail(rc);
  if (hipMemcpy(s->d_geo, host.data(), ngeo * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(IXH_ERR_HIP, "geometry upload failed"));
  const double *d = s->d_geo;
  s->geo.dxp = d; s->geo.dxu = d + gx; s->geo.dxv = d + 2 * gx; s->geo.fexp = d + 3 * gx;
  s->geo.fexu = d + 4 * gx; s->geo.xu = d + 5 * gx; s->geo.xp = d + 6 * gx;
  d += 7 * (size_t)gx;
  s->geo.dyp = d; s->geo.dyu = d + gy; s->geo.dyv = d + 2 * gy; s->geo.feyp = d + 3 * gy;
  s->geo.feyv = d + 4 * gy; s->geo.yv = d + 5 * gy; s->geo.yp = d + 6 * gy;
  // fields
  for (int k = 0; k < kNbuf; ++k) {
    if ((rc = dmalloc(&s->ub[k], s->fsize)) || (rc = dmalloc(&s->vb[k], s->fsize)) ||
        (rc = dmalloc(&s->tb[k], s->fsize)) || (rc = dmalloc(&s->pcb[k], s->fsize)) ||
        (rc = dmalloc(&s->ubT[k], s->fsize)) || (rc = dmalloc(&s->vbT[k], s->fsize)) ||
        (rc = dmalloc(&s->tbT[k], s->fsize)) || (rc = dmalloc(&s->pcbT[k], s->fsize)))
      return bail(rc);
  }
  if ((rc = dmalloc(&s->p, s->fsize)) || (rc = dmalloc(&s->bo, s->fsize)) ||
      (rc = dmalloc(&s->au, s->fsize)) || (rc = dmalloc(&s->av, s->fsize)) ||
      (rc = dmalloc(&s->resu, s->fsize)) || (rc = dmalloc(&s->pT, s->fsize)) ||
      (rc = dmalloc(&s->auT, s->fsize)) || (rc = dmalloc(&s->avT, s->fsize)))
    return bail(rc);
  {
    // a_u = a_v = 1e40 (main.f90:309-310); beyond the FP32 range: infinite
    std::vector<double> big(s->fsize, 1.e40);
    std::vector<float> bigf(s->fsize, HUGE_VALF);
    const void *src = s->f32 ? (const void *)bigf.data() : (const void *)big.data();
    const size_t bytes = s->fsize * (s->f32 ? sizeof(float) : sizeof(double));
    if (hipMemcpy(s->au, src, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->av, src, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->auT, src, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->avT, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(IXH_ERR_HIP, "a_u/a_v initialisation failed"));
  }
  s->plen = (mi > nj ? mi : nj) + 4;
  if ((rc = dmalloc(&s->part, (size_t)6 * s->plen)) || (rc = dmalloc(&s->d_out, 8))) return bail(rc);
  if (hipMalloc((void **)&s->d_slot, sizeof(int)) != hipSuccess) return bail(fail(IXH_ERR_NOMEM, "slot"));
  if ((rc = build_fast(s))) return bail(rc);
  for (int v = 0; v < 3; ++v)
    for (int sd = 0; sd < 4; ++sd) {
      const int n = side_len(s, v, sd);
      s->bclen[v][sd] = n;
      if (hipMalloc((void **)&s->bct[v][sd], n + 2) != hipSuccess) return bail(fail(IXH_ERR_NOMEM, "bc"));
      if (hipMemset(s->bct[v][sd], 0, n + 2) != hipSuccess) return bail(fail(IXH_ERR_HIP, "bc memset"));
      if ((rc = dmalloc(&s->bcv[v][sd], n + 2))) return bail(rc);
    }
  *out = s;
  return 0;
}

int ixh_destroy(ixh_state *s) {
  if (!s) return 0;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  double *ptrs[] = {s->d_geo, s->p, s->bo, s->au, s->av, s->resu, s->part, s->d_out, s->d_hist,
                    s->pT, s->auT, s->avT};
  for (double *p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < kNbuf; ++k) {
    if (s->ub[k]) (void)hipFree(s->ub[k]);
    if (s->vb[k]) (void)hipFree(s->vb[k]);
    if (s->tb[k]) (void)hipFree(s->tb[k]);
    if (s->pcb[k]) (void)hipFree(s->pcb[k]);
    double *tws[] = {s->ubT[k], s->vbT[k], s->tbT[k], s->pcbT[k]};
    for (double *p : tws)
      if (p) (void)hipFree(p);
  }
  for (int v = 0; v < 3; ++v) {
    double *srcs[] = {s->src_c[v], s->src_p[v], s->src_cT[v], s->src_pT[v]};
    for (double *p : srcs)
      if (p) (void)hipFree(p);
    for (int sd = 0; sd < 4; ++sd) {
      if (s->bct[v][sd]) (void)hipFree(s->bct[v][sd]);
      if (s->bcv[v][sd]) (void)hipFree(s->bcv[v][sd]);
    }
  }
  if (s->d_slot) (void)hipFree(s->d_slot);
  if (s->d_fast) (void)hipFree(s->d_fast);
  if (s->d_f32) (void)hipFree(s->d_f32);
  for (double *p : s->adj_ck) (void)hipFree(p);
  for (double *p : s->adj_tape) (void)hipFree(p);
  for (double *p : s->adj2_tape) (void)hipFree(p);
  for (double *p : s->adj2_w) (void)hipFree(p);
  for (double *p : s->adj_w)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 4; ++k)
    if (s->adj_g[k]) (void)hipFree(s->adj_g[k]);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->h_stage) (void)hipHostFree(s->h_stage);
  for (hipEvent_t e : s->lev) (void)hipEventDestroy(e);
  if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

int ixh_set_params(ixh_state *s, const ixh_params *params) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  CHK(check_params(params));
  if (params->precision != s->prm.precision)
    return fail(IXH_ERR_ARG, "the precision of a state is fixed when it is created");
  // the factor tables are rewritten with synchronous copies: work still
  // queued on the state's stream (possibly a non-blocking one) reads them
  HIPCHK(hipStreamSynchronize(s->stream));
  s->prm = *params;
  return build_fast(s);
}

int ixh_set_stream(ixh_state *s, void *stream) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->own_stream) HIPCHK(hipStreamDestroy(s->stream));
  s->own_stream = false;
  if (!stream) {
    HIPCHK(hipStreamCreate(&s->stream));
    s->own_stream = true;
  } else {
    s->stream = (hipStream_t)stream;
  }
  s->launch = s->stream;
  return 0;
}

int ixh_synchronize(ixh_state *s) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

static double *field_ptr(ixh_state *s, int f) {
  switch (f) {
    case IXH_U: return s->ub[s->u_cur];
    case IXH_V: return s->vb[s->v_cur];
    case IXH_T: return s->tb[s->t_cur];
    case IXH_P: return s->p;
    case IXH_U_ANT: return s->ub[s->u_ant];
    case IXH_V_ANT: return s->vb[s->v_ant];
    case IXH_T_ANT: return s->tb[s->t_ant];
    case IXH_PCORR: return s->pcb[s->pc_cur];
    case IXH_BO: return s->bo;
    case IXH_AU: return s->au;
    case IXH_AV: return s->av;
    case IXH_RESU: return s->resu;
  }
  return nullptr;
}

static double *field_ptr_T(ixh_state *s, int f) {
  switch (f) {
    case IXH_U: return s->ubT[s->u_cur];
    case IXH_V: return s->vbT[s->v_cur];
    case IXH_T: return s->tbT[s->t_cur];
    case IXH_P: return s->pT;
    case IXH_U_ANT: return s->ubT[s->u_ant];
    case IXH_V_ANT: return s->vbT[s->v_ant];
    case IXH_T_ANT: return s->tbT[s->t_ant];
    case IXH_PCORR: return s->pcbT[s->pc_cur];
    case IXH_AU: return s->auT;
    case IXH_AV: return s->avT;
  }
  return nullptr;  // b_o and Res_u live in the row layout only
}

}  // extern "C"

// fields held in the state's real type (float in the FP32-mixed path)
static bool real_field(int f) { return f != IXH_P && f != IXH_BO && f != IXH_RESU; }

// host array (extent m x n, leading dimension ld) -> device copy in the row
// layout (pitch P) or the transposed one (pitch Q), as T
template <typename T>
static int upload_as(ixh_state *s, double *dst, const double *src, int ld, int m, int n, bool tr) {
  const size_t pitch = tr ? s->Q : s->P;
  std::vector<T> t(pitch * (tr ? m : n), T(0));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      const size_t o = tr ? (size_t)j + (size_t)i * pitch : (size_t)i + (size_t)j * pitch;
      t[o] = (T)src[(size_t)i + (size_t)j * ld];
    }
  HIPCHK(hipMemcpy(dst, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

extern "C" {

int ixh_upload_field(ixh_state *s, int32_t f, const double *src, int32_t ld) {
  if (!s || !src || f < 0 || f >= IXH_NFIELDS) return fail(IXH_ERR_ARG, "bad upload arguments");
  int m, n;
  field_extent(s, f, &m, &n);
  if (ld < m) return fail(IXH_ERR_ARG, "ld %d < first extent %d", ld, m);
  // a field that shares its buffer with its lagged copy gets its own first
  CHK(unalias(s, f));
  if (f == IXH_PCORR) s->pc_zero = false;
  HIPCHK(hipStreamSynchronize(s->stream));
  const bool fl = s->f32 && real_field(f);
  if (fl)
    CHK(upload_as<float>(s, field_ptr(s, f), src, ld, m, n, false));
  else
    CHK(upload_as<double>(s, field_ptr(s, f), src, ld, m, n, false));
  if (double *dT = field_ptr_T(s, f)) {
    if (fl)
      CHK(upload_as<float>(s, dT, src, ld, m, n, true));
    else
      CHK(upload_as<double>(s, dT, src, ld, m, n, true));
  }
  return 0;
}

// a pinned staging buffer of at least `bytes` for downloads and history
// records; never smaller than one field, so that the short copies of a run's
// records do not reallocate it (a pinned allocation takes hundreds of us)
static int stage_bytes(ixh_state *s, size_t bytes) {
  if (bytes <= s->h_stage_bytes) return 0;
  bytes = std::max(bytes, s->fsize * sizeof(double));
  if (s->h_stage) HIPCHK(hipHostFree(s->h_stage));
  s->h_stage = nullptr;
  s->h_stage_bytes = 0;
  HIPCHK(hipHostMalloc(&s->h_stage, bytes, hipHostMallocDefault));
  s->h_stage_bytes = bytes;
  return 0;
}

int ixh_download_field(ixh_state *s, int32_t f, double *dst, int32_t ld) {
  if (!s || !dst || f < 0 || f >= IXH_NFIELDS) return fail(IXH_ERR_ARG, "bad download arguments");
  int m, n;
  field_extent(s, f, &m, &n);
  if (ld < m) return fail(IXH_ERR_ARG, "ld %d < first extent %d", ld, m);
  if (f == IXH_PCORR && s->pc_zero) {
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int j = 0; j < n; ++j) memset(dst + (size_t)j * ld, 0, (size_t)m * 8);
    return 0;
  }
  // the whole pitched field in one copy to pinned memory, then row by row
  // on the host (a 2-D copy to pageable memory went row by row: 20-30 ms
  // for a 301x701 field)
  const bool fl = s->f32 && real_field(f);
  const size_t cnt = (size_t)s->P * n, esz = fl ? sizeof(float) : sizeof(double);
  CHK(stage_bytes(s, cnt * esz));
  HIPCHK(hipMemcpyAsync(s->h_stage, field_ptr(s, f), cnt * esz, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (fl) {
    const float *t = static_cast<const float *>(s->h_stage);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) dst[(size_t)i + (size_t)j * ld] = t[(size_t)i + (size_t)j * s->P];
  } else {
    const double *t = static_cast<const double *>(s->h_stage);
    for (int j = 0; j < n; ++j) memcpy(dst + (size_t)j * ld, t + (size_t)j * s->P, (size_t)m * sizeof(double));
  }
  return 0;
}

int ixh_set_bc(ixh_state *s, int32_t var, int32_t side, const int32_t *type, const double *value) {
  if (!s || var < 0 || var > 2 || side < 0 || side > 3 || !type || !value)
    return fail(IXH_ERR_ARG, "bad boundary-condition arguments");
  const int n = s->bclen[var][side];
  std::vector<int8_t> t(n + 2, 0);
  std::vector<double> v(n + 2, 0.0);
  for (int k = 1; k <= n; ++k) {
    const int32_t ty = type[k - 1];
    if (ty < IXH_BC_NONE || ty > IXH_BC_NEUM) return fail(IXH_ERR_ARG, "unknown boundary type %d", ty);
    if (ty == IXH_BC_NONE && k >= 2 && k <= n - 1)
      return fail(IXH_ERR_ARG, "node %d of side %d of variable %d has no boundary condition", k, side, var);
    t[k] = (int8_t)ty;
    v[k] = value[k - 1];
  }
  // queued sweeps of the state may still read the old tables
  HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipMemcpy(s->bct[var][side], t.data(), n + 2, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(s->bcv[var][side], v.data(), (n + 2) * sizeof(double), hipMemcpyHostToDevice));
  s->bc_set[var][side] = true;
  return 0;
}

int ixh_set_sources(ixh_state *s, int32_t var, const double *sc, const double *sp) {
  if (!s || var < 0 || var > 2) return fail(IXH_ERR_ARG, "bad source arguments");
  // queued sweeps of the state may still read the old sources
  HIPCHK(hipStreamSynchronize(s->stream));
  if (!sc && !sp) {
    s->has_src[var] = false;
    return 0;
  }
  if (!sc || !sp) return fail(IXH_ERR_ARG, "Sc and Sp must both be given or both be NULL");
  int rc;
  if (!s->src_c[var] && ((rc = dmalloc(&s->src_c[var], s->fsize)) || (rc = dmalloc(&s->src_p[var], s->fsize)) ||
                          (rc = dmalloc(&s->src_cT[var], s->fsize)) || (rc = dmalloc(&s->src_pT[var], s->fsize))))
    return rc;
  const int fid = var == 0 ? IXH_U : var == 1 ? IXH_V : IXH_T;
  int m, n;
  field_extent(s, fid, &m, &n);
  // whole pitched arrays in one copy each (a 2-D copy from pageable memory
  // goes row by row)
  CHK(upload_as<double>(s, s->src_c[var], sc, m, m, n, false));
  CHK(upload_as<double>(s, s->src_p[var], sp, m, m, n, false));
  CHK(upload_as<double>(s, s->src_cT[var], sc, m, m, n, true));
  CHK(upload_as<double>(s, s->src_pT[var], sp, m, m, n, true));
  s->has_src[var] = true;
  return 0;
}

int ixh_momentum_u(ixh_state *s, int32_t dir) {
  if (!s || (dir != IXH_DIR_X && dir != IXH_DIR_Y)) return fail(IXH_ERR_ARG, "bad arguments");
  return sweep_u(s, dir);
}

int ixh_momentum_v(ixh_state *s, int32_t dir, double *err_v) {
  if (!s || (dir != IXH_DIR_X && dir != IXH_DIR_Y)) return fail(IXH_ERR_ARG, "bad arguments");
  CHK(sweep_v(s, dir));
  if (dir == IXH_DIR_X && err_v) {
    double o[8];
    CHK(read_norms(s, 1, o));
    *err_v = o[0];
  }
  return 0;
}

int ixh_pcorr_zero(ixh_state *s) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  s->pc_zero = true;
  return 0;
}

int ixh_pcorr(ixh_state *s, int32_t dir, double *err_p, double *maxbo) {
  if (!s || (dir != IXH_DIR_X && dir != IXH_DIR_Y)) return fail(IXH_ERR_ARG, "bad arguments");
  CHK(sweep_p(s, dir));
  if (dir == IXH_DIR_X && (err_p || maxbo)) {
    double o[8];
    CHK(read_norms(s, 2 | 4, o));
    if (err_p) *err_p = o[1];
    if (maxbo) *maxbo = o[2];
  }
  return 0;
}

int ixh_correct(ixh_state *s) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  return do_correct(s);
}

int ixh_energy(ixh_state *s, int32_t dir, double *err_t) {
  if (!s || (dir != IXH_DIR_X && dir != IXH_DIR_Y)) return fail(IXH_ERR_ARG, "bad arguments");
  CHK(sweep_t(s, dir));
  if (dir == IXH_DIR_X && err_t) {
    double o[8];
    CHK(read_norms(s, 8, o));
    *err_t = o[3];
  }
  return 0;
}

int ixh_residual_u(ixh_state *s, double *residuo) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  CHK(do_residual(s));
  if (residuo) {
    double o[8];
    CHK(read_norms(s, 16, o));
    *residuo = o[4];
  }
  return 0;
}

int ixh_lag(ixh_state *s) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  s->u_ant = s->u_cur;
  s->v_ant = s->v_cur;
  s->t_ant = s->t_cur;
  return 0;
}

int ixh_simple_iterations(ixh_state *s, int32_t nsteps, ixh_hist *hist, int32_t cap, int32_t *nrec) {
  if (!s || nsteps < 0 || cap < 0 || (cap > 0 && !hist)) return fail(IXH_ERR_ARG, "bad arguments");
  for (int v = 0; v < 3; ++v) CHK(check_bcs(s, v));
  const ixh_params &p = s->prm;
  int32_t count = 0;
  const bool single = p.iter_ecuaci_max < 0 && p.iter_simple_max < 0;
  if (single) {
    // every inner loop and the SIMPLE loop run exactly once per step
    CHK(ensure_hist(s, nsteps > 0 ? nsteps : 1));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s->stream, s->d_slot, 0);
    HIPCHK(hipGetLastError());
    if (p.ri_x != 0.0) {
      // a u equation that reads T: upstream order, one kernel after the
      // other (per-kernel event timing with profiling on)
      for (int k = 0; k < nsteps; ++k) CHK(fast_step(s));
    } else {
      // fused launches on one stream: work on separate streams did not
      // overlap on this stack (the cross-stream waits cost more than the
      // overlap won), and a captured HIP graph of the step ran slower than
      // direct launches
      CHK(run_fused(s, nsteps));
    }
    const int keep = nsteps < cap ? nsteps : cap;
    // the records through the pinned staging buffer (a small copy to
    // pageable memory is staged by the runtime at extra cost)
    CHK(stage_bytes(s, (size_t)8 * (keep > 0 ? keep : 1) * sizeof(double)));
    const double *h = static_cast<const double *>(s->h_stage);
    if (keep > 0)
      HIPCHK(hipMemcpyAsync(s->h_stage, s->d_hist, (size_t)8 * keep * sizeof(double), hipMemcpyDeviceToHost,
                            s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int k = 0; k < keep; ++k) {
      ixh_hist &r = hist[k];
      r.err_v = h[8 * k + 0];
      r.err_p = h[8 * k + 1];
      r.maxbo = h[8 * k + 2];
      r.err_t = h[8 * k + 3];
      r.residuo = h[8 * k + 4];
      r.passes_mom = r.passes_p = r.passes_t = 1;
      r.simple_iter = 0;
      r.logged = (r.maxbo < p.conv_paso && r.residuo < p.conv_resi) ? 1 : 0;
      r.pad_ = 0;
    }
    count = nsteps;
  } else {
    // data-dependent pass counts: read the norms back after every pass
    for (int step = 0; step < nsteps; ++step) {
      int iter_simple = 0;
      for (;;) {
        ixh_hist r;
        memset(&r, 0, sizeof r);
        int it = 0;
        for (;;) {
          CHK(sweep_u(s, IXH_DIR_Y));
          CHK(sweep_u(s, IXH_DIR_X));
          CHK(sweep_v(s, IXH_DIR_Y));
          CHK(sweep_v(s, IXH_DIR_X));
          double o[8];
          CHK(read_norms(s, 1, o));
          r.err_v = o[0];
          r.passes_mom++;
          if (r.err_v < p.conv_u || it > p.iter_ecuaci_max) break;
          it++;
        }
        it = 0;
        s->pc_zero = true;
        for (;;) {
          CHK(sweep_p(s, IXH_DIR_Y));
          CHK(sweep_p(s, IXH_DIR_X));
          double o[8];
          CHK(read_norms(s, 2 | 4, o));
          r.err_p = o[1];
          r.maxbo = o[2];
          r.passes_p++;
          if (r.err_p < p.conv_p || it > p.iter_ecuaci_max) break;
          it++;
        }
        it = 0;
        CHK(do_correct(s));
        for (;;) {
          CHK(sweep_t(s, IXH_DIR_Y));
          CHK(sweep_t(s, IXH_DIR_X));
          double o[8];
          CHK(read_norms(s, 8, o));
          r.err_t = o[3];
          r.passes_t++;
          if (r.err_t < p.conv_t || it > p.iter_ecuaci_max) break;
          it++;
        }
        CHK(do_residual(s));
        double o[8];
        CHK(read_norms(s, 16, o));
        r.residuo = o[4];
        bool done = false;
        if (r.maxbo < p.conv_paso && r.residuo < p.conv_resi) {
          iter_simple = 0;
          done = true;
          r.logged = 1;
        } else if (iter_simple > p.iter_simple_max) {
          iter_simple = 0;
          done = true;
          r.logged = 0;
        } else {
          iter_simple++;
          r.logged = 1;
        }
        r.simple_iter = iter_simple;
        if (count < cap) hist[count] = r;
        count++;
        if (done) break;
      }
      CHK(ixh_lag(s));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  if (nrec) *nrec = count;
  return 0;
}

int ixh_nusselt_y(ixh_state *s, int32_t i0, int32_t i1, double *nu0, double *nu1) {
  if (!s || !nu0 || !nu1) return fail(IXH_ERR_ARG, "bad arguments");
  if (i0 < 1 || i1 < i0 || i1 + 2 > s->mi + 1)
    return fail(IXH_ERR_ARG, "Nusselt range %d..%d outside 1..%d", i0, i1, s->mi - 1);
  NuArgs A;
  A.mi = s->mi; A.nj = s->nj; A.P = s->P; A.i0 = i0; A.i1 = i1;
  A.g = s->geo; A.t = s->tb[s->t_cur]; A.out = s->d_out; A.vertical = 0;
  if (s->f32)
    hipLaunchKernelGGL(k_nusselt<float>, dim3(1), dim3(kBlock), 0, s->stream, A);
  else
    hipLaunchKernelGGL(k_nusselt<double>, dim3(1), dim3(kBlock), 0, s->stream, A);
  HIPCHK(hipGetLastError());
  double o[2];
  HIPCHK(hipMemcpyAsync(o, s->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *nu0 = o[0];
  *nu1 = o[1];
  return 0;
}

int ixh_nusselt_x(ixh_state *s, double *nu_w, double *nu_e) {
  if (!s || !nu_w || !nu_e) return fail(IXH_ERR_ARG, "bad arguments");
  NuArgs A;
  A.mi = s->mi; A.nj = s->nj; A.P = s->P; A.i0 = 0; A.i1 = 0;
  A.g = s->geo; A.t = s->tb[s->t_cur]; A.out = s->d_out; A.vertical = 1;
  if (s->f32)
    hipLaunchKernelGGL(k_nusselt<float>, dim3(1), dim3(kBlock), 0, s->stream, A);
  else
    hipLaunchKernelGGL(k_nusselt<double>, dim3(1), dim3(kBlock), 0, s->stream, A);
  HIPCHK(hipGetLastError());
  double o[2];
  HIPCHK(hipMemcpyAsync(o, s->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *nu_w = o[0];
  *nu_e = o[1];
  return 0;
}

int ixh_tridiagonal(const double *a, double *b, const double *c, double *r, int32_t n, int32_t nsys,
                    int32_t solver) {
  CHK(check_tridiag(a, b, c, r, n, nsys, solver));
  if (nsys == 0) return 0;
  const size_t tot = (size_t)n * nsys;
  double *d = nullptr;
  HIPCHK(hipMalloc((void **)&d, 4 * tot * sizeof(double)));
  const int rc = tridiag_on(d, a, b, c, r, n, nsys, solver);
  (void)hipFree(d);
  return rc;
}

int ixh_tridiagonal_noalloc(const double *a, double *b, const double *c, double *r, int32_t n, int32_t nsys,
                            int32_t solver) {
  CHK(check_tridiag(a, b, c, r, n, nsys, solver));
  if (nsys == 0) return 0;
  if ((size_t)n * nsys > kTriScratchPerArray)
    return fail(IXH_ERR_UNSUPP, "ixh_tridiagonal_noalloc: n*nsys = %zu exceeds its scratch of %zu values per array",
                (size_t)n * nsys, kTriScratchPerArray);
  double *d = nullptr;
  CHK(tri_scratch(&d));
  return tridiag_on(d, a, b, c, r, n, nsys, solver);
}

int ixh_set_profiling(ixh_state *s, int32_t on) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  s->profiling = on != 0;
  for (int k = 0; k < K_NUM; ++k) {
    s->kms[k] = 0.0;
    s->kcount[k] = 0;
  }
  return 0;
}

int ixh_kernel_times(ixh_state *s, double *ms, int32_t *launches, int32_t cap, int32_t *nk) {
  if (!s) return fail(IXH_ERR_ARG, "state is NULL");
  for (int k = 0; k < K_NUM && k < cap; ++k) {
    if (ms) ms[k] = s->kms[k];
    if (launches) launches[k] = s->kcount[k];
  }
  if (nk) *nk = K_NUM;
  return 0;
}

}  // extern "C"

// ====================================================================== batch
// Many cavities advanced together.  Each phase of the single-pass time step
// is one launch over the lines (tiles, rows) of every cavity, through a
// workgroup -> (cavity, unit) table; per cavity the arithmetic and the phase
// order are those of ixh_simple_iterations' sequential step, so a cavity's
// fields match a run of that cavity alone bit for bit.
namespace {

enum BatchPhase { PH_U_Y = 0, PH_U_X, PH_V_Y, PH_V_X, PH_P_Y, PH_P_X, PH_T_Y, PH_T_X, PH_CORR, PH_RES, PH_NUM };
const int kPhEq[8] = {EQ_U, EQ_U, EQ_V, EQ_V, EQ_P, EQ_P, EQ_T, EQ_T};
const int kPhDir[8] = {IXH_DIR_Y, IXH_DIR_X, IXH_DIR_Y, IXH_DIR_X, IXH_DIR_Y, IXH_DIR_X, IXH_DIR_Y, IXH_DIR_X};
constexpr int kXcd = 8;          // workgroup b of a launch runs on XCD b % 8
constexpr int kRotLen = 12;      // ints of a cavity's rotation state
constexpr int kMaxCached = 16;   // distinct rotation states kept per call

// first unit and unit count of a phase on one cavity (lines, rows or tiles)
void phase_units(const ixh_state *s, int ph, int *first, int *count) {
  *first = 2;
  switch (ph) {
    case PH_U_Y: *count = s->mi - 2; break;
    case PH_U_X: case PH_P_X: case PH_T_X: case PH_RES: *count = s->nj - 1; break;
    case PH_V_X: *count = s->nj - 2; break;
    case PH_CORR: *first = 0; *count = corr_tiles_x(s) * corr_tiles_y(s); break;
    default: *count = s->mi - 1; break;  // v, p', T y-sweeps
  }
}

void rot_get(const ixh_state *s, int *r) {
  r[0] = s->u_cur; r[1] = s->u_ant; r[2] = s->v_cur; r[3] = s->v_ant; r[4] = s->t_cur; r[5] = s->t_ant;
  r[6] = s->pc_cur; r[7] = s->v_snap; r[8] = s->t_snap; r[9] = s->pc_snap; r[10] = s->pc_zero; r[11] = s->u_hold;
}

void rot_set(ixh_state *s, const int *r) {
  s->u_cur = r[0]; s->u_ant = r[1]; s->v_cur = r[2]; s->v_ant = r[3]; s->t_cur = r[4]; s->t_ant = r[5];
  s->pc_cur = r[6]; s->v_snap = r[7]; s->t_snap = r[8]; s->pc_snap = r[9]; s->pc_zero = r[10] != 0;
  s->u_hold = r[11];
}

// kernel arguments of one time step of every cavity, for one rotation state
struct BatchStep {
  std::vector<int> key, post;  // rotation states before and after the step
  char *d = nullptr;           // SweepArgs[8][n], CorrArgs[n], ResArgs[n]
  bool res_with_t = false;     // the u residual rides on the T x-sweep's launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

}  // namespace

struct ixh_batch {
  std::vector<ixh_state *> st;
  int n = 0, solver = 0;
  bool f32 = false, delta = false, sr = false;
  int ept[8] = {};
  // one 64-thread workgroup per line (PCR, every line at most 320 nodes):
  // four times as many lines resident per CU as with 256 threads
  bool wave_ok = false, wave_lines = false;
  int ept64[8] = {};
  // the u residual in the T x-sweep's launch, where no cavity's u equation
  // reads T
  bool res_with_t = false;
  static_assert(kLinesPerGroup == IXH_BATCH_LINE_GROUP, "line group size of the header");
  // consecutive lines of a cavity per workgroup in the sweep phases (one
  // wave each): 1, or kLinesPerGroup, whose other-layout stores leave as
  // runs of that many neighbouring values
  int group = 1;
  int order = IXH_ORDER_XCD;
  bool graph = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int2 *d_wg = nullptr;
  size_t wg_cap = 0;
  size_t wg_off[PH_NUM] = {};
  int nwg[PH_NUM] = {};
  std::vector<BatchStep> steps;
  FinArgs *d_fin = nullptr;
  double *d_hist = nullptr;
  size_t hist_len = 0;
  int *d_slot = nullptr;
};

namespace {

void batch_clear_steps(ixh_batch *b) {
  for (BatchStep &E : b->steps) {
    if (E.exec) (void)hipGraphExecDestroy(E.exec);
    if (E.graph) (void)hipGraphDestroy(E.graph);
    if (E.d) (void)hipFree(E.d);
  }
  b->steps.clear();
}

// workgroup tables of every phase for the current ordering
int batch_tables(ixh_batch *b) {
  const int n = b->n;
  std::vector<int> byc(n);
  for (int c = 0; c < n; ++c) byc[c] = c;
  auto cells = [&](int c) { return (long long)b->st[c]->mi * b->st[c]->nj; };
  if (b->order != IXH_ORDER_CAVITY)
    std::stable_sort(byc.begin(), byc.end(), [&](int x, int y) { return cells(x) > cells(y); });
  // XCD bins: longest-processing-time-first assignment of whole cavities
  std::vector<int> bin(n, 0);
  if (b->order == IXH_ORDER_XCD) {
    long long load[kXcd] = {};
    for (int c : byc) {
      int k = 0;
      for (int x = 1; x < kXcd; ++x)
        if (load[x] < load[k]) k = x;
      bin[c] = k;
      load[k] += cells(c);
    }
  }
  std::vector<int2> all;
  for (int ph = 0; ph < PH_NUM; ++ph) {
    b->wg_off[ph] = all.size();
    // sweep phases in line groups: one entry per group, its first line
    const int stride = ph < PH_CORR ? b->group : 1;
    if (b->order != IXH_ORDER_XCD) {
      for (int c : byc) {
        int f, m;
        phase_units(b->st[c], ph, &f, &m);
        for (int k = 0; k < m; k += stride) all.push_back(make_int2(c, f + k));
      }
    } else {
      std::vector<int2> lists[kXcd];
      for (int c : byc) {
        int f, m;
        phase_units(b->st[c], ph, &f, &m);
        for (int k = 0; k < m; k += stride) lists[bin[c]].push_back(make_int2(c, f + k));
      }
      size_t rounds = 0;
      for (int x = 0; x < kXcd; ++x) rounds = std::max(rounds, lists[x].size());
      for (size_t r = 0; r < rounds; ++r)
        for (int x = 0; x < kXcd; ++x)
          all.push_back(r < lists[x].size() ? lists[x][r] : make_int2(-1, -1));
    }
    b->nwg[ph] = (int)(all.size() - b->wg_off[ph]);
  }
  if (all.size() > b->wg_cap) {
    if (b->d_wg) HIPCHK(hipFree(b->d_wg));
    b->d_wg = nullptr;
    HIPCHK(hipMalloc((void **)&b->d_wg, all.size() * sizeof(int2)));
    b->wg_cap = all.size();
  }
  HIPCHK(hipMemcpy(b->d_wg, all.data(), all.size() * sizeof(int2), hipMemcpyHostToDevice));
  return 0;
}

// build the kernel arguments of one step of every cavity (advancing their
// rotation) and upload them
int batch_prepare(ixh_batch *b, BatchStep &E) {
  const int n = b->n;
  std::vector<SweepArgs> sw(8 * (size_t)n);
  std::vector<CorrArgs> co(n);
  std::vector<ResArgs> re(n);
  bool all_ri0 = true;
  for (int c = 0; c < n; ++c) {
    ixh_state *s = b->st[c];
    CHK(prep_u(s, IXH_DIR_Y, sw[PH_U_Y * (size_t)n + c]));
    CHK(prep_u(s, IXH_DIR_X, sw[PH_U_X * (size_t)n + c]));
    CHK(prep_v(s, IXH_DIR_Y, sw[PH_V_Y * (size_t)n + c]));
    CHK(prep_v(s, IXH_DIR_X, sw[PH_V_X * (size_t)n + c]));
    s->pc_zero = true;
    CHK(prep_p(s, IXH_DIR_Y, sw[PH_P_Y * (size_t)n + c]));
    CHK(prep_p(s, IXH_DIR_X, sw[PH_P_X * (size_t)n + c]));
    // after the sweeps u, v never share a buffer with u_ant, v_ant, so the
    // correction's unaliasing issues no copy here
    bool active = false;
    CHK(prep_correct(s, co[c], active));
    if (!active) return fail(IXH_ERR_ARG, "batch: p' is zero after its sweeps");
    CHK(prep_t(s, IXH_DIR_Y, sw[PH_T_Y * (size_t)n + c]));
    CHK(prep_t(s, IXH_DIR_X, sw[PH_T_X * (size_t)n + c]));
    prep_residual(s, re[c]);
    all_ri0 = all_ri0 && s->prm.ri_x == 0.0;
    s->u_ant = s->u_cur;
    s->v_ant = s->v_cur;
    s->t_ant = s->t_cur;
  }
  const size_t bs = sw.size() * sizeof(SweepArgs), bc = co.size() * sizeof(CorrArgs),
               br = re.size() * sizeof(ResArgs);
  HIPCHK(hipMalloc((void **)&E.d, bs + bc + br));
  HIPCHK(hipMemcpy(E.d, sw.data(), bs, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(E.d + bs, co.data(), bc, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(E.d + bs + bc, re.data(), br, hipMemcpyHostToDevice));
  E.res_with_t = b->res_with_t && all_ri0 && b->group == 1;
  return 0;
}

// One sweep phase.  The p' y-sweep reads no p' (zero at entry in every
// cavity: batch_prepare starts each step's p' from zero).  The T x-sweep
// also runs the u residual rows when re is not null.
template <int EQ, int DIR, int SOLVER, typename R>
int batch_sweep_t(ixh_batch *b, int ph, const SweepArgs *D, hipStream_t st, const ResArgs *re) {
  const int2 *wg = b->d_wg + b->wg_off[ph];
  if (b->group > 1) {
    HIPCHK((launch_sweep_group_batch_k<EQ, DIR, SOLVER, R>(st, D, wg, b->nwg[ph], b->ept64[ph])));
  } else if constexpr (EQ == EQ_P && DIR == DIR_Y) {
    HIPCHK((launch_sweep_pz_batch_k<SOLVER, R>(st, D, wg, b->nwg[ph], b->ept[ph], b->ept64[ph], b->wave_lines)));
  } else if constexpr (EQ == EQ_T && DIR == DIR_X) {
    if (re)
      HIPCHK((launch_sweep_t_res_batch_k<SOLVER, R>(st, D, wg, b->nwg[ph], re, b->d_wg + b->wg_off[PH_RES],
                                                    b->nwg[PH_RES], b->ept[ph], b->ept64[ph], b->wave_lines)));
    else
      HIPCHK((launch_sweep_batch_k<EQ, DIR, SOLVER, R>(st, D, wg, b->nwg[ph], b->ept[ph], b->ept64[ph],
                                                       b->wave_lines)));
  } else {
    HIPCHK((launch_sweep_batch_k<EQ, DIR, SOLVER, R>(st, D, wg, b->nwg[ph], b->ept[ph], b->ept64[ph],
                                                     b->wave_lines)));
  }
  return 0;
}

template <int EQ, int DIR>
int batch_sweep(ixh_batch *b, int ph, const SweepArgs *D, hipStream_t st, const ResArgs *re = nullptr) {
  if (b->nwg[ph] == 0 && !re) return 0;
  D += (size_t)ph * b->n;
  if (b->sr) return batch_sweep_t<EQ, DIR, SOLVE_DELTA_SR, float>(b, ph, D, st, re);
  if (b->f32) return batch_sweep_t<EQ, DIR, SOLVE_PCR, float>(b, ph, D, st, re);
  if (b->delta) return batch_sweep_t<EQ, DIR, SOLVE_DELTA, float>(b, ph, D, st, re);
  if (b->solver == IXH_SOLVER_THOMAS) return batch_sweep_t<EQ, DIR, SOLVE_THOMAS, double>(b, ph, D, st, re);
  if (b->solver == IXH_SOLVER_PCR_EXACT)
    return batch_sweep_t<EQ, DIR, SOLVE_PCR_EXACT, double>(b, ph, D, st, re);
  return batch_sweep_t<EQ, DIR, SOLVE_PCR, double>(b, ph, D, st, re);
}

// the launches of one step, in the sequential order of a single cavity
int batch_launch(ixh_batch *b, const BatchStep &E, hipStream_t st, int cap) {
  const int n = b->n;
  const SweepArgs *sw = (const SweepArgs *)E.d;
  const CorrArgs *co = (const CorrArgs *)(E.d + 8 * (size_t)n * sizeof(SweepArgs));
  const ResArgs *re = (const ResArgs *)((const char *)co + (size_t)n * sizeof(CorrArgs));
  CHK((batch_sweep<EQ_U, DIR_Y>(b, PH_U_Y, sw, st)));
  CHK((batch_sweep<EQ_U, DIR_X>(b, PH_U_X, sw, st)));
  CHK((batch_sweep<EQ_V, DIR_Y>(b, PH_V_Y, sw, st)));
  CHK((batch_sweep<EQ_V, DIR_X>(b, PH_V_X, sw, st)));
  CHK((batch_sweep<EQ_P, DIR_Y>(b, PH_P_Y, sw, st)));
  CHK((batch_sweep<EQ_P, DIR_X>(b, PH_P_X, sw, st)));
  const int2 *wgc = b->d_wg + b->wg_off[PH_CORR];
  if (b->f32)
    hipLaunchKernelGGL(k_correct_batch<float>, dim3(b->nwg[PH_CORR]), dim3(kBlock), 0, st, co, wgc);
  else
    hipLaunchKernelGGL(k_correct_batch<double>, dim3(b->nwg[PH_CORR]), dim3(kBlock), 0, st, co, wgc);
  HIPCHK(hipGetLastError());
  CHK((batch_sweep<EQ_T, DIR_Y>(b, PH_T_Y, sw, st)));
  CHK((batch_sweep<EQ_T, DIR_X>(b, PH_T_X, sw, st, E.res_with_t && b->nwg[PH_RES] > 0 ? re : nullptr)));
  const int2 *wgr = b->d_wg + b->wg_off[PH_RES];
  if (E.res_with_t || b->nwg[PH_RES] == 0) {
    // the residual rows ran with the T x-sweep (or there are none)
  } else if (b->wave_lines) {
    if (b->f32)
      hipLaunchKernelGGL((k_residual_u_batch<float, 64>), dim3(b->nwg[PH_RES]), dim3(64), 0, st, re, wgr);
    else
      hipLaunchKernelGGL((k_residual_u_batch<double, 64>), dim3(b->nwg[PH_RES]), dim3(64), 0, st, re, wgr);
  } else if (b->f32) {
    hipLaunchKernelGGL(k_residual_u_batch<float>, dim3(b->nwg[PH_RES]), dim3(kBlock), 0, st, re, wgr);
  } else {
    hipLaunchKernelGGL(k_residual_u_batch<double>, dim3(b->nwg[PH_RES]), dim3(kBlock), 0, st, re, wgr);
  }
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_finalize_batch, dim3(n), dim3(kBlock), 0, st, (const FinArgs *)b->d_fin,
                     (const int *)b->d_slot, cap);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_add_int, dim3(1), dim3(64), 0, st, b->d_slot, 1);
  HIPCHK(hipGetLastError());
  return 0;
}

// one time step of every cavity: look up (or build) the arguments for the
// current rotation state, then launch them directly or as a graph
int batch_step(ixh_batch *b, int cap, std::vector<int> &key) {
  const int n = b->n;
  key.resize((size_t)n * kRotLen);
  for (int c = 0; c < n; ++c) rot_get(b->st[c], &key[(size_t)c * kRotLen]);
  BatchStep *E = nullptr;
  for (BatchStep &x : b->steps)
    if (x.key == key) E = &x;
  if (E) {
    for (int c = 0; c < n; ++c) rot_set(b->st[c], &E->post[(size_t)c * kRotLen]);
  } else {
    if ((int)b->steps.size() >= kMaxCached) {
      HIPCHK(hipStreamSynchronize(b->stream));
      batch_clear_steps(b);
    }
    b->steps.emplace_back();
    E = &b->steps.back();
    E->key = key;
    CHK(batch_prepare(b, *E));
    E->post.resize(key.size());
    for (int c = 0; c < n; ++c) rot_get(b->st[c], &E->post[(size_t)c * kRotLen]);
  }
  if (!b->graph) return batch_launch(b, *E, b->stream, cap);
  if (!E->exec) {
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_launch(b, *E, b->stream, cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    CHK(rc);
    if (e != hipSuccess) return fail(IXH_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    E->graph = g;
    HIPCHK(hipGraphInstantiate(&E->exec, g, nullptr, nullptr, 0));
  }
  HIPCHK(hipGraphLaunch(E->exec, b->stream));
  return 0;
}

// wait for work queued on the cavities' own streams
int batch_join_cavities(ixh_batch *b) {
  for (ixh_state *s : b->st) HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

// per-cavity fold arguments and the history buffer for `cap` records
int batch_begin(ixh_batch *b, int cap) {
  const int n = b->n;
  const size_t per = (size_t)8 * (cap + 1);
  if (per * n > b->hist_len) {
    if (b->d_hist) HIPCHK(hipFree(b->d_hist));
    b->d_hist = nullptr;
    CHK(dmalloc(&b->d_hist, per * n));
    b->hist_len = per * n;
  }
  std::vector<FinArgs> fin(n);
  for (int c = 0; c < n; ++c) prep_finalize(b->st[c], fin[c], 31, b->d_hist + per * c, nullptr);
  HIPCHK(hipMemcpy(b->d_fin, fin.data(), n * sizeof(FinArgs), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, b->stream, b->d_slot, 0);
  HIPCHK(hipGetLastError());
  return 0;
}

// the arguments depend on each cavity's parameters, sources and boundary
// conditions, which may change between calls: rebuild them per call
int batch_run(ixh_batch *b, int nsteps, int cap) {
  CHK(batch_join_cavities(b));
  HIPCHK(hipStreamSynchronize(b->stream));
  batch_clear_steps(b);
  CHK(batch_begin(b, cap));
  std::vector<int> key;
  for (int k = 0; k < nsteps; ++k) CHK(batch_step(b, cap, key));
  return 0;
}

}  // namespace

extern "C" {

int ixh_batch_create(ixh_state *const *states, int32_t n, ixh_batch **out) {
  if (!states || !out || n <= 0) return fail(IXH_ERR_ARG, "batch: bad arguments");
  *out = nullptr;
  const int solver = states[0] ? states[0]->prm.solver : 0;
  for (int c = 0; c < n; ++c) {
    const ixh_state *s = states[c];
    if (!s) return fail(IXH_ERR_ARG, "batch: state %d is NULL", c);
    for (int d = 0; d < c; ++d)
      if (states[d] == s) return fail(IXH_ERR_ARG, "batch: state %d appears twice", c);
    if (s->prm.solver != solver) return fail(IXH_ERR_ARG, "batch: every cavity must use the same solver");
    if (s->prm.precision != states[0]->prm.precision)
      return fail(IXH_ERR_ARG, "batch: every cavity must use the same precision");
    if (s->mi + 1 > 4 * kBlock || s->nj + 1 > 4 * kBlock)
      return fail(IXH_ERR_UNSUPP, "batch: lines longer than %d nodes are not built (cavity %d)", 4 * kBlock, c);
  }
  ixh_batch *b = new ixh_batch();
  b->st.assign(states, states + n);
  b->n = n;
  b->solver = solver;
  b->f32 = states[0]->f32;
  b->delta = states[0]->delta;
  b->sr = states[0]->sr;
  int longest = 1;
  for (int ph = 0; ph < 8; ++ph) {
    int m = 1;
    for (ixh_state *s : b->st) m = std::max(m, line_len(s, kPhEq[ph], kPhDir[ph]));
    b->ept[ph] = ept_for(m);
    b->ept64[ph] = (m + 63) / 64;
    longest = std::max(longest, m);
  }
  b->wave_ok = solver == IXH_SOLVER_PCR && longest <= 5 * 64;
  b->wave_lines = b->wave_ok;
  b->group = b->wave_ok ? kLinesPerGroup : 1;
  b->res_with_t = false;
  auto bail = [&](int code) {
    ixh_batch_destroy(b);
    return code;
  };
  if (hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess)
    return bail(fail(IXH_ERR_HIP, "batch: stream/event creation failed"));
  if (hipMalloc((void **)&b->d_fin, n * sizeof(FinArgs)) != hipSuccess ||
      hipMalloc((void **)&b->d_slot, 256) != hipSuccess)
    return bail(fail(IXH_ERR_NOMEM, "batch: allocation failed"));
  if (int rc = batch_tables(b)) return bail(rc);
  *out = b;
  return 0;
}

int ixh_batch_destroy(ixh_batch *b) {
  if (!b) return 0;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_clear_steps(b);
  if (b->d_wg) (void)hipFree(b->d_wg);
  if (b->d_fin) (void)hipFree(b->d_fin);
  if (b->d_hist) (void)hipFree(b->d_hist);
  if (b->d_slot) (void)hipFree(b->d_slot);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int ixh_batch_set_schedule(ixh_batch *b, int32_t order, int32_t graph) {
  if (!b) return fail(IXH_ERR_ARG, "batch is NULL");
  if (order != IXH_ORDER_CAVITY && order != IXH_ORDER_LARGEST && order != IXH_ORDER_XCD)
    return fail(IXH_ERR_ARG, "batch: unknown order %d", order);
  HIPCHK(hipStreamSynchronize(b->stream));
  batch_clear_steps(b);
  b->order = order;
  b->graph = graph != 0;
  return batch_tables(b);
}

int ixh_batch_set_shape(ixh_batch *b, int32_t line_threads, int32_t lines_per_group, int32_t residual_with_t) {
  if (!b) return fail(IXH_ERR_ARG, "batch is NULL");
  if (line_threads != 64 && line_threads != kBlock)
    return fail(IXH_ERR_ARG, "batch: %d threads per line (64 or %d)", line_threads, kBlock);
  if (line_threads == 64 && !b->wave_ok)
    return fail(IXH_ERR_UNSUPP, "batch: one wave per line needs the pcr solver and lines of at most %d nodes",
                5 * 64);
  if (lines_per_group != 1 && lines_per_group != kLinesPerGroup)
    return fail(IXH_ERR_ARG, "batch: %d lines per group (1 or %d)", lines_per_group, kLinesPerGroup);
  if (lines_per_group > 1 && line_threads != 64)
    return fail(IXH_ERR_ARG, "batch: line groups hold one-wave lines (64 threads per line)");
  if (lines_per_group > 1 && residual_with_t)
    return fail(IXH_ERR_ARG, "batch: the residual rides on the T x-sweep of one-line workgroups only");
  HIPCHK(hipStreamSynchronize(b->stream));
  batch_clear_steps(b);
  b->wave_lines = line_threads == 64;
  b->res_with_t = residual_with_t != 0;
  if (b->group != lines_per_group) {
    b->group = lines_per_group;
    CHK(batch_tables(b));
  }
  return 0;
}

int ixh_batch_iterations(ixh_batch *b, int32_t nsteps, ixh_hist *hist, int32_t cap) {
  if (!b || nsteps < 0 || cap < 0 || (cap > 0 && !hist)) return fail(IXH_ERR_ARG, "batch: bad arguments");
  for (ixh_state *s : b->st) {
    for (int v = 0; v < 3; ++v) CHK(check_bcs(s, v));
    if (s->prm.iter_ecuaci_max >= 0 || s->prm.iter_simple_max >= 0)
      return fail(IXH_ERR_UNSUPP, "batch: every cavity needs single-pass loops "
                                  "(iter_ecuaci_max < 0 and iter_simple_max < 0)");
    if (s->prm.solver != b->solver) return fail(IXH_ERR_ARG, "batch: a cavity changed its solver");
  }
  const int keep = std::min(nsteps, cap);
  CHK(batch_run(b, nsteps, keep));
  const int n = b->n;
  const size_t per = (size_t)8 * (keep + 1);
  std::vector<double> h(per * n);
  HIPCHK(hipMemcpyAsync(h.data(), b->d_hist, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (int c = 0; c < n; ++c) {
    const ixh_params &p = b->st[c]->prm;
    for (int k = 0; k < keep; ++k) {
      const double *o = &h[per * c + 8 * (size_t)k];
      ixh_hist &r = hist[(size_t)c * cap + k];
      r.err_v = o[0];
      r.err_p = o[1];
      r.maxbo = o[2];
      r.err_t = o[3];
      r.residuo = o[4];
      r.passes_mom = r.passes_p = r.passes_t = 1;
      r.simple_iter = 0;
      r.logged = (r.maxbo < p.conv_paso && r.residuo < p.conv_resi) ? 1 : 0;
      r.pad_ = 0;
    }
  }
  return 0;
}

int ixh_batch_autotune(ixh_batch *b, int32_t trial_steps, int32_t *order, int32_t *graph, int32_t *line_threads,
                       int32_t *lines_per_group, int32_t *residual_with_t, double *ms_per_step) {
  if (!b || trial_steps < 1) return fail(IXH_ERR_ARG, "batch: bad arguments");
  bool any_ri = false;
  for (ixh_state *s : b->st) any_ri = any_ri || s->prm.ri_x != 0.0;
  if (ms_per_step)
    for (int k = 0; k < IXH_BATCH_SCHEDULES; ++k) ms_per_step[k] = -1.0;
  int best = -1;
  double best_ms = 0.0;
  // index 8 (2 order + graph) + 4 (256 threads per line) + 2 (line groups)
  // + residual with T
  for (int cfg = 0; cfg < IXH_BATCH_SCHEDULES; ++cfg) {
    const int o = cfg >> 4, g = (cfg >> 3) & 1, wide = (cfg >> 2) & 1, grp = (cfg >> 1) & 1, rt = cfg & 1;
    if (!wide && !b->wave_ok) continue;
    if (wide && grp) continue;
    if (rt && (any_ri || grp)) continue;
    CHK(ixh_batch_set_schedule(b, o, g));
    CHK(ixh_batch_set_shape(b, wide ? kBlock : 64, grp ? kLinesPerGroup : 1, rt));
    // untimed steps build the arguments (and graphs) of the rotation states
    // the timed steps meet
    CHK(batch_run(b, 2, 0));
    std::vector<int> key;
    for (int k = 0; k < 2; ++k) CHK(batch_step(b, 0, key));
    HIPCHK(hipEventRecord(b->ev0, b->stream));
    for (int k = 0; k < trial_steps; ++k) CHK(batch_step(b, 0, key));
    HIPCHK(hipEventRecord(b->ev1, b->stream));
    HIPCHK(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    const double per = ms / trial_steps;
    if (ms_per_step) ms_per_step[cfg] = per;
    if (best < 0 || per < best_ms) {
      best = cfg;
      best_ms = per;
    }
  }
  const int o = best >> 4, g = (best >> 3) & 1, lt = (best >> 2) & 1 ? kBlock : 64;
  const int lpg = (best >> 1) & 1 ? kLinesPerGroup : 1, rt = best & 1;
  CHK(ixh_batch_set_schedule(b, o, g));
  CHK(ixh_batch_set_shape(b, lt, lpg, rt));
  if (order) *order = o;
  if (graph) *graph = g;
  if (line_threads) *line_threads = lt;
  if (lines_per_group) *lines_per_group = lpg;
  if (residual_with_t) *residual_with_t = rt;
  return 0;
}

}  // extern "C"

// ============================================================ energy adjoint
// Record a run with checkpoints, then run the discrete adjoint of its energy
// steps backward, recomputing the velocities of each segment of `every`
// steps from its checkpoint.
namespace {

enum ManyOp { MANY_COPY, MANY_ZERO, MANY_ADD_CLEAR };

// one launch over whole buffers: dst = src, dst = 0, or dst += src then
// src = 0 (dst = 0 where src is null) for every listed (dst, src) pair
int many(ixh_state *s, ManyOp op, std::initializer_list<std::pair<double *, const double *>> pairs) {
  CopyList L;
  memset(&L, 0, sizeof L);
  int c = 0;
  for (const auto &pr : pairs) {
    if (c == kCopyMax) return fail(IXH_ERR_ARG, "many: more than %d buffers", kCopyMax);
    L.dst[c] = pr.first;
    L.src[c] = pr.second;
    ++c;
  }
  L.n = (int)s->fsize;
  const dim3 grid((unsigned)std::min<size_t>((s->fsize + kBlock - 1) / kBlock, 64), (unsigned)c);
  if (op == MANY_COPY)
    hipLaunchKernelGGL(k_copy_many, grid, dim3(kBlock), 0, s->stream, L);
  else if (op == MANY_ZERO)
    hipLaunchKernelGGL(k_zero_many, grid, dim3(kBlock), 0, s->stream, L);
  else
    hipLaunchKernelGGL(k_add_clear_many, grid, dim3(kBlock), 0, s->stream, L);
  HIPCHK(hipGetLastError());
  return 0;
}

int copy_many(ixh_state *s, std::initializer_list<std::pair<double *, const double *>> pairs) {
  return many(s, MANY_COPY, pairs);
}

constexpr int kCkArrays = 14;  // u, v, T, p and the lagged u, v, T, both layouts

void ck_fields(ixh_state *s, double *f[kCkArrays]) {
  double *a[kCkArrays] = {s->ub[s->u_cur], s->ubT[s->u_cur], s->vb[s->v_cur], s->vbT[s->v_cur],
                          s->tb[s->t_cur], s->tbT[s->t_cur], s->p,          s->pT,
                          s->ub[s->u_ant], s->ubT[s->u_ant], s->vb[s->v_ant], s->vbT[s->v_ant],
                          s->tb[s->t_ant], s->tbT[s->t_ant]};
  for (int k = 0; k < kCkArrays; ++k) f[k] = a[k];
}

int ck_save(ixh_state *s, double *dst) {
  double *f[kCkArrays];
  ck_fields(s, f);
  const size_t n = s->fsize;
  return many(s, MANY_COPY,
              {{dst, f[0]}, {dst + n, f[1]}, {dst + 2 * n, f[2]}, {dst + 3 * n, f[3]}, {dst + 4 * n, f[4]},
               {dst + 5 * n, f[5]}, {dst + 6 * n, f[6]}, {dst + 7 * n, f[7]}, {dst + 8 * n, f[8]},
               {dst + 9 * n, f[9]}, {dst + 10 * n, f[10]}, {dst + 11 * n, f[11]}, {dst + 12 * n, f[12]},
               {dst + 13 * n, f[13]}});
}

// restore a checkpoint as the state at a step boundary (p' zero, no snapshots)
int ck_load(ixh_state *s, const double *src) {
  s->u_ant = pick(s->u_cur, -1);
  s->v_ant = pick(s->v_cur, -1);
  s->t_ant = pick(s->t_cur, -1);
  s->pc_zero = true;
  s->v_snap = s->t_snap = s->pc_snap = -1;
  s->u_hold = -1;
  double *f[kCkArrays];
  ck_fields(s, f);
  const size_t n = s->fsize;
  return many(s, MANY_COPY,
              {{f[0], src}, {f[1], src + n}, {f[2], src + 2 * n}, {f[3], src + 3 * n}, {f[4], src + 4 * n},
               {f[5], src + 5 * n}, {f[6], src + 6 * n}, {f[7], src + 7 * n}, {f[8], src + 8 * n},
               {f[9], src + 9 * n}, {f[10], src + 10 * n}, {f[11], src + 11 * n}, {f[12], src + 12 * n},
               {f[13], src + 13 * n}});
}

int adj_ept(int n) {
  const int e = ept_for(n);
  return e;
}

template <int DIR>
int adj_sweep_launch(ixh_state *s, const AdjArgs &G, int nlines, int n) {
  dim3 grid(nlines), block(kBlock);
  switch (adj_ept(n)) {
    case 1: hipLaunchKernelGGL((k_adj_sweep<DIR, 1>), grid, block, 0, s->stream, G); break;
    case 2: hipLaunchKernelGGL((k_adj_sweep<DIR, 2>), grid, block, 0, s->stream, G); break;
    case 3: hipLaunchKernelGGL((k_adj_sweep<DIR, 3>), grid, block, 0, s->stream, G); break;
    case 4: hipLaunchKernelGGL((k_adj_sweep<DIR, 4>), grid, block, 0, s->stream, G); break;
    case 6: hipLaunchKernelGGL((k_adj_sweep<DIR, 6>), grid, block, 0, s->stream, G); break;
    default: hipLaunchKernelGGL((k_adj_sweep<DIR, 8>), grid, block, 0, s->stream, G); break;
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// adjoint of one energy step with the velocities u, v (row layout):
// lambda (dJ/dT^{k+1}) -> lambda (dJ/dT^k), boundary adjoints accumulated
int adj_step(ixh_state *s, const double *u, const double *v) {
  AdjArgs G;
  memset(&G, 0, sizeof G);
  G.S = base_args(s, IXH_DIR_X);  // row layout throughout
  set_bc(s, G.S, IXH_VAR_T, IXH_DIR_X);
  G.S.u = u;
  G.S.v = v;
  double *lam = s->adj_w[0], *lamp = s->adj_w[1], *mu = s->adj_w[2], *tant = s->adj_w[3];
  for (int k = 0; k < 3; ++k) G.cf[k] = s->adj_w[4 + k];
  const int npts = (s->mi + 1) * (s->nj + 1);
  const dim3 pg((npts + kNodeBlock - 1) / kNodeBlock), block(kNodeBlock);
  // x-sweep (applied last in the step)
  G.lam = lam; G.mu = mu; G.g_lo = s->adj_g[IXH_SIDE_A]; G.g_hi = s->adj_g[IXH_SIDE_C];
  CHK(adj_sweep_launch<DIR_X>(s, G, s->nj - 1, s->mi + 1));
  G.mu_in = mu; G.out = lamp; G.tant = tant;
  hipLaunchKernelGGL(k_adj_gather<DIR_X>, pg, block, 0, s->stream, G);
  HIPCHK(hipGetLastError());
  // y-sweep
  G.lam = lamp; G.mu = mu; G.g_lo = s->adj_g[IXH_SIDE_B]; G.g_hi = s->adj_g[IXH_SIDE_D];
  CHK(adj_sweep_launch<DIR_Y>(s, G, s->mi - 1, s->nj + 1));
  G.mu_in = mu; G.out = lam; G.tant = tant;
  hipLaunchKernelGGL(k_adj_gather<DIR_Y>, pg, block, 0, s->stream, G);
  HIPCHK(hipGetLastError());
  return 0;
}

void adj_free_ck(ixh_state *s) {
  for (double *p : s->adj_ck) (void)hipFree(p);
  s->adj_ck.clear();
}

}  // namespace

extern "C" {

int ixh_adjoint_record(ixh_state *s, int32_t nsteps, int32_t every) {
  if (!s || nsteps < 1 || every < 1) return fail(IXH_ERR_ARG, "adjoint: bad arguments");
  if (s->f32 || s->delta || s->sr) return fail(IXH_ERR_UNSUPP, "adjoint: the FP64 path only");
  if (s->prm.iter_ecuaci_max >= 0 || s->prm.iter_simple_max >= 0)
    return fail(IXH_ERR_UNSUPP, "adjoint: single-pass loops only (iter_ecuaci_max < 0, iter_simple_max < 0)");
  for (int v = 0; v < 3; ++v) CHK(check_bcs(s, v));
  HIPCHK(hipStreamSynchronize(s->stream));
  const int nck = (nsteps + every - 1) / every + 1;  // segment starts and the final state
  if ((int)s->adj_ck.size() != nck) {
    adj_free_ck(s);
    for (int k = 0; k < nck; ++k) {
      double *p = nullptr;
      CHK(dmalloc(&p, (size_t)kCkArrays * s->fsize));
      s->adj_ck.push_back(p);
    }
  }
  CHK(ensure_hist(s, nsteps));
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s->stream, s->d_slot, 0);
  HIPCHK(hipGetLastError());
  if (s->prm.ri_x == 0.0) {
    // fused steps between checkpoints (bit for bit the kernel-by-kernel
    // step); each chunk ends on a whole step
    for (int k0 = 0; k0 < nsteps; k0 += every) {
      CHK(ck_save(s, s->adj_ck[k0 / every]));
      CHK(run_fused(s, std::min(every, nsteps - k0)));
    }
  } else {
    for (int k = 0; k < nsteps; ++k) {
      if (k % every == 0) CHK(ck_save(s, s->adj_ck[k / every]));
      CHK(fast_step(s));
    }
  }
  CHK(ck_save(s, s->adj_ck[nck - 1]));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->adj_steps = nsteps;
  s->adj_every = every;
  return 0;
}

int ixh_adjoint_energy(ixh_state *s, const double *tbar, int32_t ld, double *grad_a, double *grad_b,
                       double *grad_c, double *grad_d, double *grad_t0, int32_t ld0) {
  if (!s || !tbar || ld < s->mi + 1 || (grad_t0 && ld0 < s->mi + 1))
    return fail(IXH_ERR_ARG, "adjoint: bad arguments");
  if (s->adj_steps < 1) return fail(IXH_ERR_ARG, "adjoint: no recorded run (ixh_adjoint_record)");
  const int K = s->adj_steps, C = s->adj_every;
  for (double *&p : s->adj_w)
    if (!p) CHK(dmalloc(&p, s->fsize));
  for (int k = 0; k < 4; ++k) {
    if (!s->adj_g[k]) CHK(dmalloc(&s->adj_g[k], (size_t)std::max(s->mi, s->nj) + 4));
    HIPCHK(hipMemsetAsync(s->adj_g[k], 0, ((size_t)std::max(s->mi, s->nj) + 4) * sizeof(double), s->stream));
  }
  while ((int)s->adj_tape.size() < 2 * C) {
    double *p = nullptr;
    CHK(dmalloc(&p, s->fsize));
    s->adj_tape.push_back(p);
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  CHK(upload_as<double>(s, s->adj_w[0], tbar, ld, s->mi + 1, s->nj + 1, false));
  const int nseg = (K + C - 1) / C;
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int k0 = seg * C, k1 = std::min(K, k0 + C);
    CHK(ck_load(s, s->adj_ck[seg]));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s->stream, s->d_slot, k0);
    HIPCHK(hipGetLastError());
    for (int k = k0; k < k1; ++k) {
      CHK(fast_step(s));
      HIPCHK(hipMemcpyAsync(s->adj_tape[2 * (k - k0)], s->ub[s->u_cur], s->fsize * sizeof(double),
                            hipMemcpyDeviceToDevice, s->stream));
      HIPCHK(hipMemcpyAsync(s->adj_tape[2 * (k - k0) + 1], s->vb[s->v_cur], s->fsize * sizeof(double),
                            hipMemcpyDeviceToDevice, s->stream));
    }
    for (int k = k1 - 1; k >= k0; --k) CHK(adj_step(s, s->adj_tape[2 * (k - k0)], s->adj_tape[2 * (k - k0) + 1]));
  }
  CHK(ck_load(s, s->adj_ck[nseg]));  // back to the end of the recorded run
  // results
  std::vector<double> g((size_t)std::max(s->mi, s->nj) + 4);
  double *outs[4] = {grad_a, grad_b, grad_c, grad_d};
  for (int sd = 0; sd < 4; ++sd) {
    if (!outs[sd]) continue;
    HIPCHK(hipMemcpyAsync(g.data(), s->adj_g[sd], g.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const int n = side_len(s, IXH_VAR_T, sd);
    for (int k = 1; k <= n; ++k) outs[sd][k - 1] = g[k];
  }
  if (grad_t0) {
    std::vector<double> t((size_t)s->P * (s->nj + 1));
    HIPCHK(hipMemcpyAsync(t.data(), s->adj_w[0], t.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int j = 0; j <= s->nj; ++j)
      for (int i = 0; i <= s->mi; ++i) grad_t0[(size_t)i + (size_t)j * ld0] = t[(size_t)i + (size_t)j * s->P];
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

}  // extern "C"

// ====================================================== coupled adjoint
// Adjoint of the whole single-pass step (momentum, pressure correction,
// correction, energy) over a recorded run: each segment is recomputed from
// its checkpoint phase by phase with every intermediate field on a tape, then
// the steps are differentiated backward (adjoint_device.hpp).
namespace {

enum Tape { TP_U0, TP_V0, TP_T0, TP_P0, TP_U1, TP_U2, TP_AU, TP_V1, TP_V2, TP_AV, TP_PC1, TP_PC2, TP_U3, TP_V3,
            TP_T1, TP_T2, TP_N };
enum Work { W_UB3, W_VB3, W_TB2, W_PB, W_TB1, W_TB0, W_UB2, W_VB2, W_PCB2, W_PCB1, W_AUB, W_AVB, W_VB1, W_VB0,
            W_UB1, W_UB0, W_UANTB, W_VANTB, W_TANTB, W_MU, W_FA0, W_FA1, W_FA2, W_FA3, W_CF0, W_CF1, W_CF2,
            W_N };

int dcopy(ixh_state *s, double *dst, const double *src) {
  HIPCHK(hipMemcpyAsync(dst, src, s->fsize * sizeof(double), hipMemcpyDeviceToDevice, s->stream));
  return 0;
}



// one time step in the order of fast_step, every produced field copied to
// the tape (the residual and the norms do not change fields and are
// skipped).  The copies go out in two launches, each before any of its
// buffers can be overwritten: the sweeps write buffers other than the ones
// they read (u0 is held out of the rotation for the u sweeps: after a
// checkpoint is loaded u_ant is another buffer, and the u x-sweep could
// otherwise pick u0's), so u0, v0 (the v snapshot), T0, u1, u2, v1, v2,
// a_u, a_v and p0 (which the correction updates in place) are intact after
// the v x-sweep; p'1, p'2 (which the T sweeps do not touch), u3, v3 and
// T1, T2 after the T x-sweep.
int tape_step(ixh_state *s, double *const *tp) {
  double *u0 = s->ub[s->u_cur], *v0 = s->vb[s->v_cur], *t0 = s->tb[s->t_cur];
  s->u_hold = s->u_cur;
  CHK(sweep_u(s, IXH_DIR_Y));
  double *u1 = s->ub[s->u_cur];
  CHK(sweep_u(s, IXH_DIR_X));
  s->u_hold = -1;
  double *u2 = s->ub[s->u_cur];
  CHK(sweep_v(s, IXH_DIR_Y));
  double *v1 = s->vb[s->v_cur];
  CHK(sweep_v(s, IXH_DIR_X));
  double *v2 = s->vb[s->v_cur];
  CHK(copy_many(s, {{tp[TP_U0], u0}, {tp[TP_V0], v0}, {tp[TP_T0], t0}, {tp[TP_P0], s->p}, {tp[TP_U1], u1},
                    {tp[TP_U2], u2}, {tp[TP_AU], s->au}, {tp[TP_V1], v1}, {tp[TP_V2], v2}, {tp[TP_AV], s->av}}));
  s->pc_zero = true;
  CHK(sweep_p(s, IXH_DIR_Y));
  double *pc1 = s->pcb[s->pc_cur];
  // the p' x-sweep with the correction: u3, v3 into free buffers (u0, u1,
  // v0, v1 are on the tape already), p in place, p'1 and p'2 untouched
  SweepArgs px;
  CHK(prep_p(s, IXH_DIR_X, px));
  CorrLine corr;
  prep_correct_line(s, corr);
  CHK(launch_pcorr(s, px, corr, line_len(s, EQ_P, IXH_DIR_X)));
  double *pc2 = s->pcb[s->pc_cur];
  double *u3 = s->ub[s->u_cur], *v3 = s->vb[s->v_cur];
  CHK(sweep_t(s, IXH_DIR_Y));
  double *t1 = s->tb[s->t_cur];
  CHK(sweep_t(s, IXH_DIR_X));
  // the T sweeps leave the p' buffers alone: p'1 and p'2 go with the rest
  CHK(copy_many(s, {{tp[TP_PC1], pc1}, {tp[TP_PC2], pc2}, {tp[TP_U3], u3}, {tp[TP_V3], v3}, {tp[TP_T1], t1},
                    {tp[TP_T2], s->tb[s->t_cur]}}));
  s->u_ant = s->u_cur;
  s->v_ant = s->v_cur;
  s->t_ant = s->t_cur;
  return 0;
}

template <int EQ, int DIR>
int adj2_sweep(ixh_state *s, Adj2Args &G, int nlines, int n) {
  dim3 block(kBlock), grid(nlines);
  switch (ept_for(n)) {
    case 1: hipLaunchKernelGGL((k_adj2_solve<EQ, DIR, 1>), grid, block, 0, s->stream, G); break;
    case 2: hipLaunchKernelGGL((k_adj2_solve<EQ, DIR, 2>), grid, block, 0, s->stream, G); break;
    case 3: hipLaunchKernelGGL((k_adj2_solve<EQ, DIR, 3>), grid, block, 0, s->stream, G); break;
    case 4: hipLaunchKernelGGL((k_adj2_solve<EQ, DIR, 4>), grid, block, 0, s->stream, G); break;
    case 6: hipLaunchKernelGGL((k_adj2_solve<EQ, DIR, 6>), grid, block, 0, s->stream, G); break;
    default: hipLaunchKernelGGL((k_adj2_solve<EQ, DIR, 8>), grid, block, 0, s->stream, G); break;
  }
  HIPCHK(hipGetLastError());
  const int npts = (s->mi + 1) * (s->nj + 1);
  const dim3 pg((npts + kNodeBlock - 1) / kNodeBlock);
  hipLaunchKernelGGL((k_adj2_gather<EQ, DIR>), pg, dim3(kNodeBlock), 0, s->stream, G);
  HIPCHK(hipGetLastError());
  return 0;
}

// arguments of the adjoint of a sweep of variable `var` (row layout)
Adj2Args adj2_args(ixh_state *s, int var, double *const *W) {
  Adj2Args G;
  memset(&G, 0, sizeof G);
  G.S = base_args(s, IXH_DIR_X);
  if (var >= 0) set_bc(s, G.S, var, IXH_DIR_X);
  G.mu = W[W_MU];
  for (int k = 0; k < 4; ++k) G.fa[k] = W[W_FA0 + k];
  for (int k = 0; k < 3; ++k) G.cf[k] = W[W_CF0 + k];
  return G;
}

// adjoint of one step: (ubar, vbar, Tbar, pbar) of its end, held in
// W_UB3, W_VB3, W_TB2, W_PB, become those of its start in W_UB0, W_VB0,
// W_TB0, W_PB; boundary adjoints of T accumulate in adj_g.  The
// accumulators W_AUB, W_AVB, W_UANTB, W_VANTB, W_TANTB are zero on entry
// and are left zero for the next step.
int adj2_step(ixh_state *s, double *const *tp, double *const *W) {
  const int mi = s->mi, nj = s->nj;
  // T x-sweep, then T y-sweep
  {
    Adj2Args G = adj2_args(s, IXH_VAR_T, W);
    G.S.u = tp[TP_U3]; G.S.v = tp[TP_V3];
    G.x = tp[TP_T2]; G.phi = tp[TP_T1]; G.ybar = W[W_TB2];
    G.phib = W[W_TB1]; G.velub = W[W_UB3]; G.velvb = W[W_VB3]; G.antb = W[W_TANTB];
    G.g_lo = s->adj_g[IXH_SIDE_A]; G.g_hi = s->adj_g[IXH_SIDE_C];
    CHK((adj2_sweep<EQ_T, DIR_X>(s, G, nj - 1, mi + 1)));
    G.x = tp[TP_T1]; G.phi = tp[TP_T0]; G.ybar = W[W_TB1]; G.phib = W[W_TB0];
    G.g_lo = s->adj_g[IXH_SIDE_B]; G.g_hi = s->adj_g[IXH_SIDE_D];
    CHK((adj2_sweep<EQ_T, DIR_Y>(s, G, mi - 1, nj + 1)));
  }
  // correction
  {
    AdjCorrArgs C;
    memset(&C, 0, sizeof C);
    C.mi = mi; C.nj = nj; C.P = s->P; C.g = s->geo;
    C.pc = tp[TP_PC2]; C.au = tp[TP_AU]; C.av = tp[TP_AV];
    C.ub3 = W[W_UB3]; C.vb3 = W[W_VB3]; C.pbn = W[W_PB];
    C.ub2 = W[W_UB2]; C.vb2 = W[W_VB2]; C.pcb = W[W_PCB2]; C.aub = W[W_AUB]; C.avb = W[W_AVB];
    const int npts = (mi + 1) * (nj + 1);
    hipLaunchKernelGGL(k_adj2_correct, dim3((npts + kNodeBlock - 1) / kNodeBlock), dim3(kNodeBlock), 0, s->stream, C);
    HIPCHK(hipGetLastError());
  }
  // p' x-sweep, then p' y-sweep (input zero)
  {
    Adj2Args G = adj2_args(s, -1, W);
    G.S.u = tp[TP_U2]; G.S.v = tp[TP_V2]; G.S.au = tp[TP_AU]; G.S.av = tp[TP_AV];
    G.x = tp[TP_PC2]; G.phi = tp[TP_PC1]; G.ybar = W[W_PCB2]; G.phib = W[W_PCB1];
    G.velub = W[W_UB2]; G.velvb = W[W_VB2]; G.aub = W[W_AUB]; G.avb = W[W_AVB];
    CHK((adj2_sweep<EQ_P, DIR_X>(s, G, nj - 1, mi + 1)));
    G.x = tp[TP_PC1]; G.phi = nullptr; G.ybar = W[W_PCB1]; G.phib = nullptr;
    CHK((adj2_sweep<EQ_P, DIR_Y>(s, G, mi - 1, nj + 1)));
  }
  // v x-sweep, then v y-sweep
  {
    Adj2Args G = adj2_args(s, IXH_VAR_V, W);
    G.S.u = tp[TP_U2]; G.S.v = tp[TP_V1]; G.S.t = tp[TP_T0];
    G.x = tp[TP_V2]; G.phi = tp[TP_V1]; G.ybar = W[W_VB2]; G.acb = W[W_AVB];
    G.phib = W[W_VB1]; G.velub = W[W_UB2]; G.antb = W[W_VANTB]; G.pb = W[W_PB]; G.tb = W[W_TB0];
    CHK((adj2_sweep<EQ_V, DIR_X>(s, G, nj - 2, mi + 1)));
    G.S.v = tp[TP_V0];
    G.x = tp[TP_V1]; G.phi = tp[TP_V0]; G.ybar = W[W_VB1]; G.acb = nullptr; G.phib = W[W_VB0];
    CHK((adj2_sweep<EQ_V, DIR_Y>(s, G, mi - 1, nj)));
  }
  // u x-sweep, then u y-sweep
  {
    Adj2Args G = adj2_args(s, IXH_VAR_U, W);
    G.S.u = tp[TP_U1]; G.S.v = tp[TP_V0];
    G.x = tp[TP_U2]; G.phi = tp[TP_U1]; G.ybar = W[W_UB2]; G.acb = W[W_AUB];
    G.phib = W[W_UB1]; G.velvb = W[W_VB0]; G.antb = W[W_UANTB]; G.pb = W[W_PB];
    CHK((adj2_sweep<EQ_U, DIR_X>(s, G, nj - 1, mi)));
    G.S.u = tp[TP_U0];
    G.x = tp[TP_U1]; G.phi = tp[TP_U0]; G.ybar = W[W_UB1]; G.acb = nullptr; G.phib = W[W_UB0];
    CHK((adj2_sweep<EQ_U, DIR_Y>(s, G, mi - 2, nj + 1)));
  }
  // the lagged copies are the fields of the step's start; the accumulators
  // cleared for the next step in the same launch
  CHK(many(s, MANY_ADD_CLEAR, {{W[W_UB0], W[W_UANTB]}, {W[W_VB0], W[W_VANTB]}, {W[W_TB0], W[W_TANTB]},
                               {W[W_AUB], nullptr}, {W[W_AVB], nullptr}}));
  return 0;
}

}  // namespace

extern "C" {

int ixh_adjoint_coupled(ixh_state *s, const double *tbar, int32_t ld, double *grad_a, double *grad_b,
                        double *grad_c, double *grad_d, double *grad_t0, int32_t ld0) {
  if (!s || !tbar || ld < s->mi + 1 || (grad_t0 && ld0 < s->mi + 1))
    return fail(IXH_ERR_ARG, "adjoint: bad arguments");
  if (s->adj_steps < 1) return fail(IXH_ERR_ARG, "adjoint: no recorded run (ixh_adjoint_record)");
  if (s->prm.ri_x != 0.0) return fail(IXH_ERR_UNSUPP, "coupled adjoint: the u equation must not read T (ri_x = 0)");
  const int K = s->adj_steps, C = s->adj_every;
  const size_t glen = (size_t)std::max(s->mi, s->nj) + 4;
  for (int k = 0; k < 4; ++k) {
    if (!s->adj_g[k]) CHK(dmalloc(&s->adj_g[k], glen));
    HIPCHK(hipMemsetAsync(s->adj_g[k], 0, glen * sizeof(double), s->stream));
  }
  while ((int)s->adj2_tape.size() < TP_N * C) {
    double *p = nullptr;
    CHK(dmalloc(&p, s->fsize));
    s->adj2_tape.push_back(p);
  }
  while ((int)s->adj2_w.size() < W_N) {
    double *p = nullptr;
    CHK(dmalloc(&p, s->fsize));
    s->adj2_w.push_back(p);
  }
  std::vector<double *> W(s->adj2_w.begin(), s->adj2_w.begin() + W_N);
  const size_t fb = s->fsize * sizeof(double);
  for (int k : {W_UB3, W_VB3, W_PB, W_AUB, W_AVB, W_UANTB, W_VANTB, W_TANTB})
    HIPCHK(hipMemsetAsync(W[k], 0, fb, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  CHK(upload_as<double>(s, W[W_TB2], tbar, ld, s->mi + 1, s->nj + 1, false));
  const int nseg = (K + C - 1) / C;
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int k0 = seg * C, k1 = std::min(K, k0 + C);
    CHK(ck_load(s, s->adj_ck[seg]));
    for (int k = k0; k < k1; ++k) CHK(tape_step(s, &s->adj2_tape[(size_t)TP_N * (k - k0)]));
    for (int k = k1 - 1; k >= k0; --k) {
      CHK(adj2_step(s, &s->adj2_tape[(size_t)TP_N * (k - k0)], W.data()));
      // the start of step k is the end of step k-1
      std::swap(W[W_UB3], W[W_UB0]);
      std::swap(W[W_VB3], W[W_VB0]);
      std::swap(W[W_TB2], W[W_TB0]);
    }
  }
  CHK(ck_load(s, s->adj_ck[nseg]));  // back to the end of the recorded run
  std::vector<double> g(glen);
  double *outs[4] = {grad_a, grad_b, grad_c, grad_d};
  for (int sd = 0; sd < 4; ++sd) {
    if (!outs[sd]) continue;
    HIPCHK(hipMemcpyAsync(g.data(), s->adj_g[sd], glen * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    const int n = side_len(s, IXH_VAR_T, sd);
    for (int k = 1; k <= n; ++k) outs[sd][k - 1] = g[k];
  }
  if (grad_t0) {
    std::vector<double> t((size_t)s->P * (s->nj + 1));
    HIPCHK(hipMemcpyAsync(t.data(), W[W_TB2], t.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int j = 0; j <= s->nj; ++j)
      for (int i = 0; i <= s->mi; ++i) grad_t0[(size_t)i + (size_t)j * ld0] = t[(size_t)i + (size_t)j * s->P];
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

}  // extern "C"

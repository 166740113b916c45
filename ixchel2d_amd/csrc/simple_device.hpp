// Device code of the SIMPLE iteration for gfx950 (CDNA4), FP64 path.
//
// One fused kernel per directional sweep: each workgroup owns one line (an
// x-line = fixed jj, or a y-line = fixed ii), assembles the tridiagonal rows
// of that line in registers (the ensambla_* arithmetic of
// source/mod_ec_momento.f90 / mod_ec_energia.f90), writes the boundary rows
// (impone_cond_frontera, source/mod_cond_frontera.f90:254-420), solves the
// line in LDS and writes the new field into a separate buffer (every line of
// a sweep reads the field as it was before the sweep, as upstream's
// assemble-all-then-solve-all order does).  Per-line partial sums of the
// convergence norms are written by the same kernel and folded by one small
// kernel in a fixed order, so results do not depend on scheduling.
//
// Arithmetic follows upstream term by term (same association; the library is
// built with -ffp-contract=off), so with the Thomas line solver the fields
// equal the upstream CPU build's bit for bit; the PCR solver differs only by
// the rounding of the line solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ixh {

constexpr int kBlock = 256;   // threads per line workgroup (4 wave64)
constexpr int kNmax = 2048;   // longest line the LDS solvers hold

enum Eq { EQ_U = 0, EQ_V = 1, EQ_P = 2, EQ_T = 3 };
enum Dir { DIR_Y = 0, DIR_X = 1 };
// PCR: LDS cyclic reduction with the division-free assembly below (default);
// THOMAS: upstream's Thomas order with upstream-order assembly (bit-faithful);
// PCR_EXACT: PCR with upstream-order assembly (isolates the solver rounding);
// DELTA: the FP32-refined sweeps -- FP64 fields, the line's FP64 residual
// against FP32 coefficients, and the FP32 PCR solve of the increment;
// DELTA_SR: the same increment sweeps over FP32 fields, each new value
// phi + delta (formed in FP64) stored with stochastic rounding
enum Solver { SOLVE_PCR = 0, SOLVE_THOMAS = 1, SOLVE_PCR_EXACT = 2, SOLVE_DELTA = 3, SOLVE_DELTA_SR = 4 };

// the increment (residual-correction) sweeps
constexpr bool delta_solver(int s) { return s == SOLVE_DELTA || s == SOLVE_DELTA_SR; }
// solvers run by the wave-partitioned PCR, which also fits one-wave lines
constexpr bool wave_solver(int s) { return s == SOLVE_PCR || delta_solver(s); }

// storage type of the solved fields for a sweep of real type R
template <int SOLVER, typename R>
struct StoreOf {
  using T = R;
};
template <typename R>
struct StoreOf<SOLVE_DELTA, R> {
  using T = double;
};

// ------------------------------------------------------ stochastic rounding
// Murmur3's 32-bit finaliser (Appleby, MurmurHash3, public domain): every
// input bit reaches every output bit.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}

// d rounded to FP32 stochastically: to the FP32 neighbour above |d| with
// probability (|d| - lo) / (hi - lo), else to the one below, so the rounding
// is unbiased (Connolly, Higham & Mary, SIAM J. Sci. Comput. 43 (2021)
// A566-A585).  FP32 keeps 23 of FP64's 52 fraction bits: adding a uniform
// 29-bit integer to the magnitude's bit pattern carries into the kept bits
// with exactly that probability, and truncation then picks the neighbour.
// The 29 bits come from a hash of d's own bits and of key (node and stream),
// so a run is deterministic, a cavity gives the same bits alone and in a
// batch, and a value already representable in FP32 is returned unchanged.
// Round-to-nearest would drop every increment below half an FP32 spacing:
// near a steady state the per-step changes are that small, and the
// iteration would stop short of it.
__device__ __forceinline__ float sr_f32(double d, uint32_t key) {
  const uint64_t b = (uint64_t)__double_as_longlong(d);
  const uint32_t h = fmix32(key ^ fmix32((uint32_t)b ^ fmix32((uint32_t)(b >> 32))));
  return __double2float_rz(__longlong_as_double((long long)(b + (uint64_t)(h >> 3))));
}

// key of node (ii, jj) in rounding stream s (one stream per field and writer)
__device__ __forceinline__ uint32_t sr_key(int ii, int jj, uint32_t s) {
  return (uint32_t)ii * 0x9e3779b1u ^ ((uint32_t)jj + (s << 20)) * 0x7f4a7c15u;
}

// a value formed in FP64 stored as S: stochastic rounding where sr is set
// and S is FP32, a plain conversion otherwise
template <typename S>
__device__ __forceinline__ S store_as(double d, bool sr, uint32_t key) {
  if constexpr (sizeof(S) == 4)
    return sr ? sr_f32(d, key) : (float)d;
  else
    return d;
}

// Global-memory pointers.  A pointer the kernel reads from memory (the
// argument tables of the batch kernels) is generic to the compiler, and a
// generic (flat) access counts against both the vector-memory and the
// LDS/scalar counters, so each use waits for every access in flight.  gp()
// types a pointer as global at the access; GPtr does it for the geometry
// tables, whose entries are indexed directly.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gp(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}
template <typename T>
struct GPtr {
  T *p;
  __host__ __device__ GPtr &operator=(T *q) {
    p = q;
    return *this;
  }
  __host__ __device__ operator T *() const { return p; }
  __device__ __forceinline__ __attribute__((address_space(1))) T &operator[](size_t i) const {
    return gp(p)[i];
  }
  __device__ __forceinline__ __attribute__((address_space(1))) T &operator[](int i) const {
    return gp(p)[i];
  }
};

// 1-based geometry arrays (index 0 unused), built as in
// source/mod_malla.f90:167-217; the FP32-mixed sweeps read FP32 copies
template <typename R>
struct GeoT {
  GPtr<const R> dxp, dyp, dxu, dyu, dxv, dyv;
  GPtr<const R> fexp, feyp, fexu, feyv;
  GPtr<const R> xu, yv, xp, yp;
};
using Geo = GeoT<double>;
using GeoF = GeoT<float>;

struct Phys {
  double dt, rel_pres, rel_vel, rel_ener, gm, gt, ri_x, ri_y;
};

// boundary nodes of one side: type (0 none, 1 Dirichlet, 2 Neumann) and value,
// indexed by the 1-based node number along the side
struct Side {
  const int8_t *type;
  const double *val;
};

// Geometry-only factors of the division-free assembly, one set per equation
// (0 = u, 1 = v, 2 = T): for each face the conductance G/d and the Peclet
// factor d/G (constant G).  Upstream computes G*L/d and w*d/G with divisions
// in every cell; these tables leave only multiplications (results move by a
// few ulp).  A node's two faces along one direction form one record (low:
// west or south, high: east or north), which a row reads in one access:
// fx indexed by ii, fy by jj.
template <typename R>
struct alignas(4 * sizeof(R)) FaceT {
  R c_lo, p_lo, c_hi, p_hi;
};
template <typename R>
struct FastGeoT {
  GPtr<const FaceT<R>> fx[3], fy[3];
  R inv_dt, inv_rel_vel, inv_rel_ener;
};
// record i of a face table, read as one vector (two 16-byte loads in FP64,
// one in FP32)
template <typename R>
__device__ __forceinline__ FaceT<R> face(const GPtr<const FaceT<R>> &t, int i) {
  typedef R R4 __attribute__((ext_vector_type(4)));
  const R4 v = *reinterpret_cast<const __attribute__((address_space(1))) R4 *>(&t[i]);
  return FaceT<R>{v.x, v.y, v.z, v.w};
}
using FastGeo = FastGeoT<double>;
using FastGeoF = FastGeoT<float>;

// Kernel arguments of one sweep.  The solved fields (u, v, T, p' and their
// lagged and snapshot copies) and a_u, a_v are held in the sweep's real type
// R: the pointers below are typed double and the kernels reinterpret them as
// R (float in the FP32-mixed path).  p, b_o, the sources and the partial sums
// are always FP64.
struct SweepArgs {
  int mi, nj, P;          // mesh extents and the common row pitch of all fields
  int line0, nlines;      // first line (1-based) and number of lines
  Geo g;
  Phys ph;
  FastGeo fg;
  GeoF gf;                // FP32 copies of g and fg (FP32-mixed path)
  FastGeoF ff;
  const double *u, *v, *t, *p, *uant, *vant, *tant, *pc, *au, *av;
  const double *sc, *sp;  // sources of this equation (nullptr = zero)
  Side bc[4];             // a, b, c, d of this equation
  const double *snap;     // field at the start of the pass (nullptr = zero)
  double *out;            // solved field (layout of the sweep)
  double *acoef;          // a_u / a_v written by the momentum sweeps
  int P2;                 // pitch of the other layout
  double *out2, *acoef2;  // the same results in the other layout
  double *bo;             // mass residual written by the p' x-sweep
  double *part0, *part1;  // per-line partial sums (indexed by line number)
  int in_zero;            // p' input identically zero (first pass)
};

// Outputs of the velocity and pressure correction (main.f90:931-962) when
// the p' x-sweep does it for its own line: corrected u and v into buffers of
// their own (row layout and transposed copy), p in place.  The sweep reads
// u, v, a_u, a_v and p through its SweepArgs.
struct CorrLine {
  double *u, *v, *uT, *vT, *p, *pT;
};

// a field pointer seen as R
template <typename R>
__device__ __forceinline__ const R *fld(const double *p) {
  return reinterpret_cast<const R *>(p);
}
template <typename R>
__device__ __forceinline__ R *fldw(double *p) {
  return reinterpret_cast<R *>(p);
}
template <typename R>
__device__ __forceinline__ const GeoT<R> &geo_of(const SweepArgs &A);
template <>
__device__ __forceinline__ const GeoT<double> &geo_of<double>(const SweepArgs &A) { return A.g; }
template <>
__device__ __forceinline__ const GeoT<float> &geo_of<float>(const SweepArgs &A) { return A.gf; }
template <typename R>
__device__ __forceinline__ const FastGeoT<R> &fast_of(const SweepArgs &A);
template <>
__device__ __forceinline__ const FastGeoT<double> &fast_of<double>(const SweepArgs &A) { return A.fg; }
template <>
__device__ __forceinline__ const FastGeoT<float> &fast_of<float>(const SweepArgs &A) { return A.ff; }

// quotient of the PCR line solvers: in FP64 the hardware reciprocal (about
// 2^-24 relative error on gfx950) refined by one Newton step (about 2^-48),
// times the numerator with one residual correction.  Over 4M random
// quotients spanning 60 decades this equalled IEEE division bit for bit, in
// a much shorter dependent chain.  In FP32: the hardware reciprocal (1 ulp)
// times the numerator.  The Thomas path keeps IEEE division.
__device__ __forceinline__ double qdiv(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// reciprocal of the wave solver's pivots: in FP64 the hardware reciprocal
// and two Newton steps (error about 2^-96 before rounding, so within an ulp
// of 1/b); in FP32 the hardware reciprocal (1 ulp)
__device__ __forceinline__ double rcpn(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  return fma(fma(-b, r, 1.0), r, r);
}
__device__ __forceinline__ float rcpn(float b) { return __builtin_amdgcn_rcpf(b); }
// quotient of the p' assembly, shared by every solver: IEEE division in FP64
// (the Thomas path matches upstream bit for bit), the reciprocal in FP32
__device__ __forceinline__ double adiv(double a, double b) { return a / b; }
__device__ __forceinline__ float adiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double rabs(double x) { return fabs(x); }
__device__ __forceinline__ float rabs(float x) { return fabsf(x); }

// an FP32 momentum or energy row whose source is an immersed-body penalty
// (|Sp| above this) is replaced by phi = alpha Sc / -Sp + (1 - alpha) phi_old,
// the limit of upstream's under-relaxed penalty row: the penalty itself
// (1e51) does not fit FP32
constexpr double kPenaltyRow = 1.0e30;

// Field element (i,j), 1-based.  Every field is kept twice: row layout
// (x contiguous, pitch P >= mi+1) for the x-sweeps and transposed layout
// (y contiguous, pitch >= nj+1) for the y-sweeps, so that the lanes of a wave
// always read consecutive addresses.  TR (compile time, in scope) selects the
// transposed one; P is the pitch of the layout in use.
#define IDX(i, j)                                                                 \
  (TR ? (((size_t)(j) - 1) + ((size_t)(i) - 1) * (size_t)P)                      \
      : (((size_t)(i) - 1) + ((size_t)(j) - 1) * (size_t)P))
#define AT(ptr, i, j) (gp(ptr))[IDX(i, j)]
// the other layout (pitch P2), for the second copy a sweep writes
#define AO(ptr, i, j)                                                             \
  (gp(ptr))[TR ? (((size_t)(i) - 1) + ((size_t)(j) - 1) * (size_t)P2)            \
               : (((size_t)(j) - 1) + ((size_t)(i) - 1) * (size_t)P2)]

// Optional operands without branches.  An absent source field (null) or an
// unused buoyancy term reads element 0 of a zero block, or of the field,
// instead of element (i,j) (`on` is uniform).  The load then goes out with
// all the others instead of after a branch, the term it feeds is zero, and
// x + 0 = x leaves the arithmetic of the present terms unchanged.
static __constant__ double kZeroBlock[8];
__device__ __forceinline__ const double *or_zero(const double *p) { return p ? p : kZeroBlock; }
#define ATZ(ptr, on, i, j) (gp(ptr))[(on) ? IDX(i, j) : (size_t)0]

__device__ __forceinline__ double pow5(double t) { return t * (t * (t * (t * t))); }
__device__ __forceinline__ double dmax0(double x) { return x < 0.0 ? 0.0 : x; }

// power-law neighbour coefficient with upstream's sign and association:
// -(G*L/d * max(0,(1-0.1|w d/G|)^5) + max(0, +-w L))
__device__ __forceinline__ double nbc(double g, double len, double d, double w, bool east) {
  double conv = w * len;
  if (east) conv = -conv;
  return -(g * len / d * dmax0(pow5(1.0 - 0.1 * fabs(w * d / g))) + dmax0(conv));
}

__device__ __forceinline__ double harm(double g1, double g2, double f) {
  return (g1 * g2) / (g1 * (1.0 - f) + g2 * f);
}

struct Coef {
  double ai, ad, bs, bn, ac;
};

// ---- u momentum, ensambla_velu_x/_y (source/mod_ec_momento.f90:365-724)
template <bool TR>
__device__ __forceinline__ Coef u_coef(const SweepArgs &A, int ii, int jj) {
  const int P = A.P;
  const Geo &g = A.g;
  const double gm = A.ph.gm;
  Coef c;
  double ud = g.fexu[ii] * AT(A.u, ii + 1, jj) + (1.0 - g.fexu[ii]) * AT(A.u, ii, jj);
  double ui = g.fexu[ii - 1] * AT(A.u, ii, jj) + (1.0 - g.fexu[ii - 1]) * AT(A.u, ii - 1, jj);
  double vn = g.fexp[ii] * AT(A.v, ii + 1, jj) + (1.0 - g.fexp[ii]) * AT(A.v, ii, jj);
  double vs = g.fexp[ii] * AT(A.v, ii + 1, jj - 1) + (1.0 - g.fexp[ii]) * AT(A.v, ii, jj - 1);
  double gd = harm(gm, gm, g.feyp[jj]);
  double gi = harm(gm, gm, g.feyp[jj]);
  double gn = gi * gd / (gd * (1.0 - g.fexp[ii]) + gi * g.fexp[ii]);
  gd = harm(gm, gm, g.feyp[jj - 1]);
  gi = harm(gm, gm, g.feyp[jj - 1]);
  double gs = gi * gd / (gd * (1.0 - g.fexp[ii]) + gi * g.fexp[ii]);
  double di = g.dxp[ii], dd = g.dxp[ii + 1], ds = g.dyv[jj - 1], dn = g.dyv[jj];
  double dx = g.dxu[ii], dy = g.dyu[jj];
  c.ai = nbc(gm, dy, di, ui, false);
  c.ad = nbc(gm, dy, dd, ud, true);
  c.bs = nbc(gs, dx, ds, vs, false);
  c.bn = nbc(gn, dx, dn, vn, true);
  double acc = -c.ai - c.ad - c.bs - c.bn;
  acc = acc - dx * dy * ATZ(or_zero(A.sp), A.sp != nullptr, ii, jj);
  c.ac = (acc + dx * dy / A.ph.dt) / A.ph.rel_vel;
  return c;
}

template <bool TR>
__device__ __forceinline__ double u_rhs(const SweepArgs &A, int ii, int jj, double head, double ac) {
  const int P = A.P;
  const Geo &g = A.g;
  double dx = g.dxu[ii], dy = g.dyu[jj];
  double r = head;
  const bool buoy = A.ph.ri_x != 0.0;
  double tint = g.fexp[ii] * ATZ(A.t, buoy, ii + 1, jj) + (1.0 - g.fexp[ii]) * ATZ(A.t, buoy, ii, jj);
  r = r - dx * dy * A.ph.ri_x * tint;
  r = r + dx * dy * ATZ(or_zero(A.sc), A.sc != nullptr, ii, jj);
  r = r + dx * dy * AT(A.uant, ii, jj) / A.ph.dt;
  r = r + (AT(A.p, ii, jj) - AT(A.p, ii + 1, jj)) * dy;
  r = r + ac * (1.0 - A.ph.rel_vel) * AT(A.u, ii, jj);
  return r;
}

// ---- v momentum, ensambla_velv_x/_y (source/mod_ec_momento.f90:736-1098)
template <bool TR>
__device__ __forceinline__ Coef v_coef(const SweepArgs &A, int ii, int jj) {
  const int P = A.P;
  const Geo &g = A.g;
  const double gm = A.ph.gm;
  Coef c;
  double ud = g.feyp[jj] * AT(A.u, ii, jj + 1) + (1.0 - g.feyp[jj]) * AT(A.u, ii, jj);
  double ui = g.feyp[jj] * AT(A.u, ii - 1, jj + 1) + (1.0 - g.feyp[jj]) * AT(A.u, ii - 1, jj);
  double vn = g.feyv[jj] * AT(A.v, ii, jj + 1) + (1.0 - g.feyv[jj]) * AT(A.v, ii, jj);
  double vs = g.feyv[jj - 1] * AT(A.v, ii, jj) + (1.0 - g.feyv[jj - 1]) * AT(A.v, ii, jj - 1);
  double gs = harm(gm, gm, g.fexp[ii]);
  double gn = harm(gm, gm, g.fexp[ii]);
  double gd = (gn * gs) / (gn * (1.0 - g.feyp[jj]) + gs * g.feyp[jj]);
  gs = harm(gm, gm, g.fexp[ii - 1]);
  gn = harm(gm, gm, g.fexp[ii - 1]);
  double gi = (gs * gn) / (gn * (1.0 - g.feyp[jj]) + gs * g.feyp[jj]);
  double di = g.dxu[ii - 1], dd = g.dxu[ii], ds = g.dyp[jj], dn = g.dyp[jj + 1];
  double dxv = g.dxv[ii], dyv = g.dyv[jj];
  c.ai = nbc(gi, dyv, di, ui, false);
  c.ad = nbc(gd, dyv, dd, ud, true);
  c.bs = nbc(gm, dxv, ds, vs, false);
  c.bn = nbc(gm, dxv, dn, vn, true);
  double acc = -c.ai - c.ad - c.bs - c.bn;
  acc = acc - dxv * dyv * ATZ(or_zero(A.sp), A.sp != nullptr, ii, jj);
  c.ac = (acc + dxv * dyv / A.ph.dt) / A.ph.rel_vel;
  return c;
}

template <bool TR>
__device__ __forceinline__ double v_rhs(const SweepArgs &A, int ii, int jj, double head, double ac) {
  const int P = A.P;
  const Geo &g = A.g;
  double dxv = g.dxv[ii], dyv = g.dyv[jj];
  double r = head;
  const bool buoy = A.ph.ri_y != 0.0;
  double tint = g.feyp[jj] * ATZ(A.t, buoy, ii, jj + 1) + (1.0 - g.feyp[jj]) * ATZ(A.t, buoy, ii, jj);
  r = r - dxv * dyv * A.ph.ri_y * tint;
  r = r + dxv * dyv * ATZ(or_zero(A.sc), A.sc != nullptr, ii, jj);
  r = r + dxv * dyv * AT(A.vant, ii, jj) / A.ph.dt;
  r = r + (AT(A.p, ii, jj) - AT(A.p, ii, jj + 1)) * dxv;
  r = r + ac * (1.0 - A.ph.rel_vel) * AT(A.v, ii, jj);
  return r;
}

// ---- pressure correction, ensambla_corr_pres_x/_y (mod_ec_momento.f90:145-355)
// Coefficients in the sweep's real type R; the mass residual b_o always in
// FP64 from the stored velocities.
template <typename R>
struct PCoefT {
  R ai, ad, bs, bn, ac;
  double b;
};
// S: stored type of u, v, a_u, a_v (double in the FP32-refined sweeps)
template <bool TR, typename R, typename S = R>
__device__ __forceinline__ PCoefT<R> p_coef(const SweepArgs &A, int ii, int jj) {
  const int P = A.P;
  const GeoT<R> &g = geo_of<R>(A);
  const S *u = fld<S>(A.u), *v = fld<S>(A.v), *au = fld<S>(A.au), *av = fld<S>(A.av);
  PCoefT<R> c;
  const S ud = AT(u, ii, jj), ui = AT(u, ii - 1, jj);
  const S vn = AT(v, ii, jj), vs = AT(v, ii, jj - 1);
  const R ai = R(AT(au, ii - 1, jj)), ad = R(AT(au, ii, jj));
  const R as = R(AT(av, ii, jj - 1)), an = R(AT(av, ii, jj));
  R dyp = g.dyp[jj], dxp = g.dxp[ii];
  c.ai = adiv(-dyp * dyp, ai);
  c.ad = adiv(-dyp * dyp, ad);
  c.bs = adiv(-dxp * dxp, as);
  c.bn = adiv(-dxp * dxp, an);
  c.ac = adiv(-c.ai - c.ad - c.bs - c.bn, R(A.ph.rel_pres));
  c.b = ((double)ui - (double)ud) * A.g.dyp[jj] + ((double)vs - (double)vn) * A.g.dxp[ii];
  return c;
}

// ---- energy, ensambla_energia_x/_y (mod_ec_energia.f90:119-443); the
// driver hands fexu/feyv to the routine's interpolation-weight dummies
// (main.f90:986,1057), so those weight the interface diffusivities
template <bool TR>
__device__ __forceinline__ Coef t_coef(const SweepArgs &A, int ii, int jj) {
  const int P = A.P;
  const Geo &g = A.g;
  const double gt = A.ph.gt;
  Coef c;
  double ud = AT(A.u, ii, jj), ui = AT(A.u, ii - 1, jj);
  double vn = AT(A.v, ii, jj), vs = AT(A.v, ii, jj - 1);
  double di = g.dxu[ii - 1], dd = g.dxu[ii], ds = g.dyv[jj - 1], dn = g.dyv[jj];
  double gd = harm(gt, gt, g.fexu[ii]);
  double gi = harm(gt, gt, g.fexu[ii - 1]);
  double gn = harm(gt, gt, g.feyv[jj]);
  double gs = harm(gt, gt, g.feyv[jj - 1]);
  double dxp = g.dxp[ii], dyp = g.dyp[jj];
  c.ai = nbc(gi, dyp, di, ui, false);
  c.ad = nbc(gd, dyp, dd, ud, true);
  c.bs = nbc(gs, dxp, ds, vs, false);
  c.bn = nbc(gn, dxp, dn, vn, true);
  double acc = -c.ai - c.ad - c.bs - c.bn;
  acc = acc - dxp * dyp * ATZ(or_zero(A.sp), A.sp != nullptr, ii, jj);
  c.ac = (acc + dxp * dyp / A.ph.dt) / A.ph.rel_ener;
  return c;
}

template <bool TR>
__device__ __forceinline__ double t_rhs(const SweepArgs &A, int ii, int jj, double head, double ac) {
  const int P = A.P;
  const Geo &g = A.g;
  double dxp = g.dxp[ii], dyp = g.dyp[jj];
  double r = head;
  r = r + dxp * dyp * ATZ(or_zero(A.sc), A.sc != nullptr, ii, jj);
  r = r + dxp * dyp * AT(A.tant, ii, jj) / A.ph.dt;
  r = r + ac * (1.0 - A.ph.rel_ener) * AT(A.t, ii, jj);
  return r;
}

// ---- division-free assembly (PCR path): same discretisation, geometry
// quotients from FastGeo, products instead of quotients by dt and the
// relaxation factor; the interface diffusivities of a constant G are G.
// R is the arithmetic of the assembly (double, or float in the FP32-mixed
// path).
template <typename R>
__device__ __forceinline__ R nbf(R cond, R len, R pec, R w, bool east) {
  R conv = w * len;
  if (east) conv = -conv;
  const R t = R(1.0) - R(0.1) * rabs(w * pec);
  const R t2 = t * t;
  const R pw = t2 * t2 * t;
  return -(cond * len * (pw < R(0) ? R(0) : pw) + (conv < R(0) ? R(0) : conv));
}

template <typename R>
struct CoefT {
  R ai, ad, bs, bn, ac;
  R ap, td;  // a_P before under-relaxation without the time term, and area / dt
};

// `pin` (FP32 only) receives the value of a penalised row, NaN otherwise.
// S is the stored type of the velocity fields (double in the FP32-refined
// sweeps, whose coefficients are still formed in R).
template <int EQ, bool TR, typename R, typename S = R>
__device__ __forceinline__ CoefT<R> fast_coef(const SweepArgs &A, int ii, int jj, R &area, double &pin) {
  const int P = A.P;
  const GeoT<R> &g = geo_of<R>(A);
  const FastGeoT<R> &F = fast_of<R>(A);
  const S *u = fld<S>(A.u), *v = fld<S>(A.v);
  auto ld = [&](const S *f, int i, int j) { return R(AT(f, i, j)); };
  constexpr int e = EQ == EQ_U ? 0 : EQ == EQ_V ? 1 : 2;
  R ui, ud, vs, vn, lx, ly;
  if (EQ == EQ_U) {
    ud = g.fexu[ii] * ld(u, ii + 1, jj) + (R(1.0) - g.fexu[ii]) * ld(u, ii, jj);
    ui = g.fexu[ii - 1] * ld(u, ii, jj) + (R(1.0) - g.fexu[ii - 1]) * ld(u, ii - 1, jj);
    vn = g.fexp[ii] * ld(v, ii + 1, jj) + (R(1.0) - g.fexp[ii]) * ld(v, ii, jj);
    vs = g.fexp[ii] * ld(v, ii + 1, jj - 1) + (R(1.0) - g.fexp[ii]) * ld(v, ii, jj - 1);
    lx = g.dxu[ii];
    ly = g.dyu[jj];
  } else if (EQ == EQ_V) {
    ud = g.feyp[jj] * ld(u, ii, jj + 1) + (R(1.0) - g.feyp[jj]) * ld(u, ii, jj);
    ui = g.feyp[jj] * ld(u, ii - 1, jj + 1) + (R(1.0) - g.feyp[jj]) * ld(u, ii - 1, jj);
    vn = g.feyv[jj] * ld(v, ii, jj + 1) + (R(1.0) - g.feyv[jj]) * ld(v, ii, jj);
    vs = g.feyv[jj - 1] * ld(v, ii, jj) + (R(1.0) - g.feyv[jj - 1]) * ld(v, ii, jj - 1);
    lx = g.dxv[ii];
    ly = g.dyv[jj];
  } else {
    ud = ld(u, ii, jj);
    ui = ld(u, ii - 1, jj);
    vn = ld(v, ii, jj);
    vs = ld(v, ii, jj - 1);
    lx = g.dxp[ii];
    ly = g.dyp[jj];
  }
  CoefT<R> c;
  const FaceT<R> fx = face(F.fx[e], ii), fy = face(F.fy[e], jj);
  c.ai = nbf<R>(fx.c_lo, ly, fx.p_lo, ui, false);
  c.ad = nbf<R>(fx.c_hi, ly, fx.p_hi, ud, true);
  c.bs = nbf<R>(fy.c_lo, lx, fy.p_lo, vs, false);
  c.bn = nbf<R>(fy.c_hi, lx, fy.p_hi, vn, true);
  area = lx * ly;
  R acc = -c.ai - c.ad - c.bs - c.bn;
  pin = __builtin_nan("");
  const double spv = ATZ(or_zero(A.sp), A.sp != nullptr, ii, jj);
  if constexpr (sizeof(R) == 8) {
    acc = acc - area * spv;
  } else {
    const double scv = ATZ(or_zero(A.sc), A.sc != nullptr, ii, jj);
    const bool pen = spv < -kPenaltyRow;
    // no branch (it would take the Sc load with it): an unpenalised row
    // divides by NaN
    pin = scv / (pen ? -spv : __builtin_nan(""));
    acc = pen ? acc : acc - area * R(spv);
  }
  c.ap = acc;
  c.td = area * F.inv_dt;
  c.ac = (acc + c.td) * (EQ == EQ_T ? F.inv_rel_ener : F.inv_rel_vel);
  return c;
}

template <int EQ, bool TR, typename R>
__device__ __forceinline__ R fast_rhs(const SweepArgs &A, int ii, int jj, R head, R ac, R area) {
  const int P = A.P;
  const GeoT<R> &g = geo_of<R>(A);
  const R inv_dt = fast_of<R>(A).inv_dt;
  R r = head;
  const R scv = R(ATZ(or_zero(A.sc), A.sc != nullptr, ii, jj));
  if (EQ == EQ_U) {
    const R *u = fld<R>(A.u);
    const R *t = fld<R>(A.t);
    const bool buoy = A.ph.ri_x != 0.0;
    R tint = g.fexp[ii] * ATZ(t, buoy, ii + 1, jj) + (R(1.0) - g.fexp[ii]) * ATZ(t, buoy, ii, jj);
    r = r - area * R(A.ph.ri_x) * tint;
    r = r + area * scv;
    r = r + area * AT(fld<R>(A.uant), ii, jj) * inv_dt;
    r = r + R((AT(A.p, ii, jj) - AT(A.p, ii + 1, jj)) * A.g.dyu[jj]);
    r = r + ac * R(1.0 - A.ph.rel_vel) * AT(u, ii, jj);
  } else if (EQ == EQ_V) {
    const R *v = fld<R>(A.v);
    const R *t = fld<R>(A.t);
    const bool buoy = A.ph.ri_y != 0.0;
    R tint = g.feyp[jj] * ATZ(t, buoy, ii, jj + 1) + (R(1.0) - g.feyp[jj]) * ATZ(t, buoy, ii, jj);
    r = r - area * R(A.ph.ri_y) * tint;
    r = r + area * scv;
    r = r + area * AT(fld<R>(A.vant), ii, jj) * inv_dt;
    r = r + R((AT(A.p, ii, jj) - AT(A.p, ii, jj + 1)) * A.g.dxv[ii]);
    r = r + ac * R(1.0 - A.ph.rel_vel) * AT(v, ii, jj);
  } else {
    r = r + area * scv;
    r = r + area * AT(fld<R>(A.tant), ii, jj) * inv_dt;
    r = r + ac * R(1.0 - A.ph.rel_ener) * AT(fld<R>(A.t), ii, jj);
  }
  return r;
}

// FP32-refined sweeps: the unrelaxed FP64 residual of row (ii,jj) of the
// current FP64 fields against the FP32 neighbour coefficients c,
//   r = -ai (phi_W - phi_P) - ad (phi_E - phi_P) - bs (phi_S - phi_P)
//       - bn (phi_N - phi_P) + area Sp phi_P + td (phi_ant - phi_P) + sources,
// which is the right-hand side of the row's increment equation.  a_P enters
// only through the neighbour differences (a_P = -sum a_nb - area Sp + td), so
// its FP32 rounding, about 6e-8 of a_P phi_P, does not reach the residual;
// the sources and Sp use the FP64 cell area.  The time term vanishes exactly
// where phi = phi_ant, so the iteration's fixed point is the steady state of
// the FP64 fields for these neighbour coefficients; the increments only need
// FP32 accuracy.
// S: stored type of the fields (double, or float in the stochastic-rounding
// sweeps, whose FP32 values enter the FP64 arithmetic exactly).
// pP: the field at (ii, jj), which the caller has loaded.
template <int EQ, bool TR, typename R, typename S = double>
__device__ __forceinline__ double delta_res(const SweepArgs &A, int ii, int jj, const CoefT<R> &c, double pP) {
  const int P = A.P;
  const S *phi = fld<S>(EQ == EQ_U ? A.u : EQ == EQ_V ? A.v : A.t);
  const S *ant = fld<S>(EQ == EQ_U ? A.uant : EQ == EQ_V ? A.vant : A.tant);
  const S *tf = fld<S>(A.t);
  auto at = [&](const S *f, int i, int j) { return (double)AT(f, i, j); };
  const double ar = EQ == EQ_U ? A.g.dxu[ii] * A.g.dyu[jj]
                  : EQ == EQ_V ? A.g.dxv[ii] * A.g.dyv[jj] : A.g.dxp[ii] * A.g.dyp[jj];
  double r = -(double)c.ai * (at(phi, ii - 1, jj) - pP) - (double)c.ad * (at(phi, ii + 1, jj) - pP);
  r = r - (double)c.bs * (at(phi, ii, jj - 1) - pP) - (double)c.bn * (at(phi, ii, jj + 1) - pP);
  r = r + (double)c.td * (at(ant, ii, jj) - pP);
  if (EQ == EQ_U) {
    const bool buoy = A.ph.ri_x != 0.0;
    const double tint = A.g.fexp[ii] * (double)ATZ(tf, buoy, ii + 1, jj) +
                        (1.0 - A.g.fexp[ii]) * (double)ATZ(tf, buoy, ii, jj);
    r = r - ar * A.ph.ri_x * tint;
    r = r + (AT(A.p, ii, jj) - AT(A.p, ii + 1, jj)) * A.g.dyu[jj];
  } else if (EQ == EQ_V) {
    const bool buoy = A.ph.ri_y != 0.0;
    const double tint = A.g.feyp[jj] * (double)ATZ(tf, buoy, ii, jj + 1) +
                        (1.0 - A.g.feyp[jj]) * (double)ATZ(tf, buoy, ii, jj);
    r = r - ar * A.ph.ri_y * tint;
    r = r + (AT(A.p, ii, jj) - AT(A.p, ii, jj + 1)) * A.g.dxv[ii];
  }
  r = r + ar * ATZ(or_zero(A.sc), A.sc != nullptr, ii, jj);
  r = r + ar * ATZ(or_zero(A.sp), A.sp != nullptr, ii, jj) * pP;
  return r;
}

// ---------------------------------------------------------------- solvers
// Thomas without pivoting in upstream's order (mod_solucionador.f90:25-47),
// one lane; a[0] and c[n-1] are not read.
template <typename T>
__device__ __forceinline__ void thomas_rows(T *a, T *b, T *c, T *r, int n) {
  for (int k = 1; k < n; ++k) {
    T m = a[k] / b[k - 1];
    r[k] = r[k] - m * r[k - 1];
    b[k] = b[k] - m * c[k - 1];
  }
  r[n - 1] = r[n - 1] / b[n - 1];
  for (int k = n - 2; k >= 0; --k) r[k] = (r[k] - c[k] * r[k + 1]) / b[k];
}

template <typename T>
__device__ __forceinline__ void thomas_lds(T *a, T *b, T *c, T *r, int n) {
  if (threadIdx.x == 0) thomas_rows(a, b, c, r, n);
  __syncthreads();
}

// Parallel cyclic reduction (Hockney) over the whole line in LDS; EPT rows
// per thread.  Row k couples to k-s and k+s after step s; a[0] and c[n-1]
// are set to zero first.  On exit r[k] holds x[k].
template <int EPT, typename T>
__device__ __forceinline__ void pcr_lds(T *a, T *b, T *c, T *r, int n) {
  if (threadIdx.x == 0) {
    a[0] = T(0);
    c[n - 1] = T(0);
  }
  __syncthreads();
  for (int s = 1; s < n; s <<= 1) {
    T na[EPT], nb[EPT], nc[EPT], nr[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      int k = threadIdx.x + e * kBlock;
      if (k < n) {
        T ak = a[k], bk = b[k], ck = c[k], rk = r[k];
        T ta = T(0), tc = T(0);
        if (k - s >= 0) {
          T al = -qdiv(ak, b[k - s]);
          ta = al * a[k - s];
          bk = bk + al * c[k - s];
          rk = rk + al * r[k - s];
        }
        if (k + s < n) {
          T ga = -qdiv(ck, b[k + s]);
          tc = ga * c[k + s];
          bk = bk + ga * a[k + s];
          rk = rk + ga * r[k + s];
        }
        na[e] = ta;
        nb[e] = bk;
        nc[e] = tc;
        nr[e] = rk;
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      int k = threadIdx.x + e * kBlock;
      if (k < n) {
        a[k] = na[e];
        b[k] = nb[e];
        c[k] = nc[e];
        r[k] = nr[e];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    int k = threadIdx.x + e * kBlock;
    if (k < n) r[k] = qdiv(r[k], b[k]);
  }
  __syncthreads();
}

// Wave-partitioned solve (one wave64, no workgroup barriers inside):
//  1. lane l owns rows s = l*M .. e = s+M-1 (M = ceil(n/64), rows past n are
//     identity rows) and eliminates them in registers, downward then upward,
//     so that every row reads  L_i x[s-1] + B_i x_i + C_i x[e+1] = R_i ;
//  2. the chunk end rows give two equations per lane in the chunk-boundary
//     unknowns z[2l] = x_s, z[2l+1] = x_e; combined without divisions they
//     form a tridiagonal system of 128 rows.  One cyclic-reduction step
//     removes the odd unknowns, parallel cyclic reduction across lanes with
//     shuffles solves the 64 even rows (6 steps), and each odd unknown then
//     follows from its own row;
//  3. each lane back-substitutes its interior rows from x[s-1] = z[2l-1] and
//     x[e+1] = z[2l+2].
// MMAX bounds M at compile time (registers); M itself is uniform per wave.
template <int MMAX, typename T>
__device__ __forceinline__ void wpt_body(T *a, T *b, T *c, T *r, int n, const int l) {
  {
    const int M = (n + 63) >> 6 < 2 ? 2 : (n + 63) >> 6;
    const int s0 = l * M;
    T A[MMAX], B[MMAX], C[MMAX], R[MMAX], L[MMAX], Cp[MMAX];
    // every read goes out before the first use (selects, no branches);
    // rows past the line are identity rows, slots past the chunk unused
#pragma unroll
    for (int i = 0; i < MMAX; ++i) {
      const int k = s0 + i;
      const bool in = (i < M) & (k < n);
      const int kc = in ? k : 0;
      const T av = a[kc], bv = b[kc], cv = c[kc], rv = r[kc];
      A[i] = (in & (k != 0)) ? av : T(0);
      B[i] = in ? bv : T(1);
      C[i] = (in & (k != n - 1)) ? cv : T(0);
      R[i] = in ? rv : T(0);
    }
    // The spikes are kept normalised: L_i = A_0 l_i (coefficient of x[s-1]
    // in row i) and Cp_i = C_{M-1} c_i (coefficient of x[e+1]), with l_0 = 1
    // and c_{M-1} = 1 at the start of their passes.  The reduced rows below
    // then combine rows s and e with factors of order one: multiplying by
    // the couplings A_0 or C_{M-1} themselves, as a cross-multiplied form
    // would, leaves a reduced pivot of their size, which for a vanishing
    // coupling (strongly upwinded power-law faces at high Peclet numbers)
    // underflows in FP32 and makes the whole line infinite.
    const T a0 = A[0];
    T cM = T(0);
#pragma unroll
    for (int i = 0; i < MMAX; ++i)
      if (i == M - 1) cM = C[i];
    // downward: l_i, spike on x[s-1].  The pivots B_i are final after this
    // pass, so their reciprocals serve the upward pass and the back
    // substitution too: only this pass has a reciprocal in its dependent
    // chain.
    T IB[MMAX];
    L[0] = T(1);
    IB[0] = rcpn(B[0]);
#pragma unroll
    for (int i = 1; i < MMAX; ++i) {
      if (i < M) {
        const T f = A[i] * IB[i - 1];
        B[i] = B[i] - f * C[i - 1];
        R[i] = R[i] - f * R[i - 1];
        L[i] = -f * L[i - 1];
        IB[i] = rcpn(B[i]);
      }
    }
    // upward: c_i, spike on x[e+1]
#pragma unroll
    for (int i = MMAX - 1; i >= 0; --i) {
      if (i == M - 1) {
        Cp[i] = T(1);
      } else if (i < M - 1) {
        const T g = C[i] * IB[i + 1];
        R[i] = R[i] - g * R[i + 1];
        L[i] = L[i] - g * L[i + 1];
        Cp[i] = -g * Cp[i + 1];
      }
    }
    // rows s and e:  A_0 l x[s-1] + B x + C_{M-1} c x[e+1] = R, c_e = 1
    const T Ls = L[0], Bs = B[0], Cs = Cp[0], Rs = R[0];
    T Le = T(0), Be = T(0), Re = T(0);
#pragma unroll
    for (int i = 0; i < MMAX; ++i)
      if (i == M - 1) {
        Le = L[i]; Be = B[i]; Re = R[i];
      }
    // reduced rows: p = 0 -> z[2l] = x[s] (row s - c_s row e: no x[e+1]),
    // p = 1 -> z[2l+1] = x[e] (l_s row e - l_e row s: no x[s-1])
    T ra[2], rb[2], rc[2], rr[2];
    ra[0] = a0 * (Ls - Cs * Le); rb[0] = Bs; rc[0] = -(Cs * Be); rr[0] = Rs - Cs * Re;
    ra[1] = -(Le * Bs); rb[1] = Ls * Be; rc[1] = cM * (Ls - Le * Cs); rr[1] = Ls * Re - Le * Rs;
    // The 128 reduced rows: one cyclic-reduction step removes the odd
    // unknowns z[2l+1] from the even rows, PCR solves the 64 even rows (one
    // per lane), and the odd unknowns follow from their own rows.  Each row
    // sends its coefficients already scaled by its pivot reciprocal, three
    // shuffles per neighbour.
    const T ib1 = rcpn(rb[1]);
    const T a1s = ra[1] * ib1, c1s = rc[1] * ib1, r1s = rr[1] * ib1;
    T A0, B0, C0, R0;
    {
      // even row 2l: z[2l-1] from lane l-1's odd row, z[2l+1] from its own
      const T pa = __shfl_up(a1s, 1, 64), pc = __shfl_up(c1s, 1, 64), pr = __shfl_up(r1s, 1, 64);
      const T al = l > 0 ? -ra[0] : T(0);
      const T ga = -rc[0];
      A0 = al * pa;
      C0 = ga * c1s;
      B0 = rb[0] + al * pc + ga * a1s;
      R0 = rr[0] + al * pr + ga * r1s;
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const T ib = rcpn(B0);
      const T as = A0 * ib, cs = C0 * ib, rs = R0 * ib;
      const T ua = __shfl_up(as, d, 64), uc = __shfl_up(cs, d, 64), ur = __shfl_up(rs, d, 64);
      const T da = __shfl_down(as, d, 64), dc = __shfl_down(cs, d, 64), dr = __shfl_down(rs, d, 64);
      const T al = l - d >= 0 ? -A0 : T(0);
      const T ga = l + d <= 63 ? -C0 : T(0);
      const T nA = al * ua, nC = ga * dc;
      B0 = B0 + al * uc + ga * da;
      R0 = R0 + al * ur + ga * dr;
      A0 = nA;
      C0 = nC;
    }
    const T zs = R0 * rcpn(B0);                        // z[2l] = x[s]
    const T xr = __shfl_down(zs, 1, 64);               // x[e+1] = z[2l+2]
    const T ze = (rr[1] - ra[1] * zs - (l < 63 ? rc[1] * xr : T(0))) * ib1;  // x[e]
    const T xl = __shfl_up(ze, 1, 64);                 // x[s-1]
#pragma unroll
    for (int i = 0; i < MMAX; ++i) {
      if (i < M) {
        const int k = s0 + i;
        T x;
        if (i == 0) x = zs;
        else if (i == M - 1) x = ze;
        else x = (R[i] - (l > 0 ? L[i] * (a0 * xl) : T(0)) - (l < 63 ? Cp[i] * (cM * xr) : T(0))) * IB[i];
        if (k < n) r[k] = x;
      }
    }
  }
}

// the wave solver, run by the first wave of a 256-thread line workgroup, or
// by every wave where each wave holds a line of its own (BS = 64)
template <int MMAX, typename T, int BS = kBlock>
__device__ __forceinline__ void wpt_solve(T *a, T *b, T *c, T *r, int n) {
  if (BS == 64 || threadIdx.x < 64) wpt_body<MMAX>(a, b, c, r, n, (int)(threadIdx.x & 63));
  __syncthreads();
}

// BS: threads of the line's workgroup (256, or 64 for one wave per line);
// the wave solver's chunk length is EPT * BS / 64 rows per lane
template <int SOLVER, int EPT, typename T, int BS = kBlock>
__device__ __forceinline__ void solve_line(T *a, T *b, T *c, T *r, int n) {
  constexpr int MW = EPT * BS / 64;
  if (SOLVER == SOLVE_THOMAS)
    thomas_lds(a, b, c, r, n);
  else if (MW <= 16)
    wpt_solve<(MW < 2 ? 2 : MW), T, BS>(a, b, c, r, n);  // SOLVE_PCR and SOLVE_PCR_EXACT
  else
    pcr_lds<EPT>(a, b, c, r, n);                   // 256 threads, lines longer than 1024
}

// two lines of n rows at once in a 256-thread workgroup (sets a..r and
// a2..r2): the first wave solves the first, the second wave the second;
// lines longer than 1024 rows, one after the other
template <int SOLVER, int EPT, typename T>
__device__ __forceinline__ void solve_two(T *a, T *b, T *c, T *r, T *a2, T *b2, T *c2, T *r2, int n) {
  constexpr int MW = EPT * kBlock / 64;
  if (SOLVER == SOLVE_THOMAS) {
    if (threadIdx.x == 0) thomas_rows(a, b, c, r, n);
    if (threadIdx.x == 64) thomas_rows(a2, b2, c2, r2, n);
    __syncthreads();
  } else if (MW <= 16) {
    const int w = (int)threadIdx.x >> 6, l = (int)threadIdx.x & 63;
    if (w == 0) wpt_body<(MW < 2 ? 2 : MW)>(a, b, c, r, n, l);
    if (w == 1) wpt_body<(MW < 2 ? 2 : MW)>(a2, b2, c2, r2, n, l);
    __syncthreads();
  } else {
    pcr_lds<EPT>(a, b, c, r, n);
    pcr_lds<EPT>(a2, b2, c2, r2, n);
  }
}

// deterministic block sum: butterfly within each wave64, then the four wave
// sums in a fixed order (a + b and b + a round alike, so every lane agrees)
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = x + __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ double block_sum(double x, double *scratch) {
  static_assert(kBlock == 256, "block_sum folds four waves");
  x = wave_sum(x);
  if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = x;
  __syncthreads();
  const double s = (scratch[0] + scratch[1]) + (scratch[2] + scratch[3]);
  __syncthreads();
  return s;
}

// The sum of a row's terms in the order of a 256-thread workgroup: term k
// belongs to thread k % 256, each thread adds its terms in increasing k,
// then block_sum.  One wave (BS = 64) keeps the partial of owner wave
// (k % 256) / 64 in acc[that wave]; with 256 threads only acc[0] is used.
// Every thread gets the sum; scratch: 4 doubles of LDS.
template <int BS>
__device__ __forceinline__ double owner_sum(const double acc[4], double *scratch) {
  static_assert(BS == kBlock || BS == 64, "workgroups of 256 threads or one wave");
  if (BS == kBlock) return block_sum(acc[0], scratch);
  return (wave_sum(acc[0]) + wave_sum(acc[1])) + (wave_sum(acc[2]) + wave_sum(acc[3]));
}

// Workgroup w runs on XCD w % 8; give every XCD a contiguous block of lines
// so that neighbouring lines, which read each other's cache lines, share
// one L2.  A bijection of [0, n).
__device__ __forceinline__ int xcd_line(int w, int n) {
  const int x = w & 7, k = w >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + k;
}

// boundary row of a sweep: side low (a/b) or high (c/d); an untyped node
// keeps the row passed in (a0 = a2 = 0) and returns false.  Selects, not
// branches, so both loads go out with the rest of the assembly's.
__device__ __forceinline__ bool bc_row(const Side &sd, int node, bool high, double &a0, double &a1,
                                       double &a2, double &rr) {
  // (plain selects: a nested or short-circuit condition is emitted as
  // branches, into which the loads are then sunk)
  const int t = gp(sd.type)[node];
  const double v = gp(sd.val)[node];
  const bool none = t == 0;
  const bool nlo = (t == 2) & !high, nhi = (t == 2) & high;
  const double typed_a1 = nlo ? -1.0 : 1.0;
  a0 = nhi ? -1.0 : 0.0;
  a1 = none ? a1 : typed_a1;
  a2 = nlo ? 1.0 : 0.0;
  rr = none ? rr : v;
  return !none;
}

// ---------------------------------------------------------- sweep kernel
// Extents per equation:  line length n, interior range [lo, hi] along the
// line, first/last line, and the boundary side pair.
template <int EQ, int DIR>
struct Layout {
  __device__ static int n(int mi, int nj) {
    if (DIR == DIR_X) return EQ == EQ_U ? mi : mi + 1;
    return EQ == EQ_V ? nj : nj + 1;
  }
  __device__ static int hi(int mi, int nj) {  // last interior position
    if (DIR == DIR_X) return EQ == EQ_U ? mi - 1 : mi;
    return EQ == EQ_V ? nj - 1 : nj;
  }
  __device__ static int last_line(int mi, int nj) {
    if (DIR == DIR_X) return EQ == EQ_V ? nj - 1 : nj;
    return EQ == EQ_U ? mi - 1 : mi;
  }
};

// One line of a sweep, run by one workgroup; `line` is 1-based.  R is the
// real type of the stored fields, the assembly and the line solve (double,
// or float in the FP32-mixed path, which always uses the PCR solver); rows
// are formed in FP64 registers, p' right-hand sides and norms in FP64.
// Every global load of the line (stencil operands, boundary values, the
// pass-start field for the norms, the lines carried through) is issued in
// the assembly loop, before any global store, so a workgroup waits for
// memory once; rows past the line or at its ends assemble at a clamped
// interior position and are then replaced.
// LDS of one line workgroup of BS threads, in doubles: the four line arrays
// and, for 256 threads, the reduction scratch (one wave reduces in registers;
// leaving it out raises the batch kernels' occupancy)
template <typename R, int EPT, int BS = kBlock>
constexpr int line_lds() { return (4 * EPT * BS * (int)sizeof(R)) / 8 + (BS == kBlock ? kBlock : 0); }

// A zero the compiler must treat as a per-lane value.  Indices built on it
// make uniform table lookups vector loads: vector loads return in order, so
// each can be waited for alone, while a scalar load's value is usable only
// once every outstanding scalar load has returned.
__device__ __forceinline__ int lane_zero() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// CORR (p' x-sweep only, 256 threads): the workgroup also assembles and
// solves line + 1 (its second wave, beside the first), then corrects u, v
// and p on its own line from p' of both lines, so the correction needs no
// launch of its own.  Line + 1's solve repeats the one its own workgroup
// does, on the same inputs, so both give the same bits; only the owner
// writes p' and the norms.  LDS: twice line_lds.
// PZ (p' y-sweep only): p' is zero at entry, so it is not read at all, and
// only the other (row) layout of the result is written: the p' x-sweep that
// follows reads that layout and writes both itself.
// LPW > 1 (one wave per line only): the workgroup holds LPW consecutive
// lines, wave w line line_s (= the group's first line + w), lds this wave's
// own slice of line_lds doubles, the slices of the group in wave order.  A
// wave past the cavity's last line assembles and solves that last line again
// and stores nothing.  The other-layout copies of the results (and of
// a_u / a_v) leave from LDS after the group's solves, LPW neighbouring
// lines' values per position in one run of consecutive addresses, instead of
// one value per lane, each in a cache line of its own.
template <int EQ, int DIR, int SOLVER, int EPT, typename R, int BS = kBlock, bool CORR = false, bool PZ = false,
          int LPW = 1>
__device__ __forceinline__ void sweep_line(const SweepArgs &A, const int line_s, double *lds,
                                           const CorrLine *K = nullptr) {
  static_assert(!CORR || (EQ == EQ_P && DIR == DIR_X && BS == kBlock), "the correction rides on the p' x-sweep");
  static_assert(!PZ || (EQ == EQ_P && DIR == DIR_Y), "the zero-input form is the p' y-sweep's");
  static_assert(LPW == 1 || (BS == 64 && !CORR && (LPW & (LPW - 1)) == 0), "line groups of one-wave lines");
  // this thread's position in its line's workgroup (in its wave, for LPW > 1)
  const int tid = LPW > 1 ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  const int lastl = Layout<EQ, DIR>::last_line(A.mi, A.nj);
  const bool active = LPW == 1 || line_s <= lastl;
  // The line index is made a per-lane value, so that the line's uniform
  // table lookups (geometry, face records, boundary values) go out as
  // vector loads that can be waited for one by one.  The FP32-mixed group
  // kernels of the batch keep it scalar: there the lookups go through the
  // scalar cache instead of the vector memory path the sweeps saturate
  // (7026 against 6694 Mcell-iters/s for the 256-cavity batch; FP64 4260
  // against 4302 and FP32-stochastic 5360 against 5774 the other way).
  constexpr bool kScalarLine = LPW > 1 && SOLVER == SOLVE_PCR && sizeof(R) == 4;
  const int line = (LPW > 1 ? min(line_s, lastl) : line_s) + (kScalarLine ? 0 : lane_zero());
  static_assert(sizeof(R) == 8 || wave_solver(SOLVER), "the FP32 sweeps use the PCR solver");
  static_assert(!delta_solver(SOLVER) || sizeof(R) == 4, "the increment sweeps are FP32");
  static_assert(BS == kBlock || BS == 64, "line workgroups of 256 threads or one wave");
  static_assert(BS == kBlock || SOLVER != SOLVE_THOMAS, "smaller line workgroups use the wave solver");
  using S = typename StoreOf<SOLVER, R>::T;      // stored type of the fields
  constexpr bool DELTA = delta_solver(SOLVER) && EQ != EQ_P;
  // stochastic rounding of the stored results (FP32 increment sweeps, and
  // the correction riding on their p' x-sweep)
  constexpr bool SR = SOLVER == SOLVE_DELTA_SR;
  R *sa = reinterpret_cast<R *>(lds), *sb = sa + EPT * BS, *sc_ = sb + EPT * BS, *sr = sc_ + EPT * BS;
  double *red = lds + (4 * EPT * BS * (int)sizeof(R)) / 8;
  // line + 1 (CORR): after the first set and its reduction scratch
  R *sa2 = reinterpret_cast<R *>(red + kBlock), *sb2 = sa2 + EPT * BS, *sc2 = sb2 + EPT * BS, *sr2 = sc2 + EPT * BS;
  constexpr bool TR = DIR == DIR_Y;
  constexpr bool MOM = EQ == EQ_U || EQ == EQ_V;
  constexpr bool NORM = DIR == DIR_X && EQ != EQ_U;
  const int P = A.P;
  const int P2 = A.P2;
  const int mi = A.mi, nj = A.nj;
  const int n = Layout<EQ, DIR>::n(mi, nj);
  const int hi = Layout<EQ, DIR>::hi(mi, nj);  // = n - 1
  // sides holding the two ends of this line: x-lines end on a (low) and c,
  // y-lines on b (low) and d
  const Side &lo_side = DIR == DIR_X ? A.bc[0] : A.bc[1];
  const Side &hi_side = DIR == DIR_X ? A.bc[2] : A.bc[3];
  // a_u / a_v of a boundary face (1e40, main.f90:309-310; infinite in FP32)
  // and of a penalised FP32 row (infinite)
  const R big = sizeof(R) == 8 ? R(1.e40) : R(__builtin_huge_valf());
  const S bigs = sizeof(S) == 8 ? S(1.e40) : S(__builtin_huge_valf());
  // Under-relaxation of a penalised row.  Upstream's row of an immersed-body
  // node carries the 1e51 penalty in a_P and Sc and is under-relaxed like
  // any other, so its value moves as phi <- alpha Sc/-Sp + (1 - alpha) phi
  // (source/mod_ec_momento.f90:528-537, mod_ec_energia.f90:260-266); the
  // FP32 rows replacing it follow the same update.
  const double relx = EQ == EQ_T ? A.ph.rel_ener : A.ph.rel_vel;
  S *acoef = fldw<S>(A.acoef), *acoef2 = fldw<S>(A.acoef2);
  S *out = fldw<S>(A.out), *out2 = fldw<S>(A.out2);
  // the boundary rows at both ends (p': Dirichlet 0, main.f90:789-800, 845-856)
  double lb0 = 0.0, lb1 = 1.0, lb2 = 0.0, lbr = 0.0, hb0 = 0.0, hb1 = 1.0, hb2 = 0.0, hbr = 0.0;
  bool lo_typed = false, hi_typed = false;
  if (EQ != EQ_P) {
    lo_typed = bc_row(lo_side, line, false, lb0, lb1, lb2, lbr);
    hi_typed = bc_row(hi_side, line, true, hb0, hb1, hb2, hbr);
  }
  // the first and last solved lines also carry the lines next to them
  const int other_hi = line == lastl ? lastl + 1 : 0;
  const int other = line == A.line0 ? A.line0 - 1 : other_hi;
  const S *src = fld<S>(EQ == EQ_U ? A.u : EQ == EQ_V ? A.v : EQ == EQ_T ? A.t : A.pc);
  // the field at position q of this line (increment sweeps)
  auto along = [&](int q) -> double {
    const S *f = fld<S>(EQ == EQ_U ? A.u : EQ == EQ_V ? A.v : A.t);
    return DIR == DIR_X ? (double)AT(f, q, line) : (double)AT(f, line, q);
  };

  // the p' row at (ii, jj): stencil, right-hand side, mass residual b_o;
  // clamp_n: the north p' read stays on the last row of p' (row nj + 1)
  auto p_row = [&](int ii, int jj, double &a0, double &a1, double &a2, double &rr, double &bval,
                   bool clamp_n) {
    const PCoefT<R> c = p_coef<TR, R, S>(A, ii, jj);
    // p' = 0 at the first pass: the loads still go out (the buffer is
    // allocated), their values are dropped
    const S *pcf = fld<S>(A.pc);
    const bool z = PZ || A.in_zero != 0;
    const double pcv = PZ ? 0.0 : (double)AT(pcf, ii, jj);
    double pcP = z ? 0.0 : pcv;
    double head;
    if (DIR == DIR_X) {
      a0 = c.ai; a2 = c.ad;
      const int jn = clamp_n ? min(jj + 1, nj + 1) : jj + 1;
      const double s0 = (double)AT(pcf, ii, jj - 1), n0 = (double)AT(pcf, ii, jn);
      double pS = z ? 0.0 : s0;
      double pN = z ? 0.0 : n0;
      head = -(double)c.bs * pS - (double)c.bn * pN;
      bval = c.b;
    } else {
      a0 = c.bs; a2 = c.bn;
      const double w0 = PZ ? 0.0 : (double)AT(pcf, ii - 1, jj), e0 = PZ ? 0.0 : (double)AT(pcf, ii + 1, jj);
      double pW = z ? 0.0 : w0;
      double pE = z ? 0.0 : e0;
      head = -(double)c.ai * pW - (double)c.ad * pE;
    }
    a1 = c.ac;
    rr = head + c.b + (1.0 - A.ph.rel_pres) * (double)c.ac * pcP;
    if constexpr (sizeof(R) == 4) {
      // every neighbour behind an infinite a_u, a_v (inside a body): p' = 0
      const bool shut = c.ac == R(0);
      a0 = shut ? 0.0 : a0;
      a1 = shut ? 1.0 : a1;
      a2 = shut ? 0.0 : a2;
      rr = shut ? 0.0 : rr;
    }
  };

  R keep_aw[EPT];
  S carry[EPT];
  double keep_old[EPT], keep_b[EPT];
  S keep_phi[EPT];  // the field at each row's own position (increment sweeps)
  // CORR: u, v, p, a_u, a_v and dx_v at this line's positions (pre-correction),
  // u and v of the carried boundary line
  double keep_u[EPT], keep_v[EPT], keep_p[EPT], keep_au[EPT], keep_av[EPT], keep_dxv[EPT];
  S keep_ub[EPT], keep_vb[EPT];
  const double dyu_line = CORR ? (double)A.g.dyu[line] : 0.0;
  const double u_first = CORR ? (double)AT(fld<S>(A.u), 1, line) : 0.0;
  const double v_first = CORR ? (double)AT(fld<S>(A.v), 1, line) : 0.0;
  const double v_last = CORR ? (double)AT(fld<S>(A.v), mi + 1, line) : 0.0;
  // FP32 fields read a zero block as zeros too
  const S *snap = A.snap ? fld<S>(A.snap) : reinterpret_cast<const S *>(kZeroBlock);
  const int oline = other ? other : line;  // a valid line to load from when none is carried
  // the field next to each end of the line (boundary rows of the increment sweeps)
  const double phi_2 = DELTA ? along(2) : 0.0, phi_n1 = DELTA ? along(n - 1) : 0.0;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = tid + e * BS;
    const int pos = k + 1;                                    // 1-based position along the line
    const int pc = min(max(pos, 2), hi);                     // assembled position
    const int pr = pos > n ? n : pos;                         // position for the plain loads
    const int ii = DIR == DIR_X ? pc : line;
    const int jj = DIR == DIR_X ? line : pc;
    const int iir = DIR == DIR_X ? pr : line;
    const int jjr = DIR == DIR_X ? line : pr;
    double a0 = 0.0, a1 = 1.0, a2 = 0.0, rr = 0.0, bval = 0.0;
    R aw = R(0);
    if (EQ == EQ_P) {
      p_row(ii, jj, a0, a1, a2, rr, bval, false);
    } else if constexpr (delta_solver(SOLVER)) {
      R area;
      double pin;
      const CoefT<R> c = fast_coef<EQ, TR, R, S>(A, ii, jj, area, pin);
      a0 = DIR == DIR_X ? c.ai : c.bs;
      a2 = DIR == DIR_X ? c.ad : c.bn;
      a1 = c.ac;
      // the field at the row's own position, which serves the residual and
      // the new value.  It is loaded at the unclamped position: an interior
      // row's clamped position is its own, and the residual of an end row
      // (assembled at its clamped neighbour) is replaced by its boundary row
      keep_phi[e] = DIR == DIR_X ? AT(fld<S>(EQ == EQ_U ? A.u : EQ == EQ_V ? A.v : A.t), pr, line)
                                 : AT(fld<S>(EQ == EQ_U ? A.u : EQ == EQ_V ? A.v : A.t), line, pr);
      const double phic = (double)keep_phi[e];
      rr = delta_res<EQ, TR, R, S>(A, ii, jj, c, phic);
      aw = c.ac * R(A.ph.rel_vel);
      const bool pinned = pin == pin;
      a0 = pinned ? 0.0 : a0;
      a1 = pinned ? 1.0 : a1;
      a2 = pinned ? 0.0 : a2;
      rr = pinned ? relx * (pin - phic) : rr;
      aw = pinned ? big : aw;
    } else if constexpr (SOLVER == SOLVE_PCR) {
      R area;
      double pin;
      const CoefT<R> c = fast_coef<EQ, TR, R>(A, ii, jj, area, pin);
      const R *phi = fld<R>(EQ == EQ_U ? A.u : EQ == EQ_V ? A.v : A.t);
      R head;
      if (DIR == DIR_X) {
        a0 = c.ai; a2 = c.ad;
        head = -c.bs * AT(phi, ii, jj - 1) - c.bn * AT(phi, ii, jj + 1);
      } else {
        a0 = c.bs; a2 = c.bn;
        head = -c.ai * AT(phi, ii - 1, jj) - c.ad * AT(phi, ii + 1, jj);
      }
      a1 = c.ac;
      rr = fast_rhs<EQ, TR, R>(A, ii, jj, head, c.ac, area);
      aw = c.ac * R(A.ph.rel_vel);
      if constexpr (sizeof(R) == 4) {
        const bool pinned = pin == pin;
        const double phio = (double)AT(phi, ii, jj);
        a0 = pinned ? 0.0 : a0;
        a1 = pinned ? 1.0 : a1;
        a2 = pinned ? 0.0 : a2;
        rr = pinned ? relx * pin + (1.0 - relx) * phio : rr;
        aw = pinned ? big : aw;
      }
    } else {
      const Coef c = EQ == EQ_U ? u_coef<TR>(A, ii, jj) : EQ == EQ_V ? v_coef<TR>(A, ii, jj)
                                                                       : t_coef<TR>(A, ii, jj);
      const double *phi = EQ == EQ_U ? A.u : EQ == EQ_V ? A.v : A.t;
      double head;
      if (DIR == DIR_X) {
        a0 = c.ai; a2 = c.ad;
        head = -c.bs * AT(phi, ii, jj - 1) - c.bn * AT(phi, ii, jj + 1);
      } else {
        a0 = c.bs; a2 = c.bn;
        head = -c.ai * AT(phi, ii - 1, jj) - c.ad * AT(phi, ii + 1, jj);
      }
      a1 = c.ac;
      rr = EQ == EQ_U ? u_rhs<TR>(A, ii, jj, head, c.ac)
         : EQ == EQ_V ? v_rhs<TR>(A, ii, jj, head, c.ac) : t_rhs<TR>(A, ii, jj, head, c.ac);
      aw = R(c.ac * A.ph.rel_vel);
    }
    // the boundary rows, as selects: a branch would take the boundary
    // loads into it and split the assembly into blocks
    const bool lo = pos == 1, hi_row = pos == n;
    a0 = hi_row ? hb0 : a0;
    a1 = hi_row ? hb1 : a1;
    a2 = hi_row ? hb2 : a2;
    rr = hi_row ? hbr : rr;
    a0 = lo ? lb0 : a0;
    a1 = lo ? lb1 : a1;
    a2 = lo ? lb2 : a2;
    rr = lo ? lbr : rr;
    if constexpr (DELTA) {
      // boundary rows in increments (a0 = 0 on the first row, a2 = 0 on the last)
      const double ph = (double)keep_phi[e];
      const double r_lo = rr - a1 * ph - a2 * phi_2;
      const double r_hi = rr - a0 * phi_n1 - a1 * ph;
      rr = hi_row ? r_hi : rr;
      rr = lo ? r_lo : rr;
    }
    keep_aw[e] = aw;
    keep_b[e] = bval;
    if (NORM) keep_old[e] = (double)ATZ(snap, A.snap != nullptr, iir, jjr);
    {
      const int io = DIR == DIR_X ? pr : oline;
      const int jo = DIR == DIR_X ? oline : pr;
      const S cv = PZ ? S(0) : AT(src, io, jo);
      carry[e] = (EQ == EQ_P && A.in_zero) ? S(0) : cv;
    }
    if constexpr (CORR) {
      // the correction's operands and u, v of the whole line and of the
      // boundary line it carries, which go to the new buffers unchanged.
      // Interior positions read at the row's own (ii, jj), the same loads as
      // the row's coefficients; the end positions add u(1), v(1), v(mi+1).
      const int iul = min(pos, mi), ivl = min(pos, mi + 1);
      const S *uo = fld<S>(A.u), *vo = fld<S>(A.v);
      keep_p[e] = AT(A.p, ii, line);
      const double u_in = (double)AT(uo, ii, line), v_in = (double)AT(vo, ii, line);
      keep_u[e] = pos == 1 ? u_first : u_in;
      keep_v[e] = pos == 1 ? v_first : v_in;
      keep_v[e] = pos == mi + 1 ? v_last : keep_v[e];
      keep_au[e] = (double)AT(fld<S>(A.au), ii, line);
      keep_av[e] = (double)AT(fld<S>(A.av), ii, line);
      keep_dxv[e] = A.g.dxv[ii];
      keep_ub[e] = ATZ(uo, other != 0, iul, oline);
      keep_vb[e] = ATZ(vo, other != 0, ivl, min(oline, nj));
    }
    // rows past the line fill LDS the solver does not read (no branch, so
    // the next row's loads are not held behind this one's stores)
    sa[k] = R(a0);
    sb[k] = R(a1);
    sc_[k] = R(a2);
    sr[k] = R(rr);
  }
  if constexpr (CORR) {
    // line + 1, rows as its owner builds them; past the last line (row
    // nj + 1, allocated in every field) the rows are not used, and their
    // north p' read stays in the array
    const int line2 = line + 1;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int k = tid + e * BS;
      const int pos = k + 1;
      double a0 = 0.0, a1 = 1.0, a2 = 0.0, rr = 0.0, bval = 0.0;
      p_row(min(max(pos, 2), hi), line2, a0, a1, a2, rr, bval, true);
      const bool lo = pos == 1, hi_row = pos == n;
      a0 = hi_row ? hb0 : a0;
      a1 = hi_row ? hb1 : a1;
      a2 = hi_row ? hb2 : a2;
      rr = hi_row ? hbr : rr;
      a0 = lo ? lb0 : a0;
      a1 = lo ? lb1 : a1;
      a2 = lo ? lb2 : a2;
      rr = lo ? lbr : rr;
      sa2[k] = R(a0);
      sb2[k] = R(a1);
      sc2[k] = R(a2);
      sr2[k] = R(rr);
    }
  }
  // global stores of the assembly: a_u / a_v (x-sweeps of u, v; typed
  // boundary faces in both sweeps), b_o (p' x-sweep, row layout)
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = tid + e * BS;
    const int pos = k + 1;
    if (k >= n || !active) continue;
    const int ii = DIR == DIR_X ? pos : line;
    const int jj = DIR == DIR_X ? line : pos;
    if (pos >= 2 && pos <= hi) {
      if (MOM && DIR == DIR_X) {
        AT(acoef, ii, jj) = S(keep_aw[e]);
        if (LPW == 1) AO(acoef2, ii, jj) = S(keep_aw[e]);
      }
      if (EQ == EQ_P && DIR == DIR_X) AT(A.bo, ii, jj) = keep_b[e];
    } else if (MOM && (pos == 1 ? lo_typed : hi_typed)) {
      AT(acoef, ii, jj) = bigs;
      AO(acoef2, ii, jj) = bigs;
    }
  }
  __syncthreads();
  if constexpr (CORR)
    solve_two<SOLVER, EPT, R>(sa, sb, sc_, sr, sa2, sb2, sc2, sr2, n);
  else
    solve_line<SOLVER, EPT, R, BS>(sa, sb, sc_, sr, n);

  // write the line and the carried lines, accumulate the x-sweep norms.
  // With one wave per line, row k is summed into the partial of the thread
  // of a 256-thread workgroup that would own it (wave (k % 256) / 64), so
  // the norms equal the 256-thread ones bit for bit.
  double acc0[4] = {0.0, 0.0, 0.0, 0.0}, acc1[4] = {0.0, 0.0, 0.0, 0.0};
  // LPW > 1: the slots of this wave's slice that the other-layout stores
  // read after the group's solves (sa, sb, sc are free once the line is
  // solved): the result x in sr where it is stored in the solver's type,
  // else (FP64 fields of FP32 solves) in sa-sb as FP64; a_u / a_v in sa, or
  // in sc where sa-sb hold x
  constexpr bool WIDE = sizeof(S) > sizeof(R);
  S *tile_x = WIDE ? reinterpret_cast<S *>(sa) : reinterpret_cast<S *>(sr);
  R *tile_aw = WIDE ? sc_ : sa;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = tid + e * BS;
    if (k >= n || !active) continue;
    const int pos = k + 1;
    const int ii = DIR == DIR_X ? pos : line;
    const int jj = DIR == DIR_X ? line : pos;
    S x;
    if constexpr (DELTA) x = store_as<S>((double)keep_phi[e] + (double)sr[k], SR, sr_key(ii, jj, 2 * EQ + DIR));
    else x = sr[k];
    if (!PZ) AT(out, ii, jj) = x;
    if constexpr (LPW == 1) {
      AO(out2, ii, jj) = x;
    } else {
      // the tile slot of x is this row's own (sr[k] held its increment)
      if constexpr (WIDE || DELTA) tile_x[k] = x;
      if constexpr (MOM && DIR == DIR_X) tile_aw[k] = keep_aw[e];
    }
    if (NORM && pos >= 2 && pos <= mi) {
      const int w = BS == kBlock ? 0 : (k & 255) >> 6;
      double d = (double)x - keep_old[e];
      acc0[w] = acc0[w] + d * d;
      if (EQ == EQ_P) acc1[w] = acc1[w] + keep_b[e] * keep_b[e];
    }
    if (other) {
      const int io = DIR == DIR_X ? pos : other;
      const int jo = DIR == DIR_X ? other : pos;
      if (!PZ) AT(out, io, jo) = carry[e];
      AO(out2, io, jo) = carry[e];
    }
  }
  if (NORM) {
    double s0, s1 = 0.0;
    s0 = owner_sum<BS>(acc0, red);
    if (EQ == EQ_P) s1 = owner_sum<BS>(acc1, red);
    if (EQ == EQ_P && tid == 0 && active) gp(A.part1)[line] = s1;
    if (tid == 0 && active) gp(A.part0)[line] = s0;
  }
  if constexpr (LPW > 1) {
    // the other-layout stores of the group: thread t stores line
    // first + t % LPW at positions t / LPW + 64 m, so LPW consecutive
    // threads write LPW consecutive addresses
    __syncthreads();
    constexpr int kSlice = line_lds<R, EPT, 64>();
    const int wv = (int)(threadIdx.x >> 6), t = (int)threadIdx.x;
    const int w = t & (LPW - 1), p0 = t / LPW;
    const int lw = line_s - wv + w;
    // wave w's slice, and the tile slots at their offsets in every slice
    const char *slice = reinterpret_cast<const char *>(lds + (w - wv) * kSlice);
    const char *own = reinterpret_cast<const char *>(lds);
    const S *xw = reinterpret_cast<const S *>(slice + (reinterpret_cast<const char *>(tile_x) - own));
    const R *aww = reinterpret_cast<const R *>(slice + (reinterpret_cast<const char *>(tile_aw) - own));
    if (lw <= lastl) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int k = p0 + 64 * e;
        if (k >= n) continue;
        const int pos = k + 1;
        const int ii = DIR == DIR_X ? pos : lw;
        const int jj = DIR == DIR_X ? lw : pos;
        AO(out2, ii, jj) = xw[k];
        if (MOM && DIR == DIR_X && pos >= 2 && pos <= hi) AO(acoef2, ii, jj) = S(aww[k]);
      }
    }
  }
  if constexpr (CORR) {
    // p += 0.6 p' (ii 2..mi), u += dy_u (p'_P - p'_E) / a_u (ii 2..mi-1),
    // v += dx_v (p'_P - p'_N) / a_v (ii 2..mi, jj <= nj-1), with p' as
    // stored; p in place, u and v into the new buffers, which also receive
    // every other u, v value of the line (and of the carried boundary line)
    // unchanged
    S *u1 = fldw<S>(K->u), *v1 = fldw<S>(K->v), *u1T = fldw<S>(K->uT), *v1T = fldw<S>(K->vT);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int k = tid + e * BS;
      const int pos = k + 1;
      if (pos > n) continue;
      const double pcP = (double)S(sr[k]);
      if (pos >= 2 && pos <= mi) {
        const double pn = keep_p[e] + 0.6 * pcP;
        AT(K->p, pos, line) = pn;
        AO(K->pT, pos, line) = pn;
      }
      if (pos <= mi) {
        S un = S(keep_u[e]);
        if (pos >= 2 && pos <= mi - 1) {
          const double pcE = (double)S(sr[k + 1]);
          un = store_as<S>(keep_u[e] + dyu_line * (pcP - pcE) / keep_au[e], SR, sr_key(pos, line, 8));
        }
        AT(u1, pos, line) = un;
        AO(u1T, pos, line) = un;
      }
      S vn = S(keep_v[e]);
      if (pos >= 2 && pos <= mi && line <= nj - 1) {
        const double pcN = (double)S(sr2[k]);
        vn = store_as<S>(keep_v[e] + keep_dxv[e] * (pcP - pcN) / keep_av[e], SR, sr_key(pos, line, 9));
      }
      AT(v1, pos, line) = vn;
      AO(v1T, pos, line) = vn;
      if (other) {
        if (pos <= mi) {
          AT(u1, pos, other) = keep_ub[e];
          AO(u1T, pos, other) = keep_ub[e];
        }
        if (other <= nj) {
          AT(v1, pos, other) = keep_vb[e];
          AO(v1T, pos, other) = keep_vb[e];
        }
      }
    }
  }
}

// At least 3 waves per SIMD: the FP64 sweeps of 702-node lines need 173
// registers unbounded (the wave solver's rows), which leaves room for 2
// line workgroups per CU; bounded they fit in 168 without spilling, so a
// fused launch of ~1000 lines runs in one round instead of two.
constexpr int kMinWaves = 3;

template <int EQ, int DIR, int SOLVER, int EPT, typename R, int BS = kBlock>
__global__ void __launch_bounds__(BS, kMinWaves) k_sweep(SweepArgs A) {
  __shared__ double lds[line_lds<R, EPT, BS>()];
  sweep_line<EQ, DIR, SOLVER, EPT, R, BS>(A, A.line0 + xcd_line((int)blockIdx.x, A.nlines), lds);
}

// the p' y-sweep of the fused time step: p' zero at entry, row layout out
template <int SOLVER, int EPT, typename R>
__global__ void __launch_bounds__(kBlock, kMinWaves) k_sweep_pz(SweepArgs A) {
  __shared__ double lds[line_lds<R, EPT, kBlock>()];
  sweep_line<EQ_P, DIR_Y, SOLVER, EPT, R, kBlock, false, true>(A, A.line0 + xcd_line((int)blockIdx.x, A.nlines),
                                                              lds);
}

// p' x-sweep with the velocity and pressure correction of its lines
template <int SOLVER, int EPT, typename R>
__global__ void __launch_bounds__(kBlock, kMinWaves) k_pcorr_x(SweepArgs A, CorrLine K) {
  __shared__ double lds[2 * line_lds<R, EPT>()];
  sweep_line<EQ_P, DIR_X, SOLVER, EPT, R, kBlock, true>(A, A.line0 + xcd_line((int)blockIdx.x, A.nlines), lds, &K);
}

// A batch of cavities: workgroup b runs line wg[b].y of cavity wg[b].x, whose
// arguments are D[wg[b].x] (same kernel body as the single-cavity sweep).
// Entries with wg[b].x < 0 pad a table to a multiple of the XCD count.
template <int EQ, int DIR, int SOLVER, int EPT, typename R, int BS = kBlock>
__global__ void __launch_bounds__(BS) k_sweep_batch(const SweepArgs *__restrict__ D,
                                                    const int2 *__restrict__ wg) {
  __shared__ double lds[line_lds<R, EPT, BS>()];
  const int2 w = wg[blockIdx.x];
  if (w.x < 0) return;
  sweep_line<EQ, DIR, SOLVER, EPT, R, BS>(D[w.x], w.y, lds);
}

// A batch sweep with LPW consecutive lines of one cavity per workgroup, one
// wave each (the LPW form of sweep_line): workgroup b runs lines wg[b].y ..
// wg[b].y + LPW - 1 of cavity wg[b].x.  PZ: the zero-input p' y-sweep.
template <int EQ, int DIR, int SOLVER, int EPT, typename R, int LPW, bool PZ = false>
__global__ void __launch_bounds__(64 * LPW) k_sweep_group_batch(const SweepArgs *__restrict__ D,
                                                                const int2 *__restrict__ wg) {
  constexpr int kSlice = line_lds<R, EPT, 64>();
  __shared__ double lds[LPW * kSlice];
  const int2 w = wg[blockIdx.x];
  if (w.x < 0) return;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  sweep_line<EQ, DIR, SOLVER, EPT, R, 64, false, PZ, LPW>(D[w.x], w.y + wave, lds + wave * kSlice);
}

// the p' y-sweep of a batch, every cavity's p' zero at entry (the PZ form of
// sweep_line)
template <int SOLVER, int EPT, typename R, int BS = kBlock>
__global__ void __launch_bounds__(BS) k_sweep_pz_batch(const SweepArgs *__restrict__ D,
                                                       const int2 *__restrict__ wg) {
  __shared__ double lds[line_lds<R, EPT, BS>()];
  const int2 w = wg[blockIdx.x];
  if (w.x < 0) return;
  sweep_line<EQ_P, DIR_Y, SOLVER, EPT, R, BS, false, true>(D[w.x], w.y, lds);
}

// ------------------------------------------------------------- correction
// p += 0.6 p' on ii 2..mi, jj 2..nj; u += dyu (p'_P - p'_E)/a_u on ii 2..mi-1,
// jj 2..nj; v += dxv (p'_P - p'_N)/a_v on ii 2..mi, jj 2..nj-1
// (source/main.f90:931-962), in place.  A workgroup updates a 32 x 32 tile
// in the row layout and writes the same values to the transposed copies
// through LDS, so both stores are coalesced.
struct CorrArgs {
  int mi, nj, P, Q;
  Geo g;
  double *u, *v, *p;          // row layout
  double *uT, *vT, *pT;       // transposed copies (pitch Q)
  const double *pc, *au, *av; // row layout
  int sr;                     // FP32 u, v stored with stochastic rounding
};

constexpr int kTile = 32;

template <typename R>
__device__ __forceinline__ void correct_tile(const CorrArgs &A, const int bx, const int by) {
  __shared__ R tu[kTile][kTile + 1], tv[kTile][kTile + 1];
  __shared__ double tp[kTile][kTile + 1];
  constexpr bool TR = false;
  constexpr int ROWS = kTile / (kBlock / kTile);  // rows per thread
  const int P = A.P;
  const int mi = A.mi, nj = A.nj;
  const R *pc = fld<R>(A.pc), *au = fld<R>(A.au), *av = fld<R>(A.av);
  R *u = fldw<R>(A.u), *v = fldw<R>(A.v);
  const int tx = threadIdx.x & (kTile - 1), ty = threadIdx.x >> 5;  // 32 x 8
  const int i0 = 1 + bx * kTile, j0 = 1 + by * kTile;
  // all loads first (positions clamped into the updated ranges), then stores
  double pn[ROWS];
  R un[ROWS], vn[ROWS];
#pragma unroll
  for (int m = 0; m < ROWS; ++m) {
    const int r = ty + m * (kBlock / kTile);
    const int ic = min(max(i0 + tx, 2), mi), jc = min(max(j0 + r, 2), nj);
    const int iu = min(ic, mi - 1), jv = min(jc, nj - 1);
    const double pcP = AT(pc, ic, jc);
    pn[m] = AT(A.p, ic, jc) + 0.6 * pcP;
    const double pcU = AT(pc, iu, jc);
    un[m] = store_as<R>(AT(u, iu, jc) + A.g.dyu[jc] * (pcU - (double)AT(pc, iu + 1, jc)) / (double)AT(au, iu, jc),
                        A.sr != 0, sr_key(iu, jc, 8));
    const double pcV = AT(pc, ic, jv);
    vn[m] = store_as<R>(AT(v, ic, jv) + A.g.dxv[ic] * (pcV - (double)AT(pc, ic, jv + 1)) / (double)AT(av, ic, jv),
                        A.sr != 0, sr_key(ic, jv, 9));
  }
#pragma unroll
  for (int m = 0; m < ROWS; ++m) {
    const int r = ty + m * (kBlock / kTile);
    const int ii = i0 + tx, jj = j0 + r;
    if (ii < 2 || ii > mi || jj < 2 || jj > nj) continue;
    AT(A.p, ii, jj) = pn[m];
    tp[r][tx] = pn[m];
    if (ii <= mi - 1) {
      AT(u, ii, jj) = un[m];
      tu[r][tx] = un[m];
    }
    if (jj <= nj - 1) {
      AT(v, ii, jj) = vn[m];
      tv[r][tx] = vn[m];
    }
  }
  __syncthreads();
  const int Q = A.Q;
  R *uT = fldw<R>(A.uT), *vT = fldw<R>(A.vT);
  for (int r = ty; r < kTile; r += kBlock / kTile) {
    const int ii = i0 + r, jj = j0 + tx;  // lanes run along jj: coalesced in the transposed copy
    if (ii < 2 || ii > mi || jj < 2 || jj > nj) continue;
    const size_t o = ((size_t)jj - 1) + ((size_t)ii - 1) * (size_t)Q;
    gp(A.pT)[o] = tp[tx][r];
    if (ii <= mi - 1) gp(uT)[o] = tu[tx][r];
    if (jj <= nj - 1) gp(vT)[o] = tv[tx][r];
  }
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_correct(CorrArgs A) {
  correct_tile<R>(A, (int)blockIdx.x, (int)blockIdx.y);
}

// batch: workgroup b corrects tile (wg[b].y % tiles_x, wg[b].y / tiles_x) of
// cavity wg[b].x
template <typename R>
__global__ void __launch_bounds__(kBlock) k_correct_batch(const CorrArgs *__restrict__ D,
                                                          const int2 *__restrict__ wg) {
  const int2 w = wg[blockIdx.x];
  if (w.x < 0) return;
  const CorrArgs &A = D[w.x];
  const int tiles_x = (A.mi + 1 + kTile - 1) / kTile;
  correct_tile<R>(A, w.y % tiles_x, w.y / tiles_x);
}

// ---------------------------------------------------------------- residual
// residuo_u (source/mod_ec_momento.f90:1107-1266): unrelaxed u residual on
// ii 2..mi-1, jj 2..nj; one workgroup per row, per-row sum of squares.  The
// south diffusivity helper keeps upstream's numerator Gamma(ii,jj+1)Gamma(ii,jj)
// (:1208-1209).
struct ResArgs {
  int mi, nj, P;
  Geo g;
  Phys ph;
  const double *u, *v, *t, *p, *uant, *sc, *sp;
  double *resu, *part;
};

// RB: threads per row (256, or 64 for one wave per row).  One wave sums row
// positions into the partials of the 256-thread owners, so the norm is the
// same bit for bit.  Two positions per thread per pass, all loads of a pass
// before its stores.
template <typename R, int RB = kBlock>
__device__ __forceinline__ void residual_row(const ResArgs &A, const int jj, double *red) {
  constexpr bool TR = false;
  constexpr int MR = 2;
  const int P = A.P;
  const int t = (int)threadIdx.x;
  // the stored fields in their real type, read as FP64
  const R *u = fld<R>(A.u), *v = fld<R>(A.v), *t_ = fld<R>(A.t), *uant = fld<R>(A.uant);
  const Geo &g = A.g;
  const double gm = A.ph.gm;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int base = 2; base <= A.mi - 1; base += MR * RB) {
    double rv[MR];
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int ii = min(base + t + m * RB, A.mi - 1);
      double ud = g.fexu[ii] * (double)AT(u, ii + 1, jj) + (1.0 - g.fexu[ii]) * (double)AT(u, ii, jj);
      double ui = g.fexu[ii - 1] * (double)AT(u, ii, jj) + (1.0 - g.fexu[ii - 1]) * (double)AT(u, ii - 1, jj);
      double vn = g.fexp[ii] * (double)AT(v, ii + 1, jj) + (1.0 - g.fexp[ii]) * (double)AT(v, ii, jj);
      double vs = g.fexp[ii] * (double)AT(v, ii + 1, jj - 1) + (1.0 - g.fexp[ii]) * (double)AT(v, ii, jj - 1);
      double gd = harm(gm, gm, g.feyp[jj]);
      double gi = harm(gm, gm, g.feyp[jj]);
      double gn = gi * gd / (gd * (1.0 - g.fexp[ii]) + gi * g.fexp[ii]);
      gd = harm(gm, gm, g.feyp[jj - 1]);
      gi = (gm * gm) / (gm * (1.0 - g.feyp[jj - 1]) + gm * g.feyp[jj - 1]);
      double gs = gi * gd / (gd * (1.0 - g.fexp[ii]) + gi * g.fexp[ii]);
      double di = g.dxp[ii], dd = g.dxp[ii + 1], ds = g.dyv[jj - 1], dn = g.dyv[jj];
      double dx = g.dxu[ii], dy = g.dyu[jj];
      double ai = -nbc(gm, dy, di, ui, false);
      double ad = -nbc(gm, dy, dd, ud, true);
      double bs = -nbc(gs, dx, ds, vs, false);
      double bn = -nbc(gn, dx, dn, vn, true);
      double ac = ai + ad + bs + bn;
      ac = ac - dx * dy * ATZ(or_zero(A.sp), A.sp != nullptr, ii, jj);
      ac = ac + dx * dy / A.ph.dt;
      double r = ai * (double)AT(u, ii - 1, jj) - ac * (double)AT(u, ii, jj) + ad * (double)AT(u, ii + 1, jj) +
                 bs * (double)AT(u, ii, jj - 1) + bn * (double)AT(u, ii, jj + 1);
      const bool buoy = A.ph.ri_x != 0.0;
      double tint = g.fexp[ii] * (double)ATZ(t_, buoy, ii + 1, jj) + (1.0 - g.fexp[ii]) * (double)ATZ(t_, buoy, ii, jj);
      r = r - dx * dy * A.ph.ri_x * tint;
      r = r + dx * dy * ATZ(or_zero(A.sc), A.sc != nullptr, ii, jj);
      r = r + dx * dy * (double)AT(uant, ii, jj) / A.ph.dt;
      r = r + (AT(A.p, ii, jj) - AT(A.p, ii + 1, jj)) * dy;
      rv[m] = r;
    }
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int ii = base + t + m * RB;
      if (ii > A.mi - 1) continue;
      AT(A.resu, ii, jj) = rv[m];
      const int w = RB == kBlock ? 0 : ((ii - 2) & 255) >> 6;
      acc[w] = acc[w] + rv[m] * rv[m];
    }
  }
  const double s = owner_sum<RB>(acc, red);
  if (t == 0) gp(A.part)[jj] = s;
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_residual_u(ResArgs A) {
  __shared__ double red[kBlock];
  residual_row<R>(A, 2 + xcd_line((int)blockIdx.x, A.nj - 1), red);
}

template <typename R, int BS = kBlock>
__global__ void __launch_bounds__(BS) k_residual_u_batch(const ResArgs *__restrict__ D,
                                                         const int2 *__restrict__ wg) {
  __shared__ double red[kBlock];
  const int2 w = wg[blockIdx.x];
  if (w.x < 0) return;
  residual_row<R, BS>(D[w.x], w.y, red);
}

// The T x-sweep of a batch and, in the same launch, the u residual rows of
// every cavity (workgroups nsweep on, after every line).  The residual reads
// u, v, p after the correction and u of the step before, which the T sweeps
// do not write; it reads T only where Ri != 0, and such a batch keeps the
// residual's own launch.  The rows start as the last lines drain.
template <int SOLVER, int EPT, typename R, int BS = kBlock>
__global__ void __launch_bounds__(BS) k_sweep_t_res_batch(const SweepArgs *__restrict__ D,
                                                          const int2 *__restrict__ wg, int nsweep,
                                                          const ResArgs *__restrict__ DR,
                                                          const int2 *__restrict__ wgr) {
  constexpr int kl = line_lds<R, EPT, BS>();
  __shared__ double lds[kl > kBlock ? kl : kBlock];
  const int b = (int)blockIdx.x;
  const int2 w = b < nsweep ? wg[b] : wgr[b - nsweep];
  if (w.x < 0) return;
  if (b < nsweep)
    sweep_line<EQ_T, DIR_X, SOLVER, EPT, R, BS>(D[w.x], w.y, lds);
  else
    residual_row<typename StoreOf<SOLVER, R>::T, BS>(DR[w.x], w.y, lds);
}

// ---------------------------------------------------------------- finalize
// Fold per-line partials into the five norms of one SIMPLE iteration:
// out[0] err_v (lines 2..nj-1), out[1] err_p, out[2] maxbo, out[3] err_t,
// out[4] residuo (lines 2..nj).  `which` masks the ones to fold; `slot` (if
// not null) is a device counter selecting out + 8*slot and incremented
// afterwards, so one graph can record a whole history.  Every thread loads
// its share of all five arrays at once, so the fold costs one memory latency
// and one barrier; the order of the additions is fixed.
struct FinArgs {
  int nj;
  const double *pv, *pp, *pb, *pt, *pr;
  double *out;
  int *slot;
  int which;
};

__device__ __forceinline__ void finalize_body(const FinArgs &A, double *lds) {
  double(*red)[kBlock / 64] = reinterpret_cast<double(*)[kBlock / 64]>(lds);
  const int slot = A.slot ? *A.slot : 0;  // loaded with the partials
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
  for (int j = 2 + (int)threadIdx.x; j <= A.nj; j += kBlock) {
    if ((A.which & 1) && j <= A.nj - 1) a0 = a0 + gp(A.pv)[j];
    if (A.which & 2) a1 = a1 + gp(A.pp)[j];
    if (A.which & 4) a2 = a2 + gp(A.pb)[j];
    if (A.which & 8) a3 = a3 + gp(A.pt)[j];
    if (A.which & 16) a4 = a4 + gp(A.pr)[j];
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  a2 = wave_sum(a2);
  a3 = wave_sum(a3);
  a4 = wave_sum(a4);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = a0; red[1][w] = a1; red[2][w] = a2; red[3][w] = a3; red[4][w] = a4;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __attribute__((address_space(1))) double *o = gp(A.out) + 8 * (size_t)slot;
    double r[5];
    for (int k = 0; k < 5; ++k) r[k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
    for (int k = 0; k < 5; ++k)
      if (A.which & (1 << k)) o[k] = sqrt(r[k]);
    if (A.slot) *A.slot = slot + 1;
  }
}

static __global__ void __launch_bounds__(kBlock) k_finalize(FinArgs A) {
  __shared__ double lds[5 * (kBlock / 64)];
  finalize_body(A, lds);
}

// batch: one workgroup per cavity; the norms of step *slot go to
// D[c].out + 8 * min(*slot, cap) (record `cap` is a discard row)
static __global__ void __launch_bounds__(kBlock) k_finalize_batch(const FinArgs *__restrict__ D, const int *slot,
                                                           int cap) {
  __shared__ double lds[5 * (kBlock / 64)];
  FinArgs A = D[blockIdx.x];
  const int k = *slot;
  A.out = A.out + 8 * (size_t)(k < cap ? k : cap);
  A.slot = nullptr;
  finalize_body(A, lds);
}

// advance the batch step counter once every cavity has read it
static __global__ void k_add_int(int *p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = *p + v;
}

// ---------------------------------------------------- fused time-step launch
// Independent phases of consecutive steps in one launch (one stream, no
// cross-stream waits): workgroups [off[p], off[p] + cnt[p]) run part p --
// sweep 0, sweep 1 (EQ1 >= 0), the u residual, the norm fold.  Each part
// starts at a multiple of 8 so that its lines keep their XCD mapping; padding
// workgroups return.  The parts share one LDS allocation.
struct MultiArgs {
  SweepArgs s[2];
  ResArgs r;
  FinArgs f;
  int off[4], cnt[4];
};

// BS = 64: one wave per line, for launches of sweeps only (no residual, no
// fold)
template <int EQ0, int DIR0, int EQ1, int DIR1, int SOLVER, int EPT, typename R, int BS = kBlock>
__global__ void __launch_bounds__(BS, kMinWaves) k_multi(MultiArgs M) {
  __shared__ double lds[line_lds<R, EPT, BS>()];
  const int b = (int)blockIdx.x;
  if (b < M.off[1]) {
    const int w = b - M.off[0];
    if (w < M.cnt[0])
      sweep_line<EQ0, DIR0, SOLVER, EPT, R, BS>(M.s[0], M.s[0].line0 + xcd_line(w, M.s[0].nlines), lds);
  } else if (b < M.off[2]) {
    if constexpr (EQ1 >= 0) {
      const int w = b - M.off[1];
      if (w < M.cnt[1])
        sweep_line<EQ1, DIR1, SOLVER, EPT, R, BS>(M.s[1], M.s[1].line0 + xcd_line(w, M.s[1].nlines), lds);
    }
  } else if constexpr (BS == kBlock) {
    if (b < M.off[3]) {
      const int w = b - M.off[2];
      if (w < M.cnt[2])
        residual_row<typename StoreOf<SOLVER, R>::T>(M.r, 2 + xcd_line(w, M.r.nj - 1), lds);
    } else {
      if (b == M.off[3] && M.cnt[3]) finalize_body(M.f, lds);
    }
  }
}

static __global__ void k_set_int(int *p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ----------------------------------------------------------- energy adjoint
// Discrete adjoint of one energy step T^{k+1} = X(Y(T^k)), where Y and X are
// the y- and x-sweeps of ensambla_energia_y/_x + boundary rows + line solves
// (main.f90:974-1121) with the velocities of step k.  For fixed velocities the
// step is affine in T^k, T_ant = T^k and the boundary values, so its adjoint
// is exact: per sweep, a transposed line solve A^T mu = lambda on every
// solved line (k_adj_sweep) and the transposed right-hand-side stencil
// (k_adj_gather).  Everything is in the row layout (TR = false); y-lines read
// with stride P.
struct AdjArgs {
  SweepArgs S;             // geometry, physics, boundary conditions of T, u and v of step k
  const double *lam;       // adjoint of the sweep's output (row layout)
  double *mu;              // transposed-solve result (row layout)
  const double *mu_in;     // (gather) the sweep's mu
  double *out;             // (gather) adjoint of the sweep's input
  double *tant;            // (gather) adjoint of T_ant: written by X, added by Y
  double *g_lo, *g_hi;     // boundary-value adjoints of the low / high side (per node)
  double *cf[3];           // per row, from the sweep: a_P and the two coefficients across
                           // it (bs, bn for x-sweeps, ai, ad for y-sweeps)
};

// workgroup of the adjoints' per-node kernels: one wave, so that the nodes
// of a small cavity (17k at 129^2) spread over every CU (their loads form
// long dependent chains)
constexpr int kNodeBlock = 64;

template <int DIR, int EPT>
__global__ void __launch_bounds__(kBlock) k_adj_sweep(AdjArgs G) {
  __shared__ double sa[EPT * kBlock], sb[EPT * kBlock], sc_[EPT * kBlock], sr[EPT * kBlock];
  constexpr bool TR = false;
  const SweepArgs &A = G.S;
  const int P = A.P;
  const int mi = A.mi, nj = A.nj;
  const int n = Layout<EQ_T, DIR>::n(mi, nj);
  const int hi = Layout<EQ_T, DIR>::hi(mi, nj);
  const int line = 2 + (int)blockIdx.x;  // lines 2 .. last
  const Side &lo_side = DIR == DIR_X ? A.bc[0] : A.bc[1];
  const Side &hi_side = DIR == DIR_X ? A.bc[2] : A.bc[3];
  // forward rows
  for (int k = threadIdx.x; k < n; k += kBlock) {
    const int pos = k + 1;
    const int ii = DIR == DIR_X ? pos : line;
    const int jj = DIR == DIR_X ? line : pos;
    double a0 = 0.0, a1 = 1.0, a2 = 0.0, rr = 0.0;
    if (pos >= 2 && pos <= hi) {
      double area, pin;
      const CoefT<double> c = fast_coef<EQ_T, TR, double>(A, ii, jj, area, pin);
      a0 = DIR == DIR_X ? c.ai : c.bs;
      a2 = DIR == DIR_X ? c.ad : c.bn;
      a1 = c.ac;
      AT(G.cf[0], ii, jj) = c.ac;
      AT(G.cf[1], ii, jj) = DIR == DIR_X ? c.bs : c.ai;
      AT(G.cf[2], ii, jj) = DIR == DIR_X ? c.bn : c.ad;
    } else {
      const bool high = pos != 1;
      (void)bc_row(high ? hi_side : lo_side, line, high, a0, a1, a2, rr);
    }
    sa[k] = a0;
    sb[k] = a1;
    sc_[k] = a2;
  }
  __syncthreads();
  // transpose: row k of A^T couples k-1 through c[k-1] and k+1 through a[k+1]
  double ta[EPT], tc[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = threadIdx.x + e * kBlock;
    if (k < n) {
      ta[e] = k > 0 ? sc_[k - 1] : 0.0;
      tc[e] = k < n - 1 ? sa[k + 1] : 0.0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = threadIdx.x + e * kBlock;
    if (k < n) {
      const int pos = k + 1;
      const int ii = DIR == DIR_X ? pos : line;
      const int jj = DIR == DIR_X ? line : pos;
      sa[k] = ta[e];
      sc_[k] = tc[e];
      sr[k] = AT(G.lam, ii, jj);
    }
  }
  __syncthreads();
  solve_line<SOLVE_PCR, EPT>(sa, sb, sc_, sr, n);
  for (int k = threadIdx.x; k < n; k += kBlock) {
    const int pos = k + 1;
    const int ii = DIR == DIR_X ? pos : line;
    const int jj = DIR == DIR_X ? line : pos;
    AT(G.mu, ii, jj) = sr[k];
  }
  // boundary rows take their value from the side: its adjoint is mu there
  if (threadIdx.x == 0) {
    G.g_lo[line] = G.g_lo[line] + sr[0];
    G.g_hi[line] = G.g_hi[line] + sr[n - 1];
  }
}

// adjoint of the sweep's right-hand side at node (i, j): out = d/d(input);
// X also writes tant, Y adds its own T_ant part and tant into out
template <int DIR>
__global__ void __launch_bounds__(kNodeBlock) k_adj_gather(AdjArgs G) {
  constexpr bool TR = false;
  const SweepArgs &A = G.S;
  const int P = A.P;
  const int mi = A.mi, nj = A.nj;
  const int idx = (int)(blockIdx.x * kNodeBlock + threadIdx.x);
  if (idx >= (mi + 1) * (nj + 1)) return;
  const int ii = 1 + idx % (mi + 1), jj = 1 + idx / (mi + 1);
  const double keep = 1.0 - A.ph.rel_ener;
  const double *mu = G.mu_in;
  // the rows of the sweep, and every value read without a branch: arr(i, j)
  // where `on`, else arr(ii, jj) (a valid address, value not used), so the
  // node's loads go out together
  auto row = [&](int i, int j) { return i >= 2 && i <= mi && j >= 2 && j <= nj; };
  auto at = [&](const double *arr, bool on, int i, int j) { return AT(arr, on ? i : ii, on ? j : jj); };
  const bool rP = row(ii, jj);
  const double lam = AT(G.lam, ii, jj), muP = AT(mu, ii, jj), ac = AT(G.cf[0], ii, jj);
  const double tant_old = DIR == DIR_X ? 0.0 : AT(G.tant, ii, jj);
  double o = 0.0, ta = 0.0;
  // lines the sweep does not solve are copied from its input
  const bool passthrough = DIR == DIR_X ? (jj == 1 || jj == nj + 1) : (ii == 1 || ii == mi + 1);
  // the row of (ii, jj) reads its own node through (1 - alpha) a_P phi
  o = passthrough ? lam : rP ? keep * ac * muP : o;
  ta = rP ? A.g.dxp[ii] * A.g.dyp[jj] * A.fg.inv_dt * muP : ta;  // the T cell's area, as fast_coef forms it
  // neighbouring rows read this node in their head term
  if (DIR == DIR_X) {
    const bool rN = row(ii, jj + 1), rS = row(ii, jj - 1);
    const double bs = at(G.cf[1], rN, ii, jj + 1), muN = at(mu, rN, ii, jj + 1);
    const double bn = at(G.cf[2], rS, ii, jj - 1), muS = at(mu, rS, ii, jj - 1);
    o = rN ? o - bs * muN : o;
    o = rS ? o - bn * muS : o;
    AT(G.out, ii, jj) = o;
    AT(G.tant, ii, jj) = ta;
  } else {
    const bool rE = row(ii + 1, jj), rW = row(ii - 1, jj);
    const double ai = at(G.cf[1], rE, ii + 1, jj), muE = at(mu, rE, ii + 1, jj);
    const double ad = at(G.cf[2], rW, ii - 1, jj), muW = at(mu, rW, ii - 1, jj);
    o = rE ? o - ai * muE : o;
    o = rW ? o - ad * muW : o;
    AT(G.out, ii, jj) = o + ta + tant_old;
  }
}

// ----------------------------------------------------------------- Nusselt
// nusselt_promedio_y (source/mod_postproceso.f90:100-215) and the
// vertical-wall mean Nusselt numbers; a single workgroup, sums in upstream
// order by one lane so the result equals the CPU value.
struct NuArgs {
  int mi, nj, P, i0, i1;
  Geo g;
  const double *t;
  double *out;  // [nu0, nu1] (horizontal) or [nu_w, nu_e] (vertical)
  int vertical;
};

template <typename R>
__global__ void __launch_bounds__(kBlock) k_nusselt(NuArgs A) {
  __shared__ double der[kNmax + 4];
  constexpr bool TR = false;
  const int P = A.P;
  const Geo &g = A.g;
  const R *t = fld<R>(A.t);
  if (!A.vertical) {
    for (int w = 0; w < 2; ++w) {
      const int jj = w == 0 ? 1 : A.nj - 1;
      for (int k = threadIdx.x; k < A.mi + 4; k += kBlock) der[k] = 0.0;
      __syncthreads();
      for (int ii = A.i0 + (int)threadIdx.x; ii <= A.i1; ii += kBlock) {
        double dy1 = (double)AT(t, ii, jj + 1) - (double)AT(t, ii, jj);
        double dy2 = (double)AT(t, ii, jj + 2) - (double)AT(t, ii, jj + 1);
        double dx1 = g.yp[jj + 1] - g.yp[jj];
        double dx2 = g.yp[jj + 2] - g.yp[jj + 1];
        double a = (dy2 / dx2 - dy1 / dx1) / (dx1 + dx2);
        double b = dy2 / dx2 - a * (g.yp[jj + 1] + g.yp[jj + 2]);
        double ye = w == 0 ? g.yp[jj] : g.yp[jj + 2];
        der[ii] = 2.0 * a * ye + b;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int ii = A.i0; ii <= A.i1; ii += 2) {
          double x0 = g.xp[ii], x1 = g.xp[ii + 1], x2 = g.xp[ii + 2];
          double y0 = der[ii], y1 = der[ii + 1], y2 = der[ii + 2];
          double al = ((y2 - y1) / (x2 - x1) - (y1 - y0) / (x1 - x0)) / (x2 - x0);
          double be = (y2 - y1) / (x2 - x1) - al * (x1 + x2);
          double ga = y0 - al * x0 * x0 - be * x0;
          acc = acc + al / 3.0 * (x2 * x2 * x2 - x0 * x0 * x0) + be / 2.0 * (x2 * x2 - x0 * x0) +
                ga * (x2 - x0);
        }
        A.out[w] = -acc;
      }
      __syncthreads();
    }
  } else {
    const double hgt = g.yv[A.nj] - g.yv[1];
    for (int w = 0; w < 2; ++w) {
      const int i1 = w == 0 ? 1 : A.mi - 1;
      const double xw = w == 0 ? g.xu[1] : g.xu[A.mi];
      for (int jj = 2 + (int)threadIdx.x; jj <= A.nj; jj += kBlock) {
        double x1 = g.xp[i1], x2 = g.xp[i1 + 1], x3 = g.xp[i1 + 2];
        double d1 = ((double)AT(t, i1 + 1, jj) - (double)AT(t, i1, jj)) / (x2 - x1);
        double d2 = ((double)AT(t, i1 + 2, jj) - (double)AT(t, i1 + 1, jj)) / (x3 - x2);
        double a = (d2 - d1) / (x3 - x1);
        der[jj] = (d1 + a * (2.0 * xw - x1 - x2)) * g.dyp[jj];
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        double acc = 0.0;
        for (int jj = 2; jj <= A.nj; ++jj) acc = acc + der[jj];
        A.out[w] = -acc / hgt;
      }
      __syncthreads();
    }
  }
}

// --------------------------------------------------- batched tridiagonal
template <int SOLVER, int EPT>
__global__ void __launch_bounds__(kBlock) k_tridiag(double *a, double *b, double *c, double *r, int n) {
  __shared__ double sa[EPT * kBlock], sb[EPT * kBlock], sc_[EPT * kBlock], sr[EPT * kBlock];
  const size_t base = (size_t)blockIdx.x * n;
  for (int k = threadIdx.x; k < n; k += kBlock) {
    sa[k] = a[base + k];
    sb[k] = b[base + k];
    sc_[k] = c[base + k];
    sr[k] = r[base + k];
  }
  __syncthreads();
  solve_line<SOLVER, EPT>(sa, sb, sc_, sr, n);
  for (int k = threadIdx.x; k < n; k += kBlock) {
    r[base + k] = sr[k];
    b[base + k] = sb[k];
  }
}

}  // namespace ixh

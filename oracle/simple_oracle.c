/*
 * CPU oracle for the Ixchel2D SIMPLE iteration, FP64.
 *
 * TEST INFRASTRUCTURE ONLY -- the checker for the HIP product path.  It is
 * imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * and by nothing else.
 *
 * A restatement, written from scratch in C, of what the upstream Fortran
 * computes; each routine cites the upstream file:line it follows.  Expression
 * association follows the upstream source term by term (Fortran evaluates
 * left to right, `x**5` is lowered by flang as x*(x*(x*(x*x))), and the
 * x86-64 baseline has no FMA), so with -ffp-contract=off the fields agree
 * with upstream's serial CPU build bit for bit; tests/test_oracle.py checks
 * this against the fixtures that build writes (tests/golden/).
 *
 * The OpenMP pragmas only split independent cells and independent lines, so
 * the fields do not depend on the thread count; the reductions are serial in
 * upstream's loop order.
 */
#include "simple_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct ox_state {
  int mi, nj;
  /* geometry, 1-based: g_xu[1..mi] etc.; index 0 unused */
  double *xu, *yv, *xp, *yp;
  double *dxp, *dyp, *dxu, *dyu, *dxv, *dyv, *fexp, *feyp, *fexu, *feyv;
  /* parameters */
  double dt, rel_pres, rel_vel, rel_ener, gam_m, gam_t, ri_x, ri_y;
  double conv_u, conv_t, conv_p, conv_resi, conv_paso;
  int iter_ecuaci_max, iter_simple_max;
  /* fields, Fortran layout */
  double *f[OX_NFIELDS];
  double *fv, *ft, *fpc;         /* snapshots for the inner-loop errors */
  double *sc[3], *sp[3];         /* sources u, v, T */
  double *gm, *gt;               /* constant-valued diffusivity arrays (mi+1,nj+1) */
  /* shared tridiagonal work arrays: AI..Rx (mi+1,nj+1), BS..Ry (nj+1,mi+1) */
  double *AI, *AC, *AD, *Rx, *BS, *BC, *BN, *Ry;
  /* boundary conditions per variable and side, node k at [k] (1-based) */
  int *bct[3][4];
  double *bcv[3][4];
  int bclen[3][4];
};

#define IX2(a, m, i, j) (a)[((size_t)(i) - 1) + ((size_t)(j) - 1) * (size_t)(m)]
/* shorthands inside functions that have `s` */
#define MI (s->mi)
#define NJ (s->nj)
#define Uf(i, j) IX2(s->f[OX_U], MI, i, j)
#define Vf(i, j) IX2(s->f[OX_V], MI + 1, i, j)
#define Tf(i, j) IX2(s->f[OX_T], MI + 1, i, j)
#define Pf(i, j) IX2(s->f[OX_P], MI + 1, i, j)
#define UA(i, j) IX2(s->f[OX_U_ANT], MI, i, j)
#define VA(i, j) IX2(s->f[OX_V_ANT], MI + 1, i, j)
#define TA(i, j) IX2(s->f[OX_T_ANT], MI + 1, i, j)
#define PC(i, j) IX2(s->f[OX_PCORR], MI + 1, i, j)
#define BO(i, j) IX2(s->f[OX_BO], MI + 1, i, j)
#define AUf(i, j) IX2(s->f[OX_AU], MI, i, j)
#define AVf(i, j) IX2(s->f[OX_AV], MI + 1, i, j)
#define RESU(i, j) IX2(s->f[OX_RESU], MI, i, j)
#define GM(i, j) IX2(s->gm, MI + 1, i, j)
#define GT(i, j) IX2(s->gt, MI + 1, i, j)
#define WAI(i, j) IX2(s->AI, MI + 1, i, j)
#define WAC(i, j) IX2(s->AC, MI + 1, i, j)
#define WAD(i, j) IX2(s->AD, MI + 1, i, j)
#define WRX(i, j) IX2(s->Rx, MI + 1, i, j)
#define WBS(j, i) IX2(s->BS, NJ + 1, j, i)
#define WBC(j, i) IX2(s->BC, NJ + 1, j, i)
#define WBN(j, i) IX2(s->BN, NJ + 1, j, i)
#define WRY(j, i) IX2(s->Ry, NJ + 1, j, i)

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static void field_dims(const ox_state *s, int field, int *m, int *n) {
  switch (field) {
    case OX_U: case OX_U_ANT: case OX_AU: case OX_RESU: *m = s->mi; *n = s->nj + 1; break;
    case OX_V: case OX_V_ANT: case OX_AV: *m = s->mi + 1; *n = s->nj; break;
    default: *m = s->mi + 1; *n = s->nj + 1; break;
  }
}

int64_t ox_field_size(const ox_state *s, int32_t field) {
  int m, n;
  field_dims(s, field, &m, &n);
  return (int64_t)m * n;
}

/* ---------------------------------------------------------------- setup */
ox_state *ox_create(int32_t mi, int32_t nj, const double *xu, const double *yv,
                    const double *xp, const double *yp) {
  ox_state *s = (ox_state *)calloc(1, sizeof(ox_state));
  s->mi = mi;
  s->nj = nj;
  s->xu = zalloc(mi + 2); s->yv = zalloc(nj + 2);
  s->xp = zalloc(mi + 2); s->yp = zalloc(nj + 2);
  for (int i = 1; i <= mi; ++i) s->xu[i] = xu[i - 1];
  for (int j = 1; j <= nj; ++j) s->yv[j] = yv[j - 1];
  for (int i = 1; i <= mi + 1; ++i) s->xp[i] = xp[i - 1];
  for (int j = 1; j <= nj + 1; ++j) s->yp[j] = yp[j - 1];
  s->dxp = zalloc(mi + 2); s->dyp = zalloc(nj + 2);
  s->dxu = zalloc(mi + 2); s->dyu = zalloc(nj + 2);
  s->dxv = zalloc(mi + 2); s->dyv = zalloc(nj + 2);
  s->fexp = zalloc(mi + 2); s->feyp = zalloc(nj + 2);
  s->fexu = zalloc(mi + 2); s->feyv = zalloc(nj + 2);
  /* control-volume sizes and interpolation factors (source/mod_malla.f90:167-217) */
  for (int i = 2; i <= mi; ++i) s->dxp[i] = s->xu[i] - s->xu[i - 1];
  for (int j = 2; j <= nj; ++j) s->dyp[j] = s->yv[j] - s->yv[j - 1];
  for (int i = 1; i <= mi; ++i) s->dxu[i] = s->xp[i + 1] - s->xp[i];
  for (int j = 2; j <= nj; ++j) s->dyu[j] = s->yv[j] - s->yv[j - 1];
  for (int i = 2; i <= mi; ++i) s->dxv[i] = s->xu[i] - s->xu[i - 1];
  for (int j = 1; j <= nj; ++j) s->dyv[j] = s->yp[j + 1] - s->yp[j];
  for (int i = 1; i <= mi; ++i) s->fexp[i] = (s->xp[i + 1] - s->xu[i]) / (s->xp[i + 1] - s->xp[i]);
  for (int j = 1; j <= nj; ++j) s->feyp[j] = (s->yp[j + 1] - s->yv[j]) / (s->yp[j + 1] - s->yp[j]);
  for (int i = 1; i <= mi - 1; ++i) s->fexu[i] = (s->xu[i + 1] - s->xp[i + 1]) / (s->xu[i + 1] - s->xu[i]);
  for (int j = 1; j <= nj - 1; ++j) s->feyv[j] = (s->yv[j + 1] - s->yp[j + 1]) / (s->yv[j + 1] - s->yv[j]);
  for (int k = 0; k < OX_NFIELDS; ++k) s->f[k] = zalloc((size_t)ox_field_size(s, k));
  size_t np = (size_t)(mi + 1) * (nj + 1);
  s->fv = zalloc(np); s->ft = zalloc(np); s->fpc = zalloc(np);
  s->gm = zalloc(np); s->gt = zalloc(np);
  s->AI = zalloc(np); s->AC = zalloc(np); s->AD = zalloc(np); s->Rx = zalloc(np);
  s->BS = zalloc(np); s->BC = zalloc(np); s->BN = zalloc(np); s->Ry = zalloc(np);
  s->sc[0] = zalloc((size_t)mi * (nj + 1)); s->sp[0] = zalloc((size_t)mi * (nj + 1));
  s->sc[1] = zalloc((size_t)(mi + 1) * nj); s->sp[1] = zalloc((size_t)(mi + 1) * nj);
  s->sc[2] = zalloc(np); s->sp[2] = zalloc(np);
  /* side lengths: u sides a/c run over nj+1 nodes, b/d over mi, etc. */
  int lens[3][4] = {{nj + 1, mi, nj + 1, mi}, {nj, mi + 1, nj, mi + 1}, {nj + 1, mi + 1, nj + 1, mi + 1}};
  for (int v = 0; v < 3; ++v)
    for (int sd = 0; sd < 4; ++sd) {
      s->bclen[v][sd] = lens[v][sd];
      s->bct[v][sd] = (int *)calloc((size_t)lens[v][sd] + 2, sizeof(int));
      s->bcv[v][sd] = zalloc((size_t)lens[v][sd] + 2);
    }
  /* a_u = a_v = 1e40 until assembled (source/main.f90:309-310) */
  for (int64_t k = 0; k < ox_field_size(s, OX_AU); ++k) s->f[OX_AU][k] = 1.e40;
  for (int64_t k = 0; k < ox_field_size(s, OX_AV); ++k) s->f[OX_AV][k] = 1.e40;
  s->dt = 1.0; s->rel_pres = s->rel_vel = s->rel_ener = 1.0;
  s->iter_ecuaci_max = s->iter_simple_max = -1;
  s->conv_u = s->conv_t = s->conv_p = s->conv_resi = s->conv_paso = 1e30;
  return s;
}

void ox_destroy(ox_state *s) {
  if (!s) return;
  double *ptrs[] = {s->xu, s->yv, s->xp, s->yp, s->dxp, s->dyp, s->dxu, s->dyu, s->dxv, s->dyv,
                    s->fexp, s->feyp, s->fexu, s->feyv, s->fv, s->ft, s->fpc, s->gm, s->gt,
                    s->AI, s->AC, s->AD, s->Rx, s->BS, s->BC, s->BN, s->Ry};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
  for (int k = 0; k < OX_NFIELDS; ++k) free(s->f[k]);
  for (int v = 0; v < 3; ++v) {
    free(s->sc[v]); free(s->sp[v]);
    for (int sd = 0; sd < 4; ++sd) { free(s->bct[v][sd]); free(s->bcv[v][sd]); }
  }
  free(s);
}

void ox_set_params(ox_state *s, const double *p, int32_t iter_ecuaci_max, int32_t iter_simple_max) {
  s->dt = p[0]; s->rel_pres = p[1]; s->rel_vel = p[2]; s->rel_ener = p[3];
  s->gam_m = p[4]; s->gam_t = p[5]; s->ri_x = p[6]; s->ri_y = p[7];
  s->conv_u = p[8]; s->conv_t = p[9]; s->conv_p = p[10]; s->conv_resi = p[11]; s->conv_paso = p[12];
  s->iter_ecuaci_max = iter_ecuaci_max;
  s->iter_simple_max = iter_simple_max;
  size_t np = (size_t)(s->mi + 1) * (s->nj + 1);
  for (size_t k = 0; k < np; ++k) { s->gm[k] = s->gam_m; s->gt[k] = s->gam_t; }
}

void ox_set_bc(ox_state *s, int32_t var, int32_t side, const int32_t *type, const double *value) {
  int n = s->bclen[var][side];
  for (int k = 1; k <= n; ++k) {
    s->bct[var][side][k] = type[k - 1];
    s->bcv[var][side][k] = value[k - 1];
  }
}

void ox_set_sources(ox_state *s, int32_t var, const double *sc, const double *sp) {
  int64_t n = var == 0 ? (int64_t)s->mi * (s->nj + 1)
                       : var == 1 ? (int64_t)(s->mi + 1) * s->nj : (int64_t)(s->mi + 1) * (s->nj + 1);
  for (int64_t k = 0; k < n; ++k) {
    s->sc[var][k] = sc ? sc[k] : 0.0;
    s->sp[var][k] = sp ? sp[k] : 0.0;
  }
}

void ox_set_field(ox_state *s, int32_t field, const double *src) {
  memcpy(s->f[field], src, sizeof(double) * (size_t)ox_field_size(s, field));
}

void ox_get_field(const ox_state *s, int32_t field, double *dst) {
  memcpy(dst, s->f[field], sizeof(double) * (size_t)ox_field_size(s, field));
}

/* ------------------------------------------------------------ arithmetic */
static inline double pow5(double t) { return t * (t * (t * (t * t))); }
static inline double dmax0(double x) { return x < 0.0 ? 0.0 : x; }

/* Power-law neighbour coefficient (Patankar 1980, Table 5.2) as the upstream
 * assemblies write it, sign included:  -(G*L/d * max(0,(1-0.1|w d/G|)^5)
 * + max(0, +-w L)).  `up` selects the upwind sign: +1 for west/south faces,
 * -1 for east/north faces. */
static inline double nb(double g, double len, double d, double w, int up) {
  double conv = w * len;
  if (up < 0) conv = -conv;
  return -(g * len / d * dmax0(pow5(1.0 - 0.1 * fabs(w * d / g))) + dmax0(conv));
}

/* harmonic interface diffusivity, g1 weighted by (1-f), g2 by f */
static inline double harm(double g1, double g2, double f) {
  return (g1 * g2) / (g1 * (1.0 - f) + g2 * f);
}

/* Thomas algorithm without pivoting (source/mod_solucionador.f90:25-47) */
void ox_tridiagonal(const double *a, double *b, const double *c, double *r, int32_t n) {
  for (int i = 1; i < n; ++i) {
    double m = a[i] / b[i - 1];
    r[i] = r[i] - m * r[i - 1];
    b[i] = b[i] - m * c[i - 1];
  }
  r[n - 1] = r[n - 1] / b[n - 1];
  for (int i = n - 2; i >= 0; --i) r[i] = (r[i] - c[i] * r[i + 1]) / b[i];
}

/* ------------------------------------------------- boundary-row imposition
 * impone_cond_frontera (source/mod_cond_frontera.f90:254-420): Dirichlet
 * rows [0 1 0 | val]; Neumann [0 -1 1 | val] on sides a/b and [-1 1 0 | val]
 * on sides c/d; the momentum coefficient of a boundary node is reset to the
 * 1e40 sentinel.  x-direction rows live in AI..Rx(pos, line); y-direction rows
 * in BS..Ry(pos, line).  `acoef` (NULL for T) is indexed (i,j) of extent am. */
static void bc_rows(ox_state *s, int var, int side, double *A0, double *A1, double *A2, double *R,
                    int ld, int pos, double *acoef, int am, int ydir) {
  int n = s->bclen[var][side];
  const int *t = s->bct[var][side];
  const double *val = s->bcv[var][side];
  int low = (side == 0 || side == 1);
  for (int k = 1; k <= n; ++k) {
    if (t[k] == 0) continue;
    size_t o = (size_t)(pos - 1) + (size_t)(k - 1) * ld;
    if (t[k] == 1) {
      A0[o] = 0.0; A1[o] = 1.0; A2[o] = 0.0;
    } else if (low) {
      A0[o] = 0.0; A1[o] = -1.0; A2[o] = 1.0;
    } else {
      A0[o] = -1.0; A1[o] = 1.0; A2[o] = 0.0;
    }
    R[o] = val[k];
    if (acoef) {
      if (ydir) IX2(acoef, am, k, pos) = 1.e40;
      else IX2(acoef, am, pos, k) = 1.e40;
    }
  }
}

/* ------------------------------------------------------------ u momentum
 * ensambla_velu_x / ensambla_velu_y (source/mod_ec_momento.f90:365-724) */
typedef struct { double ai, ad, bs, bn, ac; } coef_t;

static coef_t u_coefs(ox_state *s, int ii, int jj, double rel) {
  coef_t c;
  double ud = s->fexu[ii] * Uf(ii + 1, jj) + (1.0 - s->fexu[ii]) * Uf(ii, jj);
  double ui = s->fexu[ii - 1] * Uf(ii, jj) + (1.0 - s->fexu[ii - 1]) * Uf(ii - 1, jj);
  double vn = s->fexp[ii] * Vf(ii + 1, jj) + (1.0 - s->fexp[ii]) * Vf(ii, jj);
  double vs = s->fexp[ii] * Vf(ii + 1, jj - 1) + (1.0 - s->fexp[ii]) * Vf(ii, jj - 1);
  double gd = harm(GM(ii + 1, jj + 1), GM(ii + 1, jj), s->feyp[jj]);
  double gi = harm(GM(ii, jj + 1), GM(ii, jj), s->feyp[jj]);
  double gn = gi * gd / (gd * (1.0 - s->fexp[ii]) + gi * s->fexp[ii]);
  gd = harm(GM(ii + 1, jj), GM(ii + 1, jj - 1), s->feyp[jj - 1]);
  gi = harm(GM(ii, jj), GM(ii, jj - 1), s->feyp[jj - 1]);
  double gs = gi * gd / (gd * (1.0 - s->fexp[ii]) + gi * s->fexp[ii]);
  gi = GM(ii, jj);
  gd = GM(ii + 1, jj);
  double di = s->dxp[ii], dd = s->dxp[ii + 1], ds = s->dyv[jj - 1], dn = s->dyv[jj];
  double dx = s->dxu[ii], dy = s->dyu[jj];
  c.ai = nb(gi, dy, di, ui, +1);
  c.ad = nb(gd, dy, dd, ud, -1);
  c.bs = nb(gs, dx, ds, vs, +1);
  c.bn = nb(gn, dx, dn, vn, -1);
  c.ac = (-c.ai - c.ad - c.bs - c.bn - dx * dy * IX2(s->sp[0], MI, ii, jj) + dx * dy / s->dt) / rel;
  return c;
}

/* right-hand side tail after the neighbour terms, in upstream's order */
static inline double u_rhs_tail(ox_state *s, int ii, int jj, double head, double ac, double rel) {
  double dx = s->dxu[ii], dy = s->dyu[jj];
  double tint = s->fexp[ii] * Tf(ii + 1, jj) + (1.0 - s->fexp[ii]) * Tf(ii, jj);
  double r = head - dx * dy * s->ri_x * tint;
  r = r + dx * dy * IX2(s->sc[0], MI, ii, jj);
  r = r + dx * dy * UA(ii, jj) / s->dt;
  r = r + (Pf(ii, jj) - Pf(ii + 1, jj)) * dy;
  r = r + ac * (1.0 - rel) * Uf(ii, jj);
  return r;
}

void ox_momentum_u_y(ox_state *s) {
  const double rel = s->rel_vel;
  /* assembly, jj 2..nj, ii 2..mi-1 (source/main.f90:420-436) */
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI - 1; ++ii) {
      coef_t c = u_coefs(s, ii, jj, rel);
      WBS(jj, ii) = c.bs;
      WBN(jj, ii) = c.bn;
      WBC(jj, ii) = c.ac;
      WRY(jj, ii) = u_rhs_tail(s, ii, jj, -c.ai * Uf(ii - 1, jj) - c.ad * Uf(ii + 1, jj), c.ac, rel);
    }
  /* boundary rows, sides b and d (source/main.f90:445-464); for each line
   * ii the rows sit at positions 1 and nj+1 of BS..Ry(:, ii) */
  for (int side = 1; side <= 3; side += 2) {
    int pos = side == 1 ? 1 : NJ + 1;
    int n = s->bclen[0][side];
    const int *t = s->bct[0][side];
    for (int k = 1; k <= n; ++k) {
      if (!t[k]) continue;
      double *bs = &WBS(pos, k), *bc = &WBC(pos, k), *bn = &WBN(pos, k), *ry = &WRY(pos, k);
      if (t[k] == 1) { *bs = 0.0; *bc = 1.0; *bn = 0.0; }
      else if (side == 1) { *bs = 0.0; *bc = -1.0; *bn = 1.0; }
      else { *bs = -1.0; *bc = 1.0; *bn = 0.0; }
      *ry = s->bcv[0][side][k];
      AUf(k, pos) = 1.e40;
    }
  }
#pragma omp parallel for schedule(static)
  for (int ii = 2; ii <= MI - 1; ++ii)
    ox_tridiagonal(&WBS(1, ii), &WBC(1, ii), &WBN(1, ii), &WRY(1, ii), NJ + 1);
  for (int ii = 2; ii <= MI - 1; ++ii)
    for (int jj = 1; jj <= NJ + 1; ++jj) Uf(ii, jj) = WRY(jj, ii);
}

void ox_momentum_u_x(ox_state *s) {
  const double rel = s->rel_vel;
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI - 1; ++ii) {
      coef_t c = u_coefs(s, ii, jj, rel);
      WAI(ii, jj) = c.ai;
      WAD(ii, jj) = c.ad;
      WAC(ii, jj) = c.ac;
      WRX(ii, jj) = u_rhs_tail(s, ii, jj, -c.bs * Uf(ii, jj - 1) - c.bn * Uf(ii, jj + 1), c.ac, rel);
      AUf(ii, jj) = c.ac * rel;
    }
  /* sides a (position 1) and c (position mi) (source/main.f90:519-535) */
  bc_rows(s, 0, 0, s->AI, s->AC, s->AD, s->Rx, MI + 1, 1, s->f[OX_AU], MI, 0);
  bc_rows(s, 0, 2, s->AI, s->AC, s->AD, s->Rx, MI + 1, MI, s->f[OX_AU], MI, 0);
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    ox_tridiagonal(&WAI(1, jj), &WAC(1, jj), &WAD(1, jj), &WRX(1, jj), MI);
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 1; ii <= MI; ++ii) Uf(ii, jj) = WRX(ii, jj);
}

/* ------------------------------------------------------------ v momentum
 * ensambla_velv_x / ensambla_velv_y (source/mod_ec_momento.f90:736-1098) */
static coef_t v_coefs(ox_state *s, int ii, int jj, double rel) {
  coef_t c;
  double ud = s->feyp[jj] * Uf(ii, jj + 1) + (1.0 - s->feyp[jj]) * Uf(ii, jj);
  double ui = s->feyp[jj] * Uf(ii - 1, jj + 1) + (1.0 - s->feyp[jj]) * Uf(ii - 1, jj);
  double vn = s->feyv[jj] * Vf(ii, jj + 1) + (1.0 - s->feyv[jj]) * Vf(ii, jj);
  double vs = s->feyv[jj - 1] * Vf(ii, jj) + (1.0 - s->feyv[jj - 1]) * Vf(ii, jj - 1);
  double gs = harm(GM(ii + 1, jj), GM(ii, jj), s->fexp[ii]);
  double gn = harm(GM(ii + 1, jj + 1), GM(ii, jj + 1), s->fexp[ii]);
  double gd = (gn * gs) / (gn * (1.0 - s->feyp[jj]) + gs * s->feyp[jj]);
  gs = harm(GM(ii, jj), GM(ii - 1, jj), s->fexp[ii - 1]);
  gn = harm(GM(ii, jj + 1), GM(ii - 1, jj + 1), s->fexp[ii - 1]);
  double gi = (gs * gn) / (gn * (1.0 - s->feyp[jj]) + gs * s->feyp[jj]);
  gn = GM(ii, jj + 1);
  gs = GM(ii, jj);
  double di = s->dxu[ii - 1], dd = s->dxu[ii], ds = s->dyp[jj], dn = s->dyp[jj + 1];
  double dxv = s->dxv[ii], dyv = s->dyv[jj];
  c.ai = nb(gi, dyv, di, ui, +1);
  c.ad = nb(gd, dyv, dd, ud, -1);
  c.bs = nb(gs, dxv, ds, vs, +1);
  c.bn = nb(gn, dxv, dn, vn, -1);
  c.ac = (-c.ai - c.ad - c.bs - c.bn - dxv * dyv * IX2(s->sp[1], MI + 1, ii, jj) + dxv * dyv / s->dt) / rel;
  return c;
}

static inline double v_rhs_tail(ox_state *s, int ii, int jj, double head, double ac, double rel) {
  double dxv = s->dxv[ii], dyv = s->dyv[jj];
  double tint = s->feyp[jj] * Tf(ii, jj + 1) + (1.0 - s->feyp[jj]) * Tf(ii, jj);
  double r = head - dxv * dyv * s->ri_y * tint;
  r = r + dxv * dyv * IX2(s->sc[1], MI + 1, ii, jj);
  r = r + dxv * dyv * VA(ii, jj) / s->dt;
  r = r + (Pf(ii, jj) - Pf(ii, jj + 1)) * dxv;
  r = r + ac * (1.0 - rel) * Vf(ii, jj);
  return r;
}

void ox_momentum_v_y(ox_state *s) {
  const double rel = s->rel_vel;
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ - 1; ++jj)
    for (int ii = 2; ii <= MI; ++ii) {
      coef_t c = v_coefs(s, ii, jj, rel);
      WBS(jj, ii) = c.bs;
      WBN(jj, ii) = c.bn;
      WBC(jj, ii) = c.ac;
      WRY(jj, ii) = v_rhs_tail(s, ii, jj, -c.ai * Vf(ii - 1, jj) - c.ad * Vf(ii + 1, jj), c.ac, rel);
    }
  for (int side = 1; side <= 3; side += 2) {
    int pos = side == 1 ? 1 : NJ;
    int n = s->bclen[1][side];
    const int *t = s->bct[1][side];
    for (int k = 1; k <= n; ++k) {
      if (!t[k]) continue;
      double *bs = &WBS(pos, k), *bc = &WBC(pos, k), *bn = &WBN(pos, k), *ry = &WRY(pos, k);
      if (t[k] == 1) { *bs = 0.0; *bc = 1.0; *bn = 0.0; }
      else if (side == 1) { *bs = 0.0; *bc = -1.0; *bn = 1.0; }
      else { *bs = -1.0; *bc = 1.0; *bn = 0.0; }
      *ry = s->bcv[1][side][k];
      AVf(k, pos) = 1.e40;
    }
  }
#pragma omp parallel for schedule(static)
  for (int ii = 2; ii <= MI; ++ii)
    ox_tridiagonal(&WBS(1, ii), &WBC(1, ii), &WBN(1, ii), &WRY(1, ii), NJ);
  for (int ii = 2; ii <= MI; ++ii)
    for (int jj = 1; jj <= NJ; ++jj) Vf(ii, jj) = WRY(jj, ii);
}

void ox_momentum_v_x(ox_state *s) {
  const double rel = s->rel_vel;
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ - 1; ++jj)
    for (int ii = 2; ii <= MI; ++ii) {
      coef_t c = v_coefs(s, ii, jj, rel);
      WAI(ii, jj) = c.ai;
      WAD(ii, jj) = c.ad;
      WAC(ii, jj) = c.ac;
      WRX(ii, jj) = v_rhs_tail(s, ii, jj, -c.bs * Vf(ii, jj - 1) - c.bn * Vf(ii, jj + 1), c.ac, rel);
      AVf(ii, jj) = c.ac * rel;
    }
  bc_rows(s, 1, 0, s->AI, s->AC, s->AD, s->Rx, MI + 1, 1, s->f[OX_AV], MI + 1, 0);
  bc_rows(s, 1, 2, s->AI, s->AC, s->AD, s->Rx, MI + 1, MI + 1, s->f[OX_AV], MI + 1, 0);
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ - 1; ++jj)
    ox_tridiagonal(&WAI(1, jj), &WAC(1, jj), &WAD(1, jj), &WRX(1, jj), MI + 1);
  for (int jj = 2; jj <= NJ - 1; ++jj)
    for (int ii = 1; ii <= MI + 1; ++ii) Vf(ii, jj) = WRX(ii, jj);
}

/* ------------------------------------------------------ pressure correction
 * ensambla_corr_pres_x / _y (source/mod_ec_momento.f90:145-355), boundary
 * rows and the driver sequence at source/main.f90:752-925 */
void ox_pcorr_zero(ox_state *s) {
  size_t np = (size_t)(MI + 1) * (NJ + 1);
  memset(s->f[OX_PCORR], 0, np * sizeof(double));
  memset(s->fpc, 0, np * sizeof(double));
}

typedef struct { double ai, ad, bs, bn, ac, b; } pcoef_t;

static inline pcoef_t p_coefs(ox_state *s, int ii, int jj) {
  pcoef_t c;
  double ud = Uf(ii, jj), ui = Uf(ii - 1, jj), vn = Vf(ii, jj), vs = Vf(ii, jj - 1);
  double ai = AUf(ii - 1, jj), ad = AUf(ii, jj), as = AVf(ii, jj - 1), an = AVf(ii, jj);
  double dyp = s->dyp[jj], dxp = s->dxp[ii];
  c.ai = -dyp * dyp / ai;
  c.ad = -dyp * dyp / ad;
  c.bs = -dxp * dxp / as;
  c.bn = -dxp * dxp / an;
  c.ac = (-c.ai - c.ad - c.bs - c.bn) / s->rel_pres;
  c.b = (ui - ud) * dyp + (vs - vn) * dxp;
  return c;
}

void ox_pcorr_y(ox_state *s) {
  /* snapshot for the p' error (source/main.f90:762-767) */
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI; ++ii) IX2(s->fpc, MI + 1, ii, jj) = PC(ii, jj);
  const double rel = s->rel_pres;
#pragma omp parallel for schedule(static)
  for (int ii = 2; ii <= MI; ++ii)
    for (int jj = 2; jj <= NJ; ++jj) {
      pcoef_t c = p_coefs(s, ii, jj);
      WBS(jj, ii) = c.bs;
      WBN(jj, ii) = c.bn;
      WBC(jj, ii) = c.ac;
      WRY(jj, ii) = -c.ai * PC(ii - 1, jj) - c.ad * PC(ii + 1, jj) + c.b + (1.0 - rel) * c.ac * PC(ii, jj);
    }
  for (int ii = 2; ii <= MI; ++ii) {
    WBC(1, ii) = 1.0; WBN(1, ii) = 0.0; WRY(1, ii) = 0.0;
    WBC(NJ + 1, ii) = 1.0; WBS(NJ + 1, ii) = 0.0; WRY(NJ + 1, ii) = 0.0;
  }
#pragma omp parallel for schedule(static)
  for (int ii = 2; ii <= MI; ++ii)
    ox_tridiagonal(&WBS(1, ii), &WBC(1, ii), &WBN(1, ii), &WRY(1, ii), NJ + 1);
  for (int ii = 2; ii <= MI; ++ii)
    for (int jj = 1; jj <= NJ + 1; ++jj) PC(ii, jj) = WRY(jj, ii);
}

void ox_pcorr_x(ox_state *s) {
  const double rel = s->rel_pres;
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI; ++ii) {
      pcoef_t c = p_coefs(s, ii, jj);
      WAI(ii, jj) = c.ai;
      WAD(ii, jj) = c.ad;
      WAC(ii, jj) = c.ac;
      BO(ii, jj) = c.b;
      WRX(ii, jj) = -c.bs * PC(ii, jj - 1) - c.bn * PC(ii, jj + 1) + c.b + (1.0 - rel) * c.ac * PC(ii, jj);
    }
  for (int jj = 2; jj <= NJ; ++jj) {
    WAC(1, jj) = 1.0; WAD(1, jj) = 0.0; WRX(1, jj) = 0.0;
    WAI(MI + 1, jj) = 0.0; WAC(MI + 1, jj) = 1.0; WRX(MI + 1, jj) = 0.0;
  }
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    ox_tridiagonal(&WAI(1, jj), &WAC(1, jj), &WAD(1, jj), &WRX(1, jj), MI + 1);
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 1; ii <= MI + 1; ++ii) PC(ii, jj) = WRX(ii, jj);
}

/* pressure and velocity update (source/main.f90:931-962) */
void ox_correct(ox_state *s) {
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI; ++ii) Pf(ii, jj) = Pf(ii, jj) + 0.6 * PC(ii, jj);
  for (int jj = 2; jj <= NJ - 1; ++jj)
    for (int ii = 2; ii <= MI - 1; ++ii) {
      Uf(ii, jj) = Uf(ii, jj) + s->dyu[jj] * (PC(ii, jj) - PC(ii + 1, jj)) / AUf(ii, jj);
      Vf(ii, jj) = Vf(ii, jj) + s->dxv[ii] * (PC(ii, jj) - PC(ii, jj + 1)) / AVf(ii, jj);
    }
  for (int ii = 2; ii <= MI - 1; ++ii)
    Uf(ii, NJ) = Uf(ii, NJ) + s->dyu[NJ] * (PC(ii, NJ) - PC(ii + 1, NJ)) / AUf(ii, NJ);
  for (int jj = 2; jj <= NJ - 1; ++jj)
    Vf(MI, jj) = Vf(MI, jj) + s->dxv[MI] * (PC(MI, jj) - PC(MI, jj + 1)) / AVf(MI, jj);
}

/* ------------------------------------------------------------------ energy
 * ensambla_energia_x / _y (source/mod_ec_energia.f90:119-443).  The driver
 * passes the u/v-mesh factors fexu, feyv as the routine's interpolation
 * weights (source/main.f90:986,1057), so the interface diffusivities are
 * weighted with those (fexu(mi) = feyv(nj) = 0, never assigned). */
static inline coef_t t_coefs(ox_state *s, int ii, int jj) {
  coef_t c;
  double ud = Uf(ii, jj), ui = Uf(ii - 1, jj), vn = Vf(ii, jj), vs = Vf(ii, jj - 1);
  double di = s->dxu[ii - 1], dd = s->dxu[ii], ds = s->dyv[jj - 1], dn = s->dyv[jj];
  const double *wx = s->fexu, *wy = s->feyv;
  double gd = harm(GT(ii + 1, jj), GT(ii, jj), wx[ii]);
  double gi = harm(GT(ii, jj), GT(ii - 1, jj), wx[ii - 1]);
  double gn = harm(GT(ii, jj + 1), GT(ii, jj), wy[jj]);
  double gs = harm(GT(ii, jj), GT(ii, jj - 1), wy[jj - 1]);
  double dxp = s->dxp[ii], dyp = s->dyp[jj];
  c.ai = nb(gi, dyp, di, ui, +1);
  c.ad = nb(gd, dyp, dd, ud, -1);
  c.bs = nb(gs, dxp, ds, vs, +1);
  c.bn = nb(gn, dxp, dn, vn, -1);
  c.ac = (-c.ai - c.ad - c.bs - c.bn - dxp * dyp * IX2(s->sp[2], MI + 1, ii, jj) + dxp * dyp / s->dt) / s->rel_ener;
  return c;
}

static inline double t_rhs_tail(ox_state *s, int ii, int jj, double head, double ac) {
  double dxp = s->dxp[ii], dyp = s->dyp[jj];
  double r = head + dxp * dyp * IX2(s->sc[2], MI + 1, ii, jj);
  r = r + dxp * dyp * TA(ii, jj) / s->dt;
  r = r + ac * (1.0 - s->rel_ener) * Tf(ii, jj);
  return r;
}

void ox_energy_y(ox_state *s) {
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI; ++ii) IX2(s->ft, MI + 1, ii, jj) = Tf(ii, jj);
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI; ++ii) {
      coef_t c = t_coefs(s, ii, jj);
      WBS(jj, ii) = c.bs;
      WBN(jj, ii) = c.bn;
      WBC(jj, ii) = c.ac;
      WRY(jj, ii) = t_rhs_tail(s, ii, jj, -c.ai * Tf(ii - 1, jj) - c.ad * Tf(ii + 1, jj), c.ac);
    }
  for (int side = 1; side <= 3; side += 2) {
    int pos = side == 1 ? 1 : NJ + 1;
    int n = s->bclen[2][side];
    const int *t = s->bct[2][side];
    for (int k = 1; k <= n; ++k) {
      if (!t[k]) continue;
      double *bs = &WBS(pos, k), *bc = &WBC(pos, k), *bn = &WBN(pos, k), *ry = &WRY(pos, k);
      if (t[k] == 1) { *bs = 0.0; *bc = 1.0; *bn = 0.0; }
      else if (side == 1) { *bs = 0.0; *bc = -1.0; *bn = 1.0; }
      else { *bs = -1.0; *bc = 1.0; *bn = 0.0; }
      *ry = s->bcv[2][side][k];
    }
  }
#pragma omp parallel for schedule(static)
  for (int ii = 2; ii <= MI; ++ii)
    ox_tridiagonal(&WBS(1, ii), &WBC(1, ii), &WBN(1, ii), &WRY(1, ii), NJ + 1);
  for (int ii = 2; ii <= MI; ++ii)
    for (int jj = 1; jj <= NJ + 1; ++jj) Tf(ii, jj) = WRY(jj, ii);
}

void ox_energy_x(ox_state *s) {
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI; ++ii) {
      coef_t c = t_coefs(s, ii, jj);
      WAI(ii, jj) = c.ai;
      WAD(ii, jj) = c.ad;
      WAC(ii, jj) = c.ac;
      WRX(ii, jj) = t_rhs_tail(s, ii, jj, -c.bs * Tf(ii, jj - 1) - c.bn * Tf(ii, jj + 1), c.ac);
    }
  bc_rows(s, 2, 0, s->AI, s->AC, s->AD, s->Rx, MI + 1, 1, NULL, 0, 0);
  bc_rows(s, 2, 2, s->AI, s->AC, s->AD, s->Rx, MI + 1, MI + 1, NULL, 0, 0);
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    ox_tridiagonal(&WAI(1, jj), &WAC(1, jj), &WAD(1, jj), &WRX(1, jj), MI + 1);
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 1; ii <= MI + 1; ++ii) Tf(ii, jj) = WRX(ii, jj);
}

/* -------------------------------------------------------------- residuals
 * residuo_u (source/mod_ec_momento.f90:1107-1266) and its norm
 * (source/main.f90:1169-1199).  The south interface diffusivity reproduces
 * upstream's helper exactly, whose numerator pairs Gamma(ii,jj+1) with
 * Gamma(ii,jj) (:1208-1209). */
double ox_residual_u(ox_state *s) {
#pragma omp parallel for schedule(static)
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI - 1; ++ii) {
      double ud = s->fexu[ii] * Uf(ii + 1, jj) + (1.0 - s->fexu[ii]) * Uf(ii, jj);
      double ui = s->fexu[ii - 1] * Uf(ii, jj) + (1.0 - s->fexu[ii - 1]) * Uf(ii - 1, jj);
      double vn = s->fexp[ii] * Vf(ii + 1, jj) + (1.0 - s->fexp[ii]) * Vf(ii, jj);
      double vs = s->fexp[ii] * Vf(ii + 1, jj - 1) + (1.0 - s->fexp[ii]) * Vf(ii, jj - 1);
      double gd = harm(GM(ii + 1, jj + 1), GM(ii + 1, jj), s->feyp[jj]);
      double gi = harm(GM(ii, jj + 1), GM(ii, jj), s->feyp[jj]);
      double gn = gi * gd / (gd * (1.0 - s->fexp[ii]) + gi * s->fexp[ii]);
      gd = harm(GM(ii + 1, jj), GM(ii + 1, jj - 1), s->feyp[jj - 1]);
      gi = (GM(ii, jj + 1) * GM(ii, jj)) /
           (GM(ii, jj) * (1.0 - s->feyp[jj - 1]) + GM(ii, jj - 1) * s->feyp[jj - 1]);
      double gs = gi * gd / (gd * (1.0 - s->fexp[ii]) + gi * s->fexp[ii]);
      gi = GM(ii, jj);
      gd = GM(ii + 1, jj);
      double di = s->dxp[ii], dd = s->dxp[ii + 1], ds = s->dyv[jj - 1], dn = s->dyv[jj];
      double dx = s->dxu[ii], dy = s->dyu[jj];
      double tint = s->fexp[ii] * Tf(ii + 1, jj) + (1.0 - s->fexp[ii]) * Tf(ii, jj);
      /* same power-law terms, positive sign convention */
      double ai = -nb(gi, dy, di, ui, +1);
      double ad = -nb(gd, dy, dd, ud, -1);
      double bs = -nb(gs, dx, ds, vs, +1);
      double bn = -nb(gn, dx, dn, vn, -1);
      double ac = ai + ad + bs + bn - dx * dy * IX2(s->sp[0], MI, ii, jj) + dx * dy / s->dt;
      double r = ai * Uf(ii - 1, jj) - ac * Uf(ii, jj) + ad * Uf(ii + 1, jj) + bs * Uf(ii, jj - 1) +
                 bn * Uf(ii, jj + 1) - dx * dy * s->ri_x * tint;
      r = r + dx * dy * IX2(s->sc[0], MI, ii, jj);
      r = r + dx * dy * UA(ii, jj) / s->dt;
      r = r + (Pf(ii, jj) - Pf(ii + 1, jj)) * dy;
      RESU(ii, jj) = r;
    }
  double acc = 0.0;
  for (int jj = 2; jj <= NJ; ++jj)
    for (int ii = 2; ii <= MI - 1; ++ii) acc = acc + RESU(ii, jj) * RESU(ii, jj);
  return sqrt(acc);
}

/* lagged copies (source/main.f90:1251-1270) */
void ox_lag(ox_state *s) {
  memcpy(s->f[OX_T_ANT], s->f[OX_T], sizeof(double) * (size_t)ox_field_size(s, OX_T));
  memcpy(s->f[OX_U_ANT], s->f[OX_U], sizeof(double) * (size_t)ox_field_size(s, OX_U));
  memcpy(s->f[OX_V_ANT], s->f[OX_V], sizeof(double) * (size_t)ox_field_size(s, OX_V));
}

/* ------------------------------------------------------------------ driver
 * one call = nsteps time steps of source/main.f90:391-1270, with the same
 * inner-loop and SIMPLE-loop exit rules (pass limits max+2, counter reset on
 * every exit) */
static double err_v(ox_state *s) {
  double e = 0.0;
  for (int jj = 2; jj <= NJ - 1; ++jj)
    for (int ii = 2; ii <= MI; ++ii) {
      double d = Vf(ii, jj) - IX2(s->fv, MI + 1, ii, jj);
      e = e + d * d;
    }
  return sqrt(e);
}

int32_t ox_run(ox_state *s, int32_t nsteps, ox_hist *hist, int32_t cap) {
  int32_t nrec = 0;
  for (int step = 0; step < nsteps; ++step) {
    int iter_simple = 0;
    for (;;) {
      ox_hist h;
      memset(&h, 0, sizeof h);
      int it = 0;
      for (;;) { /* momentum */
        memcpy(s->fv, s->f[OX_V], sizeof(double) * (size_t)ox_field_size(s, OX_V));
        ox_momentum_u_y(s);
        ox_momentum_u_x(s);
        ox_momentum_v_y(s);
        ox_momentum_v_x(s);
        h.err_v = err_v(s);
        h.passes_mom++;
        if (h.err_v < s->conv_u) break;
        if (it > s->iter_ecuaci_max) break;
        it++;
      }
      it = 0;
      ox_pcorr_zero(s);
      for (;;) { /* pressure correction */
        ox_pcorr_y(s);
        ox_pcorr_x(s);
        double e = 0.0, b = 0.0;
        for (int jj = 2; jj <= NJ; ++jj)
          for (int ii = 2; ii <= MI; ++ii) {
            double d = PC(ii, jj) - IX2(s->fpc, MI + 1, ii, jj);
            e = e + d * d;
          }
        for (int jj = 2; jj <= NJ; ++jj)
          for (int ii = 2; ii <= MI; ++ii) b = b + BO(ii, jj) * BO(ii, jj);
        h.err_p = sqrt(e);
        h.maxbo = sqrt(b);
        h.passes_p++;
        if (h.err_p < s->conv_p) break;
        if (it > s->iter_ecuaci_max) break;
        it++;
      }
      it = 0;
      ox_correct(s);
      for (;;) { /* energy */
        ox_energy_y(s);
        ox_energy_x(s);
        double e = 0.0;
        for (int jj = 2; jj <= NJ; ++jj)
          for (int ii = 2; ii <= MI; ++ii) {
            double d = Tf(ii, jj) - IX2(s->ft, MI + 1, ii, jj);
            e = e + d * d;
          }
        h.err_t = sqrt(e);
        h.passes_t++;
        if (h.err_t < s->conv_t) break;
        if (it > s->iter_ecuaci_max) break;
        it++;
      }
      h.residuo = ox_residual_u(s);
      int done = 0;
      if (h.maxbo < s->conv_paso && h.residuo < s->conv_resi) {
        iter_simple = 0;
        done = 1;
      } else if (iter_simple > s->iter_simple_max) {
        iter_simple = 0;
        done = 1;
      } else {
        iter_simple++;
      }
      h.simple_iter = iter_simple;
      /* upstream logs to fort.102 on the converged and continue branches only */
      h.logged = done ? (h.maxbo < s->conv_paso && h.residuo < s->conv_resi) : 1;
      if (nrec < cap) hist[nrec] = h;
      nrec++;
      if (done) break;
    }
    ox_lag(s);
  }
  return nrec;
}

/* ----------------------------------------------------------------- Nusselt
 * nusselt_promedio_y (source/mod_postproceso.f90:100-215): dT/dy from the
 * quadratic through the three nodes next to each horizontal wall, evaluated
 * at the outer node, integrated over i0..i1 in panels of two with the exact
 * integral of the interpolating quadratic; derivative entries past i1 are 0
 * as upstream leaves them. */
static double nu_panel_sum(const double *xp, const double *der, int i0, int i1) {
  double acc = 0.0;
  for (int ii = i0; ii <= i1; ii += 2) {
    double x0 = xp[ii], x1 = xp[ii + 1], x2 = xp[ii + 2];
    double y0 = der[ii], y1 = der[ii + 1], y2 = der[ii + 2];
    double al = ((y2 - y1) / (x2 - x1) - (y1 - y0) / (x1 - x0)) / (x2 - x0);
    double be = (y2 - y1) / (x2 - x1) - al * (x1 + x2);
    double ga = y0 - al * x0 * x0 - be * x0;
    acc = acc + al / 3.0 * (x2 * x2 * x2 - x0 * x0 * x0) + be / 2.0 * (x2 * x2 - x0 * x0) + ga * (x2 - x0);
  }
  return acc;
}

void ox_nusselt_y(const ox_state *s, int32_t i0, int32_t i1, double *nu0, double *nu1) {
  double *der = zalloc((size_t)s->mi + 4);
  for (int w = 0; w < 2; ++w) {
    int jj = w == 0 ? 1 : s->nj - 1;
    memset(der, 0, sizeof(double) * ((size_t)s->mi + 4));
    for (int ii = i0; ii <= i1; ++ii) {
      double dy1 = Tf(ii, jj + 1) - Tf(ii, jj);
      double dy2 = Tf(ii, jj + 2) - Tf(ii, jj + 1);
      double dx1 = s->yp[jj + 1] - s->yp[jj];
      double dx2 = s->yp[jj + 2] - s->yp[jj + 1];
      double a = (dy2 / dx2 - dy1 / dx1) / (dx1 + dx2);
      double b = dy2 / dx2 - a * (s->yp[jj + 1] + s->yp[jj + 2]);
      double ye = w == 0 ? s->yp[jj] : s->yp[jj + 2];
      der[ii] = 2.0 * a * ye + b;
    }
    double v = -nu_panel_sum(s->xp, der, i0, i1);
    if (w == 0) *nu0 = v; else *nu1 = v;
  }
  free(der);
}

/* Vertical-wall Nusselt numbers (no upstream counterpart): dT/dx at the wall
 * x = xu(1) (x = xu(mi)) from the quadratic through the three nearest
 * pressure nodes of each row jj, summed over the wall faces jj = 2..nj with
 * weights dyp(jj) and divided by the wall height yv(nj)-yv(1); the sign makes
 * heat leaving the hot wall positive. */
void ox_nusselt_x(const ox_state *s, double *nu_w, double *nu_e) {
  double hgt = s->yv[s->nj] - s->yv[1];
  for (int w = 0; w < 2; ++w) {
    int i1 = w == 0 ? 1 : s->mi - 1;
    double xw = w == 0 ? s->xu[1] : s->xu[s->mi];
    double acc = 0.0;
    for (int jj = 2; jj <= s->nj; ++jj) {
      double x1 = s->xp[i1], x2 = s->xp[i1 + 1], x3 = s->xp[i1 + 2];
      double d1 = (Tf(i1 + 1, jj) - Tf(i1, jj)) / (x2 - x1);
      double d2 = (Tf(i1 + 2, jj) - Tf(i1 + 1, jj)) / (x3 - x2);
      double a = (d2 - d1) / (x3 - x1);
      double g = d1 + a * (2.0 * xw - x1 - x2);
      acc = acc + g * s->dyp[jj];
    }
    double v = -acc / hgt;
    if (w == 0) *nu_w = v; else *nu_e = v;
  }
}

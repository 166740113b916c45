// Device code of the discrete adjoint of one single-pass SIMPLE step (FP64,
// division-free assembly, row layout throughout).
//
// Forward step k (main.f90:391-1270 with one pass per loop), every field of
// it kept on a tape:
//   u1 = Y_u(u0; v0)   u2 = X_u(u1; v0) -> a_u   v1 = Y_v(v0; u2, T0)
//   v2 = X_v(v1; u2, T0) -> a_v   p'1 = Y_p(0; u2, v2, a_u, a_v)
//   p'2 = X_p(p'1; ...)   p += 0.6 p'2, u3, v3 = corrected u2, v2
//   T1 = Y_T(T0; u3, v3)   T2 = X_T(T1; u3, v3)
// Each sweep solves A(w) x = r(phi, w) line by line.  With R = A x - r and
// mu = A^-T xbar on every solved line (k_adj2_solve), the adjoint of every
// input w of the sweep is  wbar += -mu dR/dw : the right-hand-side stencil
// (neighbours, (1-alpha) a_P phi_P, phi_ant, the pressure difference, the
// buoyancy term, b_o) and the coefficients, whose power-law form depends on
// the interpolated face velocities.  After its line's solve, k_adj2_solve
// forms, per row, the adjoint of the four neighbour coefficients (through
// a_P = -sum/alpha) times their derivative with respect to their face
// velocity (adj2_row_out); k_adj2_gather collects everything per node, so each
// node is written by one thread (deterministic, no atomics).
#pragma once
#include "simple_device.hpp"

namespace ixh {

// d nbf / d w (nbf: power law + upwind, division-free form)
__device__ __forceinline__ double nbf_dw(double cond, double len, double pec, double w, bool east) {
  double conv = w * len;
  if (east) conv = -conv;
  const double wp = w * pec;
  const double t = 1.0 - 0.1 * fabs(wp);
  const double t2 = t * t;
  const double pw = t2 * t2 * t;
  double d = 0.0;
  if (pw > 0.0) {
    const double sg = wp > 0.0 ? 1.0 : (wp < 0.0 ? -1.0 : 0.0);
    d = d + cond * len * 5.0 * (t2 * t2) * (-0.1 * pec * sg);
  }
  if (conv > 0.0) d = d + (east ? -len : len);
  return -d;
}

// coefficients of row (ii, jj) of equation EQ (U, V or T), as fast_coef
// computes them, and the derivatives of ai, ad, bs, bn with respect to their
// face velocities ui, ud, vs, vn
struct CoefD {
  double ai, ad, bs, bn, ac, area;
  double dai, dad, dbs, dbn;
};

template <int EQ>
__device__ __forceinline__ CoefD coef_d(const SweepArgs &A, int ii, int jj) {
  constexpr bool TR = false;
  const int P = A.P;
  const Geo &g = A.g;
  const FastGeo &F = A.fg;
  constexpr int e = EQ == EQ_U ? 0 : EQ == EQ_V ? 1 : 2;
  const double *u = A.u, *v = A.v;
  double ui, ud, vs, vn, lx, ly;
  if (EQ == EQ_U) {
    ud = g.fexu[ii] * AT(u, ii + 1, jj) + (1.0 - g.fexu[ii]) * AT(u, ii, jj);
    ui = g.fexu[ii - 1] * AT(u, ii, jj) + (1.0 - g.fexu[ii - 1]) * AT(u, ii - 1, jj);
    vn = g.fexp[ii] * AT(v, ii + 1, jj) + (1.0 - g.fexp[ii]) * AT(v, ii, jj);
    vs = g.fexp[ii] * AT(v, ii + 1, jj - 1) + (1.0 - g.fexp[ii]) * AT(v, ii, jj - 1);
    lx = g.dxu[ii];
    ly = g.dyu[jj];
  } else if (EQ == EQ_V) {
    ud = g.feyp[jj] * AT(u, ii, jj + 1) + (1.0 - g.feyp[jj]) * AT(u, ii, jj);
    ui = g.feyp[jj] * AT(u, ii - 1, jj + 1) + (1.0 - g.feyp[jj]) * AT(u, ii - 1, jj);
    vn = g.feyv[jj] * AT(v, ii, jj + 1) + (1.0 - g.feyv[jj]) * AT(v, ii, jj);
    vs = g.feyv[jj - 1] * AT(v, ii, jj) + (1.0 - g.feyv[jj - 1]) * AT(v, ii, jj - 1);
    lx = g.dxv[ii];
    ly = g.dyv[jj];
  } else {
    ud = AT(u, ii, jj);
    ui = AT(u, ii - 1, jj);
    vn = AT(v, ii, jj);
    vs = AT(v, ii, jj - 1);
    lx = g.dxp[ii];
    ly = g.dyp[jj];
  }
  CoefD c;
  const FaceT<double> fx = face(F.fx[e], ii), fy = face(F.fy[e], jj);
  c.ai = nbf<double>(fx.c_lo, ly, fx.p_lo, ui, false);
  c.ad = nbf<double>(fx.c_hi, ly, fx.p_hi, ud, true);
  c.bs = nbf<double>(fy.c_lo, lx, fy.p_lo, vs, false);
  c.bn = nbf<double>(fy.c_hi, lx, fy.p_hi, vn, true);
  c.dai = nbf_dw(fx.c_lo, ly, fx.p_lo, ui, false);
  c.dad = nbf_dw(fx.c_hi, ly, fx.p_hi, ud, true);
  c.dbs = nbf_dw(fy.c_lo, lx, fy.p_lo, vs, false);
  c.dbn = nbf_dw(fy.c_hi, lx, fy.p_hi, vn, true);
  c.area = lx * ly;
  double acc = -c.ai - c.ad - c.bs - c.bn;
  if (A.sp) acc = acc - c.area * AT(A.sp, ii, jj);
  c.ac = (acc + c.area * F.inv_dt) * (EQ == EQ_T ? F.inv_rel_ener : F.inv_rel_vel);
  return c;
}

// area of the control volume of row (ii, jj) of equation EQ, as coef_d forms it
template <int EQ>
__device__ __forceinline__ double row_area(const Geo &g, int ii, int jj) {
  if (EQ == EQ_U) return g.dxu[ii] * g.dyu[jj];
  if (EQ == EQ_V) return g.dxv[ii] * g.dyv[jj];
  return g.dxp[ii] * g.dyp[jj];
}

// rows a sweep of EQ assembles (both directions): ii in [2, ihi], jj in [2, jhi]
template <int EQ>
__device__ __forceinline__ bool adj_row(int ii, int jj, int mi, int nj) {
  const int ihi = EQ == EQ_U ? mi - 1 : mi;
  const int jhi = EQ == EQ_V ? nj - 1 : nj;
  return ii >= 2 && ii <= ihi && jj >= 2 && jj <= jhi;
}

struct Adj2Args {
  SweepArgs S;          // geometry, physics, BCs and sources of EQ; S.u, S.v = the
                        // coefficient velocities, S.t = T (buoyancy), S.au, S.av (p')
  const double *x;      // the sweep's output
  const double *phi;    // the sweep's input (nullptr: zero, the first p' sweep)
  const double *ybar;   // adjoint of the output
  const double *acb;    // adjoint of a_u / a_v written by the x-sweep (nullptr: none)
  double *mu;           // transposed solve
  double *fa[4];        // per-row adjoints: face velocities (U, V, T) or coefficients (p')
  double *cf[3];        // per row, from the solve: a_P and the two coefficients across
                        // the sweep (bs, bn for x-sweeps, ai, ad for y-sweeps)
  double *phib;         // adjoint of the input: written (=)
  double *velub, *velvb;  // adjoints of the coefficient velocities (+=); the input's own
                          // array when it is one of them
  double *antb;         // adjoint of phi_ant (+=)
  double *pb;           // adjoint of p (+=)
  double *tb;           // adjoint of T through the buoyancy term (+=)
  double *aub, *avb;    // (p') adjoints of a_u, a_v (+=)
  double *g_lo, *g_hi;  // (T) adjoints of the boundary values (+=, per node)
};

template <int EQ, int DIR>
__device__ __forceinline__ int adj_n(int mi, int nj) { return Layout<EQ, DIR>::n(mi, nj); }

// What the row adjoints of row (ii, jj) read besides mu: the row's own x
// and phi, its four neighbour values in the order ai, ad, bs, bn (matrix
// neighbours along the sweep from x, right-hand-side neighbours from phi),
// the adjoint of its a_u / a_v store, and the derivatives of ai, ad, bs, bn
// with respect to their face velocities (U, V, T; p' uses none).
struct AdjRowIn {
  double xP, fP, nb[4], acb, d[4];
};

template <int EQ, int DIR>
__device__ __forceinline__ AdjRowIn adj2_row_in(const Adj2Args &G, int ii, int jj) {
  constexpr bool TR = false;
  const int P = G.S.P;
  const double *x = G.x, *phi = G.phi;
  AdjRowIn r;
  r.xP = AT(x, ii, jj);
  r.fP = phi ? AT(phi, ii, jj) : 0.0;
  if (DIR == DIR_X) {
    r.nb[0] = AT(x, ii - 1, jj); r.nb[1] = AT(x, ii + 1, jj);
    r.nb[2] = phi ? AT(phi, ii, jj - 1) : 0.0; r.nb[3] = phi ? AT(phi, ii, jj + 1) : 0.0;
  } else {
    r.nb[0] = phi ? AT(phi, ii - 1, jj) : 0.0; r.nb[1] = phi ? AT(phi, ii + 1, jj) : 0.0;
    r.nb[2] = AT(x, ii, jj - 1); r.nb[3] = AT(x, ii, jj + 1);
  }
  r.acb = G.acb ? AT(G.acb, ii, jj) : 0.0;
  r.d[0] = r.d[1] = r.d[2] = r.d[3] = 1.0;
  return r;
}

__device__ __forceinline__ void set_derivs(AdjRowIn &r, const CoefD &c) {
  r.d[0] = c.dai; r.d[1] = c.dad; r.d[2] = c.dbs; r.d[3] = c.dbn;
}

// the row adjoints of row (ii, jj) from mu there: the adjoints of the four
// neighbour coefficients, times their face-velocity derivative (U, V, T) or
// as they are (p')
template <int EQ, int DIR>
__device__ __forceinline__ void adj2_row_out(const Adj2Args &G, int ii, int jj, const AdjRowIn &r,
                                             double mu) {
  constexpr bool TR = false;
  const SweepArgs &A = G.S;
  const int P = A.P;
  const double alpha = EQ == EQ_P ? A.ph.rel_pres : EQ == EQ_T ? A.ph.rel_ener : A.ph.rel_vel;
  double jac = -mu * (r.xP - (1.0 - alpha) * r.fP);  // dJ/d a_P
  if (G.acb) jac = jac + alpha * r.acb;
  // a_P = (-sum a_nb ...) / alpha
  const double dsum = EQ == EQ_P ? -jac / alpha
                                 : -jac * (EQ == EQ_T ? A.fg.inv_rel_ener : A.fg.inv_rel_vel);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double jq = -mu * r.nb[q] + dsum;
    AT(G.fa[q], ii, jj) = EQ == EQ_P ? jq : jq * r.d[q];
  }
}

// mu = A^-T ybar on every solved line (lines 2 .. last of the sweep), and
// the row adjoints of the line's assembled rows
template <int EQ, int DIR, int EPT>
__global__ void __launch_bounds__(kBlock) k_adj2_solve(Adj2Args G) {
  __shared__ double sa[EPT * kBlock], sb[EPT * kBlock], sc_[EPT * kBlock], sr[EPT * kBlock];
  constexpr bool TR = false;
  const SweepArgs &A = G.S;
  const int P = A.P;
  const int mi = A.mi, nj = A.nj;
  const int n = Layout<EQ, DIR>::n(mi, nj);
  const int hi = Layout<EQ, DIR>::hi(mi, nj);
  const int line = 2 + (int)blockIdx.x;
  const Side &lo_side = DIR == DIR_X ? A.bc[0] : A.bc[1];
  const Side &hi_side = DIR == DIR_X ? A.bc[2] : A.bc[3];
  // with at most 2 rows per thread the row adjoints' reads go out with the
  // assembly's, before the solve; with more they follow the solve
  constexpr bool kEarly = EPT <= 2;
  AdjRowIn rin[kEarly ? EPT : 1];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = threadIdx.x + e * kBlock;
    if (k >= n) continue;
    const int pos = k + 1;
    const int ii = DIR == DIR_X ? pos : line;
    const int jj = DIR == DIR_X ? line : pos;
    double a0 = 0.0, a1 = 1.0, a2 = 0.0, rr = 0.0;
    if (pos >= 2 && pos <= hi) {
      if constexpr (kEarly) rin[e] = adj2_row_in<EQ, DIR>(G, ii, jj);
      double x0, x1;  // the coefficients across the sweep
      if (EQ == EQ_P) {
        const PCoefT<double> c = p_coef<TR, double>(A, ii, jj);
        a0 = DIR == DIR_X ? c.ai : c.bs;
        a2 = DIR == DIR_X ? c.ad : c.bn;
        a1 = c.ac;
        x0 = DIR == DIR_X ? c.bs : c.ai;
        x1 = DIR == DIR_X ? c.bn : c.ad;
      } else {
        const CoefD c = coef_d<EQ>(A, ii, jj);
        a0 = DIR == DIR_X ? c.ai : c.bs;
        a2 = DIR == DIR_X ? c.ad : c.bn;
        a1 = c.ac;
        x0 = DIR == DIR_X ? c.bs : c.ai;
        x1 = DIR == DIR_X ? c.bn : c.ad;
        if constexpr (kEarly) set_derivs(rin[e], c);
      }
      AT(G.cf[0], ii, jj) = a1;
      AT(G.cf[1], ii, jj) = x0;
      AT(G.cf[2], ii, jj) = x1;
    } else if (EQ != EQ_P) {
      const bool high = pos != 1;
      (void)bc_row(high ? hi_side : lo_side, line, high, a0, a1, a2, rr);
    }
    sa[k] = a0;
    sb[k] = a1;
    sc_[k] = a2;
    sr[k] = AT(G.ybar, ii, jj);
  }
  __syncthreads();
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
      sa[k] = ta[e];
      sc_[k] = tc[e];
    }
  }
  __syncthreads();
  solve_line<SOLVE_PCR, EPT>(sa, sb, sc_, sr, n);
  // mu, and the row adjoints of the line's assembled rows (positions 2 ..
  // hi of lines 2 .. last: the rows the sweep assembles), which read mu at
  // their own row only
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int k = threadIdx.x + e * kBlock;
    if (k >= n) continue;
    const int pos = k + 1;
    const int ii = DIR == DIR_X ? pos : line;
    const int jj = DIR == DIR_X ? line : pos;
    const double mu = sr[k];
    AT(G.mu, ii, jj) = mu;
    if (pos >= 2 && pos <= hi) {
      if constexpr (kEarly) {
        adj2_row_out<EQ, DIR>(G, ii, jj, rin[e], mu);
      } else {
        AdjRowIn r = adj2_row_in<EQ, DIR>(G, ii, jj);
        if (EQ != EQ_P) set_derivs(r, coef_d<EQ == EQ_P ? EQ_T : EQ>(A, ii, jj));
        adj2_row_out<EQ, DIR>(G, ii, jj, r, mu);
      }
    }
  }
  if (EQ == EQ_T && threadIdx.x == 0) {
    G.g_lo[line] = G.g_lo[line] + sr[0];
    G.g_hi[line] = G.g_hi[line] + sr[n - 1];
  }
}

// per node: adjoint of the sweep's input and of everything else the rows read.
// The coefficients come from the solve kernel (cf); every value is read
// without a branch (at the node itself where the row it belongs to does not
// exist, and then not selected), so a node's loads go out together.
template <int EQ, int DIR>
__global__ void __launch_bounds__(kNodeBlock) k_adj2_gather(Adj2Args G) {
  constexpr bool TR = false;
  constexpr int RQ = EQ == EQ_P ? EQ_T : EQ;  // row set of p' equals that of T
  const SweepArgs &A = G.S;
  const int P = A.P;
  const int mi = A.mi, nj = A.nj;
  const int idx = (int)(blockIdx.x * kNodeBlock + threadIdx.x);
  if (idx >= (mi + 1) * (nj + 1)) return;
  const int I = 1 + idx % (mi + 1), J = 1 + idx / (mi + 1);
  const Geo &g = A.g;
  const double *mu = G.mu;
  auto row = [&](int i, int j) { return adj_row<RQ>(i, j, mi, nj); };
  // arr(i, j) where `on`, else arr(I, J) (a valid address, value not used)
  auto at = [&](const double *arr, bool on, int i, int j) { return AT(arr, on ? i : I, on ? j : J); };
  const double alpha = EQ == EQ_P ? A.ph.rel_pres : EQ == EQ_T ? A.ph.rel_ener : A.ph.rel_vel;
  // extent of the swept variable
  const int im = EQ == EQ_U ? mi : mi + 1;
  const int jm = EQ == EQ_V ? nj : nj + 1;
  const bool in_phi = I <= im && J <= jm;
  const bool rP = row(I, J), rW = row(I - 1, J), rE = row(I + 1, J), rS = row(I, J - 1), rN = row(I, J + 1);
  const double muP = AT(mu, I, J);
  // old values of the accumulators this node adds to
  const double ub_old = EQ != EQ_U && G.velub ? AT(G.velub, I, J) : 0.0;
  const double vb_old = EQ != EQ_V && G.velvb ? AT(G.velvb, I, J) : 0.0;
  const double antb_old = EQ != EQ_P && G.antb ? AT(G.antb, I, J) : 0.0;
  const double pb_old = (EQ == EQ_U || EQ == EQ_V) && G.pb ? AT(G.pb, I, J) : 0.0;
  const double tb_old = EQ == EQ_V && G.tb ? AT(G.tb, I, J) : 0.0;

  // ---- input of the sweep
  double pb_in;
  {
    const int line = DIR == DIR_X ? J : I;
    const int last = Layout<EQ, DIR>::last_line(mi, nj);
    const double yb = AT(G.ybar, I, J);
    const double ac = AT(G.cf[0], I, J);
    pb_in = line == 1 || line == last + 1 ? yb : 0.0;  // carried line
    pb_in = rP ? pb_in + (1.0 - alpha) * ac * muP : pb_in;
    if (DIR == DIR_X) {
      const double bs = at(G.cf[1], rN, I, J + 1), muN = at(mu, rN, I, J + 1);
      const double bn = at(G.cf[2], rS, I, J - 1), muS = at(mu, rS, I, J - 1);
      pb_in = rN ? pb_in - bs * muN : pb_in;
      pb_in = rS ? pb_in - bn * muS : pb_in;
    } else {
      const double ai = at(G.cf[1], rE, I + 1, J), muE = at(mu, rE, I + 1, J);
      const double ad = at(G.cf[2], rW, I - 1, J), muW = at(mu, rW, I - 1, J);
      pb_in = rE ? pb_in - ai * muE : pb_in;
      pb_in = rW ? pb_in - ad * muW : pb_in;
    }
    if (!G.phi) pb_in = 0.0;
  }

  // ---- face velocities of the coefficients (U, V, T)
  double gu = 0.0, gv = 0.0;  // adjoints of u(I, J), v(I, J)
  double *const *fa = G.fa;
  if (EQ == EQ_U) {
    const bool rNW = row(I - 1, J + 1);
    const double f0P = AT(fa[0], I, J), f1P = AT(fa[1], I, J), f3P = AT(fa[3], I, J);
    const double f1W = at(fa[1], rW, I - 1, J), f0E = at(fa[0], rE, I + 1, J);
    const double f3W = at(fa[3], rW, I - 1, J), f2N = at(fa[2], rN, I, J + 1);
    const double f2NW = at(fa[2], rNW, I - 1, J + 1);
    gu = rP ? gu + (1.0 - g.fexu[I]) * f1P + g.fexu[I - 1] * f0P : gu;
    gu = rW ? gu + g.fexu[I - 1] * f1W : gu;
    gu = rE ? gu + (1.0 - g.fexu[I]) * f0E : gu;
    gv = rP ? gv + (1.0 - g.fexp[I]) * f3P : gv;
    gv = rW ? gv + g.fexp[I - 1] * f3W : gv;
    gv = rN ? gv + (1.0 - g.fexp[I]) * f2N : gv;
    gv = rNW ? gv + g.fexp[I - 1] * f2NW : gv;
  } else if (EQ == EQ_V) {
    const bool rSE = row(I + 1, J - 1);
    const double f1P = AT(fa[1], I, J), f2P = AT(fa[2], I, J), f3P = AT(fa[3], I, J);
    const double f1S = at(fa[1], rS, I, J - 1), f0E = at(fa[0], rE, I + 1, J);
    const double f0SE = at(fa[0], rSE, I + 1, J - 1);
    const double f3S = at(fa[3], rS, I, J - 1), f2N = at(fa[2], rN, I, J + 1);
    gu = rP ? gu + (1.0 - g.feyp[J]) * f1P : gu;
    gu = rS ? gu + g.feyp[J - 1] * f1S : gu;
    gu = rE ? gu + (1.0 - g.feyp[J]) * f0E : gu;
    gu = rSE ? gu + g.feyp[J - 1] * f0SE : gu;
    gv = rP ? gv + (1.0 - g.feyv[J]) * f3P + g.feyv[J - 1] * f2P : gv;
    gv = rS ? gv + g.feyv[J - 1] * f3S : gv;
    gv = rN ? gv + (1.0 - g.feyv[J]) * f2N : gv;
  } else if (EQ == EQ_T) {
    const double f1P = AT(fa[1], I, J), f3P = AT(fa[3], I, J);
    const double f0E = at(fa[0], rE, I + 1, J), f2N = at(fa[2], rN, I, J + 1);
    gu = rP ? gu + f1P : gu;
    gu = rE ? gu + f0E : gu;
    gv = rP ? gv + f3P : gv;
    gv = rN ? gv + f2N : gv;
  } else {
    // p': b_o = (u_w - u_e) dyp + (v_s - v_n) dxp and the coefficients
    // -dyp^2 / a_u, -dxp^2 / a_v
    const double muE = at(mu, rE, I + 1, J), muN = at(mu, rN, I, J + 1);
    const double f0E = at(fa[0], rE, I + 1, J), f1P = AT(fa[1], I, J);
    const double f2N = at(fa[2], rN, I, J + 1), f3P = AT(fa[3], I, J);
    const double au = AT(A.au, I, J), av = AT(A.av, I, J);
    const double aub = AT(G.aub, I, J), avb = AT(G.avb, I, J);
    gu = rE ? gu + muE * g.dyp[J] : gu;
    gu = rP ? gu - muP * g.dyp[J] : gu;
    gv = rN ? gv + muN * g.dxp[I] : gv;
    gv = rP ? gv - muP * g.dxp[I] : gv;
    if (I <= mi && J <= nj + 1) {
      double s = 0.0;
      s = rP ? s + f1P : s;
      s = rE ? s + f0E : s;
      if (s != 0.0) AT(G.aub, I, J) = aub + s * g.dyp[J] * g.dyp[J] / (au * au);
    }
    if (I <= mi + 1 && J <= nj) {
      double s = 0.0;
      s = rP ? s + f3P : s;
      s = rN ? s + f2N : s;
      if (s != 0.0) AT(G.avb, I, J) = avb + s * g.dxp[I] * g.dxp[I] / (av * av);
    }
  }

  // ---- phi_ant, pressure and buoyancy terms of the momentum and energy rows
  double ant = 0.0, pp = 0.0, tt = 0.0;
  if (EQ != EQ_P) ant = rP ? row_area<RQ>(g, I, J) * A.fg.inv_dt * muP : ant;
  if (EQ == EQ_U) {
    const double muW = at(mu, rW, I - 1, J);
    pp = rP ? pp + muP * g.dyu[J] : pp;
    pp = rW ? pp - muW * g.dyu[J] : pp;
  } else if (EQ == EQ_V) {
    const double muS = at(mu, rS, I, J - 1);
    pp = rP ? pp + muP * g.dxv[I] : pp;
    pp = rS ? pp - muS * g.dxv[I] : pp;
    if (A.ph.ri_y != 0.0) {
      tt = rP ? tt - muP * g.dxv[I] * g.dyv[J] * A.ph.ri_y * (1.0 - g.feyp[J]) : tt;
      tt = rS ? tt - muS * g.dxv[I] * g.dyv[J - 1] * A.ph.ri_y * g.feyp[J - 1] : tt;
    }
  }

  // ---- stores (one thread per node)
  if (in_phi && G.phib) {
    double v = pb_in;
    if (EQ == EQ_U) v = v + gu;
    if (EQ == EQ_V) v = v + gv;
    AT(G.phib, I, J) = v;
  }
  if (EQ != EQ_U && G.velub && I <= mi && J <= nj + 1) AT(G.velub, I, J) = ub_old + gu;
  if (EQ != EQ_V && G.velvb && I <= mi + 1 && J <= nj) AT(G.velvb, I, J) = vb_old + gv;
  if (EQ != EQ_P && G.antb && in_phi) AT(G.antb, I, J) = antb_old + ant;
  if ((EQ == EQ_U || EQ == EQ_V) && G.pb) AT(G.pb, I, J) = pb_old + pp;
  if (EQ == EQ_V && G.tb && tt != 0.0) AT(G.tb, I, J) = tb_old + tt;
}

// adjoint of the correction (main.f90:931-962):  p += 0.6 p',
// u3 = u2 + dyu (p'_P - p'_E) / a_u,  v3 = v2 + dxv (p'_P - p'_N) / a_v
struct AdjCorrArgs {
  int mi, nj, P;
  Geo g;
  const double *pc, *au, *av;       // p'2 and the a_u, a_v the correction divided by
  const double *ub3, *vb3, *pbn;    // adjoints of u3, v3 and of the new p
  double *ub2, *vb2, *pcb;          // written (=)
  double *aub, *avb;                // (+=)
};

__global__ void __launch_bounds__(kNodeBlock) k_adj2_correct(AdjCorrArgs C) {
  constexpr bool TR = false;
  const int P = C.P;
  const int mi = C.mi, nj = C.nj;
  const int idx = (int)(blockIdx.x * kNodeBlock + threadIdx.x);
  if (idx >= (mi + 1) * (nj + 1)) return;
  const int I = 1 + idx % (mi + 1), J = 1 + idx / (mi + 1);
  const Geo &g = C.g;
  auto ucor = [&](int i, int j) { return i >= 2 && i <= mi - 1 && j >= 2 && j <= nj; };
  auto vcor = [&](int i, int j) { return i >= 2 && i <= mi && j >= 2 && j <= nj - 1; };
  // arr(i, j) where `on`, else arr(I, J) (a valid address, value not used):
  // every load of the node goes out without a branch
  auto at = [&](const double *arr, bool on, int i, int j) { return AT(arr, on ? i : I, on ? j : J); };
  const bool uP = ucor(I, J), uW = ucor(I - 1, J), vP = vcor(I, J), vS = vcor(I, J - 1);
  const double ubP = AT(C.ub3, I, J), vbP = AT(C.vb3, I, J);
  const double auP = AT(C.au, I, J), avP = AT(C.av, I, J);
  const double pcP = AT(C.pc, I, J), pcE = at(C.pc, uP, I + 1, J), pcN = at(C.pc, vP, I, J + 1);
  const double aub = AT(C.aub, I, J), avb = AT(C.avb, I, J), pbn = AT(C.pbn, I, J);
  const double ubW = at(C.ub3, uW, I - 1, J), auW = at(C.au, uW, I - 1, J);
  const double vbS = at(C.vb3, vS, I, J - 1), avS = at(C.av, vS, I, J - 1);
  if (I <= mi) {
    AT(C.ub2, I, J) = ubP;
    if (uP) AT(C.aub, I, J) = aub - ubP * g.dyu[J] * (pcP - pcE) / (auP * auP);
  }
  if (J <= nj) {
    AT(C.vb2, I, J) = vbP;
    if (vP) AT(C.avb, I, J) = avb - vbP * g.dxv[I] * (pcP - pcN) / (avP * avP);
  }
  double s = 0.0;
  if (I >= 2 && I <= mi && J >= 2 && J <= nj) s = 0.6 * pbn;
  s = uP ? s + ubP * g.dyu[J] / auP : s;
  s = uW ? s - ubW * g.dyu[J] / auW : s;
  s = vP ? s + vbP * g.dxv[I] / avP : s;
  s = vS ? s - vbS * g.dxv[I] / avS : s;
  AT(C.pcb, I, J) = s;
}

// up to kCopyMax buffer copies of n doubles in one launch (grid.y = copy)
constexpr int kCopyMax = 16;
struct CopyList {
  const double *src[kCopyMax];
  double *dst[kCopyMax];
  int n;
};

__global__ void __launch_bounds__(kBlock) k_copy_many(CopyList L) {
  const double *src = L.src[blockIdx.y];
  double *dst = L.dst[blockIdx.y];
  for (int i = (int)(blockIdx.x * kBlock + threadIdx.x); i < L.n; i += (int)(gridDim.x * kBlock))
    gp(dst)[i] = gp(src)[i];
}

// dst[c] = 0 (src unused) for every listed buffer c, in one launch
__global__ void __launch_bounds__(kBlock) k_zero_many(CopyList L) {
  double *dst = L.dst[blockIdx.y];
  for (int i = (int)(blockIdx.x * kBlock + threadIdx.x); i < L.n; i += (int)(gridDim.x * kBlock))
    gp(dst)[i] = 0.0;
}

// dst[c] += src[c], then src[c] = 0, for every listed pair c, and dst[c] = 0
// where src[c] is null, in one launch (the sources are the caller's work
// arrays)
__global__ void __launch_bounds__(kBlock) k_add_clear_many(CopyList L) {
  double *src = const_cast<double *>(L.src[blockIdx.y]);
  double *dst = L.dst[blockIdx.y];
  for (int i = (int)(blockIdx.x * kBlock + threadIdx.x); i < L.n; i += (int)(gridDim.x * kBlock)) {
    if (src) {
      gp(dst)[i] = gp(dst)[i] + gp(src)[i];
      gp(src)[i] = 0.0;
    } else {
      gp(dst)[i] = 0.0;
    }
  }
}


}  // namespace ixh

// Launchers of the line-sweep kernels (k_sweep, k_multi, k_sweep_batch, ...).
// Each (solver, real type) pair is instantiated in its own translation unit
// (sweeps_*.hip), so the library's many sweep kernels compile in parallel;
// capi.hip sees the declarations only.  Every launcher returns the HIP error
// of its launch.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include "simple_device.hpp"

namespace ixh {

// where a launch goes; start and stop set: the launch records its kernel's
// start and end in them (hipExtLaunchKernelGGL)
struct LaunchCtx {
  hipStream_t stream;
  hipEvent_t start, stop;
};

// rows per thread of a 256-thread line workgroup for lines of n nodes
constexpr int ept_for(int n) {
  const int e = (n + kBlock - 1) / kBlock;
  return e <= 4 ? (e < 1 ? 1 : e) : (e <= 6 ? 6 : 8);
}

// A fused launch of two PCR sweeps whose lines have at most this many nodes
// runs one 64-thread wave per line: at 302-node x-lines the T_x + u_x launch
// takes 17.6 us that way against 22.8 us with 256 threads per line. A single
// sweep of about 700 such lines is faster with 256 threads per line
// (13.4 us against 16.6 us for v_x), so single sweeps keep that shape.
constexpr int kWaveLine = 5 * 64;

// threads per workgroup of a fused launch over lines of at most maxn
// nodes: one wave per line for sweeps only (no residual, no fold) on short
// lines with the wave solver, else 256
inline int multi_block(int solver, int maxn, bool residual, bool fold) {
  return wave_solver(solver) && !residual && !fold && maxn <= kWaveLine ? 64 : kBlock;
}

// one sweep of A.nlines lines of n nodes, one 256-thread workgroup per line
template <int EQ, int DIR, int SOLVER, typename R>
hipError_t launch_sweep_k(const LaunchCtx &c, const SweepArgs &A, int n);

// a fused launch of nwg workgroups (k_multi); maxn: longest line of its sweeps
template <int EQ0, int DIR0, int EQ1, int DIR1, int SOLVER, typename R>
hipError_t launch_multi_k(const LaunchCtx &c, const MultiArgs &M, int nwg, int ept, int maxn);

// the p' y-sweep of a fused step (k_sweep_pz): p' zero at entry, only the
// row layout written
template <int SOLVER, typename R>
hipError_t launch_sweep_pz_k(const LaunchCtx &c, const SweepArgs &A, int n);

// the p' x-sweep of A.nlines lines of n nodes with the correction of those
// lines (k_pcorr_x), one 256-thread workgroup per line
template <int SOLVER, typename R>
hipError_t launch_pcorr_k(const LaunchCtx &c, const SweepArgs &A, const CorrLine &K, int n);

// one sweep phase of a batch: nwg workgroups mapped by wg to (cavity, line);
// wave_lines: one 64-thread wave per line with ept64 rows per lane
template <int EQ, int DIR, int SOLVER, typename R>
hipError_t launch_sweep_batch_k(hipStream_t st, const SweepArgs *D, const int2 *wg, int nwg, int ept,
                                int ept64, bool wave_lines);

// the p' y-sweep phase of a batch whose p' is zero at entry in every cavity
// (k_sweep_pz_batch); shape as launch_sweep_batch_k
template <int SOLVER, typename R>
hipError_t launch_sweep_pz_batch_k(hipStream_t st, const SweepArgs *D, const int2 *wg, int nwg, int ept,
                                   int ept64, bool wave_lines);

// one sweep phase of a batch in groups of kLinesPerGroup consecutive lines
// of a cavity per workgroup, one wave per line (k_sweep_group_batch; the
// wave solver only): ngroups workgroups mapped by wg to (cavity, first line)
constexpr int kLinesPerGroup = 4;
template <int EQ, int DIR, int SOLVER, typename R>
hipError_t launch_sweep_group_batch_k(hipStream_t st, const SweepArgs *D, const int2 *wg, int ngroups, int ept64);

// the T x-sweep phase of a batch (nsweep workgroups of wg) and the u residual
// rows (nres workgroups of wgr) in one launch (k_sweep_t_res_batch); every
// cavity needs Ri = 0 in the u equation
template <int SOLVER, typename R>
hipError_t launch_sweep_t_res_batch_k(hipStream_t st, const SweepArgs *D, const int2 *wg, int nsweep,
                                      const ResArgs *DR, const int2 *wgr, int nres, int ept, int ept64,
                                      bool wave_lines);

#ifdef IXH_SWEEP_LAUNCH_DEFS

template <typename K, typename... Args>
inline void launch_on(const LaunchCtx &c, K kern, dim3 grid, dim3 block, Args... args) {
  if (c.start)
    hipExtLaunchKernelGGL(kern, grid, block, 0, c.stream, c.start, c.stop, 0, args...);
  else
    hipLaunchKernelGGL(kern, grid, block, 0, c.stream, args...);
}

template <int EQ, int DIR, int SOLVER, typename R>
hipError_t launch_sweep_k(const LaunchCtx &c, const SweepArgs &A, int n) {
  const dim3 grid(A.nlines), block(kBlock);
  switch (ept_for(n)) {
    case 1: launch_on(c, k_sweep<EQ, DIR, SOLVER, 1, R>, grid, block, A); break;
    case 2: launch_on(c, k_sweep<EQ, DIR, SOLVER, 2, R>, grid, block, A); break;
    case 3: launch_on(c, k_sweep<EQ, DIR, SOLVER, 3, R>, grid, block, A); break;
    case 4: launch_on(c, k_sweep<EQ, DIR, SOLVER, 4, R>, grid, block, A); break;
    case 6: launch_on(c, k_sweep<EQ, DIR, SOLVER, 6, R>, grid, block, A); break;
    default: launch_on(c, k_sweep<EQ, DIR, SOLVER, 8, R>, grid, block, A); break;
  }
  return hipGetLastError();
}

template <int SOLVER, typename R>
hipError_t launch_sweep_pz_k(const LaunchCtx &c, const SweepArgs &A, int n) {
  const dim3 grid(A.nlines), block(kBlock);
  switch (ept_for(n)) {
    case 1: launch_on(c, k_sweep_pz<SOLVER, 1, R>, grid, block, A); break;
    case 2: launch_on(c, k_sweep_pz<SOLVER, 2, R>, grid, block, A); break;
    case 3: launch_on(c, k_sweep_pz<SOLVER, 3, R>, grid, block, A); break;
    case 4: launch_on(c, k_sweep_pz<SOLVER, 4, R>, grid, block, A); break;
    case 6: launch_on(c, k_sweep_pz<SOLVER, 6, R>, grid, block, A); break;
    default: launch_on(c, k_sweep_pz<SOLVER, 8, R>, grid, block, A); break;
  }
  return hipGetLastError();
}

template <int SOLVER, typename R>
hipError_t launch_pcorr_k(const LaunchCtx &c, const SweepArgs &A, const CorrLine &K, int n) {
  const dim3 grid(A.nlines), block(kBlock);
  switch (ept_for(n)) {
    case 1: launch_on(c, k_pcorr_x<SOLVER, 1, R>, grid, block, A, K); break;
    case 2: launch_on(c, k_pcorr_x<SOLVER, 2, R>, grid, block, A, K); break;
    case 3: launch_on(c, k_pcorr_x<SOLVER, 3, R>, grid, block, A, K); break;
    case 4: launch_on(c, k_pcorr_x<SOLVER, 4, R>, grid, block, A, K); break;
    case 6: launch_on(c, k_pcorr_x<SOLVER, 6, R>, grid, block, A, K); break;
    default: launch_on(c, k_pcorr_x<SOLVER, 8, R>, grid, block, A, K); break;
  }
  return hipGetLastError();
}

template <int EQ0, int DIR0, int EQ1, int DIR1, int SOLVER, typename R>
hipError_t launch_multi_k(const LaunchCtx &c, const MultiArgs &M, int nwg, int ept, int maxn) {
  const dim3 grid(nwg), block(kBlock);
  const int bs = multi_block(SOLVER, maxn, M.cnt[2] > 0, M.cnt[3] > 0);
  // sweeps only, every line short: one wave per line
  if constexpr (wave_solver(SOLVER)) {
    if (bs == 64) {
      const dim3 wave(64);
      switch ((maxn + 63) / 64) {
        case 1: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 1, R, 64>, grid, wave, M); break;
        case 2: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 2, R, 64>, grid, wave, M); break;
        case 3: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 3, R, 64>, grid, wave, M); break;
        case 4: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 4, R, 64>, grid, wave, M); break;
        default: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 5, R, 64>, grid, wave, M); break;
      }
      return hipGetLastError();
    }
  }
  switch (ept) {
    case 1: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 1, R>, grid, block, M); break;
    case 2: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 2, R>, grid, block, M); break;
    case 3: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 3, R>, grid, block, M); break;
    case 4: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 4, R>, grid, block, M); break;
    case 6: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 6, R>, grid, block, M); break;
    default: launch_on(c, k_multi<EQ0, DIR0, EQ1, DIR1, SOLVER, 8, R>, grid, block, M); break;
  }
  return hipGetLastError();
}

// A batch launch of nwg workgroups of kernel K::fn<EPT, BS>: one 64-thread
// wave per line with ept64 rows per lane (wave solver only), else 256
// threads per line with ept rows per thread.
template <class K, int SOLVER, typename... Args>
hipError_t launch_batch_shaped(hipStream_t st, int nwg, int ept, int ept64, bool wave_lines, Args... args) {
  const dim3 grid(nwg), block(kBlock);
  const LaunchCtx c{st, nullptr, nullptr};
  if constexpr (wave_solver(SOLVER)) {
    if (wave_lines) {
      const dim3 wave(64);
      switch (ept64) {
        case 1: launch_on(c, K::template fn<1, 64>(), grid, wave, args...); break;
        case 2: launch_on(c, K::template fn<2, 64>(), grid, wave, args...); break;
        case 3: launch_on(c, K::template fn<3, 64>(), grid, wave, args...); break;
        case 4: launch_on(c, K::template fn<4, 64>(), grid, wave, args...); break;
        default: launch_on(c, K::template fn<5, 64>(), grid, wave, args...); break;
      }
      return hipGetLastError();
    }
  }
  switch (ept) {
    case 1: launch_on(c, K::template fn<1, kBlock>(), grid, block, args...); break;
    case 2: launch_on(c, K::template fn<2, kBlock>(), grid, block, args...); break;
    case 3: launch_on(c, K::template fn<3, kBlock>(), grid, block, args...); break;
    default: launch_on(c, K::template fn<4, kBlock>(), grid, block, args...); break;
  }
  return hipGetLastError();
}

template <int EQ, int DIR, int SOLVER, typename R>
struct SweepBatchK {
  template <int EPT, int BS>
  static auto fn() { return k_sweep_batch<EQ, DIR, SOLVER, EPT, R, BS>; }
};
template <int SOLVER, typename R>
struct SweepPzBatchK {
  template <int EPT, int BS>
  static auto fn() { return k_sweep_pz_batch<SOLVER, EPT, R, BS>; }
};
template <int SOLVER, typename R>
struct SweepTResBatchK {
  template <int EPT, int BS>
  static auto fn() { return k_sweep_t_res_batch<SOLVER, EPT, R, BS>; }
};

template <int EQ, int DIR, int SOLVER, typename R>
hipError_t launch_sweep_batch_k(hipStream_t st, const SweepArgs *D, const int2 *wg, int nwg, int ept,
                                int ept64, bool wave_lines) {
  return launch_batch_shaped<SweepBatchK<EQ, DIR, SOLVER, R>, SOLVER>(st, nwg, ept, ept64, wave_lines, D, wg);
}

template <int EQ, int DIR, int SOLVER, typename R>
hipError_t launch_sweep_group_batch_k(hipStream_t st, const SweepArgs *D, const int2 *wg, int ngroups, int ept64) {
  if constexpr (wave_solver(SOLVER)) {
    constexpr int G = kLinesPerGroup;
    constexpr bool PZ = EQ == EQ_P && DIR == DIR_Y;  // every batch step starts p' from zero
    const dim3 grid(ngroups), block(64 * G);
    const LaunchCtx c{st, nullptr, nullptr};
    switch (ept64) {
      case 1: launch_on(c, k_sweep_group_batch<EQ, DIR, SOLVER, 1, R, G, PZ>, grid, block, D, wg); break;
      case 2: launch_on(c, k_sweep_group_batch<EQ, DIR, SOLVER, 2, R, G, PZ>, grid, block, D, wg); break;
      case 3: launch_on(c, k_sweep_group_batch<EQ, DIR, SOLVER, 3, R, G, PZ>, grid, block, D, wg); break;
      case 4: launch_on(c, k_sweep_group_batch<EQ, DIR, SOLVER, 4, R, G, PZ>, grid, block, D, wg); break;
      default: launch_on(c, k_sweep_group_batch<EQ, DIR, SOLVER, 5, R, G, PZ>, grid, block, D, wg); break;
    }
    return hipGetLastError();
  } else {
    return hipErrorInvalidConfiguration;
  }
}

template <int SOLVER, typename R>
hipError_t launch_sweep_pz_batch_k(hipStream_t st, const SweepArgs *D, const int2 *wg, int nwg, int ept,
                                   int ept64, bool wave_lines) {
  return launch_batch_shaped<SweepPzBatchK<SOLVER, R>, SOLVER>(st, nwg, ept, ept64, wave_lines, D, wg);
}

template <int SOLVER, typename R>
hipError_t launch_sweep_t_res_batch_k(hipStream_t st, const SweepArgs *D, const int2 *wg, int nsweep,
                                      const ResArgs *DR, const int2 *wgr, int nres, int ept, int ept64,
                                      bool wave_lines) {
  return launch_batch_shaped<SweepTResBatchK<SOLVER, R>, SOLVER>(st, nsweep + nres, ept, ept64, wave_lines, D,
                                                                 wg, nsweep, DR, wgr);
}

// every launcher the library calls, for one (solver, real type)
#define IXH_SWEEP_ONE(EQ, DIR, SOLVER, R)                                                        \
  template hipError_t launch_sweep_k<EQ, DIR, SOLVER, R>(const LaunchCtx &, const SweepArgs &, int); \
  template hipError_t launch_sweep_batch_k<EQ, DIR, SOLVER, R>(hipStream_t, const SweepArgs *,      \
                                                               const int2 *, int, int, int, bool); \
  template hipError_t launch_sweep_group_batch_k<EQ, DIR, SOLVER, R>(hipStream_t, const SweepArgs *, \
                                                                     const int2 *, int, int);
#define IXH_MULTI_ONE(EQ0, DIR0, EQ1, DIR1, SOLVER, R)                                  \
  template hipError_t launch_multi_k<EQ0, DIR0, EQ1, DIR1, SOLVER, R>(const LaunchCtx &, \
                                                                      const MultiArgs &, int, int, int);
#define IXH_INSTANTIATE_SWEEPS(SOLVER, R)              \
  template hipError_t launch_pcorr_k<SOLVER, R>(const LaunchCtx &, const SweepArgs &, const CorrLine &, int); \
  template hipError_t launch_sweep_pz_k<SOLVER, R>(const LaunchCtx &, const SweepArgs &, int); \
  template hipError_t launch_sweep_pz_batch_k<SOLVER, R>(hipStream_t, const SweepArgs *, const int2 *, int, int, \
                                                          int, bool);                                         \
  template hipError_t launch_sweep_t_res_batch_k<SOLVER, R>(hipStream_t, const SweepArgs *, const int2 *, int,    \
                                                            const ResArgs *, const int2 *, int, int, int, bool); \
  IXH_SWEEP_ONE(EQ_U, DIR_Y, SOLVER, R)                \
  IXH_SWEEP_ONE(EQ_U, DIR_X, SOLVER, R)                \
  IXH_SWEEP_ONE(EQ_V, DIR_Y, SOLVER, R)                \
  IXH_SWEEP_ONE(EQ_V, DIR_X, SOLVER, R)                \
  IXH_SWEEP_ONE(EQ_P, DIR_Y, SOLVER, R)                \
  IXH_SWEEP_ONE(EQ_P, DIR_X, SOLVER, R)                \
  IXH_SWEEP_ONE(EQ_T, DIR_Y, SOLVER, R)                \
  IXH_SWEEP_ONE(EQ_T, DIR_X, SOLVER, R)                \
  IXH_MULTI_ONE(EQ_V, DIR_Y, -1, DIR_Y, SOLVER, R)     \
  IXH_MULTI_ONE(EQ_V, DIR_X, -1, DIR_X, SOLVER, R)     \
  IXH_MULTI_ONE(EQ_T, DIR_Y, -1, DIR_Y, SOLVER, R)     \
  IXH_MULTI_ONE(EQ_T, DIR_Y, EQ_U, DIR_Y, SOLVER, R)   \
  IXH_MULTI_ONE(EQ_T, DIR_X, -1, DIR_X, SOLVER, R)     \
  IXH_MULTI_ONE(EQ_T, DIR_X, EQ_U, DIR_X, SOLVER, R)

#endif  // IXH_SWEEP_LAUNCH_DEFS

}  // namespace ixh

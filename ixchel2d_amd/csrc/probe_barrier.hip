// Cost of a grid-wide barrier inside one cooperative launch, against the
// cost of a kernel boundary, at the grid shapes a persistent fused time step
// would run (256-thread workgroups, up to the number that fit on the GPU at
// once).  Each phase writes `per` doubles per thread (a fused sweep launch
// at 301x701 writes about 5-15 MB), then either waits at a grid barrier
// (one cooperative launch of `reps` phases) or ends the kernel (`reps`
// launches back to back on one stream).  Prints one JSON line.
//
// The barrier: thread 0 of every workgroup adds 1 to an arrival counter
// (agent scope, acquire-release) and waits until a generation word reaches
// its epoch; the last arrival publishes the epoch.  The wait is bounded: past
// `spin_limit` polls a workgroup sets the timeout flag and every workgroup
// leaves the kernel at its next check, so the launch always drains.  A
// two-level form spreads the arrivals over one counter per group of 32
// workgroups before the root counter.
//
// A second line prices the synchronisation a persistent batch without grid
// barriers would need: every workgroup that finishes a line group publishes
// it to the next phase of its cavity (an agent-scope release, then an atomic
// add on its cavity's counter), and every workgroup that starts one acquires
// what other workgroups published (an agent-scope acquire).  Here every
// workgroup writes its values, then does the arrival with the release
// fence, the acquire fence, both or neither, and the launch time is compared
// across the four forms.  Nothing waits on anything, so every form ends.
//
//   probe_barrier [reps] [per]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x)                                                                                \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) {                                                                     \
      fprintf(stderr, "%s failed: %s (line %d)\n", #x, hipGetErrorString(e_), __LINE__);       \
      exit(1);                                                                                  \
    }                                                                                           \
  } while (0)

constexpr int kThreads = 256;

__device__ __forceinline__ void phase_work(double *sink, int per, int k) {
  const size_t base = ((size_t)blockIdx.x * kThreads + threadIdx.x) * (size_t)per;
  for (int i = 0; i < per; ++i) sink[base + i] = (double)(k + i);
}

__device__ __forceinline__ bool grid_barrier(unsigned *count, unsigned *gen, int *timeout, unsigned nblocks,
                                             unsigned epoch, long spin_limit) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned arrived =
        __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (arrived == epoch * nblocks) {
      __hip_atomic_store(gen, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long spins = 0;
      while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
        if (++spins > spin_limit) {
          __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
  }
  __syncthreads();
  return __hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
}

// Two-level form: workgroups arrive in groups of kGroup on counters of their
// own (one per 256-byte line, count[64 * (1 + group)]); the last arrival of
// a group arrives at the root counter count[0], and the last root arrival
// publishes the epoch.
constexpr int kGroup = 32;

__device__ __forceinline__ bool tree_barrier(unsigned *count, unsigned *gen, int *timeout, unsigned nblocks,
                                             unsigned epoch, long spin_limit) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned g = blockIdx.x / kGroup;
    const unsigned ngroups = (nblocks + kGroup - 1) / kGroup;
    const unsigned in_group = (g + 1) * kGroup <= nblocks ? kGroup : nblocks - g * kGroup;
    const unsigned a = __hip_atomic_fetch_add(count + 64 * (1 + g), 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    bool publish = false;
    if (a == epoch * in_group) {
      const unsigned r = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
      publish = r == epoch * ngroups;
    }
    if (publish) {
      __hip_atomic_store(gen, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      long spins = 0;
      while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
        if (++spins > spin_limit) {
          __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
  }
  __syncthreads();
  return __hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
}

__global__ void __launch_bounds__(kThreads) k_phases(unsigned *count, unsigned *gen, int *timeout, double *sink,
                                                     int per, int reps, long spin_limit, int tree) {
  for (int k = 0; k < reps; ++k) {
    phase_work(sink, per, k);
    const bool ok = tree ? tree_barrier(count, gen, timeout, gridDim.x, (unsigned)k + 1u, spin_limit)
                         : grid_barrier(count, gen, timeout, gridDim.x, (unsigned)k + 1u, spin_limit);
    if (!ok) return;
  }
}

__global__ void __launch_bounds__(kThreads) k_phase(double *sink, int per, int k) { phase_work(sink, per, k); }

// mode bit 0: agent-scope release fence before the arrival, bit 1:
// agent-scope acquire fence after it; the arrival is a relaxed atomic add on
// the counter of "cavity" blockIdx.x % ncav (one 256-byte line each)
constexpr int kCavities = 256;

__global__ void __launch_bounds__(kThreads) k_publish(double *sink, int per, unsigned *counters, int mode) {
  phase_work(sink, per, 0);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (mode & 1) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(counters + 64 * (blockIdx.x % kCavities), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (mode & 2) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// times `launches` launches of k_publish of g workgroups in each mode and
// prints the per-workgroup cost of the fences
static void fence_probe(hipStream_t st, hipEvent_t e0, hipEvent_t e1, int per) {
  const int launches = 20;
  const int grids[] = {4096, 16384, 65536};
  const size_t max_n = (size_t)65536 * kThreads * (size_t)per;
  double *sink = nullptr;
  unsigned *counters = nullptr;
  CHECK(hipMalloc(&sink, max_n * sizeof(double)));
  CHECK(hipMalloc(&counters, 64 * kCavities * sizeof(unsigned)));
  CHECK(hipMemset(counters, 0, 64 * kCavities * sizeof(unsigned)));
  printf("{\"probe\": \"per-workgroup publish: agent-scope release and acquire fences\", "
         "\"launches\": %d, \"doubles_per_thread\": %d, \"cavity_counters\": %d, \"rows\": [",
         launches, per, kCavities);
  bool first = true;
  for (int g : grids) {
    float ms[4];
    for (int mode = 0; mode < 4; ++mode) {
      for (int w = 0; w < 3; ++w)
        hipLaunchKernelGGL(k_publish, dim3(g), dim3(kThreads), 0, st, sink, per, counters, mode);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(e0, st));
      for (int k = 0; k < launches; ++k)
        hipLaunchKernelGGL(k_publish, dim3(g), dim3(kThreads), 0, st, sink, per, counters, mode);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(e1, st));
      CHECK(hipEventSynchronize(e1));
      CHECK(hipEventElapsedTime(&ms[mode], e0, e1));
    }
    const double us[4] = {ms[0] * 1e3 / launches, ms[1] * 1e3 / launches, ms[2] * 1e3 / launches,
                          ms[3] * 1e3 / launches};
    printf("%s{\"workgroups\": %d, \"us_per_launch\": {\"none\": %.2f, \"release\": %.2f, \"acquire\": %.2f, "
           "\"both\": %.2f}, \"ns_per_workgroup_release\": %.2f, \"ns_per_workgroup_acquire\": %.2f, "
           "\"ns_per_workgroup_both\": %.2f}",
           first ? "" : ", ", g, us[0], us[1], us[2], us[3], (us[1] - us[0]) * 1e3 / g, (us[2] - us[0]) * 1e3 / g,
           (us[3] - us[0]) * 1e3 / g);
    first = false;
  }
  printf("]}\n");
  CHECK(hipFree(sink));
  CHECK(hipFree(counters));
}

int main(int argc, char **argv) {
  const int reps = argc > 1 ? atoi(argv[1]) : 2000;
  const int per = argc > 2 ? atoi(argv[2]) : 8;
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  int coop = 0;
  CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, 0));
  int per_cu = 0;
  CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_phases, kThreads, 0));
  const int cus = prop.multiProcessorCount;
  const int max_blocks = per_cu * cus;
  unsigned *count = nullptr, *gen = nullptr;
  int *timeout = nullptr;
  double *sink = nullptr;
  const size_t sink_n = (size_t)max_blocks * kThreads * (size_t)per;
  // the root counter and one 256-byte line per group of the two-level form
  const size_t ncount = 64 * (1 + (size_t)(max_blocks + kGroup - 1) / kGroup);
  CHECK(hipMalloc(&count, ncount * sizeof(unsigned)));
  CHECK(hipMalloc(&gen, sizeof(unsigned)));
  CHECK(hipMalloc(&timeout, sizeof(int)));
  CHECK(hipMalloc(&sink, sink_n * sizeof(double)));
  hipStream_t st;
  CHECK(hipStreamCreate(&st));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  const long spin_limit = 2000000;  // polls of about 64 cycles each: well under a second
  printf("{\"probe\": \"grid barrier vs kernel boundary\", \"cus\": %d, \"blocks_per_cu\": %d, "
         "\"cooperative\": %d, \"reps\": %d, \"doubles_per_thread_per_phase\": %d, \"rows\": [",
         cus, per_cu, coop, reps, per);
  bool first = true;
  for (int g : {256, 512, 768, 1024, 2048}) {
    if (g > max_blocks || !coop) continue;
    const double mb = (double)g * kThreads * per * 8.0 / 1e6;
    // kernel boundaries: reps launches of one phase each
    for (int w = 0; w < 50; ++w) hipLaunchKernelGGL(k_phase, dim3(g), dim3(kThreads), 0, st, sink, per, w);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e0, st));
    for (int k = 0; k < reps; ++k) hipLaunchKernelGGL(k_phase, dim3(g), dim3(kThreads), 0, st, sink, per, k);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(e1, st));
    CHECK(hipEventSynchronize(e1));
    float ms_launch = 0.f;
    CHECK(hipEventElapsedTime(&ms_launch, e0, e1));
    // one cooperative launch of reps phases separated by grid barriers,
    // flat (tree = 0) and two-level (tree = 1)
    float ms_coop[2] = {0.f, 0.f};
    int to = 0;
    for (int tree = 0; tree < 2 && !to; ++tree) {
      for (int trial = 0; trial < 2; ++trial) {
        CHECK(hipMemsetAsync(count, 0, ncount * sizeof(unsigned), st));
        CHECK(hipMemsetAsync(gen, 0, sizeof(unsigned), st));
        CHECK(hipMemsetAsync(timeout, 0, sizeof(int), st));
        int r = reps, tr = tree;
        long lim = spin_limit;
        int pv = per;
        void *args[] = {&count, &gen, &timeout, &sink, &pv, &r, &lim, &tr};
        CHECK(hipEventRecord(e0, st));
        CHECK(hipLaunchCooperativeKernel((const void *)k_phases, dim3(g), dim3(kThreads), args, 0, st));
        CHECK(hipEventRecord(e1, st));
        CHECK(hipEventSynchronize(e1));
        CHECK(hipEventElapsedTime(&ms_coop[tree], e0, e1));
        CHECK(hipMemcpy(&to, timeout, sizeof(int), hipMemcpyDeviceToHost));
        if (to) break;
      }
    }
    printf("%s{\"workgroups\": %d, \"MB_per_phase\": %.2f, \"us_per_kernel_boundary_phase\": %.3f, "
           "\"us_per_barrier_phase\": %.3f, \"us_per_two_level_barrier_phase\": %.3f, \"timeout\": %d}",
           first ? "" : ", ", g, mb, ms_launch * 1e3 / reps, ms_coop[0] * 1e3 / reps, ms_coop[1] * 1e3 / reps, to);
    first = false;
  }
  printf("]}\n");
  fflush(stdout);
  if (per == 1) fence_probe(st, e0, e1, 1);
  CHECK(hipFree(count));
  CHECK(hipFree(gen));
  CHECK(hipFree(timeout));
  CHECK(hipFree(sink));
  CHECK(hipStreamDestroy(st));
  return 0;
}

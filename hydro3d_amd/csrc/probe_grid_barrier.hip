// probe_grid_barrier.hip -- what one phase boundary costs on this GPU: a
// device-scope barrier inside one resident (persistent) kernel against a
// kernel boundary inside a captured graph.
//
// The V-cycle's levels below the finest are latency-bound: each of their
// half-sweeps, restrictions and prolongations is its own launch in the
// captured cycle graph.  A persistent kernel would keep its workgroups
// resident and separate the same phases with grid barriers instead.  This
// program times both at the workgroup counts of those launches (32, 64, 256
// and 512 workgroups of 256 threads: the 16^3 ... 128^3 levels of a 256^3
// block).  Each phase is one read of a value another workgroup wrote in the
// phase before and one write, so every phase boundary has to order real data.
//
// Barriers (one thread per workgroup arrives and waits, the others wait at
// the workgroup barrier):
//   fenced flat  -- agent-scope fences around one arrival counter and a
//                   generation word: the correct barrier (the L2 of each XCD
//                   is written back and invalidated, as at a kernel boundary)
//   unfenced flat, unfenced tree -- the same without the fences (one counter,
//                   or eight per-XCD counters under one), which does NOT
//                   order the data: a lower bound of what any barrier costs
// Every wait is bounded (0.5 s of the constant-rate wall clock): a workgroup
// that times out raises an abort word, and every workgroup leaves at its next
// barrier, so the kernel always ends.  Output: one line per case, in
// microseconds per phase.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CK(x)                                                               \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      std::fprintf(stderr, "%s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                             \
    }                                                                       \
  } while (0)

namespace {

constexpr int THREADS = 256;
constexpr int PHASES = 200;
constexpr int MAX_GROUPS = 512;

// arrival counters on their own 128-byte lines
struct Barrier {
  unsigned part[8][32];  // per-XCD counters (tree)
  unsigned top;          // the counter every arrival (flat) or every part (tree) adds to
  unsigned pad0[31];
  unsigned gen;          // phase generation: bumped by the last arrival
  unsigned pad1[31];
  int abort;
  int pad2[31];
};

__device__ __forceinline__ unsigned ld_agent(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool FENCE, bool TREE>
__device__ bool grid_barrier(Barrier *b, unsigned groups, unsigned long long deadline) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      ok = 0;
    } else {
      if (FENCE) __threadfence();
      // the generation before this arrival: it cannot move before every
      // workgroup, this one included, has arrived
      const unsigned g = ld_agent(&b->gen);
      bool last = false;
      if (TREE) {
        const unsigned x = blockIdx.x & 7u;
        const unsigned members = (groups - x + 7u) / 8u;
        const unsigned parts = groups < 8u ? groups : 8u;
        if (__hip_atomic_fetch_add(&b->part[x][0], 1u, __ATOMIC_ACQ_REL,
                                   __HIP_MEMORY_SCOPE_AGENT) == members - 1u) {
          __hip_atomic_store(&b->part[x][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (__hip_atomic_fetch_add(&b->top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
              parts - 1u) {
            __hip_atomic_store(&b->top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = true;
          }
        }
      } else if (__hip_atomic_fetch_add(&b->top, 1u, __ATOMIC_ACQ_REL,
                                        __HIP_MEMORY_SCOPE_AGENT) == groups - 1u) {
        __hip_atomic_store(&b->top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = true;
      }
      if (last) {
        __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        while (ld_agent(&b->gen) == g) {
          if (__hip_atomic_load(&b->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
              wall_clock64() > deadline) {
            __hip_atomic_store(&b->abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
        }
      }
      if (FENCE) __threadfence();
    }
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// phase ph: read the value workgroup (w + ph) % G wrote in phase ph-1, write
// this workgroup's own slot of the other half of a[2][G][THREADS]
__device__ __forceinline__ void phase_body(float *a, int ph, int groups) {
  const int w = (int)blockIdx.x, t = (int)threadIdx.x;
  const size_t half = (size_t)groups * THREADS;
  const float v = a[(size_t)(ph & 1) * half + (size_t)((w + ph) % groups) * THREADS + t];
  a[(size_t)((ph + 1) & 1) * half + (size_t)w * THREADS + t] = v + 1.0f;
}

template <bool FENCE, bool TREE>
__global__ void __launch_bounds__(THREADS)
    k_resident(Barrier *b, float *a, int phases, unsigned long long span) {
  const unsigned long long deadline = wall_clock64() + span;
  for (int ph = 0; ph < phases; ++ph) {
    phase_body(a, ph, (int)gridDim.x);
    if (!grid_barrier<FENCE, TREE>(b, gridDim.x, deadline)) return;
  }
}

__global__ void __launch_bounds__(THREADS) k_phase(float *a, int ph) {
  phase_body(a, ph, (int)gridDim.x);
}

}  // namespace

int main() {
  int rate_khz = 0;
  CK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0));
  const unsigned long long span = (unsigned long long)rate_khz * 500ull;  // 0.5 s
  Barrier *bar = nullptr;
  float *a = nullptr;
  CK(hipMalloc(&bar, sizeof(Barrier)));
  CK(hipMemset(bar, 0, sizeof(Barrier)));
  CK(hipMalloc(&a, sizeof(float) * 2 * MAX_GROUPS * THREADS));
  CK(hipMemset(a, 0, sizeof(float) * 2 * MAX_GROUPS * THREADS));
  hipStream_t s;
  CK(hipStreamCreate(&s));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  std::printf("device %s, %d CUs, %d phases per case, best of 3 runs\n", prop.gcnArchName,
              prop.multiProcessorCount, PHASES);
  std::printf("%-28s %6s %12s\n", "boundary", "groups", "us/phase");
  const int group_counts[] = {32, 64, 256, 512};
  for (int groups : group_counts) {
    // every workgroup of a resident kernel must be on the GPU at once: far
    // below the device's capacity at these counts (one 256-thread group per CU
    // or fewer than two)
    for (int variant = 0; variant < 3; ++variant) {
      float best = 1e30f;
      for (int rep = 0; rep < 3; ++rep) {
        CK(hipEventRecord(e0, s));
        if (variant == 0)
          k_resident<true, false><<<groups, THREADS, 0, s>>>(bar, a, PHASES, span);
        else if (variant == 1)
          k_resident<false, false><<<groups, THREADS, 0, s>>>(bar, a, PHASES, span);
        else
          k_resident<false, true><<<groups, THREADS, 0, s>>>(bar, a, PHASES, span);
        CK(hipGetLastError());
        CK(hipEventRecord(e1, s));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        int aborted = 0;
        CK(hipMemcpy(&aborted, &bar->abort, sizeof aborted, hipMemcpyDeviceToHost));
        if (aborted) {
          std::fprintf(stderr, "barrier wait timed out (groups %d): stopping\n", groups);
          return 2;
        }
        if (ms < best) best = ms;
      }
      const char *name = variant == 0 ? "resident, fenced flat"
                         : variant == 1 ? "resident, unfenced flat"
                                        : "resident, unfenced tree";
      std::printf("%-28s %6d %12.3f\n", name, groups, best * 1e3f / PHASES);
    }
    // the same phases as one launch each, captured in one graph
    hipGraph_t g;
    hipGraphExec_t ge;
    CK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
    for (int ph = 0; ph < PHASES; ++ph) k_phase<<<groups, THREADS, 0, s>>>(a, ph);
    CK(hipStreamEndCapture(s, &g));
    CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      CK(hipEventRecord(e0, s));
      CK(hipGraphLaunch(ge, s));
      CK(hipEventRecord(e1, s));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) best = ms;
    }
    std::printf("%-28s %6d %12.3f\n", "graph, one launch per phase", groups, best * 1e3f / PHASES);
    CK(hipGraphExecDestroy(ge));
    CK(hipGraphDestroy(g));
  }
  CK(hipFree(a));
  CK(hipFree(bar));
  return 0;
}

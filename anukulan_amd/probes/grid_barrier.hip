// Timing probe behind a design decision (not part of the library): the cost
// of a software grid barrier inside one cooperative kernel, against the
// launch floor of kernels replayed from a HIP graph. A persistent kernel that
// runs a period of PDHG steps needs two grid-wide barriers per step where
// the graph needs two kernel boundaries. All accesses are in bounds; every
// wait is bounded (spin limit sets an error flag and the barrier gives up).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>

#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

__device__ __forceinline__ void grid_barrier(unsigned* arrive, unsigned* gen_ptr, unsigned nblocks, int* err) {
  __syncthreads();
  if (threadIdx.x == 0) {
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      const unsigned gen = __hip_atomic_load(gen_ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned prev = __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (prev == nblocks - 1) {
        __hip_atomic_store(arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(gen_ptr, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        unsigned spins = 0;
        while (__hip_atomic_load(gen_ptr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
          if (++spins > (1u << 22)) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
    }
  }
  __syncthreads();
}

__global__ void barriers(unsigned* arrive, unsigned* gen, int* err, int rounds, double* buf, int work) {
  const unsigned nb = gridDim.x;
  const size_t n = static_cast<size_t>(gridDim.x) * blockDim.x;
  const size_t gid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  double acc = 0.0;
  for (int r = 0; r < rounds; ++r) {
    if (work) buf[gid] = buf[gid] + 1.0;
    grid_barrier(arrive, gen, nb, err);
    if (work) acc += buf[(gid + 4099 * (r + 1)) % n];
    grid_barrier(arrive, gen, nb, err);
  }
  if (work) buf[gid] += acc * 1e-300;
}

__global__ void empty_kernel(double* buf) {
  if (buf != nullptr && threadIdx.x == 1u << 20) buf[0] = 1.0;
}

__global__ void touch_kernel(double* buf, size_t n, int r) {
  const size_t gid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid < n) buf[gid] += buf[(gid + 4099 * (r + 1)) % n] * 1e-300 + 1.0;
}

int main() {
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  printf("CUs %d coop %d\n", prop.multiProcessorCount, prop.cooperativeLaunch);
  int per_cu = 0;
  CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, barriers, 256, 0));
  const int max_blocks = per_cu * prop.multiProcessorCount;
  printf("barrier kernel: %d blocks of 256 per CU, %d co-resident\n", per_cu, max_blocks);
  unsigned *arrive, *gen;
  int* err;
  double* buf;
  const size_t cap = static_cast<size_t>(max_blocks) * 256;
  CHECK(hipMalloc(&arrive, 4));
  CHECK(hipMalloc(&gen, 4));
  CHECK(hipMalloc(&err, 4));
  CHECK(hipMalloc(&buf, cap * sizeof(double)));
  CHECK(hipMemset(buf, 0, cap * sizeof(double)));
  hipStream_t s;
  CHECK(hipStreamCreate(&s));
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  const int rounds = 2000;
  for (int work = 0; work < 2; ++work) {
    for (int g : {64, 256, 512, 1024, 2048, 4096}) {
      if (g > max_blocks) continue;
      CHECK(hipMemset(arrive, 0, 4));
      CHECK(hipMemset(gen, 0, 4));
      CHECK(hipMemset(err, 0, 4));
      int rr = rounds;
      void* args[] = {&arrive, &gen, &err, &rr, &buf, &work};
      CHECK(hipLaunchCooperativeKernel((void*)barriers, dim3(g), dim3(256), args, 0, s));  // warm
      CHECK(hipStreamSynchronize(s));
      CHECK(hipEventRecord(a, s));
      CHECK(hipLaunchCooperativeKernel((void*)barriers, dim3(g), dim3(256), args, 0, s));
      CHECK(hipEventRecord(b, s));
      CHECK(hipEventSynchronize(b));
      float ms = 0;
      CHECK(hipEventElapsedTime(&ms, a, b));
      int herr = 0;
      CHECK(hipMemcpy(&herr, err, 4, hipMemcpyDeviceToHost));
      printf("coop work=%d blocks=%5d: %.3f us per barrier (err %d)\n", work, g, 1e3 * ms / (2.0 * rounds), herr);
      fflush(stdout);
    }
  }
  // Launch floor: a graph of 200 kernels, empty or touching the buffer.
  for (int work = 0; work < 2; ++work) {
    for (int g : {256, 1024, 4096}) {
      if (static_cast<size_t>(g) * 256 > cap) continue;
      hipGraph_t graph;
      hipGraphExec_t exec;
      CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
      for (int k = 0; k < 200; ++k) {
        if (work) touch_kernel<<<g, 256, 0, s>>>(buf, static_cast<size_t>(g) * 256, k);
        else empty_kernel<<<g, 256, 0, s>>>(buf);
      }
      CHECK(hipStreamEndCapture(s, &graph));
      CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      CHECK(hipGraphLaunch(exec, s));
      CHECK(hipStreamSynchronize(s));
      CHECK(hipEventRecord(a, s));
      for (int rep = 0; rep < 10; ++rep) CHECK(hipGraphLaunch(exec, s));
      CHECK(hipEventRecord(b, s));
      CHECK(hipEventSynchronize(b));
      float ms = 0;
      CHECK(hipEventElapsedTime(&ms, a, b));
      printf("graph work=%d blocks=%5d: %.3f us per kernel\n", work, g, 1e3 * ms / 2000.0);
      fflush(stdout);
      CHECK(hipGraphExecDestroy(exec));
      CHECK(hipGraphDestroy(graph));
    }
  }
  return 0;
}

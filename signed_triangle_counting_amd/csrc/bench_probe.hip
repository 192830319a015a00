// Read-bandwidth probe: the HBM rate this GPU gives to a sequential read and
// to reads of random short segments, the access pattern of the pull kernels'
// suffix streams (a few hundred bytes each at scattered offsets).  It gives
// those kernels' measured traffic a ceiling for their own pattern, beside
// the data-sheet peak.  Measurement only: it is built into its own library,
// libstc_bench_probe.so, which only bench.py --hbm-probe loads; no product
// header declares it and libstc.so does not contain it.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__device__ __forceinline__ uint32_t probe_mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Each wave reads `iters` segments of SEG words, each starting at a random
// 32-word (128-byte) line with the whole segment in bounds; 8 segments are in
// flight per wave.  The sum keeps the loads alive.
template <int SEG>
__global__ __launch_bounds__(256) void k_probe_segments(const uint32_t *__restrict__ a,
                                                        uint32_t nlines, uint32_t iters,
                                                        uint32_t *__restrict__ out) {
  constexpr int R = (SEG + 63) / 64;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  uint32_t acc = 0;
  for (uint32_t k0 = 0; k0 < iters; k0 += 8) {
    uint32_t v[8][R];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const uint32_t line =
          probe_mix(wid * 2654435761u + (k0 + u) * 40503u + 12345u) % (nlines - SEG / 32);
      const uint32_t *p = a + (size_t)line * 32u;
#pragma unroll
      for (int j = 0; j < R; j++) {
        const uint32_t w = j * 64u + lane;
        v[u][j] = w < (uint32_t)SEG ? p[w] : 0u;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
      for (int j = 0; j < R; j++) acc += v[u][j];
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_probe_sequential(const uint4 *__restrict__ a, size_t n,
                                                          uint32_t *__restrict__ out) {
  uint32_t acc = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint4 v = a[i];
    acc += v.x ^ v.y ^ v.z ^ v.w;
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

constexpr size_t kProbeWords = (size_t)1 << 30;  // 4 GiB: well past the 256 MB last-level cache
constexpr uint32_t kProbeLines = (uint32_t)(kProbeWords / 32);

// Best of 5 timed launches (after one untimed), in ms.
template <class F>
hipError_t best_ms(hipStream_t st, hipEvent_t e0, hipEvent_t e1, const F &launch, float &best) {
  launch();
  hipError_t e = hipStreamSynchronize(st);
  best = 1e30f;
  for (int r = 0; r < 5 && e == hipSuccess; r++) {
    e = hipEventRecord(e0, st);
    launch();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && ms < best) best = ms;
  }
  if (e == hipSuccess) e = hipGetLastError();
  return e;
}

template <int SEG>
hipError_t probe_segments(const uint32_t *a, uint32_t *out, int grid, hipStream_t st,
                          hipEvent_t e0, hipEvent_t e1, double &tbps) {
  // about 4 GB per launch, a multiple of 8 segments per wave
  const uint64_t waves = (uint64_t)grid * 4u;
  const uint32_t iters =
      (uint32_t)((((4ull << 30) / (waves * SEG * 4ull)) + 7ull) / 8ull * 8ull);
  float ms = 0.f;
  hipError_t e = best_ms(st, e0, e1, [&] {
    hipLaunchKernelGGL(k_probe_segments<SEG>, dim3(grid), dim3(256), 0, st, a, kProbeLines,
                       iters, out);
  }, ms);
  tbps = (double)waves * iters * SEG * 4.0 / (ms * 1e-3) / 1e12;
  return e;
}

}  // namespace

// out_tbps[0]: sequential read of a 4 GiB buffer; out_tbps[1..4]: random
// 128-byte-aligned segments of 128, 256, 512 and 1024 bytes, 8 in flight per
// wave.  Best of 5 launches of about 4 GB each.  Returns 0, -1 (bad
// argument), -3 (out of memory), -4 (HIP error) or -7 (no such device).
extern "C" int bench_probe_read_bandwidth(int device, double out_tbps[5]) {
  if (!out_tbps) return -1;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return -7;
  if (hipSetDevice(device) != hipSuccess) return -4;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return -4;
  const int grid = prop.multiProcessorCount * 8;  // 8 workgroups of 256 threads per CU
  uint32_t *a = nullptr, *out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, kProbeWords * 4);
  if (e == hipSuccess) e = hipMalloc(&out, (size_t)grid * 256 * 4);
  if (e == hipSuccess) e = hipStreamCreate(&st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 1, kProbeWords * 4, st);
  float ms = 0.f;
  if (e == hipSuccess)
    e = best_ms(st, e0, e1, [&] {
      hipLaunchKernelGGL(k_probe_sequential, dim3(grid), dim3(256), 0, st, (const uint4 *)a,
                         kProbeWords / 4, out);
    }, ms);
  if (e == hipSuccess) out_tbps[0] = kProbeWords * 4.0 / (ms * 1e-3) / 1e12;
  if (e == hipSuccess) e = probe_segments<32>(a, out, grid, st, e0, e1, out_tbps[1]);
  if (e == hipSuccess) e = probe_segments<64>(a, out, grid, st, e0, e1, out_tbps[2]);
  if (e == hipSuccess) e = probe_segments<128>(a, out, grid, st, e0, e1, out_tbps[3]);
  if (e == hipSuccess) e = probe_segments<256>(a, out, grid, st, e0, e1, out_tbps[4]);
  const bool oom = e == hipErrorOutOfMemory;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (out) (void)hipFree(out);
  if (a) (void)hipFree(a);
  if (e != hipSuccess) (void)hipGetLastError();
  return e == hipSuccess ? 0 : oom ? -3 : -4;
}

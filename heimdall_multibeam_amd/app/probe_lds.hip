// probe_lds: LDS read bandwidth of the MI355X by access shape.
//
// The dedispersion kernels are limited by LDS reads, and the ceiling that
// matters is the rate of the read instruction they use, not the LDS's
// nominal width.  Each block fills 64 KiB of LDS, then every wave streams
// reads of one shape over it (indices masked into the buffer, so every
// access stays in bounds and is naturally aligned) and sums what it read:
//   b32        one dword per lane, consecutive across the wave
//   read2st64  two dwords per lane 64 apart (the dedispersion inner loop)
//   b64        two consecutive dwords per lane
//   b128       four consecutive dwords per lane
// Prints bytes per clock per CU for each, from HIP event timing of a launch
// of 2 blocks of 512 threads per CU.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kDwords = 16384;  // 64 KiB
constexpr int kThreads = 512;

enum Shape { kB32 = 0, kRead2st64 = 1, kB64 = 2, kB128 = 3 };
constexpr int kBytesPerLane[4] = {4, 8, 8, 16};
const char* const kNames[4] = {"b32", "read2st64", "b64", "b128"};

template <int S>
__global__ __launch_bounds__(kThreads) void k_probe(int iters, unsigned* out) {
  __shared__ __align__(16) uint32_t L[kDwords];
  for (int i = threadIdx.x; i < kDwords; i += kThreads) L[i] = (uint32_t)i * 2654435761u;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t acc = 0;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int base = (it * 8 + u) * 64 * (kBytesPerLane[S] / 4) + wave * 977;
      if (S == kB32) {
        acc += L[(base + lane) & (kDwords - 1)];
      } else if (S == kRead2st64) {
        const int i0 = base & (kDwords - 128);
        acc += L[i0 + lane] ^ L[i0 + 64 + lane];
      } else if (S == kB64) {
        const uint2 v = *reinterpret_cast<const uint2*>(&L[(base + 2 * lane) & (kDwords - 2)]);
        acc += v.x ^ v.y;
      } else {
        const uint4 v = *reinterpret_cast<const uint4*>(&L[(base + 4 * lane) & (kDwords - 4)]);
        acc += (v.x ^ v.y) + (v.z ^ v.w);
      }
    }
  }
  out[blockIdx.x * kThreads + threadIdx.x] = acc;
}

#define CHECK(x)                                                                  \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

}  // namespace

int main() {
  hipDeviceProp_t p;
  CHECK(hipGetDeviceProperties(&p, 0));
  const int ncu = p.multiProcessorCount;
  const double hz = p.clockRate * 1e3;
  const int blocks = 2 * ncu, iters = 4096;
  unsigned* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)blocks * kThreads * sizeof(unsigned)));
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  std::printf("%s: %d CUs, %.0f MHz\n", p.gcnArchName, ncu, hz / 1e6);
  for (int s = 0; s < 4; ++s) {
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
      CHECK(hipEventRecord(a));
      if (s == kB32) hipLaunchKernelGGL(k_probe<kB32>, dim3(blocks), dim3(kThreads), 0, 0, iters, out);
      if (s == kRead2st64)
        hipLaunchKernelGGL(k_probe<kRead2st64>, dim3(blocks), dim3(kThreads), 0, 0, iters, out);
      if (s == kB64) hipLaunchKernelGGL(k_probe<kB64>, dim3(blocks), dim3(kThreads), 0, 0, iters, out);
      if (s == kB128) hipLaunchKernelGGL(k_probe<kB128>, dim3(blocks), dim3(kThreads), 0, 0, iters, out);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(b));
      CHECK(hipEventSynchronize(b));
      CHECK(hipEventElapsedTime(&ms, a, b));
    }
    const double bytes = (double)blocks * kThreads * iters * 8 * kBytesPerLane[s];
    std::printf("%-10s %8.3f ms  %6.1f TB/s  %6.1f B/clk/CU\n", kNames[s], ms,
                bytes / (ms * 1e-3) / 1e12, bytes / (ms * 1e-3) / hz / ncu);
  }
  CHECK(hipFree(out));
  return 0;
}

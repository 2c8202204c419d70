// HBM reference rates of this GPU for the roofline discussion: streaming reads
// with 8, 16 and 32 bytes per lane, a read + write copy, and a random 8-byte
// gather (which moves a whole 128-byte line per element). Each kernel runs on
// buffers far larger than every cache, timed with HIP events (best of five
// after one warm-up). Usage: hbm_probe [GiB per buffer, default 4].
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

// sum of n doubles read W at a time (W = 1, 2 or 4), four independent loads per
// thread in flight; one partial sum per thread is written (negligible bytes)
template <int W>
__global__ __launch_bounds__(256) void k_read(const double* __restrict__ a, size_t n, double* __restrict__ out) {
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t nth = (size_t)gridDim.x * blockDim.x;
  const size_t nv = n / W;
  double s = 0.0;
  size_t i = tid;
  if constexpr (W == 1) {
    for (; i + 3 * nth < nv; i += 4 * nth) s += (a[i] + a[i + nth]) + (a[i + 2 * nth] + a[i + 3 * nth]);
  } else {
    const d2* v = reinterpret_cast<const d2*>(a);
    constexpr int P = W / 2;  // d2 loads per vector
    for (; i + 3 * nth < nv; i += 4 * nth) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int p = 0; p < P; ++p) {
          const d2 x = v[(i + u * nth) * P + p];
          s += x.x + x.y;
        }
    }
  }
  for (; i < nv; i += nth)
    for (int k = 0; k < W; ++k) s += a[i * W + k];
  out[tid] = s;
}

__global__ __launch_bounds__(256) void k_copy(const d2* __restrict__ a, d2* __restrict__ b, size_t nv) {
  const size_t nth = (size_t)gridDim.x * blockDim.x;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nv; i += nth) b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_gather(const double* __restrict__ a, const uint32_t* __restrict__ idx,
                                                size_t m, double* __restrict__ out) {
  const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t nth = (size_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (size_t i = tid; i < m; i += nth) s += a[idx[i]];
  out[tid] = s;
}

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  const double gib = argc > 1 ? std::atof(argv[1]) : 4.0;
  if (!(gib >= 0.25 && gib <= 64.0)) {
    std::fprintf(stderr, "buffer size must be 0.25 .. 64 GiB\n");
    return 2;
  }
  const size_t n = size_t(gib * double(1ull << 27)) & ~size_t(63);  // doubles per buffer
  const size_t m = n / 16;                                            // gathered elements
  const int grid = 8192;
  double *a = nullptr, *b = nullptr, *out = nullptr;
  uint32_t* idx = nullptr;
  if (!ok(hipMalloc(&a, n * 8), "hipMalloc") || !ok(hipMalloc(&b, n * 8), "hipMalloc") ||
      !ok(hipMalloc(&out, sizeof(double) * grid * 256), "hipMalloc") || !ok(hipMalloc(&idx, m * 4), "hipMalloc"))
    return 1;
  if (!ok(hipMemset(a, 0, n * 8), "hipMemset") || !ok(hipMemset(b, 0, n * 8), "hipMemset")) return 1;
  {
    std::vector<uint32_t> h(m);
    uint64_t x = 88172645463325252ull;  // xorshift64
    for (size_t i = 0; i < m; ++i) {
      x ^= x << 13;
      x ^= x >> 7;
      x ^= x << 17;
      h[i] = uint32_t(x % n);
    }
    if (!ok(hipMemcpy(idx, h.data(), m * 4, hipMemcpyHostToDevice), "hipMemcpy")) return 1;
  }
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate")) return 1;
  auto best_ms = [&](auto launch) {
    float best = 1e30f;
    for (int rep = 0; rep < 6; ++rep) {
      (void)hipEventRecord(e0, nullptr);
      launch();
      (void)hipEventRecord(e1, nullptr);
      (void)hipEventSynchronize(e1);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (rep > 0 && ms < best) best = ms;
    }
    return double(best);
  };
  const double bytes = double(n) * 8.0;
  const double r8 = best_ms([&] { hipLaunchKernelGGL(k_read<1>, dim3(grid), dim3(256), 0, nullptr, a, n, out); });
  const double r16 = best_ms([&] { hipLaunchKernelGGL(k_read<2>, dim3(grid), dim3(256), 0, nullptr, a, n, out); });
  const double r32 = best_ms([&] { hipLaunchKernelGGL(k_read<4>, dim3(grid), dim3(256), 0, nullptr, a, n, out); });
  const double cp = best_ms([&] {
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, (const d2*)a, (d2*)b, n / 2);
  });
  const double ga = best_ms([&] { hipLaunchKernelGGL(k_gather, dim3(grid), dim3(256), 0, nullptr, a, idx, m, out); });
  if (!ok(hipDeviceSynchronize(), "kernels")) return 1;
  std::printf("{\"buffer_GiB\": %.3f, \"read8_GBps\": %.0f, \"read16_GBps\": %.0f, \"read32_GBps\": %.0f, "
              "\"copy_GBps\": %.0f, \"gather8_elements_per_s\": %.4g, \"gather8_line_GBps\": %.0f}\n",
              bytes / double(1ull << 30), bytes / (r8 * 1e6), bytes / (r16 * 1e6), bytes / (r32 * 1e6),
              2.0 * bytes / (cp * 1e6), double(m) / (ga * 1e-3), (double(m) * 128.0 + double(m) * 4.0) / (ga * 1e6));
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(out);
  (void)hipFree(idx);
  return 0;
}

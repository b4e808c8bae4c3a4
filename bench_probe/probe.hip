// Matrix-core throughput probe (a benchmark aid, not part of the hot path):
// independent MFMA accumulator chains with no memory traffic, timed with HIP
// events; reports the attainable FP64 and FP32 matrix-core rates of the
// device, against which the FP64 / FP32 product kernels are judged.
#include <fcntl.h>
#include <hip/hip_runtime.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <thread>
#include <vector>

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int kChains = 12;  // independent accumulators per wave

__global__ void __launch_bounds__(256) mfma_f64_probe(int iters, double seed, double* out) {
  d4 acc[kChains];
  for (int i = 0; i < kChains; ++i) acc[i] = d4{0, 0, 0, 0};
  const double a = seed + threadIdx.x * 1e-9, b = 1e-9;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < kChains; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  }
  double s = 0.0;
  for (int i = 0; i < kChains; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) mfma_f32_probe(int iters, float seed, double* out) {
  f4 acc[kChains];
  for (int i = 0; i < kChains; ++i) acc[i] = f4{0, 0, 0, 0};
  const float a = seed + threadIdx.x * 1e-6f, b = 1e-6f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < kChains; ++i) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
  }
  float s = 0.0f;
  for (int i = 0; i < kChains; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

// TFLOP/s of v_mfma_f64_16x16x4f64 and v_mfma_f32_16x16x4f32 (2048 flops
// each) over 2048 workgroups of 4 waves; 0 on success.
extern "C" int rtb_probe_mfma(double* tflops_f64, double* tflops_f32) {
  if (!tflops_f64 || !tflops_f32) return -1;
  const int wg = 2048, iters = 2000;
  const double flops = 2048.0 * kChains * iters * 4.0 * wg;
  bool ok = true;
  auto check = [&](hipError_t e) { ok = ok && e == hipSuccess; };
  double* out = nullptr;
  check(hipMalloc(&out, sizeof(double) * wg * 256));
  hipEvent_t e0{}, e1{};
  check(hipEventCreate(&e0));
  check(hipEventCreate(&e1));
  if (!ok) return -1;
  float ms = 0.0f;
  for (int pass = 0; pass < 2; ++pass) {  // pass 0 warms up
    hipLaunchKernelGGL(mfma_f64_probe, dim3(wg), dim3(256), 0, 0, iters, 1.0, out);
    if (pass == 0) check(hipEventRecord(e0, 0));
  }
  check(hipEventRecord(e1, 0));
  check(hipEventSynchronize(e1));
  check(hipEventElapsedTime(&ms, e0, e1));
  *tflops_f64 = ok && ms > 0.0f ? flops / (ms * 1e-3) / 1e12 : 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(mfma_f32_probe, dim3(wg), dim3(256), 0, 0, iters, 1.0f, out);
    if (pass == 0) check(hipEventRecord(e0, 0));
  }
  check(hipEventRecord(e1, 0));
  check(hipEventSynchronize(e1));
  check(hipEventElapsedTime(&ms, e0, e1));
  *tflops_f32 = ok && ms > 0.0f ? flops / (ms * 1e-3) / 1e12 : 0.0;
  check(hipGetLastError());
  check(hipEventDestroy(e0));
  check(hipEventDestroy(e1));
  check(hipFree(out));
  return ok ? 0 : -1;
}

// Host read ceiling for file ingestion: `threads` host threads scan equal
// parts of the memory-mapped file (newline count, vectorised), as the
// Wannier90 loader's first pass does, with nothing else per byte.  Run after
// the file is in the page cache; *gbps receives bytes / wall time.
extern "C" int rtb_probe_read(const char* path, int threads, double* gbps) {
  *gbps = 0.0;
  const int fd = open(path, O_RDONLY);
  if (fd < 0) return -1;
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size == 0) {
    close(fd);
    return -1;
  }
  const size_t S = static_cast<size_t>(st.st_size);
  void* map = mmap(nullptr, S, PROT_READ, MAP_PRIVATE, fd, 0);
  close(fd);
  if (map == MAP_FAILED) return -1;
  const unsigned char* p = static_cast<const unsigned char*>(map);
  const int T = threads > 0 ? threads : 1;
  std::vector<size_t> count(T, 0);
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      const size_t a = S / T * t, b = t == T - 1 ? S : S / T * (t + 1);
      size_t c = 0;
      for (size_t j = a; j < b; ++j) c += p[j] == '\n';
      count[t] = c;
    });
  for (auto& x : th) x.join();
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  munmap(map, S);
  size_t total = 0;
  for (size_t c : count) total += c;
  *gbps = total > 0 && dt > 0 ? S / dt / 1e9 : 0.0;
  return 0;
}

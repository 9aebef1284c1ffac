// datagen.hip — libdfrgen: seeded, counter-based synthetic columns written
// straight into device memory (the benchmark's inputs, generated on the GPU
// so that 1e8-1e9-row columns never cross PCIe).
//
// Every value is a pure function of (seed, stream, row):
//   base  = mix64(seed ^ (stream * 0xD1B54A32D192ED03))
//   u(r)  = mix64(base + (r + 1) * 0x9E3779B97F4A7C15)        (mod 2^64)
// with mix64 the SplitMix64 finaliser. Integer columns and validity bits use
// only integer arithmetic, so dataframe_mojo_amd/datagen.py reproduces them
// bit for bit on the host. Normal/lognormal values use Box-Muller from
// u(2r), u(2r+1).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t stream_base(uint64_t seed, uint64_t stream) {
  return mix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
}

__device__ __forceinline__ uint64_t draw(uint64_t base, uint64_t r) {
  return mix64(base + (r + 1) * 0x9E3779B97F4A7C15ull);
}

__global__ void k_uniform(void *out, int width, uint64_t n, uint64_t base, int64_t lo, uint64_t span) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t u = draw(base, r);
    uint64_t v = (uint64_t)lo + (span ? u % span : u);
    if (width == 4)
      ((uint32_t *)out)[r] = (uint32_t)v;
    else
      ((uint64_t *)out)[r] = v;
  }
}

// Float64 keys from integer draws: k * scale, with k = 0 as -0.0 (a signed
// zero, which groups with 0.0)
__global__ void k_f64_keys(double *out, const int64_t *in, uint64_t n, double scale) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const int64_t k = in[r];
    out[r] = k == 0 ? -0.0 : (double)k * scale;
  }
}

// value = 1 + first i with (u >> 11) < cdf[i]   (cdf scaled to 2^53)
__global__ void k_zipf(int32_t *out, uint64_t n, uint64_t base, const uint64_t *cdf, uint64_t m,
                       int64_t offset) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t x = draw(base, r) >> 11;
    uint64_t lo = 0, hi = m - 1;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (x < cdf[mid])
        hi = mid;
      else
        lo = mid + 1;
    }
    out[r] = (int32_t)((int64_t)lo + offset);
  }
}

__global__ void k_normal(double *out, uint64_t n, uint64_t base, double mu, double sigma, int lognormal) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x) {
    double u1 = ((double)(draw(base, 2 * r) >> 11) + 1.0) * 0x1p-53;  // (0, 1]
    double u2 = (double)(draw(base, 2 * r + 1) >> 11) * 0x1p-53;      // [0, 1)
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    double v = mu + sigma * z;
    out[r] = lognormal ? exp(v) : v;
  }
}

// word w holds rows [64w, 64w + 64); a row is null when (u >> 11) < threshold
__global__ void k_validity(uint64_t *words, uint64_t n, uint64_t base, uint64_t threshold) {
  uint64_t nw = (n + 63) / 64;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
       w += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t bits = 0;
    for (int b = 0; b < 64; ++b) {
      uint64_t r = w * 64 + b;
      if (r < n && (draw(base, r) >> 11) >= threshold) bits |= 1ull << b;
    }
    words[w] = bits;
  }
}

// Streaming read of n 16-byte words (the achievable HBM read rate): each
// thread folds what it reads into one xor written at the end.
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

__global__ void k_stream_read(const u64x2 *p, uint64_t n, uint64_t *sink) {
  uint64_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    u64x2 v = __builtin_nontemporal_load(p + i);
    acc ^= v.x ^ v.y;
  }
  if (acc == 0x9E3779B97F4A7C15ull) sink[0] = acc;  // keeps the loads live
}

// Streaming copy of n 16-byte words.
__global__ void k_stream_copy(const ulonglong2 *src, ulonglong2 *dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int launch_blocks(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (int)(b < 65536 ? (b ? b : 1) : 65536);
}

int finish() {
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipGetLastError();
  return e == hipSuccess ? 0 : (int)e;
}

}  // namespace

extern "C" {

// Device buffers for generated columns (the benchmark and the tests own
// them; the provider library only wraps them).
int dfrg_alloc(void **out, uint64_t bytes) {
  hipError_t e = hipMalloc(out, bytes ? bytes : 1);
  return e == hipSuccess ? 0 : (int)e;
}

int dfrg_free(void *p) { return (int)hipFree(p); }

int dfrg_copy_to_host(void *dst, const void *src, uint64_t bytes) {
  return (int)hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
}

int dfrg_copy_to_device(void *dst, const void *src, uint64_t bytes) {
  return (int)hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

int dfrg_fill_zero(void *dst, uint64_t bytes) {
  hipError_t e = hipMemset(dst, 0, bytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return (int)e;
}

// Milliseconds of a streaming read (copy when dst != NULL) of `bytes`
// (a multiple of 16), timed with HIP events; blocks = workgroups of 256.
double dfrg_stream(const void *src, void *dst, uint64_t bytes, int blocks, int reps) {
  uint64_t n = bytes / 16;
  uint64_t *sink = nullptr;
  if (hipMalloc(&sink, 8) != hipSuccess) return -1.0;
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  float best = 1e30f;
  for (int r = 0; r < reps; ++r) {
    hipEventRecord(a, 0);
    if (dst)
      hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, 0, (const ulonglong2 *)src, (ulonglong2 *)dst, n);
    else
      hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(256), 0, 0, (const u64x2 *)src, n, sink);
    hipEventRecord(b, 0);
    hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    if (ms < best) best = ms;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(sink);
  return best;
}

// Uniform integers in [lo, hi] (inclusive); width 4 or 8 bytes.
int dfrg_uniform(void *out, int width, uint64_t n, uint64_t seed, uint64_t stream, int64_t lo, int64_t hi) {
  uint64_t span = (uint64_t)hi - (uint64_t)lo + 1;  // 0 means the full 2^64 range
  hipLaunchKernelGGL(k_uniform, dim3(launch_blocks(n)), dim3(256), 0, 0, out, width, n,
                     stream_base(seed, stream), lo, span);
  return finish();
}

// out[r] = in[r] * scale as Float64 (0 as -0.0); out may alias in.
int dfrg_f64_keys(double *out, const int64_t *in, uint64_t n, double scale) {
  hipLaunchKernelGGL(k_f64_keys, dim3(launch_blocks(n)), dim3(256), 0, 0, out, in, n, scale);
  return finish();
}

int dfrg_zipf_i32(int32_t *out, uint64_t n, uint64_t seed, uint64_t stream, const uint64_t *cdf_device,
                  uint64_t m, int64_t offset) {
  hipLaunchKernelGGL(k_zipf, dim3(launch_blocks(n)), dim3(256), 0, 0, out, n, stream_base(seed, stream),
                     cdf_device, m, offset);
  return finish();
}

int dfrg_normal_f64(double *out, uint64_t n, uint64_t seed, uint64_t stream, double mu, double sigma,
                    int lognormal) {
  hipLaunchKernelGGL(k_normal, dim3(launch_blocks(n)), dim3(256), 0, 0, out, n, stream_base(seed, stream),
                     mu, sigma, lognormal);
  return finish();
}

int dfrg_validity(uint64_t *words, uint64_t n, uint64_t seed, uint64_t stream, uint64_t threshold) {
  hipLaunchKernelGGL(k_validity, dim3(launch_blocks((n + 63) / 64)), dim3(256), 0, 0, words, n,
                     stream_base(seed, stream), threshold);
  return finish();
}

}  // extern "C"

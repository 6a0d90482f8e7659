// stream_rate.hip -- achievable HBM streaming rates on this GPU: read-only, write-only and copy over 4 GiB arrays,
// far beyond the 256 MiB Infinity Cache, for 8-byte and 16-byte accesses per lane and several grid sizes
// (workgroups of 256 threads, grid-stride loops). Best of 5 runs, timed with HIP events. Every access is in
// bounds: index i < n for arrays of n elements.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
typedef double d2 __attribute__((ext_vector_type(2)));
#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("%s: %s\n", #x, hipGetErrorString(e)); return 1; } } while (0)

template <class T>
__global__ __launch_bounds__(256) void read_k(const T *__restrict__ a, size_t n, double *out) {
  double s = 0;
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
    const T v = __builtin_nontemporal_load(a + i);
    if constexpr (sizeof(T) == 16) s += v.x + v.y;
    else s += v;
  }
  out[size_t(blockIdx.x) * 256 + threadIdx.x] = s;
}

template <class T>
__global__ __launch_bounds__(256) void write_k(T *__restrict__ a, size_t n, double x) {
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) {
    if constexpr (sizeof(T) == 16) a[i] = T(x);
    else a[i] = x;
  }
}

template <class T>
__global__ __launch_bounds__(256) void copy_k(const T *__restrict__ a, T *__restrict__ b, size_t n) {
  const size_t stride = size_t(gridDim.x) * 256;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += stride) b[i] = __builtin_nontemporal_load(a + i);
}

template <class F>
float best_ms(F launch, hipEvent_t e0, hipEvent_t e1) {
  float best = 1e30f;
  for (int r = 0; r < 5; ++r) {
    hipEventRecord(e0);
    launch();
    hipEventRecord(e1);
    hipEventSynchronize(e1);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) best = ms;
  }
  return best;
}

int main() {
  const size_t bytes = size_t(4) << 30;
  const size_t n8 = bytes / 8, n16 = bytes / 16;
  double *a, *b, *out;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  const int grids[] = {cus * 4, cus * 8, cus * 16, cus * 64};
  CHECK(hipMalloc(&out, size_t(cus) * 64 * 256 * 8));
  CHECK(hipMemset(a, 0, bytes));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  printf("arrays %zu MiB each, %d CUs\n", bytes >> 20, cus);
  for (int G : grids) {
    const float r8 = best_ms([&] { read_k<double><<<G, 256>>>(a, n8, out); }, e0, e1);
    const float r16 = best_ms([&] { read_k<d2><<<G, 256>>>(reinterpret_cast<d2 *>(a), n16, out); }, e0, e1);
    const float w8 = best_ms([&] { write_k<double><<<G, 256>>>(b, n8, 1.0); }, e0, e1);
    const float w16 = best_ms([&] { write_k<d2><<<G, 256>>>(reinterpret_cast<d2 *>(b), n16, 1.0); }, e0, e1);
    const float c8 = best_ms([&] { copy_k<double><<<G, 256>>>(a, b, n8); }, e0, e1);
    const float c16 = best_ms([&] { copy_k<d2><<<G, 256>>>(reinterpret_cast<d2 *>(a), reinterpret_cast<d2 *>(b), n16); }, e0, e1);
    CHECK(hipGetLastError());
    auto gbs = [&](float ms, int k) { return double(bytes) * k / (ms * 1e-3) / 1e9; };
    printf("grid %6d: read %6.0f / %6.0f  write %6.0f / %6.0f  copy %6.0f / %6.0f GB/s (8 B / 16 B per lane)\n", G,
           gbs(r8, 1), gbs(r16, 1), gbs(w8, 1), gbs(w16, 1), gbs(c8, 2), gbs(c16, 2));
  }
  return 0;
}

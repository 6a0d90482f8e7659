// gather_rate.hip -- the rate of random FP64 gathers on this GPU, the quantity that bounds the rank kernel.
//
// 128M gathers t[idx[i]] with idx uniform in [0, K), for table sizes K from 128 KiB (L2 resident) to 512 MiB
// (beyond the Infinity Cache); each lane keeps 1, 4 or 8 gathers in flight (U1/U4/U8). "sliced-U4" regroups
// the same indices so that workgroup b only reads table lines with (line index mod 8) == (b mod 8), the
// placement that keeps each XCD's L2 on one eighth of the table. Also the streaming rate of the index array.
// idx[i] < K always and out has G*256 entries: every access is in bounds.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <vector>
#include <random>
#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("%s: %s\n", #x, hipGetErrorString(e)); return 1; } } while (0)

template <int U>
__global__ __launch_bounds__(256) void gather(const double* __restrict__ t, const uint32_t* __restrict__ idx, size_t n,
                                              double* out) {
  double s[U] = {};
  const size_t stride = size_t(gridDim.x) * 256;
  size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    uint32_t u[U];
#pragma unroll
    for (int k = 0; k < U; ++k) u[k] = __builtin_nontemporal_load(idx + i + k * stride);
#pragma unroll
    for (int k = 0; k < U; ++k) s[k] += t[u[k]];
  }
  for (; i < n; i += stride) s[0] += t[__builtin_nontemporal_load(idx + i)];
  double r = 0;
#pragma unroll
  for (int k = 0; k < U; ++k) r += s[k];
  out[size_t(blockIdx.x) * 256 + threadIdx.x] = r;
}

// sliced: block b reads only idx_s[b % 8][...] (indices whose 128B line % 8 == b % 8), n_per entries each
template <int U>
__global__ __launch_bounds__(256) void gather_sliced(const double* __restrict__ t, const uint32_t* __restrict__ idx,
                                                     size_t n_per, double* out) {
  const int x = blockIdx.x & 7;
  const uint32_t* my = idx + size_t(x) * n_per;
  double s[U] = {};
  const size_t stride = size_t(gridDim.x / 8) * 256;
  size_t i = size_t(blockIdx.x / 8) * 256 + threadIdx.x;
  for (; i + (U - 1) * stride < n_per; i += U * stride) {
    uint32_t u[U];
#pragma unroll
    for (int k = 0; k < U; ++k) u[k] = __builtin_nontemporal_load(my + i + k * stride);
#pragma unroll
    for (int k = 0; k < U; ++k) s[k] += t[u[k]];
  }
  for (; i < n_per; i += stride) s[0] += t[__builtin_nontemporal_load(my + i)];
  double r = 0;
#pragma unroll
  for (int k = 0; k < U; ++k) r += s[k];
  out[size_t(blockIdx.x) * 256 + threadIdx.x] = r;
}

__global__ __launch_bounds__(256) void stream_idx(const uint32_t* __restrict__ idx, size_t n, double* out) {
  double s = 0;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256)
    s += __builtin_nontemporal_load(idx + i);
  out[size_t(blockIdx.x) * 256 + threadIdx.x] = s;
}

int main() {
  const size_t M = size_t(1) << 27;  // 128M gathers
  const size_t KMAX = size_t(1) << 26;
  const int G = 8192;
  uint32_t *idx, *idxs;
  double *t, *out;
  CHECK(hipMalloc(&idx, M * 4));
  CHECK(hipMalloc(&idxs, M * 4));
  CHECK(hipMalloc(&t, KMAX * 8));
  CHECK(hipMemset(t, 0, KMAX * 8));
  CHECK(hipMalloc(&out, size_t(G) * 256 * 8));
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  std::vector<uint32_t> h(M), hs(M);
  std::mt19937_64 rng(1);
  auto timeit = [&](auto launch) -> float {
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
      hipEventRecord(a);
      launch();
      hipEventRecord(b);
      hipEventSynchronize(b);
      float ms;
      hipEventElapsedTime(&ms, a, b);
      best = ms < best ? ms : best;
    }
    return best;
  };
  float ms = timeit([&] { stream_idx<<<G, 256>>>(idx, M, out); });
  printf("stream idx only: %.3f ms, %.1f GB/s\n", ms, M * 4 / ms / 1e6);
  for (size_t K : {size_t(1) << 14, size_t(1) << 17, size_t(1) << 18, size_t(1) << 19, size_t(1) << 22,
                   size_t(1) << 24, size_t(1) << 26}) {
    for (size_t i = 0; i < M; ++i) h[i] = uint32_t(rng() % K);
    // sliced copy: group indices by line % 8, each group padded by repeating its own entries to M/8
    std::vector<std::vector<uint32_t>> grp(8);
    for (size_t i = 0; i < M; ++i) grp[(h[i] >> 4) & 7].push_back(h[i]);
    const size_t per = M / 8;
    for (int x = 0; x < 8; ++x)
      for (size_t i = 0; i < per; ++i) hs[size_t(x) * per + i] = grp[x].empty() ? 0 : grp[x][i % grp[x].size()];
    CHECK(hipMemcpy(idx, h.data(), M * 4, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(idxs, hs.data(), M * 4, hipMemcpyHostToDevice));
    float m1 = timeit([&] { gather<1><<<G, 256>>>(t, idx, M, out); });
    float m4 = timeit([&] { gather<4><<<G, 256>>>(t, idx, M, out); });
    float m8 = timeit([&] { gather<8><<<G, 256>>>(t, idx, M, out); });
    float s4 = timeit([&] { gather_sliced<4><<<G, 256>>>(t, idxs, per, out); });
    printf("table %8zu doubles (%7.2f MiB): U1 %6.1f  U4 %6.1f  U8 %6.1f  sliced-U4 %6.1f  G gathers/s\n", K,
           K * 8 / 1048576.0, M / m1 / 1e6, M / m4 / 1e6, M / m8 / 1e6, M / s4 / 1e6);
    fflush(stdout);
  }
  return 0;
}

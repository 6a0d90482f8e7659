// xcc_map.hip -- which XCD runs workgroup b? Each workgroup's thread 0 writes its XCC_ID hardware register to
// out[b] (b < G, in bounds); the host reports how often XCD == b mod 8, for small and large grids, with and
// without enough work per workgroup for workgroups to overlap. The sliced rank plan relies on this placement.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
__global__ void xcc_kernel(unsigned *out, int spin) {
  unsigned id;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(id));
  if (threadIdx.x == 0) out[blockIdx.x] = id & 0xF;
  // some work so that workgroups overlap
  volatile float x = threadIdx.x;
  for (int i = 0; i < spin; ++i) x = x * 1.0001f + 0.5f;
}
int main() {
  const int sizes[] = {64, 4096, 131072};
  for (int G : sizes) {
    unsigned *d;
    hipMalloc(&d, G * sizeof(unsigned));
    hipMemset(d, 0xFF, G * sizeof(unsigned));
    for (int spin : {0, 2000}) {
      xcc_kernel<<<G, 256>>>(d, spin);
      if (hipDeviceSynchronize() != hipSuccess) { printf("error\n"); return 1; }
      std::vector<unsigned> h(G);
      hipMemcpy(h.data(), d, G * sizeof(unsigned), hipMemcpyDeviceToHost);
      int match = 0, hist[16] = {0};
      for (int b = 0; b < G; ++b) { match += (h[b] == unsigned(b % 8)); if (h[b] < 16) hist[h[b]]++; }
      printf("G=%d spin=%d: xcc == b%%8 for %d of %d; first 16:", G, spin, match, G);
      for (int b = 0; b < 16; ++b) printf(" %u", h[b]);
      printf(" | per-xcc:");
      for (int x = 0; x < 8; ++x) printf(" %d", hist[x]);
      printf("\n");
    }
    hipFree(d);
  }
  return 0;
}

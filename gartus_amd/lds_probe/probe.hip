// LDS access-pattern probe for gfx950. Each dispatch runs one list of
// per-lane LDS addresses; rocprofv3's SQ_LDS_BANK_CONFLICT and
// SQ_LDS_IDX_ACTIVE per dispatch then give the bank-conflict cycles of that
// pattern (patterns.py writes the lists, summarize.py reads the counters).
//
// patterns.bin: int32 ndisp, then per dispatch: int32 kind, int32 npat,
// npat*64 uint32 byte addresses (0xffffffff: lane inactive). Kinds:
// 0 ds_read_b128, 1 ds_write_b64, 2 ds_write_b16, 3 ds_read_b64,
// 4 ds_read_b64_tr_b16, 5 ds_write_b128. Every address is checked on the
// host (aligned to the access width, inside the 64 KiB array) before launch.
// Then one wave reads ds_read_b64_tr_b16 at the 64 addresses of trdump.bin
// from an LDS holding element index i at element i, and writes what each lane
// received to the file named by the first argument.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <vector>
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef short v4s __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;
__global__ __launch_bounds__(256) void k_pat(const uint32_t* __restrict__ tab, int kind, int npat, int iters, uint32_t* out) {
  __shared__ u32x4 s[4096];  // 64 KiB
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) s[i] = u32x4{uint32_t(i), 0u, 0u, 0u};
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const uint32_t base = uint32_t(reinterpret_cast<uintptr_t>((lds_void*)s));
  uint32_t acc = 0;
  for (int it = 0; it < iters; ++it) {
    for (int k = 0; k < npat; ++k) {
      const uint32_t t = tab[k * 64 + lane];
      if (t == 0xffffffffu) continue;
      const uint32_t a = base + t;
      if (kind == 0) {
        u32x4 x;
        asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(x) : "v"(a) : "memory");
        acc += x[0] + x[3];
      } else if (kind == 1) {
        u32x2 v = {acc, uint32_t(it)};
        asm volatile("ds_write_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" ::"v"(a), "v"(v) : "memory");
      } else if (kind == 2) {
        asm volatile("ds_write_b16 %0, %1\n\ts_waitcnt lgkmcnt(0)" ::"v"(a), "v"(acc) : "memory");
      } else if (kind == 3) {
        u32x2 x;
        asm volatile("ds_read_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(x) : "v"(a) : "memory");
        acc += x[0] + x[1];
      } else if (kind == 4) {
        u32x2 x;
        asm volatile("ds_read_b64_tr_b16 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(x) : "v"(a) : "memory");
        acc += x[0] + x[1];
      } else {
        u32x4 v = {acc, uint32_t(it), 0u, 1u};
        asm volatile("ds_write_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" ::"v"(a), "v"(v) : "memory");
      }
    }
  }
  if (acc == 0xdeadbeefu) out[0] = acc;
}
__global__ void k_trdump(const uint32_t* __restrict__ addr, uint32_t* out) {
  __shared__ uint16_t s[8192];
  for (int i = threadIdx.x; i < 8192; i += 64) s[i] = uint16_t(i);
  __syncthreads();
  const uint32_t a = addr[threadIdx.x];
  v4s r = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) v4s*)((__attribute__((address_space(3))) char*)s + a));
  for (int e = 0; e < 4; ++e) out[threadIdx.x * 4 + e] = uint16_t(r[e]);
}
int main(int argc, char** argv) {
  FILE* f = fopen("patterns.bin", "rb");
  if (!f) { printf("no pattern file\n"); return 1; }
  int nd = 0;
  if (fread(&nd, 4, 1, f) != 1) return 1;
  uint32_t *out, *tab;
  if (hipMalloc(&out, 4096) != hipSuccess || hipMalloc(&tab, 4u << 20) != hipSuccess) return 2;
  const int width[6] = {16, 8, 2, 8, 8, 16};
  for (int d = 0; d < nd; ++d) {
    int kind, npat;
    if (fread(&kind, 4, 1, f) != 1 || fread(&npat, 4, 1, f) != 1) return 3;
    if (kind < 0 || kind > 5 || npat <= 0 || npat * 64 * 4 > (4 << 20)) { printf("bad entry %d\n", d); return 4; }
    std::vector<uint32_t> h(size_t(npat) * 64);
    if (fread(h.data(), 4, h.size(), f) != h.size()) return 5;
    for (uint32_t a : h)
      if (a != 0xffffffffu && ((a % width[kind]) || a > 65536u - width[kind])) { printf("bad address %u in entry %d\n", a, d); return 6; }
    if (hipMemcpy(tab, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return 7;
    const int iters = std::max(1, 256 / npat);
    hipLaunchKernelGGL(k_pat, dim3(64), dim3(256), 0, 0, tab, kind, npat, iters, out);
    if (hipDeviceSynchronize() != hipSuccess) { printf("launch %d failed\n", d); return 8; }
  }
  FILE* g = fopen("trdump.bin", "rb");
  uint32_t addr[64];
  if (!g || fread(addr, 4, 64, g) != 64) { printf("no trdump.bin\n"); return 9; }
  for (uint32_t a : addr) if ((a & 7) || a > 16384 - 8) { printf("bad tr address %u\n", a); return 10; }
  if (hipMemcpy(tab, addr, 256, hipMemcpyHostToDevice) != hipSuccess) return 11;
  hipLaunchKernelGGL(k_trdump, dim3(1), dim3(64), 0, 0, tab, out);
  uint32_t res[256];
  if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(res, out, 1024, hipMemcpyDeviceToHost) != hipSuccess) return 12;
  FILE* o = fopen(argc > 1 ? argv[1] : "trdump_out.txt", "w");
  for (int l = 0; l < 64; ++l) fprintf(o, "%d %u %u %u %u %u\n", l, addr[l], res[4 * l], res[4 * l + 1], res[4 * l + 2], res[4 * l + 3]);
  fclose(o);
  printf("probe4: %d dispatches and the transpose dump done\n", nd);
  return 0;
}

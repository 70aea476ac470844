// denoise.hip — kernels of the learned denoiser for gfx950.
//
// The network is a convolutional encoder-decoder over NHWC bf16 tensors.
// Every 3x3 convolution is an implicit GEMM on the matrix cores:
//   forward      M = pixels, N = Cout, K = 9 * Cin   (v_mfma_f32_16x16x32_bf16)
//   data grad    M = pixels, N = Cin,  K = 9 * Cout  (taps mirrored; same kernel)
//   weight grad  M = Cout,   N = Cin,  K = pixels    (v_mfma_f32_16x16x32_bf16, all 9 taps per pass)
// Operand lanes: lane l holds A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16] for j < 8, and D[4 (l / 16) + r][l % 16] for
// r < 4. Channel counts are multiples of 16, so every fragment is one
// aligned 16-byte load.
// Accumulation is fp32; weight gradients are summed per split and reduced
// in a fixed order, so training steps are deterministic.
#include "denoise.h"

#include <algorithm>
#include <atomic>

namespace gdn {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even; NaN stays NaN
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}


// ---- convolution, forward and data gradient --------------------------------
// A block owns a tile of kTR image rows x kTC columns of one image and NT*16
// output channels; wave w computes row w of the tile: 4 pixel fragments of
// 16 columns x NT channel fragments. The input channels run in chunks of
// kCK: per chunk the block stages the tile's input halo ((kTR+2) x (kTC+2)
// pixels, read once for all 9 taps) and the chunk's weights in LDS, then
// every wave runs K = 9 * chunk through v_mfma_f32_16x16x32_bf16. One K=32
// step covers two 16-channel sub-slices (tap, channel group) in tap-major
// order: lanes 0-31 take the first, lanes 32-63 the second (zero past the
// end). DGRAD reads the input at p - off_t (the transposed convolution) and
// masks it by the ReLU of the forward output while staging.
constexpr int kTR = 4;               // tile rows, one per wave
constexpr int kTC = 64;              // tile columns, 4 fragments of 16
constexpr int kHR = kTR + 2, kHC = kTC + 2;
constexpr int kCK = 32;              // input channels per staged chunk
constexpr int kLS = kCK;             // LDS row (pixel or weight row) in bf16: 64 B, pieces swizzled by swz<2>

// A 16-byte LDS read (ds_read_b128) is served 16 lanes at a time: lanes
// 0-31 and 32-63 are separate halves, and within a half the quads (4 lanes)
// whose index l >> 2 has an even number of set bits go in one pass, the others
// in the next. Two lanes of a pass conflict when their addresses differ but
// agree in bits 4-7 (the same four of the 64 banks). An MFMA operand read
// takes pixels P .. P + 15 (lane & 15) of piece q = (lane >> 4) & 1 and, in
// lanes 32-63, the next sub-slice. With 32-byte pixels (16 channels, G = 1)
// every pass then covers the 16 bank groups once as they are; with 64-byte
// pixels (G = 2) piece q of pixel p goes to slot q ^ swz<2>(p), and the
// weight rows the same way by row. Both layouts were checked conflict free
// against SQ_LDS_BANK_CONFLICT for every read of a chunk, mirrored taps and
// 16-channel chunks in 64-byte rows included.
template <int G>
__device__ __forceinline__ int swz(int p) {
  return G == 2 ? (p >> 1) & 3 : 0;
}

template <int NT>
constexpr int conv_out_stride() {  // fp32 row stride of the output staging: the 4 lane groups' rows land in distinct banks
  return NT * 16 + 4;
}
template <int NT>
constexpr size_t conv_lds_bytes() {  // the halo and weights of a chunk, or the output staging
  return std::max(size_t(kHR * kHC + 9 * NT * 16) * kLS * sizeof(uint16_t),
                  size_t(4 * 64 * conv_out_stride<NT>()) * sizeof(float));
}

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v4f mfma32(uint4 a, uint4 b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                 0);
}


// Zeroes the bf16 elements of a whose mask element is not positive.
__device__ __forceinline__ uint4 mask8(uint4 a, uint4 m) {
  uint32_t av[4] = {a.x, a.y, a.z, a.w}, mv[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t lo = bf2f(uint16_t(mv[k])) > 0.f ? 0x0000ffffu : 0u;
    const uint32_t hi = bf2f(uint16_t(mv[k] >> 16)) > 0.f ? 0xffff0000u : 0u;
    av[k] &= lo | hi;
  }
  return make_uint4(av[0], av[1], av[2], av[3]);
}

// The MFMA steps of one staged chunk of G 16-channel groups.
template <int NT, bool DGRAD, int G>
__device__ __forceinline__ void conv_chunk(const uint16_t* halo, const uint16_t* wts, int lane, int wave,
                                           v4f (&acc)[4][NT]) {
  const int half = (lane >> 4) & 1;
  const int col = lane & 15;
#pragma unroll
  for (int i = 0; i < (9 * G + 1) / 2; ++i) {
    const int sub = 2 * i + (lane >> 5);
    const bool valid = sub < 9 * G;
    const int tap = valid ? sub / G : 0;
    const int q = (sub % G) * 2 + half;  // 16-byte piece of the pixel or weight row
    int dy = tap / 3 - 1, dx = tap % 3 - 1;
    if (DGRAD) dy = -dy, dx = -dx;
    uint4 b[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int r = tap * NT * 16 + j * 16 + col;
      b[j] = valid ? *reinterpret_cast<const uint4*>(wts + r * kLS + 8 * (q ^ swz<2>(r))) : make_uint4(0, 0, 0, 0);
    }
    const int p0 = (wave + 1 + dy) * kHC + col + 1 + dx;  // pixels p0 + 16 f share the swizzle of p0
    const uint16_t* arow = halo + p0 * kLS + 8 * (q ^ swz<2>(p0));
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const uint4 a = valid ? *reinterpret_cast<const uint4*>(arow + f * 16 * kLS) : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[f][j] = mfma32(a, b[j], acc[f][j]);
    }
  }
}

// Stages one chunk (G 16-channel groups from channel c0) of the tile's input
// halo and of the weights into LDS. Each batch of loads is issued before
// its LDS stores, halo first, then weights, so a block waits for memory twice
// per chunk rather than once per piece. When xa is
// set the input is the decoder's concatenation [up2(xa) (ca channels), x]:
// channels below ca come from xa at half resolution (nearest), the rest from x.
template <int NT, bool DGRAD, int G>
__device__ __forceinline__ void conv_stage(uint16_t* halo, uint16_t* wts, const uint16_t* __restrict__ x,
                                           const uint16_t* __restrict__ mask, int cin,
                                           const uint16_t* __restrict__ xa, int ca,
                                           const uint16_t* __restrict__ wgt, int co0, int c0, int n, int h0,
                                           int w0, int H, int W, int tid) {
  constexpr int QP = 2 * G;  // 16-byte pieces per pixel or weight row
  constexpr int HP = kHR * kHC * QP, HI = (HP + 255) / 256;
  constexpr int WP = 9 * NT * 16 * QP, WI = (WP + 255) / 256;
  const int cx = xa ? cin - ca : cin, cs = xa ? ca : 0;  // channels of x, first channel from x
  {
    uint4 hv[HI];
#pragma unroll
    for (int k = 0; k < HI; ++k) {
      const int i = tid + 256 * k;
      hv[k] = make_uint4(0, 0, 0, 0);
      if (i < HP) {
        const int px = i / QP, c = c0 + 8 * (i % QP);
        const int hh = h0 - 1 + px / kHC, ww = w0 - 1 + px % kHC;
        if (hh >= 0 && hh < H && ww >= 0 && ww < W) {
          if (c < cs) {
            hv[k] = *reinterpret_cast<const uint4*>(
                xa + ((size_t(n) * (H / 2) + hh / 2) * (W / 2) + ww / 2) * ca + c);
          } else {
            const size_t off = ((size_t(n) * H + hh) * W + ww) * cx + (c - cs);
            hv[k] = *reinterpret_cast<const uint4*>(x + off);
            if (DGRAD && mask) hv[k] = mask8(hv[k], *reinterpret_cast<const uint4*>(mask + off));
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < HI; ++k) {
      const int i = tid + 256 * k;
      if (i < HP) *reinterpret_cast<uint4*>(halo + (i / QP) * kLS + 8 * ((i % QP) ^ swz<2>(i / QP))) = hv[k];
    }
  }
  __builtin_amdgcn_sched_barrier(0);  // keep the two batches apart: fewer live registers
  {
    uint4 wv[WI];
#pragma unroll
    for (int k = 0; k < WI; ++k) {
      const int i = tid + 256 * k;
      wv[k] = make_uint4(0, 0, 0, 0);
      if (i < WP) {
        const int r = i / QP, co = r % (NT * 16), tap = r / (NT * 16);
        wv[k] = *reinterpret_cast<const uint4*>(wgt + (size_t(co0 + co) * 9 + tap) * cin + c0 + 8 * (i % QP));
      }
    }
#pragma unroll
    for (int k = 0; k < WI; ++k) {
      const int i = tid + 256 * k;
      if (i < WP) *reinterpret_cast<uint4*>(wts + (i / QP) * kLS + 8 * ((i % QP) ^ swz<2>(i / QP))) = wv[k];
    }
  }
}

template <int NT, bool DGRAD, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv(const uint16_t* __restrict__ x, const uint16_t* __restrict__ mask,
                                              int cin, const uint16_t* __restrict__ xa, int ca,
                                              const uint16_t* __restrict__ wgt, const float* __restrict__ bias,
                                              void* __restrict__ y, uint16_t* __restrict__ pooled, int cout,
                                              int co_first, int H, int W, int nrt, int nct, int relu,
                                              const uint16_t* __restrict__ omask) {
  extern __shared__ uint4 lds4[];
  uint16_t* halo = reinterpret_cast<uint16_t*>(lds4);
  uint16_t* wts = halo + kHR * kHC * kLS;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int ct = blockIdx.x % nct, rt = (blockIdx.x / nct) % nrt, n = blockIdx.x / (nct * nrt);
  const int h0 = rt * kTR, w0 = ct * kTC;
  const int co0 = co_first + blockIdx.y * NT * 16;
  v4f acc[4][NT];
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[f][j] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int c0 = 0; c0 < cin; c0 += kCK) {
    __syncthreads();  // the previous chunk's reads are done
    if (cin - c0 >= kCK) {
      conv_stage<NT, DGRAD, 2>(halo, wts, x, mask, cin, xa, ca, wgt, co0, c0, n, h0, w0, H, W, tid);
      __syncthreads();
      conv_chunk<NT, DGRAD, 2>(halo, wts, lane, wave, acc);
    } else {
      conv_stage<NT, DGRAD, 1>(halo, wts, x, mask, cin, xa, ca, wgt, co0, c0, n, h0, w0, H, W, tid);
      __syncthreads();
      conv_chunk<NT, DGRAD, 1>(halo, wts, lane, wave, acc);
    }
  }
  // Epilogue through LDS: each wave writes its 64 pixels x NT*16 channels
  // (bias and ReLU applied) as fp32 rows, then stores whole 16-byte pieces of
  // consecutive channels, instead of 2-byte values scattered over 4 pixels.
  constexpr int OS = conv_out_stride<NT>();
  __syncthreads();  // every wave is done with the staged halo and weights
  float* stage = reinterpret_cast<float*>(lds4) + wave * 64 * OS;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int cl = j * 16 + (lane & 15);
    const float b = (!DGRAD && bias) ? bias[co0 + cl] : 0.f;
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = acc[f][j][r] + b;
        if (relu) v = fmaxf(v, 0.f);
        stage[(f * 16 + (lane >> 4) * 4 + r) * OS + cl] = v;
      }
  }
  __syncthreads();
  if (!OUT_F32 && pooled) {  // also the 2x2 max pool of the tile: 2 rows of 32 pixels
    // rounding to bf16 is monotonic, so the max of the rounded values is the rounded max
    const int H2 = H / 2, W2 = W / 2;
    constexpr int G = NT * 2;  // 8-channel pieces per pixel
    for (int i = tid; i < 2 * 32 * G; i += 256) {
      const int g = i % G, pc = (i / G) % 32, pr = i / (G * 32);
      const int h2 = h0 / 2 + pr, w2 = w0 / 2 + pc;
      if (h2 >= H2 || w2 >= W2) continue;
      const float* a = reinterpret_cast<const float*>(lds4) + (2 * pr) * 64 * OS + (2 * pc) * OS + 8 * g;
      uint32_t o[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float m[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int c = 2 * k + e;
          m[e] = fmaxf(fmaxf(a[c], a[OS + c]), fmaxf(a[64 * OS + c], a[64 * OS + OS + c]));
        }
        o[k] = uint32_t(f2bf(m[0])) | (uint32_t(f2bf(m[1])) << 16);
      }
      *reinterpret_cast<uint4*>(pooled + ((size_t(n) * H2 + h2) * W2 + w2) * cout + co0 + 8 * g) =
          make_uint4(o[0], o[1], o[2], o[3]);
    }
  }
  const int row = h0 + wave;
  if (row >= H) return;
  const size_t pix0 = (size_t(n) * H + row) * W;
  if (OUT_F32) {
    constexpr int G = NT * 4;  // 4-channel pieces per pixel
    for (int i = lane; i < 64 * G; i += 64) {
      const int p = i / G, g = i % G;
      if (w0 + p >= W) continue;
      const float4 v = *reinterpret_cast<const float4*>(stage + p * OS + 4 * g);
      *reinterpret_cast<float4*>(static_cast<float*>(y) + (pix0 + w0 + p) * cout + co0 + 4 * g) = v;
    }
  } else {
    constexpr int G = NT * 2;  // 8-channel pieces per pixel
    for (int i = lane; i < 64 * G; i += 64) {
      const int p = i / G, g = i % G;
      if (w0 + p >= W) continue;
      const float4 a = *reinterpret_cast<const float4*>(stage + p * OS + 8 * g);
      const float4 b = *reinterpret_cast<const float4*>(stage + p * OS + 8 * g + 4);
      const uint4 o = make_uint4(uint32_t(f2bf(a.x)) | (uint32_t(f2bf(a.y)) << 16),
                                 uint32_t(f2bf(a.z)) | (uint32_t(f2bf(a.w)) << 16),
                                 uint32_t(f2bf(b.x)) | (uint32_t(f2bf(b.y)) << 16),
                                 uint32_t(f2bf(b.z)) | (uint32_t(f2bf(b.w)) << 16));
      const size_t off = (pix0 + w0 + p) * cout + co0 + 8 * g;
      *reinterpret_cast<uint4*>(static_cast<uint16_t*>(y) + off) =
          omask ? mask8(o, *reinterpret_cast<const uint4*>(omask + off)) : o;
    }
  }
}

// ---- convolution, forward: persistent blocks, halos copied straight to LDS --
// The same tiles and MFMA steps as k_conv, for layers of up to 48 output
// channels (one group). A block keeps the layer's weights in LDS for its
// whole life and walks a run of tiles. A tile's input channels run in chunks
// of 16*GMAX that never straddle the two parts of a decoder concatenation;
// a (tile, chunk) pair is one stage. The halo of stage s + 1 is copied from
// global memory into the second of two LDS buffers by buffer loads that
// write LDS directly (no registers), while the waves run the MFMA steps of
// stage s; border pixels are read at an offset past the buffer's range and
// land as zeros. The halo has no padding: piece q (16 bytes) of pixel p sits
// at slot p * 2G + (q ^ swz<G>(p)), and the weights likewise by row. Blocks are dealt
// round-robin to the 8 XCDs; each XCD's blocks walk one contiguous eighth of
// the tiles, so tiles that share halo rows are read through the same L2.
typedef __attribute__((address_space(3))) void lds_void;

constexpr uint32_t kOutOfRange = 0x80000000u;  // a buffer offset past every tensor this kernel reads
constexpr int kWaitVm0 = 0x0F70;    // s_waitcnt vmcnt(0)
constexpr int kWaitLgkm0 = 0xC07F;  // s_waitcnt lgkmcnt(0)

// A workgroup barrier for LDS data only: this wave's LDS reads and writes are
// complete, then the waves meet. Unlike __syncthreads it does not also wait
// for the LDS copies in flight; those are waited for per buffer.
__device__ __forceinline__ void lds_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// The halo buffers hold a chunk of up to 16*GMAX channels in whole 64-slot
// copies. The output tile is staged in the buffer it was computed from, in
// passes of at most 32 channels (bf16; 16 channels fp32 for the network
// output): 256 pixels x the pass's channels, rows padded by 16 bytes when the
// padded pass still fits.
template <int NT, bool OUT_F32>
constexpr int pipe_pass_channels() {
  return OUT_F32 ? 16 : std::min(NT * 16, 32);
}
template <int NT, bool OUT_F32>
constexpr int pipe_passes() {
  return (NT * 16 + pipe_pass_channels<NT, OUT_F32>() - 1) / pipe_pass_channels<NT, OUT_F32>();
}
template <int NT, bool OUT_F32>
constexpr int pipe_raw_row_bytes() {  // one pass
  return pipe_pass_channels<NT, OUT_F32>() * (OUT_F32 ? 4 : 2);
}
template <int NT, bool OUT_F32, int GMAX>
constexpr int pipe_buf_slots() {
  return std::max((kHR * kHC * 2 * GMAX + 63) / 64 * 64, 256 * pipe_raw_row_bytes<NT, OUT_F32>() / 16);
}
template <int NT, bool OUT_F32, int GMAX>
constexpr int pipe_row_bytes() {
  constexpr int raw = pipe_raw_row_bytes<NT, OUT_F32>();
  return 256 * (raw + 16) <= pipe_buf_slots<NT, OUT_F32, GMAX>() * 16 ? raw + 16 : raw;
}
template <int NT, bool OUT_F32>
constexpr int pipe_row_pieces() {  // 16-byte pieces of one output pixel
  return OUT_F32 ? NT * 4 : NT * 2;
}
constexpr int pipe_pool_per_thread(int nt) { return (2 * 32 * nt * 2 + 255) / 256; }

// Chunks of a layer's input: channels [0, ca) of the half-resolution part
// first, then [ca, cin) of the full-resolution part, each in runs of 16*GMAX
// channels and a last run of 16 when GMAX is 2 and the part has an odd
// number of 16-channel groups.
struct PipeChunk {
  int c0, g;
  bool half_res;
};
template <int GMAX>
__device__ __forceinline__ int pipe_nchunks(int cin, int ca) {
  constexpr int CS = 16 * GMAX;
  return (ca + CS - 1) / CS + (cin - ca + CS - 1) / CS;
}
template <int GMAX>
__device__ __forceinline__ PipeChunk pipe_chunk(int k, int cin, int ca) {
  constexpr int CS = 16 * GMAX;
  const int na = (ca + CS - 1) / CS;
  if (k < na) return {CS * k, std::min(GMAX, (ca - CS * k) / 16), true};
  const int c0 = ca + CS * (k - na);
  return {c0, std::min(GMAX, (cin - c0) / 16), false};
}

typedef int32_t i32x4 __attribute__((ext_vector_type(4)));

// A raw buffer resource over [base, base + bytes): reads past the end return
// zeros.
__device__ __forceinline__ i32x4 buffer_rsrc(const void* base, uint32_t bytes) {
  const uint64_t a = reinterpret_cast<uint64_t>(base);
  i32x4 r;
  r.x = __builtin_amdgcn_readfirstlane(int32_t(uint32_t(a)));
  r.y = __builtin_amdgcn_readfirstlane(int32_t(uint32_t(a >> 32) & 0xffffu));  // stride 0
  r.z = __builtin_amdgcn_readfirstlane(int32_t(bytes));
  r.w = 0x00020000;
  return r;
}

// Copies 16 bytes per lane from byte offset off of the buffer src to LDS:
// lane l writes dst + 16 l, dst being the same LDS address in every lane (it
// goes to M0). Written in assembly so that the compiler does not take the
// copy in flight for a write to every LDS array and wait for it before the
// reads of the other buffer; the kernel waits for it itself (vmcnt) before it
// reads the buffer. The copy writes no register. The compiler keeps nothing in
// M0 across these blocks in this kernel (no LDS-copy builtins, message sends
// or relative register moves), so clobbering it is safe here.
__device__ __forceinline__ void lds_copy16(const uint4* dst, i32x4 src, uint32_t off) {
  const uint32_t m0 = __builtin_amdgcn_readfirstlane(uint32_t(reinterpret_cast<uintptr_t>((const lds_void*)dst)));
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tbuffer_load_dwordx4 %1, %2, 0 offen lds" ::"s"(m0), "v"(off), "s"(src)
               : "memory", "m0");
#pragma clang diagnostic pop
}

// Issues the copies of one chunk's halo into buf: wave w copies slots
// [64 i, 64 i + 64) for i = w, w + 4, ... Channel c of the source (src_c
// channels per pixel) is the chunk's first.
template <int G>
__device__ __forceinline__ void pipe_copy_halo(uint4* buf, i32x4 src, int src_c, int c,
                                               bool half_res, int n, int h0, int w0, int H, int W, int wave,
                                               int lane) {
  constexpr int QP = 2 * G, SL = kHR * kHC * QP, NI = (SL + 63) / 64;
  const int Hs = half_res ? H / 2 : H, Ws = half_res ? W / 2 : W;
#pragma unroll
  for (int i0 = 0; i0 < NI; i0 += 4) {
    const int i = i0 + wave;
    if (i < NI) {
      const int s = i * 64 + lane, px = s / QP;
      const int q = (s % QP) ^ swz<G>(px);
      const int hh = h0 - 1 + px / kHC, ww = w0 - 1 + px % kHC;
      uint32_t off = kOutOfRange;
      if (s < SL && hh >= 0 && hh < H && ww >= 0 && ww < W) {
        const int sh = half_res ? hh >> 1 : hh, sw = half_res ? ww >> 1 : ww;
        off = uint32_t((((size_t(n) * Hs + sh) * Ws + sw) * src_c + c + 8 * q) * 2);
      }
      lds_copy16(buf + i * 64, src, off);
    }
  }
}

// The MFMA steps of one chunk (G 16-channel groups) from the swizzled halo
// and the chunk's weights from slot wb of wts (rows tap * NT*16 + co, 2G
// pieces each, swizzled the same way by row). Both arrays are indexed from
// their bases in one step.
template <int NT, int G, bool MIRROR>
__device__ __forceinline__ void pipe_chunk_mfma(const uint4* halo, const uint4* wts, int wb, int lane, int wave,
                                                v4f (&acc)[4][NT]) {
  constexpr int QP = 2 * G;
  const int half = (lane >> 4) & 1, col = lane & 15;
#pragma unroll
  for (int i = 0; i < (9 * G + 1) / 2; ++i) {
    const int sub = 2 * i + (lane >> 5);
    const bool valid = sub < 9 * G;
    const int tap = valid ? sub / G : 0;
    const int q = (sub % G) * 2 + half;
    const int dy = MIRROR ? 1 - tap / 3 : tap / 3 - 1, dx = MIRROR ? 1 - tap % 3 : tap % 3 - 1;
    uint4 b[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int r = tap * NT * 16 + j * 16 + col;
      b[j] = valid ? wts[wb + r * QP + (q ^ swz<G>(r))] : make_uint4(0, 0, 0, 0);
    }
    const int p0 = (wave + 1 + dy) * kHC + col + 1 + dx;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const int p = p0 + 16 * f;
      const uint4 a = valid ? halo[p * QP + (q ^ swz<G>(p))] : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[f][j] = mfma32(a, b[j], acc[f][j]);
    }
  }
}

// A finished tile's output, read back from its LDS staging into registers:
// this wave's row (pieces lane, lane + 64, ...) and this thread's pieces of
// the 2x2 max pool. The stores are issued one step later, right after the
// next halo copy, so that the wait for that copy does not also wait for them.
template <int NT, bool OUT_F32>
struct PipeOut {
  u32x4 row[pipe_row_pieces<NT, OUT_F32>()];  // vector types, not uint4: they stay in registers
  u32x4 pool[pipe_pool_per_thread(NT)];
  u32x4 mrow[OUT_F32 ? 1 : pipe_row_pieces<NT, OUT_F32>()];  // the output mask of the row pieces
  int tile = -1;
};

// Bias and ReLU of the accumulators, staged through buf (bf16, or fp32 for the
// network output) pass by pass and read back into out; clears acc for the
// next tile. With rad (the network output only) lane l keeps instead the
// radiance expm1(max(z, 0)) of channels 0..2 of its pixel, the wave's row's
// pixel l.
template <int NT, bool OUT_F32, int GMAX>
__device__ __forceinline__ void pipe_finish(uint4* buf, v4f (&acc)[4][NT], const float (&bj)[NT], int relu,
                                            bool pooled, bool rad, PipeOut<NT, OUT_F32>& out, int t, int tid,
                                            int lane, int wave, const uint16_t* omask, int cout, int co0, int H,
                                            int W, int nrt, int nct) {
  if (!OUT_F32 && omask) {
    // the output mask of this lane's row pieces, loaded now: the next step's wait for its halo
    // copy (issued before these loads) covers them, so the stores wait for nothing more
    constexpr int GP = pipe_row_pieces<NT, OUT_F32>();
    const int ct = t % nct, rt = (t / nct) % nrt, n = t / (nct * nrt);
    const int row = rt * kTR + wave, w0 = ct * kTC;
#pragma unroll
    for (int k = 0; k < GP; ++k) {
      const int i = lane + 64 * k, p = i / GP, g = i % GP;
      out.mrow[k] = u32x4{0u, 0u, 0u, 0u};
      if (row < H && w0 + p < W)
        out.mrow[k] = *reinterpret_cast<const u32x4*>(omask + ((size_t(n) * H + row) * W + w0 + p) * cout + co0 + 8 * g);
    }
  }
  constexpr int RB = pipe_row_bytes<NT, OUT_F32, GMAX>(), GP = pipe_row_pieces<NT, OUT_F32>();
  constexpr int PC = pipe_pass_channels<NT, OUT_F32>(), PJ = PC / 16;  // channels, fragments per pass
  constexpr int CP = OUT_F32 ? 4 : 8;                                  // channels per 16-byte piece
  char* st = reinterpret_cast<char*>(buf);
#pragma unroll
  for (int pass = 0; pass < pipe_passes<NT, OUT_F32>(); ++pass) {
    lds_barrier();  // every wave is done with buf: the halo, or the previous pass
#pragma unroll
    for (int jj = 0; jj < PJ; ++jj) {
      const int j = pass * PJ + jj;
      if (j >= NT) break;
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float v = acc[f][j][r] + bj[j];
          if (relu) v = fmaxf(v, 0.f);
          const int px = wave * 64 + f * 16 + (lane >> 4) * 4 + r, c = jj * 16 + (lane & 15);
          if (OUT_F32) *reinterpret_cast<float*>(st + px * RB + 4 * c) = v;
          else *reinterpret_cast<uint16_t*>(st + px * RB + 2 * c) = f2bf(v);
          acc[f][j][r] = 0.f;
        }
    }
    lds_barrier();
    const int g0 = pass * PC / CP, g1 = std::min(GP, (pass + 1) * PC / CP);  // this pass's pieces
    if (OUT_F32 && rad) {
      const float* z = reinterpret_cast<const float*>(st + (wave * 64 + lane) * RB);
      out.row[0] = u32x4{__float_as_uint(expm1f(fmaxf(z[0], 0.f))), __float_as_uint(expm1f(fmaxf(z[1], 0.f))),
                         __float_as_uint(expm1f(fmaxf(z[2], 0.f))), 0u};
    } else {
#pragma unroll
      for (int k = 0; k < GP; ++k) {
        const int i = lane + 64 * k, p = i / GP, g = i % GP;
        if (g >= g0 && g < g1) out.row[k] = *reinterpret_cast<const u32x4*>(st + (wave * 64 + p) * RB + 16 * (g - g0));
      }
    }
    if (!OUT_F32 && pooled) {  // the staged values are already bf16: the max of them is the pooled value
      constexpr int G8 = NT * 2;
#pragma unroll
      for (int k = 0; k < pipe_pool_per_thread(NT); ++k) {
        const int i = tid + 256 * k;
        const int g = i % G8, pc = (i / G8) % 32, pr = i / (G8 * 32);
        if (i >= 2 * 32 * G8 || g < g0 || g >= g1) continue;
        const char* a = st + ((2 * pr) * 64 + 2 * pc) * RB + 16 * (g - g0);
        const u32x4 v[4] = {*reinterpret_cast<const u32x4*>(a), *reinterpret_cast<const u32x4*>(a + RB),
                            *reinterpret_cast<const u32x4*>(a + 64 * RB),
                            *reinterpret_cast<const u32x4*>(a + 65 * RB)};
        u32x4 o;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t w[4] = {v[0][e], v[1][e], v[2][e], v[3][e]};
          float lo = bf2f(uint16_t(w[0])), hi = bf2f(uint16_t(w[0] >> 16));
#pragma unroll
          for (int m = 1; m < 4; ++m) {
            lo = fmaxf(lo, bf2f(uint16_t(w[m])));
            hi = fmaxf(hi, bf2f(uint16_t(w[m] >> 16)));
          }
          o[e] = uint32_t(f2bf(lo)) | (uint32_t(f2bf(hi)) << 16);
        }
        out.pool[k] = o;
      }
    }
  }
  out.tile = t;
}

// The global stores of a finished tile held in out. With rad: radiance
// [n][rh][rw][3] of the pixels inside rh x rw (the caller's image in the
// padded tensor), instead of y.
template <int NT, bool OUT_F32>
__device__ __forceinline__ void pipe_store(PipeOut<NT, OUT_F32>& out, void* y, uint16_t* pooled, float* rad, int rh,
                                           int rw, const uint16_t* omask, int cout, int co0, int H, int W, int nrt,
                                           int nct, int tid, int lane, int wave) {
  constexpr int GP = pipe_row_pieces<NT, OUT_F32>();
  const int t = out.tile;
  out.tile = -1;
  const int ct = t % nct, rt = (t / nct) % nrt, n = t / (nct * nrt);
  const int h0 = rt * kTR, w0 = ct * kTC;
  const int row = h0 + wave;
  if (OUT_F32 && rad) {
    if (row < rh && w0 + lane < rw) {
      float* o = rad + ((size_t(n) * rh + row) * rw + w0 + lane) * 3;
      o[0] = __uint_as_float(out.row[0][0]);
      o[1] = __uint_as_float(out.row[0][1]);
      o[2] = __uint_as_float(out.row[0][2]);
    }
  } else if (row < H) {
    const size_t pix0 = (size_t(n) * H + row) * W;
#pragma unroll
    for (int k = 0; k < GP; ++k) {
      const int i = lane + 64 * k, p = i / GP, g = i % GP;
      if (w0 + p >= W) continue;
      if (OUT_F32) {
        *reinterpret_cast<u32x4*>(static_cast<float*>(y) + (pix0 + w0 + p) * cout + co0 + 4 * g) = out.row[k];
      } else {
        const size_t off = (pix0 + w0 + p) * cout + co0 + 8 * g;
        u32x4 v = out.row[k];
        if (omask) {  // zero where the mask is not positive (the ReLU of the tensor this gradient belongs to)
          const u32x4 m = out.mrow[k];
#pragma unroll
          for (int e = 0; e < 4; ++e)
            v[e] &= (bf2f(uint16_t(m[e])) > 0.f ? 0x0000ffffu : 0u) | (bf2f(uint16_t(m[e] >> 16)) > 0.f ? 0xffff0000u : 0u);
        }
        *reinterpret_cast<u32x4*>(static_cast<uint16_t*>(y) + off) = v;
      }
    }
  }
  if (!OUT_F32 && pooled) {
    constexpr int G8 = NT * 2;
    const int H2 = H / 2, W2 = W / 2;
#pragma unroll
    for (int k = 0; k < pipe_pool_per_thread(NT); ++k) {
      const int i = tid + 256 * k;
      if (i >= 2 * 32 * G8) break;
      const int g = i % G8, pc = (i / G8) % 32, pr = i / (G8 * 32);
      const int h2 = h0 / 2 + pr, w2 = w0 / 2 + pc;
      if (h2 < H2 && w2 < W2)
        *reinterpret_cast<u32x4*>(pooled + ((size_t(n) * H2 + h2) * W2 + w2) * cout + co0 + 8 * g) = out.pool[k];
    }
  }
}

template <int NT, bool OUT_F32, int CIN, int GMAX, bool MIRROR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_conv_pipe(const uint16_t* __restrict__ x,
                                                   const uint16_t* __restrict__ xa, int ca,
                                                   const uint16_t* __restrict__ wgt, const float* __restrict__ bias,
                                                   void* __restrict__ y, uint16_t* __restrict__ pooled, int cout,
                                                   int co_first, int N, int H, int W, int nrt, int nct, int relu,
                                                   float* __restrict__ rad, int rh, int rw,
                                                   const uint16_t* __restrict__ omask) {
  __shared__ uint4 bufA[pipe_buf_slots<NT, OUT_F32, GMAX>()];
  __shared__ uint4 bufB[pipe_buf_slots<NT, OUT_F32, GMAX>()];
  __shared__ uint4 wts4[9 * NT * 16 * CIN / 8];
  constexpr int cin = CIN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int co0 = co_first + blockIdx.y * NT * 16;
  const int cx = cin - ca;
  const i32x4 rx = buffer_rsrc(x, uint32_t(size_t(N) * H * W * cx * 2));
  const i32x4 ra = buffer_rsrc(xa ? xa : x, xa ? uint32_t(size_t(N) * (H / 2) * (W / 2) * ca * 2) : 0u);

  // this block's tiles: the XCD's eighth of them, every per-th from slot0
  const int ntiles = N * nrt * nct;
  const int xcd = blockIdx.x & 7, per = gridDim.x >> 3, slot0 = blockIdx.x >> 3;
  const int t_begin = int(int64_t(ntiles) * xcd / 8), t_end = int(int64_t(ntiles) * (xcd + 1) / 8);
  const int mine = t_end - t_begin > slot0 ? (t_end - t_begin - slot0 + per - 1) / per : 0;
  const int nch = pipe_nchunks<GMAX>(cin, ca);
  const int S = mine * nch;
  if (S == 0) return;  // uniform over the block: no barrier is skipped by part of it

  // the weights, once: piece c/8 of row (tap, co) of the chunk holding channel c
  const int npc = cin / 8;
  for (int i = tid; i < 9 * NT * 16 * npc; i += 256) {
    const int cp = i % npc, rt = i / npc, co = rt / 9, tap = rt % 9, c = 8 * cp;
    constexpr int CS = 16 * GMAX;
    const int c0 = c < ca ? c / CS * CS : ca + (c - ca) / CS * CS;
    const int g = std::min(GMAX, ((c < ca ? ca : cin) - c0) / 16);
    const int r = tap * NT * 16 + co, q = (c - c0) >> 3;
    const int sw = g == 2 ? swz<2>(r) : swz<1>(r);
    wts4[9 * NT * 16 * (c0 / 8) + r * 2 * g + (q ^ sw)] =
        *reinterpret_cast<const uint4*>(wgt + (size_t(co0 + co) * 9 + tap) * cin + c);
  }
  float bj[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) bj[j] = bias ? bias[co0 + j * 16 + (lane & 15)] : 0.f;

  auto copy = [&](uint4* buf, int s) {
    const int k = s / nch, ch = s - k * nch;
    const int t = t_begin + slot0 + k * per;
    const int ct = t % nct, rt = (t / nct) % nrt, n = t / (nct * nrt);
    const PipeChunk c = pipe_chunk<GMAX>(ch, cin, ca);
    const i32x4 src = c.half_res ? ra : rx;
    const int src_c = c.half_res ? ca : cx, c_src = c.half_res ? c.c0 : c.c0 - ca;
    if (GMAX == 2 && c.g == 2)
      pipe_copy_halo<2>(buf, src, src_c, c_src, c.half_res, n, rt * kTR, ct * kTC, H, W, wave, lane);
    else pipe_copy_halo<1>(buf, src, src_c, c_src, c.half_res, n, rt * kTR, ct * kTC, H, W, wave, lane);
  };
  v4f acc[4][NT];
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[f][j] = v4f{0.f, 0.f, 0.f, 0.f};
  PipeOut<NT, OUT_F32> out;
  auto step = [&](uint4* cur, uint4* nxt, int s) {
    __builtin_amdgcn_s_waitcnt(kWaitVm0);  // this wave's copies into cur (and its last stores) are done
    lds_barrier();  // ... and every wave's; nxt was last read in the previous step
    if (s + 1 < S) copy(nxt, s + 1);
    if (out.tile >= 0)
      pipe_store<NT, OUT_F32>(out, y, pooled, rad, rh, rw, omask, cout, co0, H, W, nrt, nct, tid, lane, wave);
    const int k = s / nch, ch = s - k * nch;
    const PipeChunk c = pipe_chunk<GMAX>(ch, cin, ca);
    const int wb = 9 * NT * 16 * (c.c0 / 8);
    if (GMAX == 2 && c.g == 2) pipe_chunk_mfma<NT, 2, MIRROR>(cur, wts4, wb, lane, wave, acc);
    else pipe_chunk_mfma<NT, 1, MIRROR>(cur, wts4, wb, lane, wave, acc);
    if (ch == nch - 1)
      pipe_finish<NT, OUT_F32, GMAX>(cur, acc, bj, relu, pooled != nullptr, rad != nullptr, out,
                                     t_begin + slot0 + k * per, tid, lane, wave, omask, cout, co0, H, W, nrt, nct);
  };
  copy(bufA, 0);
  for (int s = 0; s < S; s += 2) {
    step(bufA, bufB, s);
    if (s + 1 < S) step(bufB, bufA, s + 1);
  }
  if (out.tile >= 0)
    pipe_store<NT, OUT_F32>(out, y, pooled, rad, rh, rw, omask, cout, co0, H, W, nrt, nct, tid, lane, wave);
}

// ---- convolution, weight gradient -----------------------------------------
// dW[co][tap][ci] = sum over pixels p of dPre[p][co] * x[p + off(tap)][ci],
// with dPre the output gradient (masked by the ReLU of the forward output by
// the kernel that made it). A block owns kWBands bands of kWR image rows of
// one image (its split) and NCO x NCI 16-channel fragments of (co, ci); wave w
// takes row w of each band. The four waves' sums are added in wave order
// through LDS, and the blocks of the first ci group also sum dPre against
// ones (the bias).
constexpr int kWR = 4;
constexpr size_t kWRed = 3 * 36 * 64 * sizeof(float);  // waves 1..3 hand wave 0 one (a, b) pair of sums
// Bands of kWR rows per block: a block's sums cover kWBands bands (its split),
// which halves the partial sums written and reduced against one band per block.
constexpr int kWBands = 2;

// The end of a weight-gradient block: waves 1..3 hand their sums to wave 0
// through LDS (red, kWRed bytes), one (a, b) pair at a time, in wave order,
// and wave 0 writes the block's row of partial sums (out: weights, then the
// bias at nw). The caller has synchronised the block after its last LDS read.
template <int NCO, int NCI>
__device__ __forceinline__ void wgrad_finish(float* red, v4f (&acc)[9][NCO][NCI], v4f (&bacc)[NCO], bool bias,
                                             float* __restrict__ out, size_t nw, int cin, int co0, int ci0, int wave,
                                             int lane) {
#pragma unroll
  for (int a = 0; a < NCO; ++a)
#pragma unroll
    for (int b = 0; b < NCI; ++b) {
      __syncthreads();
      if (wave > 0) {
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) red[((wave - 1) * 36 + t * 4 + r) * 64 + lane] = acc[t][a][b][r];
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float v = acc[t][a][b][r];
            v += red[(0 * 36 + t * 4 + r) * 64 + lane];
            v += red[(1 * 36 + t * 4 + r) * 64 + lane];
            v += red[(2 * 36 + t * 4 + r) * 64 + lane];
            const int co = co0 + a * 16 + (lane >> 4) * 4 + r;
            const int ci = ci0 + b * 16 + (lane & 15);
            out[(size_t(co) * 9 + t) * cin + ci] = v;
          }
      }
    }
  if (bias) {
    __syncthreads();
    if (wave > 0) {
#pragma unroll
      for (int a = 0; a < NCO; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[((wave - 1) * 4 * NCO + a * 4 + r) * 64 + lane] = bacc[a][r];
    }
    __syncthreads();
    if (wave == 0 && (lane & 15) == 0) {  // column 0 of the all-ones product
#pragma unroll
      for (int a = 0; a < NCO; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float v = bacc[a][r];
          v += red[((0 * NCO + a) * 4 + r) * 64 + lane];
          v += red[((1 * NCO + a) * 4 + r) * 64 + lane];
          v += red[((2 * NCO + a) * 4 + r) * 64 + lane];
          out[nw + co0 + a * 16 + (lane >> 4) * 4 + r] = v;
        }
    }
  }
}

// The kernel walks its bands in tiles of TC columns. dPre (kWR rows x TC
// columns) and x (kWR + 2 rows x TC + 2 columns: one beyond the tile on each
// side) lie in LDS as they lie in HBM, one 32-byte row of 16 channels per
// pixel, in one region per 16-channel fragment, copied straight from global
// memory by buffer loads that write LDS (each region from its own tensor, so
// a block's channels may straddle the decoder concatenation); the next tile's
// copy runs while the waves work on this one. The MFMA
// operands come from ds_read_b64_tr_b16: over 16 lanes, lane l names 4
// channels of pixel row (l >> 2) & 3 of a 4 x 16 block and receives channel
// l's 4 pixels, so two reads give a lane its 8 K values (pixels) and a column
// tap is an offset of the pixel index. Slots are swizzled by bit 3 of the
// pixel, which puts the two 16-lane groups of a read pass (pixels 8 apart) in
// different banks. The buffer resources start at the first row the block's
// bands read, so offsets stay small for tensors of any size (kWBands * kWR + 2
// rows of the widest tensor must stay below 2 GB).
template <int TC>  // tile columns
constexpr int wgt_dslots() {  // 16-byte slots of one dPre fragment region
  return (kWR * TC * 2 + 63) / 64 * 64;
}
template <int TC>
constexpr int wgt_xslots() {  // ... of one x fragment region (TC + 2 columns)
  return ((kWR + 2) * (TC + 2) * 2 + 63) / 64 * 64;
}

__device__ __forceinline__ int wgt_slot(int p, int k) {  // piece k (0, 1) of pixel p
  return (2 * p + k) ^ (((p >> 3) & 1) << 3);
}
__device__ __forceinline__ int wgt_pixel(int slot) {  // the pixel of a slot (wgt_slot is its own inverse on bit 3)
  const int q = slot >> 1;
  return q ^ (((q >> 3) & 1) << 2);
}

template <int NCO, int NCI, int TC>
constexpr size_t wgrad_tr_lds_bytes() {  // two buffers
  return std::max(size_t(2) * (NCO * wgt_dslots<TC>() + NCI * wgt_xslots<TC>()) * 16, kWRed);
}

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef short v4s __attribute__((ext_vector_type(4)));

// The 8 K values of this lane for a 16 x 32 operand whose K runs over the
// pixels p0 + 0 .. 31 of a fragment region (byte address base).
__device__ __forceinline__ uint4 wgt_operand(const char* base, int p0, int lane) {
  const int g = lane >> 4, r = (lane >> 2) & 3, c = lane & 3;
  u32x2 v[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int p = p0 + 8 * g + 4 * h + r;
    const char* a = base + wgt_slot(p, c >> 1) * 16 + (c & 1) * 8;
    v[h] = __builtin_bit_cast(u32x2, __builtin_amdgcn_ds_read_tr16_b64_v4i16(
                                         (__attribute__((address_space(3))) v4s*)((const lds_void*)a)));
  }
  return make_uint4(v[0].x, v[0].y, v[1].x, v[1].y);
}

template <int NCO, int NCI, int TC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_wgrad_tr(
    const uint16_t* __restrict__ x, int cin, const uint16_t* __restrict__ xa, int ca, const uint16_t* __restrict__ dy,
    int cout, int N, int H, int W, int nsplit, int ncog, int ci_first, float* __restrict__ partial) {
  constexpr int kWDS = wgt_dslots<TC>(), kWXS = wgt_xslots<TC>(), kWXC = TC + 2;
  constexpr int BUF = NCO * kWDS + NCI * kWXS;  // slots of one buffer: dPre regions, then x regions
  extern __shared__ uint4 wlds4[];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n = blockIdx.x / nsplit, band0 = (blockIdx.x % nsplit) * kWBands;
  const int co0 = (blockIdx.y % ncog) * NCO * 16, ci0 = ci_first + (blockIdx.y / ncog) * NCI * 16;
  const bool bias = ci0 == 0;  // uniform over the block
  const int cx = cin - ca;
  // rows [hb, he) hold everything the block reads (x's halo row above and below included)
  const int hb = max(band0 * kWR - 1, 0), he = min((band0 + kWBands) * kWR + 1, H);
  const i32x4 rd = buffer_rsrc(dy + (size_t(n) * H + hb) * W * cout, uint32_t(size_t(he - hb) * W * cout * 2));
  const i32x4 rx = buffer_rsrc(x + (size_t(n) * H + hb) * W * cx, uint32_t(size_t(he - hb) * W * cx * 2));
  const int hb2 = hb / 2, he2 = (he + 1) / 2;  // the same rows of the half-resolution part
  const i32x4 ra = xa ? buffer_rsrc(xa + (size_t(n) * (H / 2) + hb2) * (W / 2) * ca,
                                    uint32_t(size_t(he2 - hb2) * (W / 2) * ca * 2))
                      : buffer_rsrc(x, 0u);
  v4f acc[9][NCO][NCI], bacc[NCO];
#pragma unroll
  for (int a = 0; a < NCO; ++a) {
    bacc[a] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
      for (int b = 0; b < NCI; ++b) acc[t][a][b] = v4f{0.f, 0.f, 0.f, 0.f};
  }
  // the copies of the tile at rows r0.., column c0 into buffer buf: wave w issues slots
  // [64 i, 64 i + 64) of each region for i = w, w + 4, ...
  auto copy = [&](uint4* buf, int r0, int c0) {
#pragma unroll
    for (int a = 0; a < NCO; ++a) {
#pragma unroll
      for (int i0 = 0; i0 < kWDS / 64; i0 += 4) {
        const int i = i0 + wave, sl = i * 64 + lane, p = wgt_pixel(sl);
        const int h = r0 + p / TC, w = c0 + p % TC;
        uint32_t off = kOutOfRange;
        if (p < kWR * TC && h < H && w < W)
          off = uint32_t((((size_t(h) - hb) * W + w) * cout + co0 + 16 * a + 8 * (sl & 1)) * 2);
        lds_copy16(buf + a * kWDS + i * 64, rd, off);
      }
    }
#pragma unroll
    for (int b = 0; b < NCI; ++b) {
      const int c = ci0 + 16 * b;  // the fragment's first channel: below ca it comes from xa at half resolution
      const bool half_res = c < ca;
#pragma unroll
      for (int i0 = 0; i0 < kWXS / 64; i0 += 4) {
        const int i = i0 + wave;
        if (i < kWXS / 64) {
          const int sl = i * 64 + lane, p = wgt_pixel(sl);
          const int h = r0 - 1 + p / kWXC, w = c0 - 1 + p % kWXC;
          uint32_t off = kOutOfRange;
          if (p < (kWR + 2) * kWXC && h >= 0 && h < H && w >= 0 && w < W)
            off = half_res ? uint32_t((((size_t(h) / 2 - hb2) * (W / 2) + w / 2) * ca + c + 8 * (sl & 1)) * 2)
                           : uint32_t((((size_t(h) - hb) * W + w) * cx + c - ca + 8 * (sl & 1)) * 2);
          lds_copy16(buf + NCO * kWDS + b * kWXS + i * 64, half_res ? ra : rx, off);
        }
      }
    }
  };
  const uint4 ones = make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u);  // bf16 1.0 x 8
  // steps: the tiles of band band0, then of band0 + 1, ...
  const int ntiles = (W + TC - 1) / TC;
  const int nb = min(kWBands, (H + kWR - 1) / kWR - band0), S = nb * ntiles;
  copy(wlds4, band0 * kWR, 0);
  for (int t = 0; t < S; ++t) {
    uint4* cur = wlds4 + (t & 1) * BUF;
    __builtin_amdgcn_s_waitcnt(kWaitVm0);  // this wave's copies of step t are done
    lds_barrier();                        // ... and every wave's; the other buffer's reads are done
    if (t + 1 < S) copy(wlds4 + ((t + 1) & 1) * BUF, (band0 + (t + 1) / ntiles) * kWR, (t + 1) % ntiles * TC);
    if ((band0 + t / ntiles) * kWR + wave >= H) continue;  // this wave's row is past the image: its dPre is zero
    const char* dT = reinterpret_cast<const char*>(cur);
    const char* xT = reinterpret_cast<const char*>(cur + NCO * kWDS);
#pragma unroll
    for (int kc = 0; kc < TC / 32; ++kc) {
      uint4 A[NCO];
#pragma unroll
      for (int a = 0; a < NCO; ++a) A[a] = wgt_operand(dT + a * kWDS * 16, wave * TC + 32 * kc, lane);
#pragma unroll
      for (int b = 0; b < NCI; ++b) {
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
          const int dr = tap / 3, dx = tap % 3 - 1;
          const uint4 B = wgt_operand(xT + b * kWXS * 16, (wave + dr) * kWXC + 32 * kc + 1 + dx, lane);
#pragma unroll
          for (int a = 0; a < NCO; ++a) acc[tap][a][b] = mfma32(A[a], B, acc[tap][a][b]);
        }
      }
      if (bias) {
#pragma unroll
        for (int a = 0; a < NCO; ++a) bacc[a] = mfma32(A[a], ones, bacc[a]);
      }
    }
  }
  __syncthreads();  // every wave is done with the tiles
  const size_t nw = size_t(cout) * 9 * cin;
  wgrad_finish<NCO, NCI>(reinterpret_cast<float*>(wlds4), acc, bacc, bias,
                         partial + size_t(blockIdx.x) * (nw + cout), nw, cin, co0, ci0, wave, lane);
}

// out[i] = sum over s < splits of partial[s * stride + i], i < n. A block of 256 threads
// owns 32 consecutive outputs; its 8 thread groups sum the splits congruent
// to their group index mod 8, in increasing order, and group sums are added
// in group order: a fixed order, so results are deterministic.
__global__ __launch_bounds__(256) void k_reduce(const float* __restrict__ partial, int splits, size_t n,
                                                size_t stride, float* __restrict__ out) {
  __shared__ float part[8][33];
  const int o = threadIdx.x & 31, g = threadIdx.x >> 5;
  const size_t i = blockIdx.x * size_t(32) + o;
  float acc = 0.f;
  if (i < n) {
    int k = g;
    for (; k + 24 < splits; k += 32) {  // four independent loads in flight
      const float a = partial[size_t(k) * stride + i], b = partial[size_t(k + 8) * stride + i];
      const float c = partial[size_t(k + 16) * stride + i], d = partial[size_t(k + 24) * stride + i];
      acc += a;
      acc += b;
      acc += c;
      acc += d;
    }
    for (; k < splits; k += 8) acc += partial[size_t(k) * stride + i];
  }
  part[g][o] = acc;
  __syncthreads();
  if (g == 0 && i < n) {
    float t = part[0][o];
#pragma unroll
    for (int j = 1; j < 8; ++j) t += part[j][o];
    out[i] = t;
  }
}

// ---- pooling, upsampling, concatenation -----------------------------------

// Pieces of 8 channels (16 bytes): every channel count and offset here is a
// multiple of 16.
__device__ __forceinline__ void unpack8(uint4 v, float (&f)[8]) {
  const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f[2 * k] = bf2f(uint16_t(u[k]));
    f[2 * k + 1] = bf2f(uint16_t(u[k] >> 16));
  }
}
__device__ __forceinline__ uint4 pack8(const float (&f)[8]) {
  uint32_t u[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) u[k] = uint32_t(f2bf(f[2 * k])) | (uint32_t(f2bf(f[2 * k + 1])) << 16);
  return make_uint4(u[0], u[1], u[2], u[3]);
}

__global__ void k_pool_bwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
                           const uint16_t* __restrict__ add, int add_c, int add_off, uint16_t* __restrict__ dx, int c,
                           int H, int W, size_t n8, int relu_mask) {
  // one 2x2 window and 8 channels per thread: n8 = windows * c / 8
  const int H2 = H / 2, W2 = W / 2, c8 = c / 8;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n8; i += size_t(gridDim.x) * blockDim.x) {
    const int ch = int(i % c8) * 8;
    const size_t q = i / c8;
    const int w2 = int(q % W2);
    const size_t t = q / W2;
    const int h2 = int(t % H2);
    const size_t b = t / H2;
    const size_t p0 = (b * H + 2 * h2) * W + 2 * w2;
    const size_t px[4] = {p0, p0 + 1, p0 + W, p0 + W + 1};  // scan order
    float xv[4][8];
#pragma unroll
    for (int k = 0; k < 4; ++k) unpack8(*reinterpret_cast<const uint4*>(x + px[k] * c + ch), xv[k]);
    // per channel, the window's first maximum in scan order takes the gradient
    int arg[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float m = xv[0][e];
      arg[e] = 0;
#pragma unroll
      for (int k = 1; k < 4; ++k)
        if (xv[k][e] > m) { m = xv[k][e]; arg[e] = k; }
    }
    float up[8];
    unpack8(*reinterpret_cast<const uint4*>(dy + q * c + ch), up);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      float g[8], ad[8];
      if (add) unpack8(*reinterpret_cast<const uint4*>(add + px[k] * add_c + add_off + ch), ad);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        g[e] = k == arg[e] ? up[e] : 0.f;
        if (add) g[e] += ad[e];
        if (relu_mask && !(xv[k][e] > 0.f)) g[e] = 0.f;  // x is the ReLU output
      }
      *reinterpret_cast<uint4*>(dx + px[k] * c + ch) = pack8(g);
    }
  }
}

__global__ void k_upcat_bwd(const uint16_t* __restrict__ dout, int ca, int cb, uint16_t* __restrict__ da,
                            const uint16_t* __restrict__ da_mask, int H, int W, size_t n8) {
  // 8 channels of one half-resolution pixel per thread: n8 = pixels * ca / 8
  const int c = ca + cb, H2 = H / 2, W2 = W / 2, ca8 = ca / 8;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n8; i += size_t(gridDim.x) * blockDim.x) {
    const int ch = int(i % ca8) * 8;
    size_t q = i / ca8;
    const int w2 = int(q % W2);
    q /= W2;
    const int h2 = int(q % H2);
    const size_t b = q / H2;
    const size_t base = ((b * H + 2 * h2) * W + 2 * w2) * c + ch;
    float v0[8], v1[8], v2[8], v3[8], o[8];
    unpack8(*reinterpret_cast<const uint4*>(dout + base), v0);
    unpack8(*reinterpret_cast<const uint4*>(dout + base + c), v1);
    unpack8(*reinterpret_cast<const uint4*>(dout + base + size_t(W) * c), v2);
    unpack8(*reinterpret_cast<const uint4*>(dout + base + size_t(W) * c + c), v3);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float s = v0[e] + v1[e];
      s += v2[e] + v3[e];
      o[e] = s;
    }
    const size_t off = (i / ca8) * ca + ch;
    uint4 r = pack8(o);
    if (da_mask) r = mask8(r, *reinterpret_cast<const uint4*>(da_mask + off));
    *reinterpret_cast<uint4*>(da + off) = r;
  }
}

// ---- input, loss, output ---------------------------------------------------
constexpr float kHalvesClip = 2.f;  // bound of the halves' log difference (channels 9-11)

__global__ void k_pack(const float* __restrict__ beauty, const float* __restrict__ beauty_b,
                       const float* __restrict__ albedo, const float* __restrict__ normal, int h, int w,
                       uint16_t* __restrict__ x, int Hp, int Wp, size_t npx) {
  for (size_t p = blockIdx.x * size_t(blockDim.x) + threadIdx.x; p < npx; p += size_t(gridDim.x) * blockDim.x) {
    const int xx = int(p % Wp);
    const size_t t = p / Wp;
    const int yy = int(t % Hp);
    const size_t b = t / Hp;
    const size_t s = ((b * h + min(yy, h - 1)) * w + min(xx, w - 1)) * 3;
    uint16_t v[16];
#pragma unroll
    for (int c = 9; c < 16; ++c) v[c] = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float a = beauty[s + c];
      if (beauty_b) {
        const float b2 = beauty_b[s + c];
        v[c] = f2bf(log1pf(fmaxf(0.5f * (a + b2), 0.f)));
        // clipped: a firefly's log ratio (up to ~10) would take the network far outside
        // what it was trained on, while a ratio of e^2 already marks one half as an outlier
        v[9 + c] = f2bf(fminf(fmaxf(log1pf(fmaxf(a, 0.f)) - log1pf(fmaxf(b2, 0.f)), -kHalvesClip), kHalvesClip));
      } else {
        v[c] = f2bf(log1pf(fmaxf(a, 0.f)));
      }
      v[3 + c] = f2bf(albedo[s + c]);
      v[6 + c] = f2bf(normal[s + c]);
    }
    uint32_t u[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) u[k] = uint32_t(v[2 * k]) | (uint32_t(v[2 * k + 1]) << 16);
    uint4* o = reinterpret_cast<uint4*>(x + p * 16);  // the pixel's 32 bytes in two 16-byte stores
    o[0] = make_uint4(u[0], u[1], u[2], u[3]);
    o[1] = make_uint4(u[4], u[5], u[6], u[7]);
  }
}

__device__ __forceinline__ float tone(float z) { return z >= 0.f ? 1.f - expf(-z) : z; }
__device__ __forceinline__ float dtone(float z) { return z >= 0.f ? expf(-z) : 1.f; }
constexpr float kRelL2Eps = 0.01f;

// One element's loss and d loss / d z.
template <int KIND>
__device__ __forceinline__ float loss_elem(float zc, float tc, float* dl) {
  if (KIND == GT_DENOISE_LOSS_REL_L2) {
    const float y = zc >= 0.f ? expm1f(zc) : zc;
    const float dy = zc >= 0.f ? expf(zc) : 1.f;
    const float yc = fmaxf(y, 0.f);
    const float inv = 1.f / (yc * yc + kRelL2Eps);  // held constant: the gradient's root is the mean target
    const float d = y - tc;
    *dl = 2.f * d * inv * dy;
    return d * d * inv;
  }
  const float d = tone(zc) - tc / (1.f + tc);
  *dl = (d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f)) * dtone(zc);
  return fabsf(d);
}

template <int KIND>
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ z, const float* __restrict__ target, int h,
                                              int w, uint16_t* __restrict__ dz, float* __restrict__ partial, int Hp,
                                              int Wp, size_t npx, float inv_count) {
  __shared__ float red[256];
  float acc = 0.f;
  for (size_t p = blockIdx.x * size_t(blockDim.x) + threadIdx.x; p < npx; p += size_t(gridDim.x) * blockDim.x) {
    const int xx = int(p % Wp);
    const size_t t = p / Wp;
    const int yy = int(t % Hp);
    const size_t b = t / Hp;
    uint4* g = reinterpret_cast<uint4*>(dz + p * 16);  // channels 0-7, 8-15; only 0-2 carry a gradient
    g[1] = make_uint4(0, 0, 0, 0);
    if (yy >= h || xx >= w) {
      g[0] = make_uint4(0, 0, 0, 0);
      continue;
    }
    const float* tg = target + ((b * h + yy) * w + xx) * 3;
    const float4 zp = *reinterpret_cast<const float4*>(z + p * 16);
    const float zc[3] = {zp.x, zp.y, zp.z};
    uint16_t gc[3];
    for (int c = 0; c < 3; ++c) {
      float dl;
      acc += loss_elem<KIND>(zc[c], fmaxf(tg[c], 0.f), &dl);
      gc[c] = f2bf(dl * inv_count);
    }
    g[0] = make_uint4(uint32_t(gc[0]) | (uint32_t(gc[1]) << 16), uint32_t(gc[2]), 0, 0);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_unpack(const float* __restrict__ z, float* __restrict__ out, int h, int w, int Hp, int Wp,
                         size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const int c = int(i % 3);
    const size_t p = i / 3;
    const int xx = int(p % w);
    const size_t t = p / w;
    const int yy = int(t % h);
    const size_t b = t / h;
    out[i] = expm1f(fmaxf(z[((b * Hp + yy) * Wp + xx) * 16 + c], 0.f));
  }
}

// ---- optimiser ---------------------------------------------------------------
__global__ void k_adam_all(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                           const AdamSeg* __restrict__ segs, int nseg) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    for (int k = 0; k < nseg; ++k) {  // the segment holding i, if any (biases have no copies)
      const AdamSeg sg = segs[k];
      const size_t cin = size_t(sg.cin), cout = size_t(sg.cout);
      if (i < sg.w_off || i - sg.w_off >= cout * 9 * cin) continue;
      const size_t j = i - sg.w_off;
      const uint16_t b = f2bf(pi);
      sg.w16[j] = b;
      const size_t ci = j % cin, t = (j / cin) % 9, co = j / (cin * 9);
      sg.wt16[(ci * 9 + t) * cout + co] = b;
      break;
    }
  }
}

__global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ w16, uint16_t* __restrict__ wt16,
                          int cin, int cout, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const uint16_t b = f2bf(p[i]);
    w16[i] = b;
    const size_t ci = i % cin, t = (i / cin) % 9, co = i / (size_t(cin) * 9);
    wt16[(ci * 9 + t) * cout + co] = b;
  }
}

unsigned grid_for(size_t n, unsigned block = 256) {
  size_t g = (n + block - 1) / block;
  return unsigned(g < 65535u * 8 ? (g ? g : 1) : 65535u * 8);
}

template <bool DGRAD, bool F32, int NT>
void launch_nt(hipStream_t s, const uint16_t* x, const uint16_t* mask, int cin, const uint16_t* xa, int ca,
               const uint16_t* w, const float* bias, void* y, uint16_t* pooled, int cout, int co_first, int groups,
               Dims d, int relu, const uint16_t* omask) {
  const int nrt = (d.h + kTR - 1) / kTR, nct = (d.w + kTC - 1) / kTC;
  constexpr size_t lds = conv_lds_bytes<NT>();
  static const hipError_t attr =  // once per instantiation: allow more than 64 KiB of LDS
      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_conv<NT, DGRAD, F32>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  (void)attr;
  hipLaunchKernelGGL((k_conv<NT, DGRAD, F32>), dim3(d.n * nrt * nct, groups), dim3(256), lds, s, x, mask, cin, xa,
                     ca, w, bias, y, pooled, cout, co_first, d.h, d.w, nrt, nct, relu, omask);
}

// Output channels in groups of 64 (NT = 4), then one launch for the rest.
int device_cus() {
  static std::atomic<int> cus{0};
  int n = cus.load(std::memory_order_relaxed);
  if (!n) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
    cus.store(n, std::memory_order_relaxed);
  }
  return n;
}

// The persistent forward convolution, for the layer shapes where it is faster
// than k_conv on one MI355X: the default network's full-resolution layers and
// level 1's 32 -> 48 and 48 -> 48, tensors addressable by 31-bit buffer
// offsets. These fit two blocks per CU. False for the others, which take
// k_conv: the decoder's 64 + 48 -> 48 keeps 97 KB of weights resident, one
// block fits a CU, and it ran 40 % slower than k_conv.
template <int NT, bool F32, int CIN, int GMAX, bool MIRROR = false>
bool launch_pipe_nt(hipStream_t s, const uint16_t* x, const uint16_t* xa, int ca, const uint16_t* w,
                    const float* bias, void* y, uint16_t* pooled, int cout, Dims d, int relu, float* rad = nullptr,
                    int rh = 0, int rw = 0, const uint16_t* omask = nullptr, int co_first = 0) {
  if (size_t(d.n) * d.h * d.w * (CIN - ca) * 2 >= kOutOfRange) return false;
  if (xa && size_t(d.n) * (d.h / 2) * (d.w / 2) * ca * 2 >= kOutOfRange) return false;
  static const int per_cu = [] {  // resident blocks per CU
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_conv_pipe<NT, F32, CIN, GMAX, MIRROR>, 256, 0) !=
            hipSuccess ||
        n <= 0)
      n = 1;
    return n;
  }();
  const int nrt = (d.h + kTR - 1) / kTR, nct = (d.w + kTC - 1) / kTC;
  const int ntiles = d.n * nrt * nct;
  const int nb = (std::min(ntiles, device_cus() * per_cu) + 7) / 8 * 8;  // a multiple of the 8 XCDs
  hipLaunchKernelGGL((k_conv_pipe<NT, F32, CIN, GMAX, MIRROR>), dim3(nb, 1), dim3(256), 0, s, x, xa, ca, w, bias, y,
                     pooled, cout, co_first, d.n, d.h, d.w, nrt, nct, relu, rad, rh, rw, omask);
  return true;
}

// The data gradient (taps mirrored, input already masked) through the same
// kernel, for the full-resolution shapes that fit one group; omask zeroes the
// output where the forward tensor it belongs to is not positive.
bool launch_pipe_dgrad(hipStream_t s, const uint16_t* x, int cin, const uint16_t* w, void* y, int cout, Dims d,
                       const uint16_t* omask) {
  // wider outputs (the gradient of the decoder's concatenation, 80 or 112 channels from 32 or 48):
  // launches of 48 output channels, then one of the rest
  if (cout > 48 && cout % 16 == 0 && (cin == 32 || cin == 48)) {
    if (size_t(d.n) * d.h * d.w * std::max(cin, cout) * 2 >= kOutOfRange) return false;
    for (int c = 0; c < cout;) {
      const int nt = std::min(3, (cout - c) / 16);
      if (cin == 32) {
        if (nt == 3) launch_pipe_nt<3, false, 32, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0,
                                                            nullptr, 0, 0, omask, c);
        else if (nt == 2) launch_pipe_nt<2, false, 32, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0,
                                                                 nullptr, 0, 0, omask, c);
        else launch_pipe_nt<1, false, 32, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0, nullptr, 0,
                                                    0, omask, c);
      } else {
        if (nt == 3) launch_pipe_nt<3, false, 48, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0,
                                                            nullptr, 0, 0, omask, c);
        else if (nt == 2) launch_pipe_nt<2, false, 48, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0,
                                                                 nullptr, 0, 0, omask, c);
        else launch_pipe_nt<1, false, 48, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0, nullptr, 0,
                                                    0, omask, c);
      }
      c += 16 * nt;
    }
    return true;
  }
  switch ((cout / 16) * 1000 + cin) {
    case 2016:
      return launch_pipe_nt<2, false, 16, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0, nullptr, 0, 0,
                                                   omask);
    case 2032:
      return launch_pipe_nt<2, false, 32, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0, nullptr, 0, 0,
                                                   omask);
    case 2048:
      return launch_pipe_nt<2, false, 48, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0, nullptr, 0, 0,
                                                   omask);
    case 3048:
      return launch_pipe_nt<3, false, 48, 1, true>(s, x, nullptr, 0, w, nullptr, y, nullptr, cout, d, 0, nullptr, 0, 0,
                                                   omask);
    default: return false;
  }
}

template <bool F32>
bool launch_pipe(hipStream_t s, const uint16_t* x, int cin, const uint16_t* xa, int ca, const uint16_t* w,
                 const float* bias, void* y, uint16_t* pooled, int cout, Dims d, int relu) {
  const int key = (cout / 16) * 1000 + cin;
  if (F32) {
    if (key == 1032) return launch_pipe_nt<1, true, 32, 1>(s, x, xa, ca, w, bias, y, pooled, cout, d, relu);
    return false;
  }
  switch (key) {
    case 2016: return launch_pipe_nt<2, false, 16, 1>(s, x, xa, ca, w, bias, y, pooled, cout, d, relu);
    case 2032: return launch_pipe_nt<2, false, 32, 1>(s, x, xa, ca, w, bias, y, pooled, cout, d, relu);
    case 2080: return launch_pipe_nt<2, false, 80, 1>(s, x, xa, ca, w, bias, y, pooled, cout, d, relu);
    case 3032: return launch_pipe_nt<3, false, 32, 1>(s, x, xa, ca, w, bias, y, pooled, cout, d, relu);
    case 3048: return launch_pipe_nt<3, false, 48, 1>(s, x, xa, ca, w, bias, y, pooled, cout, d, relu);
    default: return false;
  }
}

template <bool DGRAD, bool F32>
void launch_groups(hipStream_t s, int nt, int first_frag, int groups, const uint16_t* x, const uint16_t* mask, int cin,
                   const uint16_t* xa, int ca, const uint16_t* w, const float* bias, void* y, int cout, Dims d,
                   int relu, uint16_t* pooled, const uint16_t* omask) {
  const int c0 = first_frag * 16;
  switch (nt) {
    case 1: launch_nt<DGRAD, F32, 1>(s, x, mask, cin, xa, ca, w, bias, y, pooled, cout, c0, groups, d, relu, omask); break;
    case 2: launch_nt<DGRAD, F32, 2>(s, x, mask, cin, xa, ca, w, bias, y, pooled, cout, c0, groups, d, relu, omask); break;
    case 3: launch_nt<DGRAD, F32, 3>(s, x, mask, cin, xa, ca, w, bias, y, pooled, cout, c0, groups, d, relu, omask); break;
    default: launch_nt<DGRAD, F32, 4>(s, x, mask, cin, xa, ca, w, bias, y, pooled, cout, c0, groups, d, relu, omask); break;
  }
}

// Output channels in groups of NT 16-channel fragments, then one launch for
// the rest. Levels with few tiles take narrower groups (more blocks, each
// staging the same halo for fewer channels), so that the launch still has
// blocks for most CUs.
template <bool DGRAD, bool F32>
void launch_conv(hipStream_t s, const uint16_t* x, const uint16_t* mask, int cin, const uint16_t* xa, int ca,
                 const uint16_t* w, const float* bias, void* y, int cout, Dims d, int relu,
                 uint16_t* pooled = nullptr, const uint16_t* omask = nullptr) {
  if (!DGRAD && launch_pipe<F32>(s, x, cin, xa, ca, w, bias, y, pooled, cout, d, relu)) return;
  if (DGRAD && !F32 && !mask && !xa && d.n * ((d.h + kTR - 1) / kTR) * ((d.w + kTC - 1) / kTC) >= 1024 &&
      launch_pipe_dgrad(s, x, cin, w, y, cout, d, omask))
    return;
  const int tiles = d.n * ((d.h + kTR - 1) / kTR) * ((d.w + kTC - 1) / kTC);
  const int nt = tiles >= 1024 ? 4 : tiles >= 256 ? 2 : 1;
  const int frags = cout / 16, full = frags / nt, rest = frags % nt;
  if (full) launch_groups<DGRAD, F32>(s, nt, 0, full, x, mask, cin, xa, ca, w, bias, y, cout, d, relu, pooled, omask);
  if (rest)
    launch_groups<DGRAD, F32>(s, rest, full * nt, 1, x, mask, cin, xa, ca, w, bias, y, cout, d, relu, pooled, omask);
}

// Weight-gradient splits: one per kWBands bands of kWR rows of one image.
int wgrad_splits(Dims d) { return d.n * (((d.h + kWR - 1) / kWR + kWBands - 1) / kWBands); }

template <int NCO, int NCI, int TC>
void launch_wgrad_tr(hipStream_t s, const uint16_t* x, int cin, const uint16_t* xa, int ca, const uint16_t* dy,
                     int cout, Dims d, float* partial, int ci_first, int ncig) {
  constexpr size_t lds = wgrad_tr_lds_bytes<NCO, NCI, TC>();
  static const hipError_t attr =  // once per instantiation: allow more than 64 KiB of LDS
      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wgrad_tr<NCO, NCI, TC>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  (void)attr;
  const int ncog = cout / (NCO * 16);
  const int nsplit = wgrad_splits(d) / d.n;
  hipLaunchKernelGGL((k_wgrad_tr<NCO, NCI, TC>), dim3(d.n * nsplit, ncog * ncig), dim3(256), lds, s, x, cin, xa, ca,
                     dy, cout, d.n, d.h, d.w, nsplit, ncog, ci_first, partial);
}


}  // namespace

void conv_fwd(hipStream_t s, const uint16_t* x, int cin, const uint16_t* w, const float* bias, void* y, int cout,
              Dims d, bool relu, bool out_f32, uint16_t* pooled) {
  if (out_f32) launch_conv<false, true>(s, x, nullptr, cin, nullptr, 0, w, bias, y, cout, d, relu ? 1 : 0);
  else launch_conv<false, false>(s, x, nullptr, cin, nullptr, 0, w, bias, y, cout, d, relu ? 1 : 0, pooled);
}

bool conv_fwd_radiance(hipStream_t s, const uint16_t* x, int cin, const uint16_t* w, const float* bias, float* out,
                       int h, int w_img, Dims d) {
  if (cin != 32) return false;
  return launch_pipe_nt<1, true, 32, 1>(s, x, nullptr, 0, w, bias, nullptr, nullptr, 16, d, 0, out, h, w_img);
}

void conv_fwd_upcat(hipStream_t s, const uint16_t* a, int ca, const uint16_t* b, int cb, const uint16_t* w,
                    const float* bias, uint16_t* y, int cout, Dims d, bool relu) {
  launch_conv<false, false>(s, b, nullptr, ca + cb, a, ca, w, bias, y, cout, d, relu ? 1 : 0);
}

void conv_dgrad(hipStream_t s, const uint16_t* dy, const uint16_t* y, int cout, const uint16_t* wt, uint16_t* dx,
                int cin, Dims d, const uint16_t* dx_mask) {
  // the transposed problem: input dy (cout channels), output dx (cin channels)
  launch_conv<true, false>(s, dy, y, cout, nullptr, 0, wt, nullptr, dx, cin, d, 0, nullptr, dx_mask);
}

size_t wgrad_scratch_floats(int cin, int cout, Dims d) {
  const size_t splits = size_t(wgrad_splits(d));
  return splits * (size_t(cout) * 9 * cin + cout);
}

void conv_wgrad(hipStream_t s, const uint16_t* x, int cin, const uint16_t* dy, int cout, Dims d, float* dw, float* db,
                float* scratch, size_t scratch_floats, const uint16_t* xa, int ca) {
  const int splits = wgrad_splits(d);
  const size_t nw = size_t(cout) * 9 * cin;
  if (scratch_floats < size_t(splits) * (nw + cout)) return;  // sized by wgrad_scratch_floats
  // a block's buffer resources span kWBands * kWR + 2 rows of the widest tensor (the network's
  // widths and 256 channels at most keep this far below 2 GB)
  if (size_t(kWBands * kWR + 2) * d.w * std::max(cin, cout) * 2 >= kOutOfRange) return;
  // Two output- and two input-channel fragments per block where the layer has them, so that
  // fewer blocks stage the same dPre and x tiles; an odd last input fragment takes a launch of its
  // own (it writes its own columns of the same partial rows).
  const int pairs = cin / 32, odd = (cin / 16) % 2;
  // Double-buffered tiles of 32 columns (two blocks of 2 x 2 fragments fit a CU with both buffers
  // only at that size; 1 x 2 and 1 x 1 also measured faster) or 64 (2 x 1).
  if ((cout / 16) % 2 == 0) {
    if (pairs) launch_wgrad_tr<2, 2, 32>(s, x, cin, xa, ca, dy, cout, d, scratch, 0, pairs);
    if (odd) launch_wgrad_tr<2, 1, 64>(s, x, cin, xa, ca, dy, cout, d, scratch, 32 * pairs, 1);
  } else {
    if (pairs) launch_wgrad_tr<1, 2, 32>(s, x, cin, xa, ca, dy, cout, d, scratch, 0, pairs);
    if (odd) launch_wgrad_tr<1, 1, 32>(s, x, cin, xa, ca, dy, cout, d, scratch, 32 * pairs, 1);
  }
  const size_t row = nw + cout;  // a split's partial sums: weights, then bias
  if (db == dw + nw) {  // the usual parameter layout: one reduction for both
    hipLaunchKernelGGL(k_reduce, dim3(unsigned((row + 31) / 32)), dim3(256), 0, s, scratch, splits, row, row, dw);
  } else {
    hipLaunchKernelGGL(k_reduce, dim3(unsigned((nw + 31) / 32)), dim3(256), 0, s, scratch, splits, nw, row, dw);
    hipLaunchKernelGGL(k_reduce, dim3(unsigned((size_t(cout) + 31) / 32)), dim3(256), 0, s, scratch + nw, splits,
                       size_t(cout), row, db);
  }
}

void pool_bwd(hipStream_t s, const uint16_t* x, const uint16_t* dy, const uint16_t* add, int add_c, int add_off,
              uint16_t* dx, int c, Dims d, bool relu_mask) {
  const size_t n8 = size_t(d.n) * (d.h / 2) * (d.w / 2) * (c / 8);  // windows x 8-channel pieces
  hipLaunchKernelGGL(k_pool_bwd, dim3(grid_for(n8)), dim3(256), 0, s, x, dy, add, add_c, add_off, dx, c, d.h, d.w,
                     n8, relu_mask ? 1 : 0);
}

void upcat_bwd(hipStream_t s, const uint16_t* dout, int ca, int cb, uint16_t* da, Dims d, const uint16_t* da_mask) {
  const size_t n8 = size_t(d.n) * (d.h / 2) * (d.w / 2) * (ca / 8);
  hipLaunchKernelGGL(k_upcat_bwd, dim3(grid_for(n8)), dim3(256), 0, s, dout, ca, cb, da, da_mask, d.h, d.w, n8);
}

void pack_input(hipStream_t s, const float* beauty, const float* beauty_b, const float* albedo, const float* normal,
                int h, int w, uint16_t* x, Dims dp) {
  const size_t n = size_t(dp.pixels());
  hipLaunchKernelGGL(k_pack, dim3(grid_for(n)), dim3(256), 0, s, beauty, beauty_b, albedo, normal, h, w, x, dp.h,
                     dp.w, n);
}

int loss_blocks(Dims dp) {
  const size_t n = size_t(dp.pixels());
  return int(std::min<size_t>((n + 255) / 256, 1024));
}

void loss(hipStream_t s, int kind, const float* z, const float* target, int h, int w, uint16_t* dz, float* partial,
          Dims dp) {
  const size_t n = size_t(dp.pixels());
  const float inv = 1.0f / float(size_t(dp.n) * h * w * 3);
  if (kind == GT_DENOISE_LOSS_REL_L2)
    hipLaunchKernelGGL(k_loss<GT_DENOISE_LOSS_REL_L2>, dim3(loss_blocks(dp)), dim3(256), 0, s, z, target, h, w, dz,
                       partial, dp.h, dp.w, n, inv);
  else
    hipLaunchKernelGGL(k_loss<GT_DENOISE_LOSS_L1_TONEMAP>, dim3(loss_blocks(dp)), dim3(256), 0, s, z, target, h, w,
                       dz, partial, dp.h, dp.w, n, inv);
}

void unpack_output(hipStream_t s, const float* z, float* out, int h, int w, Dims dp) {
  const size_t n = size_t(dp.n) * h * w * 3;
  hipLaunchKernelGGL(k_unpack, dim3(grid_for(n)), dim3(256), 0, s, z, out, h, w, dp.h, dp.w, n);
}

void adam_all(hipStream_t s, float* p, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
              float eps, float c1, float c2, const AdamSeg* segs, int nseg) {
  hipLaunchKernelGGL(k_adam_all, dim3(grid_for(n)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps, c1, c2, segs,
                     nseg);
}

void to_bf16(hipStream_t s, const float* p, uint16_t* w16, uint16_t* wt16, int cin, int cout) {
  const size_t n = size_t(cout) * 9 * cin;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(n)), dim3(256), 0, s, p, w16, wt16, cin, cout, n);
}

}  // namespace gdn

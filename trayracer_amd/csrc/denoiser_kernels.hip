// U-Net denoiser kernels for gfx950: BF16 activations/weights, FP32 accumulation on
// MFMA (v_mfma_f32_32x32x16_bf16), NHWC tensors.
//
// Replaces the OptiX AI denoiser the reference invokes (Tracer/Optix/Denoiser.cpp:36-42,
// 101-112; closed vendor code) with a project-trained U-Net.
//
// Convolution = implicit GEMM: M = output pixels, N = output channels, K = 9 taps x Cin.
// MFMA operand layouts (checked on the device): for 32x32x16, lane l holds A row l%32,
// k = 8*(l/32)..+7 and B column l%32, k = 8*(l/32)..+7; accumulator register r holds
// row 8*(r/4) + 4*(l/32) + r%4, column l%32.
#include <hip/hip_runtime.h>

#include "denoiser.h"

namespace trt {
namespace dn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (no NaNs occur here)
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(r >> 16);
}

constexpr int kTileH = 8, kTileW = 16;          // output pixels per workgroup
constexpr int kPatchH = kTileH + 2, kPatchW = kTileW + 2;
constexpr int kPatch = kPatchH * kPatchW;        // 180 staged input positions

// ------------------------------------------------------------------------------------
// 3x3 'same' convolution as an implicit GEMM (M = pixels, N = output channels,
// K = 9 taps x Cin), input and weights staged through LDS per 32-channel chunk.
//
// Input: `x` [N][H][W][cx] bf16, optionally preceded (in channel order) by `coarse`
// [N][H/2][W/2][c1] bf16 read with nearest upsampling, so a decoder convolution reads
// [upsample(coarse) | skip] without materialising the concatenation. Cin = c1 + cx;
// c1, cx multiples of 32, or Cin = 16 without coarse. Weights [Cout][9][Cin] bf16.
//
// Workgroup = 4 waves over a (8 MB) x 16 pixel tile and 32 NB output channels (one
// channel group); wave w owns MB m-blocks of 2 x 16 pixels (rows 2(w MB + m) + 0..1)
// and all NB n-blocks, so every A fragment feeds NB MFMAs and every B fragment MB.
// The next chunk's global loads are issued before the current chunk's MFMAs.
struct ConvArgs {
  const uint16_t* x;
  const uint16_t* coarse;
  int c1, cx;
  const uint16_t* w;
  const float* bias;  // nullable
  void* y;            // bf16 or fp32 [N][H][W][cout]; unused by kEpiRadiance
  int cout;
  uint16_t* pooled;     // kEpiReluPool: 2x2 average of the bf16 output [N][H/2][W/2][cout]
  const uint16_t* feat; // kEpiRadiance: input features [N][H][W][16]
  float4* radiance;     // kEpiRadiance: expm1(max(feat[0..2] + y[0..2], 0)), alpha 1
  int N, H, W;
  int groups;  // cout / (32 NB)
};

typedef unsigned v4u32_t __attribute__((ext_vector_type(4)));
// third dword of a raw buffer resource on gfx9 parts (32-bit data format, no swizzle)
constexpr int kBufFlags = 0x00020000;

// two floats -> packed bf16 (round to nearest even), low half first
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  const f32x2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// LDS layouts, in 16-byte slots. A 128-bit LDS read serves each half-wave (32 lanes) at
// once and is conflict-free when their 32 slots are distinct modulo 32. A-fragment
// reads take 16 consecutive patch positions of two consecutive rows: 5 slots per
// position (odd, so 16 positions are distinct mod 32) and a row pitch of 112 slots
// (= 16 mod 32, so the second row fills the other 16 residues). B-fragment reads take
// 32 consecutive output channels 37 slots apart (odd).
constexpr int kPosStride = 40;           // bf16 per staged position (32 + 8 padding)
constexpr int kRowPitch = 112 * 8;       // bf16 per staged patch row (18 positions + padding)
constexpr int kWStride = 9 * 32 + 8;     // bf16 per staged output channel

template <int MB, int NB, int EPI, bool SC, bool C16>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a) {
  constexpr int TH = 8 * MB, PH = TH + 2, PW = kTileW + 2, NPOS = PH * PW;
  constexpr int PLOADS = (NPOS * 4 + 255) / 256;
  constexpr int WLOADS = (32 * NB * 9 * 4 + 255) / 256;
  constexpr int OS = 32 * NB + 8;                                 // bf16 per pixel in the output staging
  constexpr int OWAVE = (EPI == kEpiF32 ? 8 : MB * 32 * OS);      // output staging per wave (bf16)
  constexpr int PARTS = 4 * NB;                                   // 16-byte parts per staged pixel
  constexpr int FLUSH = (MB * 32 * PARTS) / 64;                   // 16-byte stores per lane per tile
  constexpr int kOOB = 0x7fffffff;                                // buffer offset past every range
  // one spare 16-byte slot at the end of patch and wl takes the stores of idle threads
  __shared__ __align__(16) uint16_t patch[PH * kRowPitch + 8];
  __shared__ __align__(16) uint16_t wl[32 * NB * kWStride + 8];
  __shared__ __align__(16) uint16_t ob[4 * OWAVE];
  __shared__ __align__(16) float bl[32 * NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int H = a.H, W = a.W, Cin = a.c1 + a.cx, Cout = a.cout;
  constexpr int gs = C16 ? 1 : 2;    // log2 of 16-byte groups per staged position
  constexpr int cc = 8 << gs;        // channels per chunk (Cin = 16 is one 16-channel chunk)
  constexpr int KS = cc / 16;        // MFMA k-steps per tap
  constexpr int NK = 9 * KS;         // MFMA k-steps per chunk
  const int nchunks = Cin / cc;      // 1 when SC (single chunk)
  const int tiles_x = W / kTileW, tiles_y = (H + TH - 1) / TH;
  const int total = tiles_x * tiles_y * a.N * a.groups;
  // Persistent workgroups. Workgroups are dealt round-robin over the 8 XCDs, so XCD
  // x = blockIdx % 8 gets one contiguous range of items (neighbouring tiles share
  // their halo rows in that XCD's L2); its workgroups stride through the range by
  // `per`, a multiple of the group count, so one workgroup keeps one channel group.
  // Item order: channel group fastest, then tile column, tile row, image.
  const int per = static_cast<int>(gridDim.x) >> 3, xcd = static_cast<int>(blockIdx.x) & 7;
  const int q = total >> 3, rem = total & 7;
  const int begin = xcd * q + (xcd < rem ? xcd : rem);
  const int end = begin + q + (xcd < rem ? 1 : 0);
  const int first = begin + (static_cast<int>(blockIdx.x) >> 3);
  if (first >= end) return;
  struct Step {
    int it, ch;  // item, input-channel chunk
    int x0, y0, co0, n;
  };
  auto set_item = [&](Step& s, int it) __attribute__((always_inline)) {
    s.it = it;
    it = it < end ? it : end - 1;  // past the end: stage the last item again (discarded)
    const int g = it % a.groups;
    it /= a.groups;
    const int tx = it % tiles_x;
    it /= tiles_x;
    s.x0 = tx * kTileW;
    s.y0 = (it % tiles_y) * TH;
    s.n = it / tiles_y;
    s.co0 = g * 32 * NB;
  };
  auto advance = [&](Step s) __attribute__((always_inline)) {
    if (++s.ch == nchunks) {
      s.ch = 0;
      set_item(s, s.it + per);
    }
    return s;
  };

  f32x16 acc[MB][NB];
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[m][j] = f32x16{};

  // Staging: global -> registers -> LDS, the patch two steps ahead (two register sets),
  // the weights (reloaded per step unless SC) and features one step ahead, through
  // buffer loads: a buffer resource per image, per-thread byte offsets fixed for the
  // kernel (relative to the tile's corner), and out-of-range offsets reading zero (the
  // padding of the 'same' convolution). Every thread issues the same loads on every
  // step and none sits in a branch, so the wait before each LDS store counts exactly
  // the loads issued after it. Idle threads store to the spare slot.
  int pyx[PLOADS];  // patch position: row | column << 8
  int prel[PLOADS], crel[PLOADS], plds[PLOADS];
  uint32_t pidle = 0;
#pragma unroll
  for (int k = 0; k < PLOADS; ++k) {
    const int i = tid + k * 256;
    const int pos = i >> gs, gi = i & ((1 << gs) - 1), py = pos / PW, px = pos - py * PW;
    const bool ok = i < (NPOS << gs);
    pidle |= (ok ? 1u : 0u) << k;
    pyx[k] = py | (px << 8);
    prel[k] = ((py * W + px) * a.cx + gi * 8) * 2;
    crel[k] = (((py + 1) >> 1) * (W >> 1) + ((px + 1) >> 1)) * a.c1 * 2 + (gi * 8) * 2;
    plds[k] = ok ? py * kRowPitch + px * kPosStride + gi * 8 : PH * kRowPitch;
  }
  u32x4 prA[PLOADS], prB[PLOADS], wr[WLOADS];
  float br = 0.f;
  uint2 fr_next = make_uint2(0u, 0u), fr_cur = make_uint2(0u, 0u);  // kEpiRadiance features
  const uint32_t img_bytes = static_cast<uint32_t>(H) * W * a.cx * 2;
  const uint32_t cimg_bytes = static_cast<uint32_t>(H >> 1) * (W >> 1) * a.c1 * 2;
  auto load_patch = [&](u32x4(&pr)[PLOADS], const Step& s) __attribute__((always_inline)) {
    const int c0 = s.ch * cc;
    const bool coarse = c0 < a.c1;  // chunk of the nearest-upsampled coarse input
    const uint16_t* base = coarse ? a.coarse + static_cast<size_t>(s.n) * (H >> 1) * (W >> 1) * a.c1 + c0
                                  : a.x + static_cast<size_t>(s.n) * H * W * a.cx + (c0 - a.c1);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint16_t*>(base), static_cast<short>(0), coarse ? cimg_bytes : img_bytes, kBufFlags);
    // byte offset of the patch corner (row y0 - 1, column x0 - 1)
    const int corner = coarse ? (((s.y0 >> 1) - 1) * (W >> 1) + ((s.x0 >> 1) - 1)) * a.c1 * 2
                              : ((s.y0 - 1) * W + (s.x0 - 1)) * a.cx * 2;
    const bool interior = __builtin_amdgcn_readfirstlane(
        s.y0 >= 1 && s.y0 + PH - 1 <= H && s.x0 >= 1 && s.x0 + PW - 1 <= W ? 1 : 0);
    uint32_t okm = pidle;  // bit k: load k reads image data (else the zero padding)
    if (!interior) {       // uniform branch; it only narrows the mask
      uint32_t m = 0;
#pragma unroll
      for (int k = 0; k < PLOADS; ++k) {
        const int yy = s.y0 + (pyx[k] & 0xff) - 1, xx = s.x0 + ((pyx[k] >> 8) & 0xff) - 1;
        const bool ok = static_cast<unsigned>(yy) < static_cast<unsigned>(H) &&
                        static_cast<unsigned>(xx) < static_cast<unsigned>(W);
        m |= (ok ? 1u : 0u) << k;
      }
      okm &= m;
    }
#pragma unroll
    for (int k = 0; k < PLOADS; ++k) {
      int r0 = prel[k], r1 = crel[k];
      asm volatile("" : "+v"(r0), "+v"(r1));  // keep the per-load select a register select
      const int off = (okm >> k) & 1u ? corner + (coarse ? r1 : r0) : kOOB;
      pr[k] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
    }
  };
  int wrel[WLOADS];
#pragma unroll
  for (int k = 0; k < WLOADS; ++k) {
    const int i = tid + k * 256;
    const int qq = i >> gs, co = qq / 9, tap = qq - co * 9;
    wrel[k] = i < (32 * NB * 9 << gs) ? ((co * 9 + tap) * Cin + (i & ((1 << gs) - 1)) * 8) * 2 : kOOB;
  }
  const uint32_t w_bytes = static_cast<uint32_t>(Cout) * 9 * Cin * 2;
  auto load_weights = [&](const Step& s) __attribute__((always_inline)) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint16_t*>(a.w), static_cast<short>(0), w_bytes, kBufFlags);
    const int base = (s.co0 * 9 * Cin + s.ch * cc) * 2;
#pragma unroll
    for (int k = 0; k < WLOADS; ++k)
      wr[k] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, wrel[k] == kOOB ? kOOB : base + wrel[k], 0, 0));
    br = a.bias ? a.bias[s.co0 + (tid & (32 * NB - 1))] : 0.f;
  };
  auto load_feat = [&](const Step& s) __attribute__((always_inline)) {  // kEpiRadiance: features 0..3
    const int pi = lane & 31, yy = s.y0 + 2 * (wave * MB + (lane >> 5) % MB) + (pi >> 4);
    const size_t p = (static_cast<size_t>(s.n) * H + (yy < H ? yy : 0)) * W + s.x0 + (pi & 15);
    fr_next = *reinterpret_cast<const uint2*>(a.feat + p * 16);
  };
  auto store_patch = [&](const u32x4(&pr)[PLOADS]) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < PLOADS; ++k) *reinterpret_cast<u32x4*>(patch + plds[k]) = pr[k];
  };
  auto store_weights = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < WLOADS; ++k) {
      const int i = tid + k * 256;
      const int qq = i >> gs, co = qq / 9, tap = qq - co * 9;
      const int o = i < (32 * NB * 9 << gs) ? co * kWStride + tap * 32 + (i & ((1 << gs) - 1)) * 8 : 32 * NB * kWStride;
      *reinterpret_cast<u32x4*>(wl + o) = wr[k];
    }
    if (tid < 32 * NB) bl[tid] = br;
  };
  // Output of a finished tile, staged per wave in LDS by the epilogue ([pixel][channel]
  // bf16) and written one step later with 16-byte buffer stores (rows past the image
  // fall outside the buffer and are dropped), so that the stores are issued before the
  // following loads (one in-order counter tracks both) and never delay the wait for
  // them. The 2x2 pooling reads the staged bf16 values back.
  uint16_t* obw = ob + wave * OWAVE;
  int orel[FLUSH];
#pragma unroll
  for (int k = 0; k < FLUSH; ++k) {
    const int c = lane + 64 * k, px = c / PARTS, part = c - px * PARTS, pi = px & 31;
    orel[k] = (((2 * (wave * MB + (px >> 5)) + (pi >> 4)) * W + (pi & 15)) * Cout + part * 8) * 2;
  }
  const uint32_t out_bytes = static_cast<uint32_t>(H) * W * Cout * 2;
  auto flush = [&](const Step& s) __attribute__((always_inline)) {
    if (EPI == kEpiRadiance) {
      if (lane < 32 * MB) {
        const int pi = lane & 31, yy = s.y0 + 2 * (wave * MB + (lane >> 5)) + (pi >> 4);
        const float4 y = *reinterpret_cast<const float4*>(reinterpret_cast<const float*>(obw) + lane * 4);
        if (yy < H) {
          float4 o;
          o.x = expm1f(fmaxf(bf2f(static_cast<uint16_t>(fr_cur.x & 0xffffu)) + y.x, 0.f));
          o.y = expm1f(fmaxf(bf2f(static_cast<uint16_t>(fr_cur.x >> 16)) + y.y, 0.f));
          o.z = expm1f(fmaxf(bf2f(static_cast<uint16_t>(fr_cur.y & 0xffffu)) + y.z, 0.f));
          o.w = 1.f;
          a.radiance[(static_cast<size_t>(s.n) * H + yy) * W + s.x0 + (pi & 15)] = o;
        }
      }
      return;
    }
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        static_cast<uint16_t*>(a.y) + static_cast<size_t>(s.n) * H * W * Cout + s.co0, static_cast<short>(0),
        out_bytes, kBufFlags);
    const int corner = (s.y0 * W + s.x0) * Cout * 2;
#pragma unroll
    for (int k = 0; k < FLUSH; ++k) {
      const int c = lane + 64 * k, px = c / PARTS, part = c - px * PARTS;
      const u32x4 v = *reinterpret_cast<const u32x4*>(obw + px * OS + part * 8);
      __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32_t, v), rs, corner + orel[k], 0, 0);
    }
    if (EPI == kEpiReluPool) {  // H even: pooled rows are whole
      const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc(
          a.pooled + static_cast<size_t>(s.n) * (H >> 1) * (W >> 1) * Cout + s.co0, static_cast<short>(0),
          out_bytes >> 2, kBufFlags);
      for (int c = lane; c < MB * 8 * PARTS; c += 64) {
        const int pq = c / PARTS, part = c - pq * PARTS, m = pq >> 3, px = 2 * (pq & 7);
        const int yy = s.y0 + 2 * (wave * MB + m);
        const uint16_t* src = obw + (m * 32 + px) * OS + part * 8;
        const u32x4 v00 = *reinterpret_cast<const u32x4*>(src), v01 = *reinterpret_cast<const u32x4*>(src + OS);
        const u32x4 v10 = *reinterpret_cast<const u32x4*>(src + 16 * OS);
        const u32x4 v11 = *reinterpret_cast<const u32x4*>(src + 17 * OS);
        u32x4 o;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          // two channels per dword; top pair + bottom pair as in a stand-alone pooling
          const float lo = (__uint_as_float(v00[d] << 16) + __uint_as_float(v01[d] << 16)) +
                           (__uint_as_float(v10[d] << 16) + __uint_as_float(v11[d] << 16));
          const float hi = (__uint_as_float(v00[d] & 0xffff0000u) + __uint_as_float(v01[d] & 0xffff0000u)) +
                           (__uint_as_float(v10[d] & 0xffff0000u) + __uint_as_float(v11[d] & 0xffff0000u));
          o[d] = pack_bf16(0.25f * lo, 0.25f * hi);
        }
        const int off = (((yy >> 1) * (W >> 1) + (s.x0 >> 1) + (px >> 1)) * Cout + part * 8) * 2;
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32_t, o), rp, off, 0, 0);
      }
    }
  };

  const int l32 = lane & 31, khalf = (lane >> 5) * 8;
  bool pend = false;
  Step pend_step;
  // one step on the patch staged in `pr`: stage it, write the previous tile, issue the
  // loads of the following steps, then MFMAs and the epilogue of a finished tile
  auto step = [&](u32x4(&pr)[PLOADS], const Step& s0, const Step& s1, const Step& s2, bool valid)
                  __attribute__((always_inline)) {
    __syncthreads();
    store_patch(pr);
    if (!SC) store_weights();
    __syncthreads();
    if (pend) {
      flush(pend_step);
      pend = false;
    }
    if (EPI == kEpiRadiance) fr_cur = fr_next;
    if (!SC) load_weights(s1);
    if (EPI == kEpiRadiance) load_feat(s1);  // single-chunk: step s1 is the next item
    // issue order = wait order: nothing of the next step's loads may move behind these
    __builtin_amdgcn_sched_barrier(0);
    load_patch(pr, s2);
    // MFMA with the weights as A (rows = output channels) and the patch as B (columns =
    // pixels): accumulator register r of lane l then holds pixel l % 32 of the m-block
    // and channel 8 (r / 4) + 4 (l / 32) + r % 4 of the n-block, i.e. four runs of four
    // consecutive channels of one pixel. Fragments are read from LDS two k-steps ahead
    // of the MFMAs that use them (three register sets), so LDS latency overlaps MFMAs.
    bf16x8 af[3][MB], bf[3][NB];
    auto frag = [&](int k, int buf) __attribute__((always_inline)) {
      const int tap = k / KS, ks = (k % KS) * 16, ky = tap / 3, kx = tap % 3;
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int row = 2 * (wave * MB + m) + (l32 >> 4) + ky, col = (l32 & 15) + kx;
        af[buf][m] = *reinterpret_cast<const bf16x8*>(patch + row * kRowPitch + col * kPosStride + ks + khalf);
      }
#pragma unroll
      for (int j = 0; j < NB; ++j)
        bf[buf][j] = *reinterpret_cast<const bf16x8*>(wl + (j * 32 + l32) * kWStride + tap * 32 + ks + khalf);
    };
    frag(0, 0);
    frag(1, 1);
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      if (k + 2 < NK) frag(k + 2, (k + 2) % 3);
#pragma unroll
      for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int j = 0; j < NB; ++j)
          acc[m][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bf[k % 3][j], af[k % 3][m], acc[m][j], 0, 0, 0);
    }
    if (valid && s0.ch == nchunks - 1) {
      const int pi = l32, h4 = 4 * (lane >> 5);
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int yy = s0.y0 + 2 * (wave * MB + m) + (pi >> 4);
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int qd = 0; qd < 4; ++qd) {
            const int col = j * 32 + 8 * qd + h4;  // first of four channels
            const float4 bb = *reinterpret_cast<const float4*>(bl + col);
            float v0 = acc[m][j][4 * qd] + bb.x, v1 = acc[m][j][4 * qd + 1] + bb.y;
            float v2 = acc[m][j][4 * qd + 2] + bb.z, v3 = acc[m][j][4 * qd + 3] + bb.w;
            if (EPI == kEpiRelu || EPI == kEpiReluPool) {
              v0 = fmaxf(v0, 0.f);
              v1 = fmaxf(v1, 0.f);
              v2 = fmaxf(v2, 0.f);
              v3 = fmaxf(v3, 0.f);
            }
            if (EPI == kEpiF32) {
              if (yy < H)
                *reinterpret_cast<float4*>(static_cast<float*>(a.y) +
                                           ((static_cast<size_t>(s0.n) * H + yy) * W + s0.x0 + (pi & 15)) * Cout +
                                           s0.co0 + col) = make_float4(v0, v1, v2, v3);
            } else if (EPI == kEpiRadiance) {
              if (j == 0 && qd == 0 && h4 == 0)
                *reinterpret_cast<float4*>(reinterpret_cast<float*>(obw) + (m * 32 + pi) * 4) =
                    make_float4(v0, v1, v2, v3);
            } else {
              uint2 o;
              o.x = pack_bf16(v0, v1);
              o.y = pack_bf16(v2, v3);
              *reinterpret_cast<uint2*>(obw + (m * 32 + pi) * OS + col) = o;
            }
          }
#pragma unroll
        for (int j = 0; j < NB; ++j) acc[m][j] = f32x16{};
      }
      if (EPI != kEpiF32) {
        pend = true;
        pend_step = s0;
      }
    }
  };

  // step cursors: s0 is computed now, s1 and s2 follow. The loop runs pairs of steps
  // (register sets A, B) with its only exit test at the end; an odd last step is a B
  // step that computes a repeat of the last staged data and discards it.
  const int n_steps = ((end - first + per - 1) / per) * nchunks;
  Step s0;
  s0.ch = 0;
  set_item(s0, first);
  Step s1 = advance(s0), s2 = advance(s1);
  if (SC) {  // weights and bias once
    load_weights(s0);
    __syncthreads();
    store_weights();
  } else {
    load_weights(s0);
  }
  if (EPI == kEpiRadiance) load_feat(s0);
  __builtin_amdgcn_sched_barrier(0);
  load_patch(prA, s0);
  __builtin_amdgcn_sched_barrier(0);
  load_patch(prB, s1);
  for (int k = 0; k < n_steps; k += 2) {
    step(prA, s0, s1, s2, true);
    s0 = s1;
    s1 = s2;
    s2 = advance(s2);
    step(prB, s0, s1, s2, k + 1 < n_steps);
    s0 = s1;
    s1 = s2;
    s2 = advance(s2);
  }
  if (pend) flush(pend_step);
}

// ------------------------------------------------------------------------------------
// Weight gradient: dW[co][tap][ci] = sum_p dZ[p][co] X[p + s(tap)][ci] (X optionally
// [upsample(coarse) | X] as in the forward convolution), as an MFMA GEMM
// with M = co (32-block), N = ci (32-block), K = pixels. Each workgroup covers `tiles`
// consecutive 8x16 pixel tiles of one image row band and writes its partial sums to
// partial[chunk][co][tap][ci]; a fixed-order reduction over chunks follows.
// Grid: (n_chunks, Cout/32 * ceil(Cin/32)); wave w handles taps w, w+4, w+8.
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t* __restrict__ X, const uint16_t* __restrict__ coarse,
                                                int c1, const uint16_t* __restrict__ dZ, float* __restrict__ partial,
                                                int N, int H, int W, int Cin, int Cout, int tiles_per_chunk) {
  __shared__ uint16_t xs[32 * (kPatch + 8)];   // [ci][pos]
  __shared__ uint16_t zs[32 * (kTileH * kTileW + 8)];  // [co][pix]
  constexpr int XS = kPatch + 8, ZS = kTileH * kTileW + 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci_blocks = (Cin + 31) / 32;
  const int cb = blockIdx.y / ci_blocks, ib = blockIdx.y - cb * ci_blocks;
  const int tiles_x = W / kTileW, tiles_y = (H + kTileH - 1) / kTileH;
  const int tiles_img = tiles_x * tiles_y;
  const int total_tiles = tiles_img * N;
  f32x16 acc[3];
  for (int j = 0; j < 3; ++j)
    for (int r = 0; r < 16; ++r) acc[j][r] = 0.f;
  const int ntap = wave == 0 ? 3 : 2;  // taps {0,4,8}, {1,5}, {2,6}, {3,7}
  const int khalf = (lane >> 5) * 8;
  for (int t = 0; t < tiles_per_chunk; ++t) {
    const int tile = blockIdx.x * tiles_per_chunk + t;
    if (tile >= total_tiles) break;
    const int n = tile / tiles_img, tr = tile - n * tiles_img;
    const int ty = tr / tiles_x, tx = tr - ty * tiles_x;
    const int y0 = ty * kTileH, x0 = tx * kTileW;
    const size_t img = static_cast<size_t>(n) * H * W;
    const size_t img_c = static_cast<size_t>(n) * (H >> 1) * (W >> 1);
    const int cx = Cin - c1;
    __syncthreads();
    for (int i = threadIdx.x; i < 32 * kPatch; i += 256) {
      const int ci = i / kPatch, pos = i - ci * kPatch;
      const int py = pos / kPatchW, px = pos - py * kPatchW;
      const int yy = y0 + py - 1, xx = x0 + px - 1, c = ib * 32 + ci;
      uint16_t v = 0;
      if (c < Cin && yy >= 0 && yy < H && xx >= 0 && xx < W)
        v = c < c1 ? coarse[(img_c + static_cast<size_t>(yy >> 1) * (W >> 1) + (xx >> 1)) * c1 + c]
                   : X[(img + static_cast<size_t>(yy) * W + xx) * cx + (c - c1)];
      xs[ci * XS + pos] = v;
    }
    for (int i = threadIdx.x; i < 32 * kTileH * kTileW; i += 256) {
      const int co = i / (kTileH * kTileW), pix = i - co * (kTileH * kTileW);
      const int yy = y0 + pix / kTileW, xx = x0 + (pix % kTileW);
      uint16_t v = 0;
      if (yy < H) v = dZ[(img + static_cast<size_t>(yy) * W + xx) * Cout + cb * 32 + co];
      zs[co * ZS + pix] = v;
    }
    __syncthreads();
    for (int row = 0; row < kTileH; ++row) {  // K-step = 16 pixels of one tile row
      bf16x8 a;
      {
        const uint16_t* src = zs + (lane & 31) * ZS + row * kTileW + khalf;
        for (int v = 0; v < 8; ++v) a[v] = __builtin_bit_cast(__bf16, src[v]);
      }
      for (int j = 0; j < ntap; ++j) {
        const int tap = wave + 4 * j;
        const int ky = tap / 3, kx = tap - ky * 3;
        bf16x8 bv;
        const uint16_t* src = xs + (lane & 31) * XS + (row + ky) * kPatchW + kx + khalf;
        for (int v = 0; v < 8; ++v) bv[v] = __builtin_bit_cast(__bf16, src[v]);
        acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc[j], 0, 0, 0);
      }
    }
  }
  // write partial: rows = co (A rows), cols = ci (B cols)
  const size_t wsize = static_cast<size_t>(Cout) * 9 * Cin;
  float* out = partial + blockIdx.x * wsize;
  for (int j = 0; j < ntap; ++j) {
    const int tap = wave + 4 * j;
    const int ci = ib * 32 + (lane & 31);
    if (ci >= Cin) continue;
    for (int r = 0; r < 16; ++r) {
      const int co = cb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      out[(static_cast<size_t>(co) * 9 + tap) * Cin + ci] = acc[j][r];
    }
  }
}

// out[i] = sum over chunks c (in order) of partial[c][i]
__global__ void k_reduce_chunks(const float* __restrict__ partial, int n_chunks, size_t n, float* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float t = 0.f;
  for (int c = 0; c < n_chunks; ++c) t += partial[static_cast<size_t>(c) * n + i];
  out[i] = t;
}

// bias gradient partials: block b sums pixels [b*P/B, (b+1)*P/B) of dZ (bf16) per channel
__global__ void k_bias_partial(const uint16_t* __restrict__ dZ, size_t P, int C, int n_blocks,
                               float* __restrict__ partial) {
  const int c = threadIdx.x;
  if (c >= C) return;
  const size_t p0 = P * blockIdx.x / n_blocks, p1 = P * (blockIdx.x + 1) / n_blocks;
  float t = 0.f;
  for (size_t p = p0; p < p1; ++p) t += bf2f(dZ[p * C + c]);
  partial[static_cast<size_t>(blockIdx.x) * C + c] = t;
}

// dZ = bf16(dY * (Y > 0))  (ReLU backward; Y is the post-activation output)
__global__ void k_relu_bwd(const float* __restrict__ dY, const uint16_t* __restrict__ Y, uint16_t* __restrict__ dZ,
                           size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dZ[i] = bf2f(Y[i]) > 0.f ? f2bf(dY[i]) : static_cast<uint16_t>(0);
}

__global__ void k_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, size_t n) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

// dX (fine, fp32) (+)= dY(coarse)/4 broadcast; accumulate selects += vs =
__global__ void k_pool_bwd(const float* __restrict__ dY, float* __restrict__ dX, int N, int H, int W, int C,
                           int accumulate) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t total = static_cast<size_t>(N) * H * W * C;
  if (i >= total) return;
  const int c = static_cast<int>(i % C);
  size_t r = i / C;
  const int x = static_cast<int>(r % W);
  r /= W;
  const int y = static_cast<int>(r % H);
  const int n = static_cast<int>(r / H);
  const float g = 0.25f * dY[((static_cast<size_t>(n) * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
  dX[i] = accumulate ? dX[i] + g : g;
}

// coarse gradient: d_coarse[n][y][x][c] = sum of the 2x2 fine gradients of channel c < C1
__global__ void k_upcat_bwd_coarse(const float* __restrict__ dout, float* __restrict__ dcoarse, int N, int Hc,
                                   int Wc, int C1, int C2) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t total = static_cast<size_t>(N) * Hc * Wc * C1;
  if (i >= total) return;
  const int C = C1 + C2;
  const int c = static_cast<int>(i % C1);
  size_t r = i / C1;
  const int x = static_cast<int>(r % Wc);
  r /= Wc;
  const int y = static_cast<int>(r % Hc);
  const int n = static_cast<int>(r / Hc);
  const int W = 2 * Wc;
  const size_t b = ((static_cast<size_t>(n) * 2 * Hc + 2 * y) * W + 2 * x) * C + c;
  dcoarse[i] = (dout[b] + dout[b + C]) + (dout[b + static_cast<size_t>(W) * C] + dout[b + static_cast<size_t>(W) * C + C]);
}

// skip gradient: dskip (+)= dout[..][C1 + c]
__global__ void k_upcat_bwd_skip(const float* __restrict__ dout, float* __restrict__ dskip, size_t P, int C1, int C2,
                                 int accumulate) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= P * C2) return;
  const size_t p = i / C2;
  const int c = static_cast<int>(i % C2);
  const float g = dout[p * (C1 + C2) + C1 + c];
  dskip[i] = accumulate ? dskip[i] + g : g;
}

// Input features (16 channels, 10 used): log(1+radiance) rgb, albedo rgb (clamped to
// [0,1]), normal xyz, log(1+view depth); from the renderer's per-pixel buffers.
__global__ void k_features(const float4* __restrict__ color, const float4* __restrict__ albedo,
                           const float4* __restrict__ normal, uint16_t* __restrict__ out, size_t P) {
  const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float4 c = color[p], a = albedo[p], nd = normal[p];
  float f[16];
  f[0] = log1pf(fmaxf(c.x, 0.f));
  f[1] = log1pf(fmaxf(c.y, 0.f));
  f[2] = log1pf(fmaxf(c.z, 0.f));
  f[3] = fminf(fmaxf(a.x, 0.f), 1.f);
  f[4] = fminf(fmaxf(a.y, 0.f), 1.f);
  f[5] = fminf(fmaxf(a.z, 0.f), 1.f);
  f[6] = nd.x;
  f[7] = nd.y;
  f[8] = nd.z;
  f[9] = log1pf(fmaxf(nd.w, 0.f));
  for (int k = 10; k < 16; ++k) f[k] = 0.f;
  uint16_t* o = out + p * 16;
  for (int k = 0; k < 16; ++k) o[k] = f2bf(f[k]);
}

// denoised = (1 - b) denoised + b noisy (OptiX's blend factor, Tracer/Optix/Denoiser.cpp:108)
__global__ void k_blend(float4* __restrict__ den, const float4* __restrict__ color, float b, size_t P) {
  const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float4 d = den[p], c = color[p];
  den[p] = make_float4(fmaf(b, c.x - d.x, d.x), fmaf(b, c.y - d.y, d.y), fmaf(b, c.z - d.z, d.z), d.w);
}

// training target: log(1+radiance) of the reference image
__global__ void k_target(const float4* __restrict__ color, float* __restrict__ out, size_t P) {
  const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float4 c = color[p];
  out[p * 3 + 0] = log1pf(fmaxf(c.x, 0.f));
  out[p * 3 + 1] = log1pf(fmaxf(c.y, 0.f));
  out[p * 3 + 2] = log1pf(fmaxf(c.z, 0.f));
}

// crop gather: batch item b copies a (S x S) window at (x0, y0) of dataset image idx
__global__ void k_gather_crop(const uint16_t* __restrict__ feat, const float* __restrict__ tgt, int Hs, int Ws,
                              const int4* __restrict__ crops, int S, uint16_t* __restrict__ fo, float* __restrict__ to) {
  const int b = blockIdx.y;
  const int4 cr = crops[b];  // (image, x0, y0, flip)
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(S) * S) return;
  const int y = static_cast<int>(i / S), x = static_cast<int>(i % S);
  const int sx = cr.w ? cr.y + (S - 1 - x) : cr.y + x;
  const size_t src = (static_cast<size_t>(cr.x) * Hs + cr.z + y) * Ws + sx;
  const size_t dst = static_cast<size_t>(b) * S * S + i;
  const uint4* fs = reinterpret_cast<const uint4*>(feat + src * 16);
  uint4* fd = reinterpret_cast<uint4*>(fo + dst * 16);
  fd[0] = fs[0];
  fd[1] = fs[1];
  if (cr.w) {  // mirrored crop: flip the normal's x component sign bit (channel 6)
    uint16_t* h = fo + dst * 16;
    h[6] ^= 0x8000u;
  }
  to[dst * 3 + 0] = tgt[src * 3 + 0];
  to[dst * 3 + 1] = tgt[src * 3 + 1];
  to[dst * 3 + 2] = tgt[src * 3 + 2];
}

// L1 loss in log space. The network predicts a correction added to the noisy input
// (feature channels 0..2 = log(1 + radiance)): prediction = x + y over the first 3 of
// 32 output channels. Partial sums of |x + y - t| per block and dOut = sign(x + y - t) /
// (3 P) (zero on padding channels). L1 keeps the firefly pixels of many-emitter scenes
// from dominating the gradient (with an L2 loss the network trained on the soup and the
// instanced scene reached relMSE 0.0125 at 8 spp on a held-out soup view; with L1,
// 0.0085).
__global__ void k_loss(const float* __restrict__ Y, const uint16_t* __restrict__ X, const float* __restrict__ T,
                       size_t P, float* __restrict__ dY, float* __restrict__ partial) {
  __shared__ float red[256];
  const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (p < P) {
    const float inv = 1.f / (3.f * static_cast<float>(P));
    for (int c = 0; c < 32; ++c) {
      float g = 0.f;
      if (c < 3) {
        const float d = (bf2f(X[p * 16 + c]) + Y[p * 32 + c]) - T[p * 3 + c];
        l += fabsf(d);
        g = (d > 0.f ? inv : (d < 0.f ? -inv : 0.f));
      }
      dY[p * 32 + c] = g;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) on fp32 master weights; bf16 copy refreshed; `mask` zeroes
// updates of padding entries (nullable)
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                       uint16_t* __restrict__ wb, const uint8_t* __restrict__ mask, size_t n, float lr, float b1,
                       float b2, float c1, float c2, float eps) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (mask && !mask[i]) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

// Wd[ci][tap][co] = W[co][8 - tap][ci]: the data-gradient convolution is a 3x3
// convolution of dZ with the spatially flipped, channel-transposed kernel
__global__ void k_flip(const uint16_t* __restrict__ w, uint16_t* __restrict__ wd, int cin, int cout) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t n = static_cast<size_t>(cin) * 9 * cout;
  if (i >= n) return;
  const int ci = static_cast<int>(i / (9 * cout));
  const int r = static_cast<int>(i % (9 * cout));
  const int tap = r / cout, co = r % cout;
  wd[i] = w[(static_cast<size_t>(co) * 9 + (8 - tap)) * cin + ci];
}

inline unsigned blocks_for(size_t n, unsigned b = 256) { return static_cast<unsigned>((n + b - 1) / b); }

}  // namespace dn

using namespace dn;

// Persistent grid: as many workgroups as fit on the device at once (occupancy of the
// instantiation x CUs), a multiple of 8 (one share per XCD) whose per-XCD share is a
// multiple of the channel-group count, so a workgroup keeps one group's weights.
template <typename K>
static unsigned persistent_blocks(K kernel, long items, int groups) {
  static thread_local int cus = 0, dev = -1;
  int d = 0;
  (void)hipGetDevice(&d);
  if (d != dev) {
    dev = d;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || cus <= 0) cus = 256;
  }
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, 256, 0) != hipSuccess || occ <= 0) occ = 1;
  long per = long(cus) * occ / 8;
  per = per / groups * groups;
  if (per < groups) per = groups;
  const long need = (items + 7) / 8;
  if (per > need) per = (need + groups - 1) / groups * groups;
  return static_cast<unsigned>(per * 8);
}

template <int MB, int NB, int EPI>
static void launch_conv_epi(const ConvArgs& a, long items, hipStream_t s) {
  const int cin = a.c1 + a.cx;
  if (cin == 16) {
    const unsigned blocks = persistent_blocks(k_conv<MB, NB, EPI, true, true>, items, a.groups);
    hipLaunchKernelGGL((k_conv<MB, NB, EPI, true, true>), dim3(blocks), dim3(256), 0, s, a);
  } else if (cin == 32) {
    const unsigned blocks = persistent_blocks(k_conv<MB, NB, EPI, true, false>, items, a.groups);
    hipLaunchKernelGGL((k_conv<MB, NB, EPI, true, false>), dim3(blocks), dim3(256), 0, s, a);
  } else {
    const unsigned blocks = persistent_blocks(k_conv<MB, NB, EPI, false, false>, items, a.groups);
    hipLaunchKernelGGL((k_conv<MB, NB, EPI, false, false>), dim3(blocks), dim3(256), 0, s, a);
  }
}

template <int MB, int NB>
static void launch_conv(const ConvArgs& a, int epi, long items, hipStream_t s) {
  switch (epi) {
    case kEpiRelu: launch_conv_epi<MB, NB, kEpiRelu>(a, items, s); break;
    case kEpiF32: launch_conv_epi<MB, NB, kEpiF32>(a, items, s); break;
    case kEpiReluPool: launch_conv_epi<MB, NB, kEpiReluPool>(a, items, s); break;
    default: launch_conv_epi<MB, NB, kEpiRadiance>(a, items, s); break;
  }
}

// Tile shape per layer: 16-row tiles (MB = 2) unless that leaves < 1024 tiles, then
// 8-row tiles with 64-channel groups (NB = 2) when Cout is a multiple of 64.
bool dn_conv(const DnConv& c, hipStream_t s) {
  const int Cin = c.c1 + c.cx;
  const bool shape_ok = (Cin == 16 && c.c1 == 0) || (c.c1 % 32 == 0 && c.cx % 32 == 0 && Cin > 0);
  if (!shape_ok || c.cout % 32 || c.W % kTileW || c.H <= 0 || c.N <= 0) return false;
  if ((c.epi == kEpiReluPool || c.c1) && c.H % 2) return false;
  if (c.epi == kEpiRadiance && c.cout < 4) return false;
  if (c.c1 && !c.coarse) return false;
  // element offsets inside one image are 32-bit
  const long widest = c.cout > c.cx ? c.cout : c.cx;
  if (long(c.H) * c.W * widest >= (1L << 31)) return false;
  if (c.epi == kEpiRadiance && Cin > 32) return false;  // feature prefetch assumes one chunk
  const long tiles16 = long(c.W / kTileW) * ((c.H + 15) / 16) * c.N;
  const int MB = tiles16 >= 1024 ? 2 : 1;
  const long tiles = long(c.W / kTileW) * ((c.H + 8 * MB - 1) / (8 * MB)) * c.N;
  const int NB = MB == 1 && c.cout % 64 == 0 ? 2 : 1;  // MB = NB = 2 would need > 100 KB of LDS
  ConvArgs a;
  a.x = c.x;
  a.coarse = c.coarse;
  a.c1 = c.c1;
  a.cx = c.cx;
  a.w = c.w;
  a.bias = c.bias;
  a.y = c.y;
  a.cout = c.cout;
  a.pooled = c.pooled;
  a.feat = c.feat;
  a.radiance = c.radiance;
  a.N = c.N;
  a.H = c.H;
  a.W = c.W;
  a.groups = c.cout / (32 * NB);
  const long items = tiles * a.groups;
  if (MB == 2) {
    launch_conv<2, 1>(a, c.epi, items, s);
  } else {
    if (NB == 2) launch_conv<1, 2>(a, c.epi, items, s); else launch_conv<1, 1>(a, c.epi, items, s);
  }
  return true;
}

bool dn_conv3x3(const uint16_t* X, const uint16_t* Wt, const float* bias, void* Y, int N, int H, int W, int Cin,
                int Cout, bool relu, bool out_f32, hipStream_t s) {
  if (relu == out_f32) return false;  // bf16 + ReLU or fp32 linear
  DnConv c;
  c.x = X;
  c.cx = Cin;
  c.w = Wt;
  c.bias = bias;
  c.y = Y;
  c.cout = Cout;
  c.N = N;
  c.H = H;
  c.W = W;
  c.epi = relu ? kEpiRelu : kEpiF32;
  return dn_conv(c, s);
}

int dn_wgrad_chunks(int N, int H, int W, int tiles_per_chunk) {
  const int tiles = (W / kTileW) * ((H + kTileH - 1) / kTileH) * N;
  return (tiles + tiles_per_chunk - 1) / tiles_per_chunk;
}

void dn_wgrad(const uint16_t* X, const uint16_t* coarse, int c1, const uint16_t* dZ, float* partial, float* dW, int N,
              int H, int W, int Cin, int Cout, int tiles_per_chunk, hipStream_t s) {
  const int chunks = dn_wgrad_chunks(N, H, W, tiles_per_chunk);
  const dim3 grid(chunks, (Cout / 32) * ((Cin + 31) / 32));
  hipLaunchKernelGGL(k_wgrad, grid, dim3(256), 0, s, X, coarse, c1, dZ, partial, N, H, W, Cin, Cout,
                     tiles_per_chunk);
  const size_t n = static_cast<size_t>(Cout) * 9 * Cin;
  hipLaunchKernelGGL(k_reduce_chunks, dim3(blocks_for(n)), dim3(256), 0, s, partial, chunks, n, dW);
}

void dn_bias_grad(const uint16_t* dZ, size_t P, int C, float* partial, float* db, hipStream_t s) {
  const int nb = kBiasBlocks;
  hipLaunchKernelGGL(k_bias_partial, dim3(nb), dim3(256), 0, s, dZ, P, C, nb, partial);
  hipLaunchKernelGGL(k_reduce_chunks, dim3(blocks_for(C)), dim3(256), 0, s, partial, nb, static_cast<size_t>(C), db);
}

void dn_relu_bwd(const float* dY, const uint16_t* Y, uint16_t* dZ, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(n)), dim3(256), 0, s, dY, Y, dZ, n);
}
void dn_to_bf16(const float* a, uint16_t* b, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(n)), dim3(256), 0, s, a, b, n);
}
void dn_pool_bwd(const float* dY, float* dX, int N, int H, int W, int C, bool accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(static_cast<size_t>(N) * H * W * C)), dim3(256), 0, s, dY, dX, N, H,
                     W, C, accumulate ? 1 : 0);
}
void dn_upcat_bwd(const float* dout, float* dcoarse, float* dskip, int N, int H, int W, int C1, int C2,
                  bool accumulate_skip, hipStream_t s) {
  hipLaunchKernelGGL(k_upcat_bwd_coarse, dim3(blocks_for(static_cast<size_t>(N) * (H / 2) * (W / 2) * C1)),
                     dim3(256), 0, s, dout, dcoarse, N, H / 2, W / 2, C1, C2);
  const size_t P = static_cast<size_t>(N) * H * W;
  hipLaunchKernelGGL(k_upcat_bwd_skip, dim3(blocks_for(P * C2)), dim3(256), 0, s, dout, dskip, P, C1, C2,
                     accumulate_skip ? 1 : 0);
}
void dn_features(const float4* color, const float4* albedo, const float4* normal, uint16_t* out, size_t P,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_features, dim3(blocks_for(P)), dim3(256), 0, s, color, albedo, normal, out, P);
}
void dn_blend(float4* den, const float4* color, float b, size_t P, hipStream_t s) {
  hipLaunchKernelGGL(k_blend, dim3(blocks_for(P)), dim3(256), 0, s, den, color, b, P);
}
void dn_target(const float4* color, float* out, size_t P, hipStream_t s) {
  hipLaunchKernelGGL(k_target, dim3(blocks_for(P)), dim3(256), 0, s, color, out, P);
}
void dn_gather_crop(const uint16_t* feat, const float* tgt, int Hs, int Ws, const int4* crops, int B, int S,
                    uint16_t* fo, float* to, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_crop, dim3(blocks_for(static_cast<size_t>(S) * S), B), dim3(256), 0, s, feat, tgt, Hs,
                     Ws, crops, S, fo, to);
}
int dn_loss(const float* Y, const uint16_t* X, const float* T, size_t P, float* dY, float* partial, hipStream_t s) {
  const unsigned nb = blocks_for(P);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, Y, X, T, P, dY, partial);
  return static_cast<int>(nb);
}
void dn_adam(float* w, float* m, float* v, const float* g, uint16_t* wb, const uint8_t* mask, size_t n, float lr,
             float b1, float b2, float c1, float c2, float eps, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, s, w, m, v, g, wb, mask, n, lr, b1, b2, c1, c2, eps);
}

void dn_flip_weights(const uint16_t* w, uint16_t* wd, int cin, int cout, hipStream_t s) {
  hipLaunchKernelGGL(k_flip, dim3(blocks_for(static_cast<size_t>(cin) * 9 * cout)), dim3(256), 0, s, w, wd, cin, cout);
}

}  // namespace trt

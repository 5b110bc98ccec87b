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
constexpr int kChunk = 32;                       // input channels staged per step
constexpr int kPosStride = kChunk + 8;           // bf16 per LDS position (padding vs bank conflicts)

// ------------------------------------------------------------------------------------
// 3x3 'same' convolution. X [N][H][W][Cin] bf16, Wt [Cout][9][Cin] bf16, bias [Cout]
// fp32 (nullable). Output bf16 (optionally ReLU) or fp32. Cin % 16 == 0, Cout = 32*NB.
// Grid: (W/16, ceil(H/8), N); block 256 = 4 waves, wave w owns tile rows 2w, 2w+1.
template <int NB, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt,
                                                  const float* __restrict__ bias, void* __restrict__ Y, int H, int W,
                                                  int Cin) {
  __shared__ uint16_t patch[kPatch * kPosStride];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x0 = blockIdx.x * kTileW, y0 = blockIdx.y * kTileH, n = blockIdx.z;
  const int Cout = NB * 32;
  const size_t img = static_cast<size_t>(n) * H * W;
  f32x16 acc[NB];
  for (int b = 0; b < NB; ++b)
    for (int r = 0; r < 16; ++r) acc[b][r] = 0.f;
  // this lane's A row: pixel p = lane % 32 of the wave's 2x16 block
  const int p = lane & 31;
  const int prow = 2 * wave + (p >> 4), pcol = p & 15;
  const int khalf = (lane >> 5) * 8;  // k offset of this lane inside a 16-wide step

  for (int c0 = 0; c0 < Cin; c0 += kChunk) {
    const int cc = Cin - c0 < kChunk ? Cin - c0 : kChunk;  // 16 or 32
    __syncthreads();
    // stage the (10 x 18) x cc patch, 8 channels (16 bytes) per thread-iteration
    const int groups = cc / 8;
    for (int i = threadIdx.x; i < kPatch * groups; i += 256) {
      const int pos = i / groups, g = i - pos * groups;
      const int py = pos / kPatchW, px = pos - py * kPatchW;
      const int yy = y0 + py - 1, xx = x0 + px - 1;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (yy >= 0 && yy < H && xx >= 0 && xx < W)
        v = *reinterpret_cast<const uint4*>(X + (img + static_cast<size_t>(yy) * W + xx) * Cin + c0 + g * 8);
      *reinterpret_cast<uint4*>(patch + pos * kPosStride + g * 8) = v;
    }
    __syncthreads();
    for (int tap = 0; tap < 9; ++tap) {
      const int ky = tap / 3, kx = tap - ky * 3;
      const uint16_t* arow = patch + ((prow + ky) * kPatchW + (pcol + kx)) * kPosStride;
      for (int ks = 0; ks < cc; ks += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(arow + ks + khalf);
        for (int b = 0; b < NB; ++b) {
          const int co = b * 32 + (lane & 31);
          const bf16x8 bv =
              *reinterpret_cast<const bf16x8*>(Wt + (static_cast<size_t>(co) * 9 + tap) * Cin + c0 + ks + khalf);
          acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bv, acc[b], 0, 0, 0);
        }
      }
    }
  }
  // epilogue
  for (int b = 0; b < NB; ++b) {
    const int co = b * 32 + (lane & 31);
    const float bb = bias ? bias[co] : 0.f;
    for (int r = 0; r < 16; ++r) {
      const int pi = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      const int yy = y0 + 2 * wave + (pi >> 4), xx = x0 + (pi & 15);
      if (yy >= H) continue;
      float v = acc[b][r] + bb;
      if (RELU) v = v > 0.f ? v : 0.f;
      const size_t o = (img + static_cast<size_t>(yy) * W + xx) * Cout + co;
      if (OUT_F32)
        static_cast<float*>(Y)[o] = v;
      else
        static_cast<uint16_t*>(Y)[o] = f2bf(v);
    }
  }
}

// ------------------------------------------------------------------------------------
// Weight gradient: dW[co][tap][ci] = sum_p dZ[p][co] X[p + s(tap)][ci], as an MFMA GEMM
// with M = co (32-block), N = ci (32-block), K = pixels. Each workgroup covers `tiles`
// consecutive 8x16 pixel tiles of one image row band and writes its partial sums to
// partial[chunk][co][tap][ci]; a fixed-order reduction over chunks follows.
// Grid: (n_chunks, Cout/32 * ceil(Cin/32)); wave w handles taps w, w+4, w+8.
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t* __restrict__ X, const uint16_t* __restrict__ dZ,
                                                float* __restrict__ partial, int N, int H, int W, int Cin, int Cout,
                                                int tiles_per_chunk) {
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
    __syncthreads();
    for (int i = threadIdx.x; i < 32 * kPatch; i += 256) {
      const int ci = i / kPatch, pos = i - ci * kPatch;
      const int py = pos / kPatchW, px = pos - py * kPatchW;
      const int yy = y0 + py - 1, xx = x0 + px - 1, c = ib * 32 + ci;
      uint16_t v = 0;
      if (c < Cin && yy >= 0 && yy < H && xx >= 0 && xx < W) v = X[(img + static_cast<size_t>(yy) * W + xx) * Cin + c];
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

// 2x2 average pooling, bf16 -> bf16 (H, W even)
__global__ void k_pool(const uint16_t* __restrict__ X, uint16_t* __restrict__ Y, int N, int H, int W, int C) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  const size_t total = static_cast<size_t>(N) * Ho * Wo * C;
  if (i >= total) return;
  const int c = static_cast<int>(i % C);
  size_t r = i / C;
  const int x = static_cast<int>(r % Wo);
  r /= Wo;
  const int y = static_cast<int>(r % Ho);
  const int n = static_cast<int>(r / Ho);
  const size_t b = ((static_cast<size_t>(n) * H + 2 * y) * W + 2 * x) * C + c;
  const float s = (bf2f(X[b]) + bf2f(X[b + C])) + (bf2f(X[b + static_cast<size_t>(W) * C]) +
                                                    bf2f(X[b + static_cast<size_t>(W) * C + C]));
  Y[i] = f2bf(0.25f * s);
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

// out[n][y][x] = [ coarse[n][y/2][x/2][0:C1] | skip[n][y][x][0:C2] ]
__global__ void k_upcat(const uint16_t* __restrict__ coarse, const uint16_t* __restrict__ skip,
                        uint16_t* __restrict__ out, int N, int H, int W, int C1, int C2) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int C = C1 + C2;
  const size_t total = static_cast<size_t>(N) * H * W * C;
  if (i >= total) return;
  const int c = static_cast<int>(i % C);
  size_t r = i / C;
  const int x = static_cast<int>(r % W);
  r /= W;
  const int y = static_cast<int>(r % H);
  const int n = static_cast<int>(r / H);
  out[i] = c < C1 ? coarse[((static_cast<size_t>(n) * (H / 2) + y / 2) * (W / 2) + x / 2) * C1 + c]
                  : skip[((static_cast<size_t>(n) * H + y) * W + x) * C2 + (c - C1)];
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

// L2 loss in log space. The network predicts a correction added to the noisy input
// (feature channels 0..2 = log(1 + radiance)): prediction = x + y over the first 3 of
// 32 output channels. Partial sums per block and dOut = 2 (x + y - t) / (3 P) (zero
// on padding channels).
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
        l += d * d;
        g = 2.f * d * inv;
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

// network output (log-space correction of the input) -> radiance
__global__ void k_output(const float* __restrict__ Y, const uint16_t* __restrict__ X, float4* __restrict__ out,
                         size_t P) {
  const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float v[3];
  for (int c = 0; c < 3; ++c) v[c] = expm1f(fmaxf(bf2f(X[p * 16 + c]) + Y[p * 32 + c], 0.f));
  out[p] = make_float4(v[0], v[1], v[2], 1.f);
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

#define TRT_CONV_CASE(NBV)                                                                                       \
  case NBV:                                                                                                      \
    if (out_f32)                                                                                                 \
      hipLaunchKernelGGL((k_conv3x3<NBV, false, true>), grid, dim3(256), 0, s, X, Wt, bias, Y, H, W, Cin);     \
    else if (relu)                                                                                               \
      hipLaunchKernelGGL((k_conv3x3<NBV, true, false>), grid, dim3(256), 0, s, X, Wt, bias, Y, H, W, Cin);     \
    else                                                                                                         \
      hipLaunchKernelGGL((k_conv3x3<NBV, false, false>), grid, dim3(256), 0, s, X, Wt, bias, Y, H, W, Cin);    \
    break;

bool dn_conv3x3(const uint16_t* X, const uint16_t* Wt, const float* bias, void* Y, int N, int H, int W, int Cin,
                int Cout, bool relu, bool out_f32, hipStream_t s) {
  if (Cin % 16 || Cout % 32 || W % kTileW || Cout > 224) return false;
  const dim3 grid(W / kTileW, (H + kTileH - 1) / kTileH, N);
  switch (Cout / 32) {
    TRT_CONV_CASE(1)
    TRT_CONV_CASE(2)
    TRT_CONV_CASE(3)
    TRT_CONV_CASE(4)
    TRT_CONV_CASE(5)
    TRT_CONV_CASE(6)
    TRT_CONV_CASE(7)
    default:
      return false;
  }
  return true;
}

int dn_wgrad_chunks(int N, int H, int W, int tiles_per_chunk) {
  const int tiles = (W / kTileW) * ((H + kTileH - 1) / kTileH) * N;
  return (tiles + tiles_per_chunk - 1) / tiles_per_chunk;
}

void dn_wgrad(const uint16_t* X, const uint16_t* dZ, float* partial, float* dW, int N, int H, int W, int Cin, int Cout,
              int tiles_per_chunk, hipStream_t s) {
  const int chunks = dn_wgrad_chunks(N, H, W, tiles_per_chunk);
  const dim3 grid(chunks, (Cout / 32) * ((Cin + 31) / 32));
  hipLaunchKernelGGL(k_wgrad, grid, dim3(256), 0, s, X, dZ, partial, N, H, W, Cin, Cout, tiles_per_chunk);
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
void dn_pool(const uint16_t* X, uint16_t* Y, int N, int H, int W, int C, hipStream_t s) {
  hipLaunchKernelGGL(k_pool, dim3(blocks_for(static_cast<size_t>(N) * (H / 2) * (W / 2) * C)), dim3(256), 0, s, X, Y,
                     N, H, W, C);
}
void dn_pool_bwd(const float* dY, float* dX, int N, int H, int W, int C, bool accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_pool_bwd, dim3(blocks_for(static_cast<size_t>(N) * H * W * C)), dim3(256), 0, s, dY, dX, N, H,
                     W, C, accumulate ? 1 : 0);
}
void dn_upcat(const uint16_t* coarse, const uint16_t* skip, uint16_t* out, int N, int H, int W, int C1, int C2,
              hipStream_t s) {
  hipLaunchKernelGGL(k_upcat, dim3(blocks_for(static_cast<size_t>(N) * H * W * (C1 + C2))), dim3(256), 0, s, coarse,
                     skip, out, N, H, W, C1, C2);
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
void dn_output(const float* Y, const uint16_t* X, float4* out, size_t P, hipStream_t s) {
  hipLaunchKernelGGL(k_output, dim3(blocks_for(P)), dim3(256), 0, s, Y, X, out, P);
}
void dn_adam(float* w, float* m, float* v, const float* g, uint16_t* wb, const uint8_t* mask, size_t n, float lr,
             float b1, float b2, float c1, float c2, float eps, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, s, w, m, v, g, wb, mask, n, lr, b1, b2, c1, c2, eps);
}

void dn_flip_weights(const uint16_t* w, uint16_t* wd, int cin, int cout, hipStream_t s) {
  hipLaunchKernelGGL(k_flip, dim3(blocks_for(static_cast<size_t>(cin) * 9 * cout)), dim3(256), 0, s, w, wd, cin, cout);
}

}  // namespace trt

// U-Net denoiser: kernel launchers (denoiser_kernels.hip) and the host-side network
// (denoiser.cpp). Replaces the reference's Denoiser (Tracer/Optix/Denoiser.h:14-38),
// which wraps the closed OptiX AI denoiser.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace trt {

constexpr int kBiasBlocks = 256;
constexpr int kFeatures = 16;  // input channels (10 used: radiance, albedo, normal, depth)

// One 3x3 convolution (denoiser_kernels.hip). Input [upsample(coarse) | x] (coarse
// nullable, c1 = 0), weights [cout][9][c1 + cx], bias nullable. Epilogues:
//   kEpiRelu      y = bf16(relu(conv + bias))
//   kEpiF32       y = fp32(conv + bias)
//   kEpiReluPool  as kEpiRelu, plus pooled = bf16(2x2 mean of the bf16 y)
//   kEpiRadiance  radiance[p] = (expm1(max(feat[p][c] + conv + bias, 0)) for c < 3, 1)
enum DnEpi { kEpiRelu = 0, kEpiF32 = 1, kEpiReluPool = 2, kEpiRadiance = 3 };
struct DnConv {
  const uint16_t* x = nullptr;
  const uint16_t* coarse = nullptr;
  int c1 = 0, cx = 0;
  const uint16_t* w = nullptr;
  const float* bias = nullptr;
  void* y = nullptr;
  int cout = 0;
  uint16_t* pooled = nullptr;
  const uint16_t* feat = nullptr;
  float4* radiance = nullptr;
  int N = 0, H = 0, W = 0;
  int epi = kEpiRelu;
};
// false when the shape is unsupported (channels % 32 or Cin = 16, Cout % 32, W % 16,
// odd H with pooling or a coarse input)
bool dn_conv(const DnConv& c, hipStream_t s);
bool dn_conv3x3(const uint16_t* X, const uint16_t* Wt, const float* bias, void* Y, int N, int H, int W, int Cin,
                int Cout, bool relu, bool out_f32, hipStream_t s);
int dn_wgrad_chunks(int N, int H, int W, int tiles_per_chunk);
void dn_wgrad(const uint16_t* X, const uint16_t* coarse, int c1, const uint16_t* dZ, float* partial, float* dW, int N,
              int H, int W, int Cin, int Cout, int tiles_per_chunk, hipStream_t s);
void dn_bias_grad(const uint16_t* dZ, size_t P, int C, float* partial, float* db, hipStream_t s);
void dn_relu_bwd(const float* dY, const uint16_t* Y, uint16_t* dZ, size_t n, hipStream_t s);
void dn_to_bf16(const float* a, uint16_t* b, size_t n, hipStream_t s);
void dn_pool_bwd(const float* dY, float* dX, int N, int H, int W, int C, bool accumulate, hipStream_t s);
void dn_upcat_bwd(const float* dout, float* dcoarse, float* dskip, int N, int H, int W, int C1, int C2,
                  bool accumulate_skip, hipStream_t s);
void dn_features(const float4* color, const float4* albedo, const float4* normal, uint16_t* out, size_t P,
                 hipStream_t s);
void dn_target(const float4* color, float* out, size_t P, hipStream_t s);
// den = (1 - b) den + b color (rgb; alpha kept)
void dn_blend(float4* den, const float4* color, float b, size_t P, hipStream_t s);
void dn_gather_crop(const uint16_t* feat, const float* tgt, int Hs, int Ws, const int4* crops, int B, int S,
                    uint16_t* fo, float* to, hipStream_t s);
int dn_loss(const float* Y, const uint16_t* X, const float* T, size_t P, float* dY, float* partial, hipStream_t s);
void dn_adam(float* w, float* m, float* v, const float* g, uint16_t* wb, const uint8_t* mask, size_t n, float lr,
             float b1, float b2, float c1, float c2, float eps, hipStream_t s);
void dn_flip_weights(const uint16_t* w, uint16_t* wd, int cin, int cout, hipStream_t s);

// Network: 4-level U-Net, 32/64/96/128 channels, 3x3 convolutions with ReLU, 2x2
// average pooling, nearest upsampling + skip concatenation, a final 3x3 convolution to
// 3 channels (padded to 32). It predicts a log(1 + radiance) correction that is added
// to the noisy input's log(1 + radiance), so an untrained (near-zero) output layer is
// the identity and more input samples keep improving the result.
class Unet {
 public:
  struct Conv {
    int cin, cout;
    bool relu;
    size_t woff, boff;  // offsets into the parameter vector (weights [cout][9][cin], bias [cout])
  };
  Unet();
  ~Unet();
  size_t num_params() const { return n_params_; }
  const std::vector<Conv>& convs() const { return convs_; }
  hipError_t init(uint32_t seed, hipStream_t s);
  hipError_t set_params(const float* host, hipStream_t s);
  hipError_t get_params(float* host, hipStream_t s);
  // forward on features [N][H][W][16] (device). With radiance == nullptr: output fp32
  // [N][H][W][32] in output(), activations kept for backward(). Otherwise the last layer
  // writes the denoised radiance (x + y mapped back from log space) and output() and
  // backward() are not available until the next such forward.
  hipError_t forward(const uint16_t* x, int N, int H, int W, hipStream_t s, float4* radiance = nullptr);
  const float* output() const { return out_; }
  // training step on a batch; returns the loss through *loss (synchronises)
  hipError_t train_step(const uint16_t* x, const float* target, int N, int H, int W, float lr, hipStream_t s,
                        float* loss);
  // gradient only (no parameter update), for tests: dOut given as dL/dY of the fp32 output
  hipError_t backward(const float* dout, int N, int H, int W, hipStream_t s);
  hipError_t get_grads(float* host, hipStream_t s);
  // intermediate activation by conv index (bf16, device), for tests
  const uint16_t* activation(int conv) const { return act_[conv]; }
  // element count of conv output `conv` for the last forward (0 for the fp32 output layer)
  size_t activation_elems(int conv) const;
  bool shape_ok(int H, int W) const { return H % 8 == 0 && W % 128 == 0; }
  uint32_t adam_steps() const { return t_; }

 private:
  hipError_t ensure(int N, int H, int W);
  void free_acts();
  std::vector<Conv> convs_;
  size_t n_params_ = 0;
  float *w_ = nullptr, *m_ = nullptr, *v_ = nullptr, *g_ = nullptr;  // fp32 master, Adam moments, grads
  uint16_t *wb_ = nullptr, *wd_ = nullptr;                              // bf16 weights, flipped for dgrad
  // activations (bf16) per conv output, plus the pooled encoder outputs
  std::vector<uint16_t*> act_;
  uint16_t *p0_ = nullptr, *p1_ = nullptr, *p2_ = nullptr;
  float* out_ = nullptr;
  bool have_output_ = false;  // out_ holds the last forward's output
  const uint16_t* x_ = nullptr;  // input of the last forward (read by the first layer's wgrad)
  // backward scratch
  float *gA_ = nullptr, *gB_ = nullptr, *gs0_ = nullptr, *gs1_ = nullptr, *gs2_ = nullptr, *dout_ = nullptr;
  uint16_t* dz_ = nullptr;
  float *wpart_ = nullptr, *bpart_ = nullptr, *lpart_ = nullptr;
  size_t wpart_elems_ = 0;
  int N_ = 0, H_ = 0, W_ = 0;
  uint32_t t_ = 0;
  std::vector<void*> allocs_;
};

}  // namespace trt

// Host side of the U-Net denoiser: parameters, activations, forward, backward, Adam.
// Every reduction (weight/bias gradients, loss) runs in a fixed order, so training is
// bitwise reproducible for a given seed.
#include "denoiser.h"

#include <cmath>
#include <cstring>
#include <random>

namespace trt {

namespace {
// conv index -> role
enum { E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2A, D2B, D1A, D1B, D0A, D0B, OUT, NCONV };
constexpr int kWgradTiles = 8;
}  // namespace

Unet::Unet() {
  const int spec[NCONV][3] = {{kFeatures, 32, 1}, {32, 32, 1},  {32, 64, 1},  {64, 64, 1}, {64, 96, 1},
                              {96, 96, 1},        {96, 128, 1}, {128, 128, 1}, {224, 96, 1}, {96, 96, 1},
                              {160, 64, 1},       {64, 64, 1},  {96, 32, 1},  {32, 32, 1}, {32, 32, 0}};
  size_t off = 0;
  for (auto& c : spec) {
    Conv cv;
    cv.cin = c[0];
    cv.cout = c[1];
    cv.relu = c[2] != 0;
    cv.woff = off;
    off += size_t(cv.cout) * 9 * cv.cin;
    cv.boff = off;
    off += size_t(cv.cout);
    convs_.push_back(cv);
  }
  n_params_ = off;
  act_.assign(NCONV, nullptr);
}

Unet::~Unet() {
  free_acts();
  for (void* p : {static_cast<void*>(w_), static_cast<void*>(m_), static_cast<void*>(v_), static_cast<void*>(g_),
                  static_cast<void*>(wb_), static_cast<void*>(wd_)})
    if (p) (void)hipFree(p);
}

void Unet::free_acts() {
  for (void* p : allocs_) (void)hipFree(p);
  allocs_.clear();
  N_ = H_ = W_ = 0;
}

hipError_t Unet::init(uint32_t seed, hipStream_t s) {
  if (!w_) {
    hipError_t e;
    if ((e = hipMalloc(&w_, n_params_ * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&m_, n_params_ * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&v_, n_params_ * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&g_, n_params_ * 4)) != hipSuccess) return e;
    if ((e = hipMalloc(&wb_, n_params_ * 2)) != hipSuccess) return e;
    if ((e = hipMalloc(&wd_, n_params_ * 2)) != hipSuccess) return e;
  }
  // He-normal initialisation over the real fan-in (10 input features, 3 outputs)
  std::vector<float> h(n_params_, 0.f);
  std::mt19937 rng(seed);
  for (size_t li = 0; li < convs_.size(); ++li) {
    const Conv& c = convs_[li];
    const int fan_cin = li == E0A ? 10 : c.cin;
    const int real_cout = li == OUT ? 3 : c.cout;
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * fan_cin)) * (li == OUT ? 0.1f : 1.f));
    for (int co = 0; co < real_cout; ++co)
      for (int t = 0; t < 9; ++t)
        for (int ci = 0; ci < (li == E0A ? 10 : c.cin); ++ci) h[c.woff + (size_t(co) * 9 + t) * c.cin + ci] = nd(rng);
  }
  return set_params(h.data(), s);
}

hipError_t Unet::set_params(const float* host, hipStream_t s) {
  hipError_t e;
  if ((e = hipMemcpyAsync(w_, host, n_params_ * 4, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(m_, 0, n_params_ * 4, s)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(v_, 0, n_params_ * 4, s)) != hipSuccess) return e;
  dn_to_bf16(w_, wb_, n_params_, s);
  for (size_t li = 1; li < convs_.size(); ++li)
    dn_flip_weights(wb_ + convs_[li].woff, wd_ + convs_[li].woff, convs_[li].cin, convs_[li].cout, s);
  t_ = 0;
  return hipStreamSynchronize(s);
}

hipError_t Unet::get_params(float* host, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(host, w_, n_params_ * 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t Unet::get_grads(float* host, hipStream_t s) {
  hipError_t e = hipMemcpyAsync(host, g_, n_params_ * 4, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t Unet::ensure(int N, int H, int W) {
  if (N == N_ && H == H_ && W == W_) return hipSuccess;
  free_acts();
  auto alloc = [&](size_t bytes, void** p) {
    hipError_t e = hipMalloc(p, bytes < 256 ? 256 : bytes);
    if (e == hipSuccess) allocs_.push_back(*p);
    return e;
  };
  const size_t P0 = size_t(N) * H * W, P1 = P0 / 4, P2 = P0 / 16, P3 = P0 / 64;
  const size_t lvl[NCONV] = {P0, P0, P1, P1, P2, P2, P3, P3, P2, P2, P1, P1, P0, P0, P0};
  hipError_t e;
  for (int i = 0; i < OUT; ++i)
    if ((e = alloc(lvl[i] * convs_[i].cout * 2, reinterpret_cast<void**>(&act_[i]))) != hipSuccess) return e;
  act_[OUT] = nullptr;
  if ((e = alloc(P1 * 32 * 2, reinterpret_cast<void**>(&p0_))) != hipSuccess) return e;
  if ((e = alloc(P2 * 64 * 2, reinterpret_cast<void**>(&p1_))) != hipSuccess) return e;
  if ((e = alloc(P3 * 96 * 2, reinterpret_cast<void**>(&p2_))) != hipSuccess) return e;
  if ((e = alloc(P0 * 32 * 4, reinterpret_cast<void**>(&out_))) != hipSuccess) return e;
  // backward scratch: the largest per-pixel gradient tensor is 96 ch at full resolution
  const size_t gmax = P0 * 96;
  if ((e = alloc(gmax * 4, reinterpret_cast<void**>(&gA_))) != hipSuccess) return e;
  if ((e = alloc(gmax * 4, reinterpret_cast<void**>(&gB_))) != hipSuccess) return e;
  if ((e = alloc(P0 * 32 * 4, reinterpret_cast<void**>(&gs0_))) != hipSuccess) return e;
  if ((e = alloc(P1 * 64 * 4, reinterpret_cast<void**>(&gs1_))) != hipSuccess) return e;
  if ((e = alloc(P2 * 96 * 4, reinterpret_cast<void**>(&gs2_))) != hipSuccess) return e;
  if ((e = alloc(P0 * 32 * 4, reinterpret_cast<void**>(&dout_))) != hipSuccess) return e;
  if ((e = alloc(gmax * 2, reinterpret_cast<void**>(&dz_))) != hipSuccess) return e;
  size_t wp = 0;
  for (int i = 0; i < NCONV; ++i) {
    const int hh = (i <= E0B || i >= D0A) ? H : (i <= E1B || i >= D1A) ? H / 2 : (i <= E2B || i >= D2A) ? H / 4 : H / 8;
    const int ww = W * hh / H;
    const size_t chunks = size_t(dn_wgrad_chunks(N, hh, ww, kWgradTiles));
    const size_t need = chunks * convs_[i].cout * 9 * convs_[i].cin;
    if (need > wp) wp = need;
  }
  wpart_elems_ = wp;
  if ((e = alloc(wp * 4, reinterpret_cast<void**>(&wpart_))) != hipSuccess) return e;
  if ((e = alloc(size_t(kBiasBlocks) * 256 * 4, reinterpret_cast<void**>(&bpart_))) != hipSuccess) return e;
  if ((e = alloc((P0 / 256 + 2) * 4, reinterpret_cast<void**>(&lpart_))) != hipSuccess) return e;
  N_ = N;
  H_ = H;
  W_ = W;
  return hipSuccess;
}

size_t Unet::activation_elems(int i) const {
  if (i < 0 || i >= OUT || N_ == 0) return 0;
  const size_t P0 = size_t(N_) * H_ * W_;
  const size_t lvl[NCONV] = {P0, P0, P0 / 4, P0 / 4, P0 / 16, P0 / 16, P0 / 64, P0 / 64,
                             P0 / 16, P0 / 16, P0 / 4, P0 / 4, P0, P0, P0};
  return lvl[i] * size_t(convs_[size_t(i)].cout);
}

hipError_t Unet::forward(const uint16_t* x, int N, int H, int W, hipStream_t s, float4* radiance) {
  if (!w_ || !shape_ok(H, W)) return hipErrorInvalidValue;
  hipError_t e = ensure(N, H, W);
  if (e != hipSuccess) return e;
  x_ = x;
  have_output_ = false;
  const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
  bool ok = true;
  // conv i on [upsample(coarse) | in] at h x w; pooled: fused 2x2 pooling of the output
  auto conv = [&](int i, const uint16_t* coarse, int c1, const uint16_t* in, int h, int w, uint16_t* pooled) {
    const Conv& c = convs_[i];
    DnConv d;
    d.x = in;
    d.coarse = coarse;
    d.c1 = c1;
    d.cx = c.cin - c1;
    d.w = wb_ + c.woff;
    d.bias = w_ + c.boff;
    d.y = act_[i];
    d.cout = c.cout;
    d.pooled = pooled;
    d.N = N;
    d.H = h;
    d.W = w;
    d.epi = pooled ? kEpiReluPool : kEpiRelu;
    ok = ok && dn_conv(d, s);
  };
  conv(E0A, nullptr, 0, x, H, W, nullptr);
  conv(E0B, nullptr, 0, act_[E0A], H, W, p0_);
  conv(E1A, nullptr, 0, p0_, H1, W1, nullptr);
  conv(E1B, nullptr, 0, act_[E1A], H1, W1, p1_);
  conv(E2A, nullptr, 0, p1_, H2, W2, nullptr);
  conv(E2B, nullptr, 0, act_[E2A], H2, W2, p2_);
  conv(E3A, nullptr, 0, p2_, H3, W3, nullptr);
  conv(E3B, nullptr, 0, act_[E3A], H3, W3, nullptr);
  conv(D2A, act_[E3B], 128, act_[E2B], H2, W2, nullptr);
  conv(D2B, nullptr, 0, act_[D2A], H2, W2, nullptr);
  conv(D1A, act_[D2B], 96, act_[E1B], H1, W1, nullptr);
  conv(D1B, nullptr, 0, act_[D1A], H1, W1, nullptr);
  conv(D0A, act_[D1B], 64, act_[E0B], H, W, nullptr);
  conv(D0B, nullptr, 0, act_[D0A], H, W, nullptr);
  const Conv& c = convs_[OUT];
  DnConv d;
  d.x = act_[D0B];
  d.cx = c.cin;
  d.w = wb_ + c.woff;
  d.bias = w_ + c.boff;
  d.cout = c.cout;
  d.N = N;
  d.H = H;
  d.W = W;
  if (radiance) {
    d.feat = x;
    d.radiance = radiance;
    d.epi = kEpiRadiance;
  } else {
    d.y = out_;
    d.epi = kEpiF32;
  }
  ok = ok && dn_conv(d, s);
  if (!ok) return hipErrorInvalidValue;
  have_output_ = radiance == nullptr;
  return hipGetLastError();
}

hipError_t Unet::backward(const float* dout, int N, int H, int W, hipStream_t s) {
  if (N != N_ || H != H_ || W != W_ || !x_ || !have_output_) return hipErrorInvalidValue;
  const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4, H3 = H / 8, W3 = W / 8;
  // one conv layer: gradient in (fp32, post-activation) -> dW, db, gradient out (fp32)
  auto layer = [&](int i, const uint16_t* coarse, int c1, const uint16_t* in, int h, int w, const float* gin,
                   float* gout) {
    const Conv& c = convs_[i];
    const size_t P = size_t(N) * h * w;
    if (c.relu)
      dn_relu_bwd(gin, act_[i], dz_, P * c.cout, s);
    else
      dn_to_bf16(gin, dz_, P * c.cout, s);
    dn_wgrad(in, coarse, c1, dz_, wpart_, g_ + c.woff, N, h, w, c.cin, c.cout, kWgradTiles, s);
    dn_bias_grad(dz_, P, c.cout, bpart_, g_ + c.boff, s);
    if (gout) dn_conv3x3(dz_, wd_ + c.woff, nullptr, gout, N, h, w, c.cout, c.cin, false, true, s);
  };
  layer(OUT, nullptr, 0, act_[D0B], H, W, dout, gA_);
  layer(D0B, nullptr, 0, act_[D0A], H, W, gA_, gB_);
  layer(D0A, act_[D1B], 64, act_[E0B], H, W, gB_, gA_);  // gA = d u0 (96 ch @H)
  dn_upcat_bwd(gA_, gB_, gs0_, N, H, W, 64, 32, false, s);  // gB = d d1b (64 @H1), gs0 = d e0b (skip part)
  layer(D1B, nullptr, 0, act_[D1A], H1, W1, gB_, gA_);
  layer(D1A, act_[D2B], 96, act_[E1B], H1, W1, gA_, gB_);  // gB = d u1 (160 @H1)
  dn_upcat_bwd(gB_, gA_, gs1_, N, H1, W1, 96, 64, false, s);  // gA = d d2b (96 @H2), gs1 = d e1b
  layer(D2B, nullptr, 0, act_[D2A], H2, W2, gA_, gB_);
  layer(D2A, act_[E3B], 128, act_[E2B], H2, W2, gB_, gA_);  // gA = d u2 (224 @H2)
  dn_upcat_bwd(gA_, gB_, gs2_, N, H2, W2, 128, 96, false, s);  // gB = d e3b (128 @H3), gs2 = d e2b
  layer(E3B, nullptr, 0, act_[E3A], H3, W3, gB_, gA_);
  layer(E3A, nullptr, 0, p2_, H3, W3, gA_, gB_);  // gB = d p2 (96 @H3)
  dn_pool_bwd(gB_, gs2_, N, H2, W2, 96, true, s);  // gs2 += pool backward
  layer(E2B, nullptr, 0, act_[E2A], H2, W2, gs2_, gA_);
  layer(E2A, nullptr, 0, p1_, H2, W2, gA_, gB_);  // gB = d p1 (64 @H2)
  dn_pool_bwd(gB_, gs1_, N, H1, W1, 64, true, s);
  layer(E1B, nullptr, 0, act_[E1A], H1, W1, gs1_, gA_);
  layer(E1A, nullptr, 0, p0_, H1, W1, gA_, gB_);  // gB = d p0 (32 @H1)
  dn_pool_bwd(gB_, gs0_, N, H, W, 32, true, s);
  layer(E0B, nullptr, 0, act_[E0A], H, W, gs0_, gA_);
  layer(E0A, nullptr, 0, x_, H, W, gA_, nullptr);  // the input gradient of the network is not needed
  return hipGetLastError();
}

hipError_t Unet::train_step(const uint16_t* x, const float* target, int N, int H, int W, float lr, hipStream_t s,
                            float* loss) {
  hipError_t e = forward(x, N, H, W, s);
  if (e != hipSuccess) return e;
  const size_t P = size_t(N) * H * W;
  const int nb = dn_loss(out_, x, target, P, dout_, lpart_, s);
  e = backward(dout_, N, H, W, s);
  if (e != hipSuccess) return e;
  ++t_;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, float(t_)), c2 = 1.f - std::pow(b2, float(t_));
  dn_adam(w_, m_, v_, g_, wb_, nullptr, n_params_, lr, b1, b2, c1, c2, 1e-8f, s);
  for (size_t li = 1; li < convs_.size(); ++li)
    dn_flip_weights(wb_ + convs_[li].woff, wd_ + convs_[li].woff, convs_[li].cin, convs_[li].cout, s);
  std::vector<float> lp(static_cast<size_t>(nb));
  if ((e = hipMemcpyAsync(lp.data(), lpart_, size_t(nb) * 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  double t = 0.0;
  for (float v : lp) t += v;
  if (loss) *loss = float(t / (3.0 * double(P)));
  return hipSuccess;
}

}  // namespace trt

// Counter-based PCG32 keyed by (seed, pixel, sample, depth).
//
// Replaces the reference's TEA-2 + 24-bit LCG seeding (thirdparty/CUDA/random.h:31-67,
// seeded at Optix/GPU/optix.cu:247 and CUDA/GPU/Shade/PathTracing.h:103). Every path
// vertex owns an independent PCG32 stream (O'Neill 2014, pcg32 "XSH RR" output,
// 64-bit LCG state, odd increment), initialised from a splitmix64 hash of its key, so
// any kernel (forward, shadow, adjoint replay) can recreate the same random numbers
// from the key alone, with no per-path RNG state in memory.
#pragma once
#include "common.h"

namespace trt {

struct Pcg32 {
  uint64_t state;
  uint64_t inc;
};

TRT_FI uint32_t pcg32_next(Pcg32& r) {
  const uint64_t old = r.state;
  r.state = old * 6364136223846793005ULL + r.inc;
  const uint32_t xorshifted = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = static_cast<uint32_t>(old >> 59u);
  return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

// pcg32_srandom_r of the reference C implementation (state 0, inc = seq<<1|1, step,
// add initstate, step)
TRT_FI void pcg32_seed(Pcg32& r, uint64_t initstate, uint64_t initseq) {
  r.state = 0u;
  r.inc = (initseq << 1u) | 1u;
  pcg32_next(r);
  r.state += initstate;
  pcg32_next(r);
}

// splitmix64 finalizer (Steele, Lea, Flood 2014)
TRT_FI uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Depth code of the camera stream (the camera draws jitter and lens samples from it).
constexpr uint32_t kCameraDepth = 0xffffu;

// Stream for one path vertex: seed (render seed), pixel (y*W+x), sample (global
// sample index of that pixel, counting over frames), depth (bounce, or kCameraDepth).
TRT_FI Pcg32 rng_for(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t depth) {
  Pcg32 r;
  const uint64_t k0 = mix64((static_cast<uint64_t>(pixel) << 32) | sample);
  const uint64_t k1 = mix64((static_cast<uint64_t>(seed) << 32) | depth);
  pcg32_seed(r, k0, k1 ^ 0x9e3779b97f4a7c15ULL);
  return r;
}

// uniform float in [0,1) with 24 random bits
TRT_FI float rng_float(Pcg32& r) { return static_cast<float>(pcg32_next(r) >> 8) * 0x1p-24f; }

}  // namespace trt

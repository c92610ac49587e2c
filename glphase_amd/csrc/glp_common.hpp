// glp_common.hpp -- definitions shared by host and device code of the
// GLPhase MH surrogate-parent engine for gfx950.
//
// The random streams here replace curand XORWOW (hmmLike.cu:163,188,196,
// 220-226) and the host GSL draws of HMMLike::RunHMMOnSamples
// (hmmLike.cpp:91-102) with a counter-based Philox4x32-10 generator, so a
// proposal's draws depend only on (seed, generation, window, sample, proposal)
// and not on the order in which lanes evaluate proposals.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GLP_HD __host__ __device__ __forceinline__

#if defined(__HIPCC__)
namespace glp {
// A device-memory pointer in the global address space.  Loads through it are
// global loads, counted by vmcnt alone; through a generic pointer (as read
// from a descriptor struct) they are flat loads, which also count against
// lgkmcnt, so every wait for an LDS read would wait for them as well.
template <class T>
using gptr_t = const __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr_t<T> gmem(const T *p) {
  return (gptr_t<T>)p;
}
template <class T>
using gmut_t = __attribute__((address_space(1))) T *;

// 2 and 4 floats at one aligned global address, as one load or store (the
// HIP vector types cannot bind global-address-space references; the
// compiler's vector types can)
typedef float glp_f2v __attribute__((ext_vector_type(2)));
typedef float glp_f4v __attribute__((ext_vector_type(4)));
typedef uint32_t glp_u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 load_f2(gptr_t<float> p) {
  const glp_f2v v = *reinterpret_cast<gptr_t<glp_f2v>>(p);
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 load_f4(gptr_t<float> p) {
  const glp_f4v v = *reinterpret_cast<gptr_t<glp_f4v>>(p);
  return make_float4(v.x, v.y, v.z, v.w);
}
typedef unsigned glp_u4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 load_u4(gptr_t<uint32_t> p) {
  const glp_u4v v = *reinterpret_cast<gptr_t<glp_u4v>>(p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store_f4(float *p, float a, float b, float c, float d) {
  glp_f4v v;
  v.x = a;
  v.y = b;
  v.z = c;
  v.w = d;
  *(gmut_t<glp_f4v>)p = v;  // a C cast: the one that may change address space
}

// 1.0f / x rounded to nearest-even, for 2^-149 <= x < 2^-20.  The operand is
// scaled by 2^64 (exact), inverted by v_rcp_f32 plus one Newton step with
// fused multiply-adds, and scaled back (exact, or the same overflow to inf
// as the division).  For x = 0 it gives NaN where the division gives inf.
__device__ __forceinline__ float recip_small(float x) {
  const float xs = x * 0x1p64f;
  const float r = __builtin_amdgcn_rcpf(xs);
  const float e = __builtin_fmaf(-xs, r, 1.0f);
  return __builtin_fmaf(e, r, r) * 0x1p64f;
}

// The same without the scaling, two instructions shorter: rounded to
// nearest-even for 2^-126 <= x <= 2^126 (normal x and result), NaN for
// subnormal x and for 0 (v_rcp_f32 flushes a subnormal input), 0 past
// 2^126.  Callers detect the NaN (it carries through their recurrences) and
// redo the work with an exact division.  glp_selftest_recip checks both
// reciprocals against the IEEE division for every float of these ranges.
__device__ __forceinline__ float recip_normal(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  const float e = __builtin_fmaf(-x, r, 1.0f);
  return __builtin_fmaf(e, r, r);
}
}  // namespace glp
#endif

namespace glp {

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11).
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

struct U4 {
  uint32_t x, y, z, w;
};

GLP_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

GLP_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    const uint32_t hi0 = mulhi32(kPhiloxM0, c.x), lo0 = kPhiloxM0 * c.x;
    const uint32_t hi1 = mulhi32(kPhiloxM1, c.z), lo1 = kPhiloxM1 * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

// Counter layout: {k, sample, window, generation}; key = 64-bit seed.
// k < 2^31 is MH proposal k; kInitTag draws the 4 fresh parents of a
// generation; kWorkTag | (site >> 2) draws hmm_work's per-site uniforms;
// kHapInitTag, with the site in place of the window and the window in place
// of the generation, draws a sample's initial alleles at one site.
constexpr uint32_t kInitTag = 0xFFFFFFFFu;
constexpr uint32_t kWorkTag = 0x80000000u;
constexpr uint32_t kHapInitTag = 0xFFFFFFFEu;

GLP_HD U4 draw(uint64_t seed, uint32_t k, uint32_t sample, uint32_t window,
               uint32_t gen) {
  return philox4x32_10(U4{k, sample, window, gen}, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}

// floor(x * n / 2^32): a uniform integer of [0, n)
GLP_HD uint32_t uniform_int(uint32_t x, uint32_t n) { return mulhi32(x, n); }

// UnifSampler::SampleHap (sampler.cpp:29-39): uniform over [0, H) minus the
// sample's own haplotypes 2s and 2s+1, in closed form.
GLP_HD uint32_t unif_excl(uint32_t x, uint32_t nhaps, uint32_t sample) {
  const uint32_t r = uniform_int(x, nhaps - 2u);
  return r + (r >= 2u * sample ? 2u : 0u);
}

// a double of [0, 1) with 32 random bits (the resolution of gsl_rng_uniform)
GLP_HD double u01(uint32_t x) { return (double)x * (1.0 / 4294967296.0); }

// Proposal-selection modes.
enum PropMode : int {
  kPropUniform = 0,  // candidates uniform over all other haplotypes
  kPropKList = 1,    // candidates uniform over a per-sample list of K haps
  kPropHost = 2,     // initial parents and candidates supplied by the host
};

enum Precision : int {
  kPrecF32 = 0,    // FP32 recurrence and FP32 log accumulation (reference)
  kPrecMixed = 1,  // FP32 recurrence, FP64 log accumulation and MH ratio
  kPrecMixedFB = 2,  // the same arithmetic over two half-window passes joined in FP64
};

}  // namespace glp

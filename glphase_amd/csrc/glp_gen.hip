// glp_gen.hip -- the per-generation bookkeeping around the MH update on
// gfx950, and its part of the C ABI (include/glphase_amd.h):
//  - k_init_haps: Insti::initialize's random initial haplotypes from the GLs
//    (insti.cpp:1316-1365), GSL's draws replaced by Philox blocks;
//  - k_replace: Impute::replace's running allele tallies of the sampling
//    generations (impute.cpp:564-597), which Impute::result turns into the
//    allele probabilities of the output (impute.cpp:599-613).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/glphase_amd.h"
#include "glp_common.hpp"
#include "glp_internal.hpp"

namespace glp {

// One lane per (sample i, word w), i fastest, so a wavefront's GL reads of a
// site are contiguous.  Per site m of the word, with t = (max(1 - Het - AA, 0),
// max(Het, 0), max(AA, 0)) in float and the Philox block
// {kHapInitTag, i, m, window}: u0 < t0 leaves 0|0; else u1 < t1 / (t1 + t2)
// makes a heterozygote, the alternative allele on hap a when u2 < 0.5 and on
// hap b otherwise; else 1|1.  The uniforms are doubles of [0, 1) with 32
// random bits, compared with the float thresholds widened to double, as
// gsl_rng_uniform(rng) < t[0] compares them.
__global__ void __launch_bounds__(256)
    k_init_haps(const float2 *__restrict__ het_aa, unsigned N, unsigned M, unsigned wn,
                uint64_t seed, unsigned window, uint64_t *__restrict__ haps) {
  const size_t t = (size_t)blockIdx.x * 256u + threadIdx.x;
  if (t >= (size_t)N * wn) return;
  const unsigned i = (unsigned)(t % N), w = (unsigned)(t / N);
  const unsigned m1 = 64u * w + 64u < M ? 64u * w + 64u : M;
  uint64_t a = 0, b = 0;
  for (unsigned m = 64u * w; m < m1; ++m) {
    const float2 p = het_aa[(size_t)m * N + i];
    const float r = 1.0f - p.x - p.y;
    const float t0 = r < 0.0f ? 0.0f : r;
    const float t1 = p.x < 0.0f ? 0.0f : p.x;
    const float t2 = p.y < 0.0f ? 0.0f : p.y;
    const U4 x = draw(seed, kHapInitTag, i, m, window);
    const uint64_t bit = 1ull << (m & 63u);
    if (u01(x.x) < (double)t0) {
      // reference homozygote
    } else if (u01(x.y) < (double)(t1 / (t1 + t2))) {
      if (u01(x.z) < 0.5)
        a |= bit;
      else
        b |= bit;
    } else {
      a |= bit;
      b |= bit;
    }
  }
  haps[(size_t)(2u * i) * wn + w] = a;
  haps[(size_t)(2u * i + 1u) * wn + w] = b;
}

// One wavefront per sample: sis / tra = agreement of the sample's current
// haplotypes (a, b) with its tallies (ha, hb) as they stand and swapped; the
// tallies then count the pair in the orientation that agrees more (a to ha
// when sis > tra, else b to ha).  Integer sums, so any order gives the
// reference's values.  VEC (M a multiple of 4, tallies 16-byte aligned): a
// lane takes 4 consecutive sites of each 256-site chunk through one 16-byte
// load and store per tally row; otherwise one site per lane and 64-site word.
template <bool VEC>
__global__ void __launch_bounds__(256)
    k_replace(const uint64_t *__restrict__ haps, unsigned N, unsigned M, unsigned wn,
              uint32_t *__restrict__ hsum) {
  const unsigned i = blockIdx.x * 4u + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
  if (i >= N) return;
  const uint64_t *oa = haps + (size_t)(2u * i) * wn, *ob = oa + wn;
  uint32_t *ha = hsum + (size_t)(2u * i) * M, *hb = ha + M;
  unsigned sis = 0, tra = 0;
  if (VEC) {
    // chunk c, lane l: sites 256 c + 4 l .. + 3, bits 4 (l & 15) .. + 3 of
    // allele word 4 c + l / 16
    uint4 *ha4 = reinterpret_cast<uint4 *>(ha), *hb4 = reinterpret_cast<uint4 *>(hb);
    const unsigned nc = (M + 255u) >> 8, sh = 4u * (lane & 15u);
    for (unsigned c = 0; c < nc; ++c) {
      const unsigned m0 = 256u * c + 4u * lane;
      if (m0 < M) {
        const unsigned w = 4u * c + (lane >> 4);
        const unsigned xa = (unsigned)(oa[w] >> sh) & 15u, xb = (unsigned)(ob[w] >> sh) & 15u;
        const uint4 sa = ha4[m0 >> 2], sb = hb4[m0 >> 2];
        const unsigned a0 = xa & 1u, a1 = (xa >> 1) & 1u, a2 = (xa >> 2) & 1u, a3 = xa >> 3;
        const unsigned b0 = xb & 1u, b1 = (xb >> 1) & 1u, b2 = (xb >> 2) & 1u, b3 = xb >> 3;
        sis += a0 * sa.x + a1 * sa.y + a2 * sa.z + a3 * sa.w + b0 * sb.x + b1 * sb.y +
               b2 * sb.z + b3 * sb.w;
        tra += a0 * sb.x + a1 * sb.y + a2 * sb.z + a3 * sb.w + b0 * sa.x + b1 * sa.y +
               b2 * sa.z + b3 * sa.w;
      }
    }
    for (unsigned o = 32; o; o >>= 1) {
      sis += (unsigned)__shfl_xor((int)sis, (int)o, 64);
      tra += (unsigned)__shfl_xor((int)tra, (int)o, 64);
    }
    const bool keep = sis > tra;
    for (unsigned c = 0; c < nc; ++c) {
      const unsigned m0 = 256u * c + 4u * lane;
      if (m0 < M) {
        const unsigned w = 4u * c + (lane >> 4);
        const unsigned xa = (unsigned)(oa[w] >> sh) & 15u, xb = (unsigned)(ob[w] >> sh) & 15u;
        const unsigned ka = keep ? xa : xb, kb = keep ? xb : xa;
        uint4 sa = ha4[m0 >> 2], sb = hb4[m0 >> 2];
        sa.x += ka & 1u;
        sa.y += (ka >> 1) & 1u;
        sa.z += (ka >> 2) & 1u;
        sa.w += ka >> 3;
        sb.x += kb & 1u;
        sb.y += (kb >> 1) & 1u;
        sb.z += (kb >> 2) & 1u;
        sb.w += kb >> 3;
        ha4[m0 >> 2] = sa;
        hb4[m0 >> 2] = sb;
      }
    }
    return;
  }
  for (unsigned w = 0; w < wn; ++w) {
    const unsigned m = 64u * w + lane;
    const uint64_t wa = oa[w], wb = ob[w];
    if (m < M) {
      const unsigned x = (unsigned)(wa >> lane) & 1u, y = (unsigned)(wb >> lane) & 1u;
      const unsigned sa = ha[m], sb = hb[m];
      sis += x * sa + y * sb;
      tra += x * sb + y * sa;
    }
  }
  for (unsigned o = 32; o; o >>= 1) {
    sis += (unsigned)__shfl_xor((int)sis, (int)o, 64);
    tra += (unsigned)__shfl_xor((int)tra, (int)o, 64);
  }
  const bool keep = sis > tra;
  for (unsigned w = 0; w < wn; ++w) {
    const unsigned m = 64u * w + lane;
    const uint64_t wa = oa[w], wb = ob[w];
    if (m < M) {
      const unsigned x = (unsigned)(wa >> lane) & 1u, y = (unsigned)(wb >> lane) & 1u;
      ha[m] += keep ? x : y;
      hb[m] += keep ? y : x;
    }
  }
}

}  // namespace glp

using namespace glp;

extern "C" int glp_init_haps(const float *het_aa, unsigned num_samples, unsigned num_sites,
                             uint64_t seed, unsigned window, uint64_t *haps, void *stream) {
  if (!het_aa || !haps || num_samples == 0 || num_sites == 0)
    return set_error(GLP_ERR_ARG, "null pointer or empty window");
  const unsigned wn = (num_sites + 63) / 64;
  const size_t lanes = (size_t)num_samples * wn;
  hipLaunchKernelGGL(k_init_haps, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<const float2 *>(het_aa),
                     num_samples, num_sites, wn, seed, window, haps);
  return hip_check(hipGetLastError(), "k_init_haps launch");
}

extern "C" int glp_replace(const uint64_t *haps, unsigned num_samples, unsigned num_sites,
                           uint32_t *hsum, void *stream) {
  if (!haps || !hsum || num_samples == 0 || num_sites == 0)
    return set_error(GLP_ERR_ARG, "null pointer or empty window");
  const unsigned wn = (num_sites + 63) / 64;
  const bool vec = num_sites % 4u == 0u && (reinterpret_cast<uintptr_t>(hsum) & 15u) == 0u;
  if (vec)
    hipLaunchKernelGGL(k_replace<true>, dim3((num_samples + 3) / 4), dim3(256), 0,
                       static_cast<hipStream_t>(stream), haps, num_samples, num_sites, wn, hsum);
  else
    hipLaunchKernelGGL(k_replace<false>, dim3((num_samples + 3) / 4), dim3(256), 0,
                       static_cast<hipStream_t>(stream), haps, num_samples, num_sites, wn, hsum);
  return hip_check(hipGetLastError(), "k_replace launch");
}

// glp_kernels.hip -- gfx950 kernels of the GLPhase MH surrogate-parent update
// and the device-resident half of the C ABI (include/glphase_amd.h).
//
// Arithmetic contract: the forward recurrence keeps the operation order of
// Impute::hmm_like (impute.cpp:327-375) and this file is compiled with
// -ffp-contract=off, so every float product and sum is rounded exactly as in
// the reference CPU build; logf is the correctly rounded natural log (see
// cr_logf).  The results are therefore bit-identical to the CPU oracle, and
// accept decisions and chosen parents match it exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/glphase_amd.h"
#include "glp_common.hpp"
#include "glp_internal.hpp"

namespace glp {

// ---------------------------------------------------------------------------
// correctly rounded logf
// ---------------------------------------------------------------------------
// (float)log((double)x) is the correctly rounded logf of x for every positive
// float except these 20, whose log lies so close to a float rounding midpoint
// that a 2-ulp change of the double log flips the float (exhaustive search
// over all positive floats; values from an 80-digit decimal log).
__constant__ uint32_t c_logf_hard[20][2] = {
    {0x0dc8bba4u, 0xc289bac4u}, {0x111c87f8u, 0xc28085dfu},
    {0x1a8446cbu, 0xc24d0a45u}, {0x1f116ab8u, 0xc233b53eu},
    {0x29fd22f8u, 0xc1ee8859u}, {0x2c4c24b7u, 0xc1d48710u},
    {0x38dcbe38u, 0xc1128ba4u}, {0x3bf86ef0u, 0xc09c399eu},
    {0x3c413d3au, 0xc08e158fu}, {0x41178febu, 0x400fe5e7u},
    {0x464d5b2bu, 0x4117bcf1u}, {0x4665a9a6u, 0x41198725u},
    {0x4c5d65a5u, 0x418f034bu}, {0x4d604ebeu, 0x419a352cu},
    {0x4e85f412u, 0x41a6b811u}, {0x5ee8984eu, 0x422e4a21u},
    {0x65d890d3u, 0x4254d1f9u}, {0x66a8c860u, 0x42595e46u},
    {0x6f31a8ecu, 0x42845a89u}, {0x79e7ec37u, 0x42a1ffb7u}};

__device__ __noinline__ float cr_logf(float x) {
  const uint32_t b = __float_as_uint(x);
  float r = (float)log((double)x);
#pragma unroll
  for (int q = 0; q < 20; ++q)
    if (c_logf_hard[q][0] == b) r = __uint_as_float(c_logf_hard[q][1]);
  return r;
}

// ---------------------------------------------------------------------------
// forward pass
// ---------------------------------------------------------------------------
// Emission table in LDS: entry (code << 4) | nib holds the 4 emissions
// (e00, e01, e10, e11) of a site whose code is `code` and whose 4 parent
// alleles are nib = f0 | f1 << 1 | m0 << 2 | m1 << 3, where
// e_xy = FillEmit(code)[(f_x << 1) | m_y] (hmmLike.cu:124-127).
struct FwdState {
  float l00, l01, l10, l11;
  float fscore;
  double dscore;
};

// One 32-site chunk of the recurrence.  h[] are the 32-bit halves of the 4
// parents' haplotype words covering the chunk; codes points at the chunk's
// first code word of this sample (stride = words between sites groups);
// tr = tran + 3 * (first site of the chunk).
template <int BITS, bool MIXED, bool FIRST, bool TAIL>
__device__ __forceinline__ void fwd_chunk(FwdState &st, const uint32_t (&h)[4],
                                          const uint32_t *__restrict__ codes,
                                          size_t stride,
                                          const float *__restrict__ tr,
                                          unsigned n, const float4 *lut,
                                          float norm) {
  constexpr int CPW = 32 / BITS;  // codes per word
  constexpr int NW = 32 / CPW;    // code words per chunk
  constexpr uint32_t MASK = (1u << BITS) - 1u;
  uint32_t cw[NW];
#pragma unroll
  for (int q = 0; q < NW; ++q)
    cw[q] = (!TAIL || (unsigned)(q * CPW) < n) ? codes[q * stride] : 0u;
#pragma unroll
  for (int s = 0; s < 32; ++s) {
    if (TAIL && (unsigned)s >= n) break;
    const uint32_t code = (cw[s / CPW] >> (32 - BITS * (s % CPW + 1))) & MASK;
    const uint32_t nib = ((h[0] >> s) & 1u) | (((h[1] >> s) & 1u) << 1) |
                         (((h[2] >> s) & 1u) << 2) | (((h[3] >> s) & 1u) << 3);
    const float4 e = lut[(code << 4) | nib];
    if (FIRST && s == 0) {
      // impute.cpp:344-347
      st.l00 = 0.25f * e.x;
      st.l01 = 0.25f * e.y;
      st.l10 = 0.25f * e.z;
      st.l11 = 0.25f * e.w;
      continue;
    }
    const float t0 = tr[3 * s], t1 = tr[3 * s + 1], t2 = tr[3 * s + 2];
    const float l00 = st.l00, l01 = st.l01, l10 = st.l10, l11 = st.l11;
    // impute.cpp:354-366
    const float b00 = l00 * t0 + (l01 + l10) * t1 + l11 * t2;
    const float b01 = l01 * t0 + (l00 + l11) * t1 + l10 * t2;
    const float b10 = l10 * t0 + (l00 + l11) * t1 + l01 * t2;
    const float b11 = l11 * t0 + (l01 + l10) * t1 + l00 * t2;
    float n00 = b00 * e.x, n01 = b01 * e.y, n10 = b10 * e.z, n11 = b11 * e.w;
    float sum = n00 + n01 + n10 + n11;
    // impute.cpp:367-372: rescale when the probabilities become too small
    if (sum < norm) {
      sum = 1.0f / sum;
      if (MIXED)
        st.dscore -= log((double)sum);
      else
        st.fscore -= cr_logf(sum);
      n00 *= sum;
      n01 *= sum;
      n10 *= sum;
      n11 *= sum;
    }
    st.l00 = n00;
    st.l01 = n01;
    st.l10 = n10;
    st.l11 = n11;
  }
}

// Log-likelihood of one sample given 4 parents (father pair f0, f1, mother
// pair m0, m1), rows viewed as 32-bit words.
template <int BITS, bool MIXED>
__device__ double forward_pass(const uint32_t *__restrict__ f0,
                               const uint32_t *__restrict__ f1,
                               const uint32_t *__restrict__ m0,
                               const uint32_t *__restrict__ m1,
                               const uint32_t *__restrict__ codes,
                               size_t stride, const float *__restrict__ tran,
                               unsigned M, const float4 *lut, float norm) {
  constexpr int NW = BITS;  // code words per 32 sites
  FwdState st;
  st.fscore = 0.0f;
  st.dscore = 0.0;
  const unsigned nfull = M >> 5, rem = M & 31u;
  if (nfull > 0) {
    {
      const uint32_t h[4] = {f0[0], f1[0], m0[0], m1[0]};
      fwd_chunk<BITS, MIXED, true, false>(st, h, codes, stride, tran, 32, lut,
                                          norm);
    }
    for (unsigned ch = 1; ch < nfull; ++ch) {
      const uint32_t h[4] = {f0[ch], f1[ch], m0[ch], m1[ch]};
      fwd_chunk<BITS, MIXED, false, false>(st, h, codes + ch * NW * stride,
                                           stride, tran + 96 * ch, 32, lut,
                                           norm);
    }
    if (rem) {
      const uint32_t h[4] = {f0[nfull], f1[nfull], m0[nfull], m1[nfull]};
      fwd_chunk<BITS, MIXED, false, true>(st, h, codes + nfull * NW * stride,
                                          stride, tran + 96 * nfull, rem, lut,
                                          norm);
    }
  } else {
    const uint32_t h[4] = {f0[0], f1[0], m0[0], m1[0]};
    fwd_chunk<BITS, MIXED, true, true>(st, h, codes, stride, tran, rem, lut,
                                       norm);
  }
  const float last = st.l00 + st.l01 + st.l10 + st.l11;
  if (MIXED) return st.dscore + log((double)last);
  return (double)(st.fscore + cr_logf(last));
}

__device__ __forceinline__ const uint32_t *hap_row(const glp_window &W,
                                                   unsigned h) {
  return reinterpret_cast<const uint32_t *>(W.panel + (size_t)h * W.words_per_hap);
}

template <int BITS, bool MIXED>
__device__ __forceinline__ double score_set(const glp_window &W,
                                            const unsigned (&P)[4],
                                            unsigned sample, unsigned N,
                                            const float4 *lut, float norm) {
  return forward_pass<BITS, MIXED>(hap_row(W, P[0]), hap_row(W, P[1]),
                                   hap_row(W, P[2]), hap_row(W, P[3]),
                                   W.packed + sample, N, W.tran, W.num_sites,
                                   lut, norm);
}

__device__ __forceinline__ void load_lut(float4 *s_lut,
                                         const float4 *__restrict__ g_lut,
                                         unsigned n) {
  for (unsigned q = threadIdx.x; q < n; q += blockDim.x) s_lut[q] = g_lut[q];
  __syncthreads();
}

// MHSampler::SampleHap (MHSampler.hpp:82-96): accept a better or equal
// proposal, else accept with probability exp((prop - curr) * pen) in FP64.
__device__ __forceinline__ bool mh_accept(double prop, double curr, double pen,
                                          uint32_t ubits) {
  return prop >= curr || u01(ubits) < exp((prop - curr) * pen);
}

// fresh parents of a generation (insti.cpp:1493-1494 / hmmLike.cpp:93-96)
__device__ __forceinline__ void initial_parents(const KArgs &a,
                                                const glp_window &W,
                                                unsigned w, unsigned i,
                                                unsigned (&P)[4]) {
  if (a.prop_mode == kPropHost) {
#pragma unroll
    for (int q = 0; q < 4; ++q) P[q] = W.parents[4 * (size_t)i + q];
    return;
  }
  const U4 x = draw(a.seed, kInitTag, i, w, a.gen);
  const uint32_t xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
  for (int q = 0; q < 4; ++q)
    P[q] = a.prop_mode == kPropKList
               ? a.klist[((size_t)w * a.N + i) * a.K + uniform_int(xs[q], a.K)]
               : unif_excl(xs[q], a.H, i);
}

// the draws of proposal c: replace slot (insti.cpp:1506, hmmLike.cu:188),
// candidate (sampler.cpp:29-39 / hmmLike.cpp:101-102) and accept uniform
__device__ __forceinline__ void proposal_draws(const KArgs &a, unsigned w,
                                               unsigned i, unsigned c,
                                               unsigned &rp, unsigned &cand,
                                               uint32_t &ubits) {
  const U4 x = draw(a.seed, a.k0 + c, i, w, a.gen);
  rp = x.x & 3u;
  ubits = x.z;
  if (a.prop_mode == kPropHost)
    cand = a.extra[((size_t)w * a.C + c) * a.N + i];
  else if (a.prop_mode == kPropKList)
    cand = a.klist[((size_t)w * a.N + i) * a.K + uniform_int(x.y, a.K)];
  else
    cand = unif_excl(x.y, a.H, i);
}

// ---------------------------------------------------------------------------
// plain path: one lane per (window, sample), one launch per MH iteration
// ---------------------------------------------------------------------------
template <int BITS, bool MIXED>
__global__ void __launch_bounds__(256)
    k_plain(const glp_window *__restrict__ wins, KArgs a, unsigned step) {
  extern __shared__ float4 s_lut[];
  load_lut(s_lut, reinterpret_cast<const float4 *>(a.lut), a.lut_entries);
  const unsigned w = blockIdx.y;
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const glp_window W = wins[w];
  const size_t row = (size_t)w * a.N + i;
  unsigned P[4];
  if (step == 0) {
    initial_parents(a, W, w, i, P);
    const double v = score_set<BITS, MIXED>(W, P, i, a.N, s_lut, a.norm);
#pragma unroll
    for (int q = 0; q < 4; ++q) W.parents[4 * (size_t)i + q] = P[q];
    W.like[i] = v;
    if (a.trace_like) a.trace_like[row * (a.C + 1)] = v;
    return;
  }
  const unsigned c = step - 1;
#pragma unroll
  for (int q = 0; q < 4; ++q) P[q] = W.parents[4 * (size_t)i + q];
  const double curr = W.like[i];
  unsigned rp, cand;
  uint32_t ub;
  proposal_draws(a, w, i, c, rp, cand, ub);
  const unsigned oh = P[rp];
  P[rp] = cand;
  const double prop = score_set<BITS, MIXED>(W, P, i, a.N, s_lut, a.norm);
  const bool acc = mh_accept(prop, curr, (double)a.pen, ub);
  if (acc) {
    W.parents[4 * (size_t)i + rp] = cand;
    W.like[i] = prop;
  }
  (void)oh;
  if (a.trace_like) a.trace_like[row * (a.C + 1) + 1 + c] = prop;
  if (a.trace_acc) a.trace_acc[row * a.C + c] = acc ? 1 : 0;
}

// ---------------------------------------------------------------------------
// batched form: persistent work-queue kernel with speculative proposal windows
// ---------------------------------------------------------------------------
// Each wavefront holds 64/Q chain slots of Q lanes.  A slot owns one
// (window, sample) task.  In a round, the slot's lanes score the next Q
// undecided proposals (the first lane of a fresh task scores its initial
// parents), all from the slot's current parent set S.  Then every lane of the
// slot replays the serial MH walk over those Q values in proposal order.  A
// value is the serial chain's value exactly when every slot of S changed in
// this walk so far equals the proposal's own replace slot; the walk stops at
// the first proposal for which that fails, and the next round re-scores from
// there.  The result equals the serial loop of insti.cpp:1505-1539 draw for
// draw; speculation only spends extra lane time.
template <int BITS, bool MIXED, int Q>
__global__ void __launch_bounds__(256)
    k_batched(const glp_window *__restrict__ wins,
              const uint2 *__restrict__ tasks, unsigned ntasks,
              unsigned long long *__restrict__ ctrs, KArgs a) {
  extern __shared__ float4 s_lut[];
  load_lut(s_lut, reinterpret_cast<const float4 *>(a.lut), a.lut_entries);
  unsigned *queue = reinterpret_cast<unsigned *>(ctrs);
  const unsigned lane = threadIdx.x & 63u;
  const unsigned j = lane % Q;
  const unsigned sbase = lane - j;
  const double pen = (double)a.pen;

  bool active = false, fresh = false;
  unsigned w = 0, i = 0, kp = 0;
  unsigned S[4] = {0, 0, 0, 0};
  double cur = 0.0;
  glp_window W{};
  unsigned long long n_eval = 0, n_useful = 0;

  for (;;) {
    // 1. idle slots take the next task (longest windows first)
    unsigned t = 0xFFFFFFFFu;
    if (j == 0 && !active) t = atomicAdd(queue, 1u);
    t = __shfl(t, sbase, 64);
    if (!active && t < ntasks) {
      const uint2 tk = tasks[t];
      w = tk.x;
      i = tk.y;
      W = wins[w];
      initial_parents(a, W, w, i, S);
      active = true;
      fresh = true;
      kp = 0;
    }
    if (!__any(active)) break;

    // 2. this lane's job: initial set or one speculative proposal
    bool work = false;
    unsigned rp = 4, cand = 0;
    uint32_t ub = 0;
    unsigned P[4] = {S[0], S[1], S[2], S[3]};
    if (active) {
      const int job = fresh ? (int)j - 1 : (int)j;
      if (job < 0) {
        work = true;
      } else if (kp + (unsigned)job < a.C) {
        work = true;
        proposal_draws(a, w, i, kp + (unsigned)job, rp, cand, ub);
        P[rp] = cand;
      }
    }
    double v = 0.0;
    if (work) {
      v = score_set<BITS, MIXED>(W, P, i, a.N, s_lut, a.norm);
      ++n_eval;
    }

    // 3. serial MH walk over the slot's values (every lane of the slot
    //    replays it, so the slot state stays replicated)
    const double v0 = __shfl(v, sbase, 64);
    const int jstart = fresh ? 1 : 0;
    if (active && fresh) {
      cur = v0;
      if (j == 0 && a.trace_like)
        a.trace_like[((size_t)w * a.N + i) * (a.C + 1)] = v0;
    }
    unsigned changed = 0, decided = 0;
    bool stop = !active;
#pragma unroll
    for (int jj = 0; jj < Q; ++jj) {
      const unsigned src = sbase + (unsigned)jj;
      const double vj = __shfl(v, src, 64);
      const unsigned rpj = __shfl(rp, src, 64);
      const unsigned cj = __shfl(cand, src, 64);
      const uint32_t uj = __shfl(ub, src, 64);
      if (!stop && jj >= jstart) {
        const unsigned kk = kp + (unsigned)(jj - jstart);
        if (kk >= a.C || (changed & ~(1u << rpj)) != 0u) {
          stop = true;
        } else {
          const bool acc = mh_accept(vj, cur, pen, uj);
          if (acc) {
            S[0] = rpj == 0 ? cj : S[0];
            S[1] = rpj == 1 ? cj : S[1];
            S[2] = rpj == 2 ? cj : S[2];
            S[3] = rpj == 3 ? cj : S[3];
            cur = vj;
            changed |= 1u << rpj;
          }
          ++decided;
          if (j == (unsigned)jj) {
            const size_t row = (size_t)w * a.N + i;
            if (a.trace_like) a.trace_like[row * (a.C + 1) + 1 + kk] = vj;
            if (a.trace_acc) a.trace_acc[row * a.C + kk] = acc ? 1 : 0;
          }
        }
      }
    }
    if (active) {
      kp += decided;
      fresh = false;
      if (j == 0) n_useful += decided + (jstart ? 1u : 0u);
      if (kp >= a.C) {
        if (j == 0) {
#pragma unroll
          for (int q = 0; q < 4; ++q) W.parents[4 * (size_t)i + q] = S[q];
          W.like[i] = cur;
        }
        active = false;
      }
    }
  }
  // per-wave statistics
  for (int off = 32; off > 0; off >>= 1) {
    n_eval += __shfl_down(n_eval, off, 64);
    n_useful += __shfl_down(n_useful, off, 64);
  }
  if (lane == 0) {
    atomicAdd(&ctrs[1], n_eval);
    atomicAdd(&ctrs[2], n_useful);
  }
}

__global__ void k_reset_ctrs(unsigned long long *ctrs) {
  if (threadIdx.x < 4) ctrs[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------------------
// single evaluations: Impute::hmm_like batched
// ---------------------------------------------------------------------------
template <int BITS>
__global__ void __launch_bounds__(256)
    k_eval(const glp_window *__restrict__ wins, KArgs a,
           const unsigned *__restrict__ samples,
           const unsigned *__restrict__ parents, unsigned n, float *out) {
  extern __shared__ float4 s_lut[];
  load_lut(s_lut, reinterpret_cast<const float4 *>(a.lut), a.lut_entries);
  const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const glp_window W = wins[0];
  const unsigned P[4] = {parents[4 * (size_t)e], parents[4 * (size_t)e + 1],
                         parents[4 * (size_t)e + 2], parents[4 * (size_t)e + 3]};
  out[e] = (float)score_set<BITS, false>(W, P, samples[e], a.N, s_lut, a.norm);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static size_t lut_smem(const KArgs &a) { return (size_t)a.lut_entries * 16; }

int launch_plain(const KArgs &a, const glp_window *d_wins, unsigned nwin,
                 unsigned step, hipStream_t stream) {
  const dim3 block(256), grid((a.N + 255) / 256, nwin);
  const size_t smem = lut_smem(a);
  const bool mixed = a.precision != kPrecF32;
  if (a.bits == 8) {
    if (mixed)
      hipLaunchKernelGGL((k_plain<8, true>), grid, block, smem, stream, d_wins, a, step);
    else
      hipLaunchKernelGGL((k_plain<8, false>), grid, block, smem, stream, d_wins, a, step);
  } else {
    if (mixed)
      hipLaunchKernelGGL((k_plain<4, true>), grid, block, smem, stream, d_wins, a, step);
    else
      hipLaunchKernelGGL((k_plain<4, false>), grid, block, smem, stream, d_wins, a, step);
  }
  return hip_check(hipGetLastError(), "k_plain launch");
}

template <int BITS, bool MIXED>
static void *batched_fn(unsigned Q) {
  switch (Q) {
    case 1: return reinterpret_cast<void *>(&k_batched<BITS, MIXED, 1>);
    case 2: return reinterpret_cast<void *>(&k_batched<BITS, MIXED, 2>);
    case 4: return reinterpret_cast<void *>(&k_batched<BITS, MIXED, 4>);
    case 8: return reinterpret_cast<void *>(&k_batched<BITS, MIXED, 8>);
    case 16: return reinterpret_cast<void *>(&k_batched<BITS, MIXED, 16>);
    case 32: return reinterpret_cast<void *>(&k_batched<BITS, MIXED, 32>);
    case 64: return reinterpret_cast<void *>(&k_batched<BITS, MIXED, 64>);
  }
  return nullptr;
}

void *batched_kernel(unsigned bits, bool mixed, unsigned Q) {
  if (bits == 8) return mixed ? batched_fn<8, true>(Q) : batched_fn<8, false>(Q);
  return mixed ? batched_fn<4, true>(Q) : batched_fn<4, false>(Q);
}

int launch_batched(const KArgs &a, const glp_window *d_wins, const uint2 *d_tasks,
                   unsigned ntasks, unsigned long long *d_ctrs, unsigned Q,
                   unsigned grid, hipStream_t stream) {
  hipLaunchKernelGGL(k_reset_ctrs, dim3(1), dim3(64), 0, stream, d_ctrs);
  void *fn = batched_kernel(a.bits, a.precision != kPrecF32, Q);
  if (!fn) return set_error(GLP_ERR_ARG, "lanes_per_chain must be a power of two <= 64");
  void *args[] = {(void *)&d_wins, (void *)&d_tasks, (void *)&ntasks, (void *)&d_ctrs,
                  (void *)&a};
  hipError_t err = hipLaunchKernel(fn, dim3(grid), dim3(256), args, lut_smem(a), stream);
  return hip_check(err, "k_batched launch");
}

int batched_occupancy(unsigned bits, bool mixed, unsigned Q, size_t smem, int *blocks_per_cu) {
  void *fn = batched_kernel(bits, mixed, Q);
  if (!fn) return set_error(GLP_ERR_ARG, "bad lanes_per_chain");
  return hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, 256, smem),
                   "occupancy query");
}

int launch_eval(const KArgs &a, const glp_window *d_wins, const unsigned *d_samples,
                const unsigned *d_parents, unsigned n, float *d_out, hipStream_t stream) {
  const dim3 block(256), grid((n + 255) / 256);
  if (a.bits == 8)
    hipLaunchKernelGGL((k_eval<8>), grid, block, lut_smem(a), stream, d_wins, a, d_samples,
                       d_parents, n, d_out);
  else
    hipLaunchKernelGGL((k_eval<4>), grid, block, lut_smem(a), stream, d_wins, a, d_samples,
                       d_parents, n, d_out);
  return hip_check(hipGetLastError(), "k_eval launch");
}

}  // namespace glp

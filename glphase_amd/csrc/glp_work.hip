// glp_work.hip -- gfx950 kernel of the haplotype resampling step that
// follows each generation's MH update: Impute::hmm_work (impute.cpp:380-485)
// for every sample, and its part of the C ABI (include/glphase_amd.h).
//
// Arithmetic contract: every float operation of impute.cpp:380-485 and of the
// emission / genotype-probability builder of Insti::initialize
// (insti.cpp:1345-1378) is kept in the reference's order, compiled with
// -ffp-contract=off, and 1.0f / x is the IEEE division; powf(x, S) is the
// float of the FP64 power (x itself at S = 1).  The sampled haplotypes equal
// the CPU oracle's bit for bit, given the same Philox stream.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <algorithm>
#include <cstring>

#include "../../include/glphase_amd.h"
#include "glp_common.hpp"
#include "glp_internal.hpp"

namespace glp {

struct WorkArgs {
  unsigned N, H;
  float pen;
  uint64_t seed;
  unsigned gen;
  float pc[16];
};

// Genotype probabilities t and the emission of one state from the input pair
// (Het, AA) and the state's row of the mutation matrix (insti.cpp:1349-1378).
__device__ __forceinline__ float3 work_prob(float2 q) {
  const float t0r = 1 - q.x - q.y;
  return make_float3(t0r > 0.0f ? t0r : 0.0f, q.x > 0.0f ? q.x : 0.0f, q.y > 0.0f ? q.y : 0.0f);
}

__device__ __forceinline__ float work_emit(const float4 row, const float3 t) {
  return row.x * t.x + row.y * t.y + row.z * t.y + row.w * t.z;
}

__device__ __forceinline__ unsigned hbit64(uint64_t w, unsigned m) {
  return (unsigned)(w >> (m & 63u)) & 1u;
}

constexpr int kWorkBlock = 8;  // sites per prefetch block

// Genotype choice of impute.cpp:472-484 from the weights and the site's
// uniform word.
__device__ __forceinline__ unsigned work_choose(float c00, float c01, float c10, float c11,
                                                uint32_t xm) {
  const float u = (float)(u01(xm) * (double)(c00 + c01 + c10 + c11));
  if (u < c00) return 0u;
  if (u < c00 + c01) return 1u;
  if (u < c00 + c01 + c10) return 2u;
  return 3u;
}

__device__ __forceinline__ void load_pc_rows(float4 *s_pc, const WorkArgs &a) {
  if (threadIdx.x < 4)
    s_pc[threadIdx.x] = make_float4(a.pc[4 * threadIdx.x], a.pc[4 * threadIdx.x + 1],
                                    a.pc[4 * threadIdx.x + 2], a.pc[4 * threadIdx.x + 3]);
  __syncthreads();
}

// One block of the recursions: sites 8 blk .. 8 blk + 7 (aligned, so a block
// lies inside one 64-site allele word), their GL pairs and transitions.
struct WorkBlock {
  float2 q[kWorkBlock];
  float t[3 * kWorkBlock];
};

// Loads block blk's inputs for sites in [lo, hi] (others zero).
__device__ __forceinline__ void load_block(WorkBlock &b, int blk, int lo, int hi,
                                           gptr_t<float> q2, gptr_t<float> tran, unsigned N,
                                           unsigned i) {
#pragma unroll
  for (int k = 0; k < kWorkBlock; ++k) {
    const int m = kWorkBlock * blk + k;
    const bool in = blk >= 0 && m >= lo && m <= hi;
    b.q[k] = in ? load_f2(q2 + 2 * ((size_t)m * N + i)) : make_float2(0, 0);
    b.t[3 * k] = in ? tran[3 * m] : 0.0f;
    b.t[3 * k + 1] = in ? tran[3 * m + 1] : 0.0f;
    b.t[3 * k + 2] = in ? tran[3 * m + 2] : 0.0f;
  }
}

// Emissions e00, e01, e10, e11 of site m (alleles from the block's word).
__device__ __forceinline__ float4 site_emit(const float4 *s_pc, const uint64_t (&hw)[4],
                                            unsigned m, float2 q) {
  const unsigned f0 = hbit64(hw[0], m), f1 = hbit64(hw[1], m), m0 = hbit64(hw[2], m),
                 m1 = hbit64(hw[3], m);
  const float3 t = work_prob(q);
  return make_float4(work_emit(s_pc[(f0 << 1) | m0], t), work_emit(s_pc[(f0 << 1) | m1], t),
                     work_emit(s_pc[(f1 << 1) | m0], t), work_emit(s_pc[(f1 << 1) | m1], t));
}

// The two recursions of impute.cpp:398-470, one lane per (window, sample):
// the scaled backward pass writes beta to the workspace ([M][N] float4,
// site-major, so a wavefront's stores coalesce), then the scaled forward
// pass replaces each site's beta with the state posteriors
// (l00 b00, l01 b01, l10 b10, l11 b11) that the sampling step weighs.  Sites
// go in aligned blocks of 8 whose GL pairs, transitions (and, forward, beta)
// are loaded a block ahead; a block's emissions are computed before its
// recursion steps, off their dependency chain.  Parent indices past the
// panel are clamped (the MH update never produces them).
__global__ void __launch_bounds__(256)
    k_hmm_work_pass(const glp_work_window *__restrict__ wins, const size_t *__restrict__ beta_off,
                    float4 *__restrict__ ws, WorkArgs a) {
  __shared__ float4 s_pc[4];  // rows pc[s][0..3]
  load_pc_rows(s_pc, a);
  const unsigned w = blockIdx.y;
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const glp_work_window W = wins[w];
  const unsigned M = W.num_sites, wn = W.words_per_hap, N = a.N;
  float *beta = reinterpret_cast<float *>(ws + beta_off[w]);  // [M][N] x 4
  const auto gbeta = gmem(beta);
  const auto q2 = gmem(W.het_aa), tran = gmem(W.tran);
  gptr_t<uint64_t> rows[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned h = gmem(W.parents)[4 * (size_t)i + q];
    h = h < a.H ? h : a.H - 1;
    rows[q] = gmem(W.panel) + (size_t)h * wn;
  }
  constexpr int B = kWorkBlock;
  const int top_blk = (int)(M - 1) / B;

  // backward recursion (impute.cpp:398-416): site m = M-1 .. 1 turns beta_m
  // into beta_{m-1}
  float b0 = 1, b1 = 1, b2 = 1, b3 = 1;
  store_f4(beta + 4 * ((size_t)(M - 1) * N + i), 1, 1, 1, 1);
  uint64_t hw[4];
  WorkBlock cur, nxt;
  load_block(nxt, top_blk, 1, (int)M - 1, q2, tran, N, i);
  for (int blk = top_blk; blk >= 0; --blk) {
    cur = nxt;
    load_block(nxt, blk - 1, 1, (int)M - 1, q2, tran, N, i);
    const unsigned word = (unsigned)(B * blk) >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) hw[q] = rows[q][word];
    float4 e[B];
#pragma unroll
    for (int k = 0; k < B; ++k) e[k] = site_emit(s_pc, hw, (unsigned)(B * blk + k), cur.q[k]);
#pragma unroll
    for (int k = B - 1; k >= 0; --k) {
      const int mi = B * blk + k;
      if (mi > (int)M - 1 || mi < 1) continue;
      const unsigned m = (unsigned)mi;
      const float bb00 = b0 * e[k].x, bb01 = b1 * e[k].y, bb10 = b2 * e[k].z, bb11 = b3 * e[k].w;
      const float t0 = cur.t[3 * k], t1 = cur.t[3 * k + 1], t2 = cur.t[3 * k + 2];
      const float n0 = bb00 * t0 + (bb01 + bb10) * t1 + bb11 * t2;
      const float n1 = bb01 * t0 + (bb00 + bb11) * t1 + bb10 * t2;
      const float n2 = bb10 * t0 + (bb00 + bb11) * t1 + bb01 * t2;
      const float n3 = bb11 * t0 + (bb01 + bb10) * t1 + bb00 * t2;
      const float sum = 1.0f / (n0 + n1 + n2 + n3);
      b0 = n0 * sum;
      b1 = n1 * sum;
      b2 = n2 * sum;
      b3 = n3 * sum;
      store_f4(beta + 4 * ((size_t)(m - 1) * N + i), b0, b1, b2, b3);
    }
  }

  // forward recursion (impute.cpp:418-455), sites 0 .. M-1; each site's beta
  // is read (a block ahead) before its slot takes the posterior
  float l0 = 0, l1 = 0, l2 = 0, l3 = 0;
  float4 bn[B];
  load_block(nxt, 0, 0, (int)M - 1, q2, tran, N, i);
#pragma unroll
  for (int k = 0; k < B; ++k)
    bn[k] = (unsigned)k < M ? load_f4(gbeta + 4 * ((size_t)k * N + i)) : make_float4(0, 0, 0, 0);
  for (int blk = 0; blk <= top_blk; ++blk) {
    cur = nxt;
    float4 bc[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      bc[k] = bn[k];
      const unsigned m = (unsigned)(B * (blk + 1) + k);
      bn[k] = m < M ? load_f4(gbeta + 4 * ((size_t)m * N + i)) : make_float4(0, 0, 0, 0);
    }
    load_block(nxt, blk + 1, 0, (int)M - 1, q2, tran, N, i);
    const unsigned word = (unsigned)(B * blk) >> 6;
#pragma unroll
    for (int q = 0; q < 4; ++q) hw[q] = rows[q][word];
    float4 e[B];
#pragma unroll
    for (int k = 0; k < B; ++k) e[k] = site_emit(s_pc, hw, (unsigned)(B * blk + k), cur.q[k]);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const unsigned m = (unsigned)(B * blk + k);
      if (m >= M) break;
      if (m) {
        const float t0 = cur.t[3 * k], t1 = cur.t[3 * k + 1], t2 = cur.t[3 * k + 2];
        const float n0 = l0 * t0 + l1 * t1 + l2 * t1 + l3 * t2;
        const float n1 = l0 * t1 + l1 * t0 + l2 * t2 + l3 * t1;
        const float n2 = l0 * t1 + l1 * t2 + l2 * t0 + l3 * t1;
        const float n3 = l0 * t2 + l1 * t1 + l2 * t1 + l3 * t0;
        l0 = n0 * e[k].x;
        l1 = n1 * e[k].y;
        l2 = n2 * e[k].z;
        l3 = n3 * e[k].w;
      } else {
        l0 = 0.25f * e[k].x;
        l1 = 0.25f * e[k].y;
        l2 = 0.25f * e[k].z;
        l3 = 0.25f * e[k].w;
      }
      const float sum = 1.0f / (l0 + l1 + l2 + l3);
      l0 *= sum;
      l1 *= sum;
      l2 *= sum;
      l3 *= sum;
      const float4 b = bc[k];
      store_f4(beta + 4 * ((size_t)m * N + i), l0 * b.x, l1 * b.y, l2 * b.z, l3 * b.w);
    }
  }
}

// The sampling of impute.cpp:457-484, one lane per (window, sample, 64-site
// word), sample fastest so a wavefront's reads of a site coalesce: from the
// posteriors p (written by k_hmm_work_pass) and the GL pair, the genotype
// weights c_g = (t_g * sum_s p_s pc[s][g])^S, one Philox word per site, the
// genotype choice, and the two new haplotype words.  POW: S != 1 (powf is
// the identity at S = 1).
template <bool POW>
__global__ void __launch_bounds__(256)
    k_hmm_work_sample(const glp_work_window *__restrict__ wins, const size_t *__restrict__ beta_off,
                      const float4 *__restrict__ ws, WorkArgs a) {
  __shared__ float4 s_pc[4];
  load_pc_rows(s_pc, a);
  const unsigned w = blockIdx.y;
  const glp_work_window W = wins[w];
  const unsigned M = W.num_sites, wn = W.words_per_hap, N = a.N;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * wn) return;
  const unsigned i = (unsigned)(t % N), wd = (unsigned)(t / N);
  const auto post = gmem(reinterpret_cast<const float *>(ws + beta_off[w]));
  const auto q2 = gmem(W.het_aa);
  uint64_t hw[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned h = gmem(W.parents)[4 * (size_t)i + q];
    h = h < a.H ? h : a.H - 1;
    hw[q] = gmem(W.panel)[(size_t)h * wn + wd];
  }
  const float S = a.pen;
  const unsigned m1 = 64u * wd + 64u < M ? 64u * wd + 64u : M;
  uint64_t wa = 0, wb = 0;
  U4 x{0, 0, 0, 0};
  for (unsigned m = 64u * wd; m < m1; ++m) {
    const float4 p = load_f4(post + 4 * ((size_t)m * N + i));
    const float3 tq = work_prob(load_f2(q2 + 2 * ((size_t)m * N + i)));
    const unsigned f0 = hbit64(hw[0], m), f1 = hbit64(hw[1], m), m0 = hbit64(hw[2], m),
                   m1b = hbit64(hw[3], m);
    const float4 r00 = s_pc[(f0 << 1) | m0], r01 = s_pc[(f0 << 1) | m1b],
                 r10 = s_pc[(f1 << 1) | m0], r11 = s_pc[(f1 << 1) | m1b];
    float c00 = tq.x * (p.x * r00.x + p.y * r01.x + p.z * r10.x + p.w * r11.x);
    float c01 = tq.y * (p.x * r00.y + p.y * r01.y + p.z * r10.y + p.w * r11.y);
    float c10 = tq.y * (p.x * r00.z + p.y * r01.z + p.z * r10.z + p.w * r11.z);
    float c11 = tq.z * (p.x * r00.w + p.y * r01.w + p.z * r10.w + p.w * r11.w);
    if (POW) {
      c00 = (float)pow((double)c00, (double)S);
      c01 = (float)pow((double)c01, (double)S);
      c10 = (float)pow((double)c10, (double)S);
      c11 = (float)pow((double)c11, (double)S);
    }
    if ((m & 3u) == 0) x = draw(a.seed, kWorkTag | (m >> 2), i, w, a.gen);
    const uint32_t xm = (m & 3u) == 0 ? x.x : (m & 3u) == 1 ? x.y : (m & 3u) == 2 ? x.z : x.w;
    const unsigned g = work_choose(c00, c01, c10, c11, xm);
    wa |= (uint64_t)(g >> 1) << (m & 63u);
    wb |= (uint64_t)(g & 1u) << (m & 63u);
  }
  const auto hnew = (gmut_t<uint64_t>)W.hnew;
  hnew[(size_t)(2 * i) * wn + wd] = wa;
  hnew[(size_t)(2 * i + 1) * wn + wd] = wb;
}

}  // namespace glp

using namespace glp;

namespace {

int check_work_desc(const glp_work_desc *d) {
  if (!d || !d->windows || d->num_windows == 0)
    return set_error(GLP_ERR_ARG, "descriptor has no windows");
  if (d->num_samples == 0) return set_error(GLP_ERR_ARG, "num_samples is 0");
  if (d->num_haps == 0) return set_error(GLP_ERR_ARG, "num_haps is 0");
  for (unsigned w = 0; w < d->num_windows; ++w) {
    const glp_work_window &W = d->windows[w];
    if (W.num_sites == 0) return set_error(GLP_ERR_ARG, "window with 0 sites");
    if (W.words_per_hap != (W.num_sites + 63) / 64)
      return set_error(GLP_ERR_ARG, "words_per_hap must be ceil(num_sites / 64)");
    if (!W.panel || !W.parents || !W.tran || !W.het_aa || !W.hnew)
      return set_error(GLP_ERR_ARG, "window with a null pointer");
  }
  return GLP_OK;
}

}  // namespace

extern "C" size_t glp_work_workspace_bytes(const glp_work_desc *d) {
  if (!d || !d->windows) return 0;
  size_t b = 0;
  for (unsigned w = 0; w < d->num_windows; ++w)
    b += (size_t)d->windows[w].num_sites * d->num_samples * sizeof(float4);
  return b;
}

extern "C" int glp_hmm_work(const glp_work_desc *d, void *stream) {
  int rc = check_work_desc(d);
  if (rc != GLP_OK) return rc;
  if (!d->workspace) return set_error(GLP_ERR_ARG, "null workspace");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // window table and per-window workspace offsets, in one short-lived block
  const size_t nw = d->num_windows;
  const size_t tbl = sizeof(glp_work_window) * nw, offb = sizeof(size_t) * nw;
  char *dev = nullptr;
  rc = hip_check(hipMallocAsync((void **)&dev, tbl + offb, s), "hipMallocAsync");
  if (rc != GLP_OK) return rc;
  char *host = new char[tbl + offb];
  std::memcpy(host, d->windows, tbl);
  size_t *off = reinterpret_cast<size_t *>(host + tbl);
  size_t o = 0;
  for (unsigned w = 0; w < d->num_windows; ++w) {
    off[w] = o;
    o += (size_t)d->windows[w].num_sites * d->num_samples;
  }
  // a copy from pageable memory has left the host buffer when it returns
  rc = hip_check(hipMemcpyAsync(dev, host, tbl + offb, hipMemcpyHostToDevice, s), "table copy");
  delete[] host;
  if (rc == GLP_OK) {
    WorkArgs a;
    a.N = d->num_samples;
    a.H = d->num_haps;
    a.pen = d->pen;
    a.seed = d->seed;
    a.gen = d->gen;
    for (int k = 0; k < 16; ++k) a.pc[k] = d->pc16[k];
    unsigned max_wn = 0;
    for (unsigned w = 0; w < d->num_windows; ++w)
      max_wn = std::max(max_wn, d->windows[w].words_per_hap);
    const dim3 grid((d->num_samples + 255) / 256, d->num_windows), block(256);
    const size_t words = (size_t)d->num_samples * max_wn;
    const dim3 grid2((unsigned)((words + 255) / 256), d->num_windows);
    const auto *wt = reinterpret_cast<const glp_work_window *>(dev);
    const auto *wo = reinterpret_cast<const size_t *>(dev + tbl);
    float4 *ws = static_cast<float4 *>(d->workspace);
    hipLaunchKernelGGL(k_hmm_work_pass, grid, block, 0, s, wt, wo, ws, a);
    if (d->pen == 1.0f)
      hipLaunchKernelGGL((k_hmm_work_sample<false>), grid2, block, 0, s, wt, wo, ws, a);
    else
      hipLaunchKernelGGL((k_hmm_work_sample<true>), grid2, block, 0, s, wt, wo, ws, a);
    rc = hip_check(hipGetLastError(), "hmm_work launch");
  }
  (void)hipFreeAsync(dev, s);
  return rc;
}

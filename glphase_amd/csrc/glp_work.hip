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

// One lane per (window, sample).  beta: this window's workspace slice,
// [M][N] float4 (site-major, so a wavefront's stores coalesce).  GL pairs
// and beta are loaded a block of sites ahead of their use.  POW: S != 1
// (powf is the identity at S = 1).
template <bool POW>
__global__ void __launch_bounds__(256)
    k_hmm_work(const glp_work_window *__restrict__ wins, const size_t *__restrict__ beta_off,
               float4 *__restrict__ ws, WorkArgs a) {
  __shared__ float4 s_pc[4];  // rows pc[s][0..3]
  if (threadIdx.x < 4)
    s_pc[threadIdx.x] = make_float4(a.pc[4 * threadIdx.x], a.pc[4 * threadIdx.x + 1],
                                    a.pc[4 * threadIdx.x + 2], a.pc[4 * threadIdx.x + 3]);
  __syncthreads();
  const unsigned w = blockIdx.y;
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N) return;
  const glp_work_window W = wins[w];
  const unsigned M = W.num_sites, wn = W.words_per_hap, N = a.N;
  float4 *beta = ws + beta_off[w];
  const float2 *q2 = reinterpret_cast<const float2 *>(W.het_aa);
  // parent rows; indices past the panel are clamped (the MH update never
  // produces them)
  const uint64_t *rows[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned h = W.parents[4 * (size_t)i + q];
    h = h < a.H ? h : a.H - 1;
    rows[q] = W.panel + (size_t)h * wn;
  }
  constexpr int B = kWorkBlock;

  // backward recursion (impute.cpp:398-416), sites M-1 .. 1 in blocks
  float b0 = 1, b1 = 1, b2 = 1, b3 = 1;
  beta[(size_t)(M - 1) * N + i] = make_float4(1, 1, 1, 1);
  // allele words of the current 64-site word and (prefetched) the next one
  // in the walk's direction
  unsigned wcur = (M - 1) >> 6;
  uint64_t hw[4], hwn[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    hw[q] = rows[q][wcur];
    hwn[q] = wcur > 0 ? rows[q][wcur - 1] : 0ull;
  }
  float2 qn[B];
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const int m = (int)M - 1 - k;
    qn[k] = m >= 1 ? q2[(size_t)m * N + i] : make_float2(0, 0);
  }
  for (int top = (int)M - 1; top >= 1; top -= B) {
    float2 qc[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      qc[k] = qn[k];
      const int m = top - B - k;
      qn[k] = m >= 1 ? q2[(size_t)m * N + i] : make_float2(0, 0);
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int mi = top - k;
      if (mi < 1) break;
      const unsigned m = (unsigned)mi;
      if ((m >> 6) != wcur) {
        wcur = m >> 6;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          hw[q] = hwn[q];
          hwn[q] = wcur > 0 ? rows[q][wcur - 1] : 0ull;
        }
      }
      const unsigned f0 = hbit64(hw[0], m), f1 = hbit64(hw[1], m), m0 = hbit64(hw[2], m),
                     m1 = hbit64(hw[3], m);
      const float3 t = work_prob(qc[k]);
      const float e00 = work_emit(s_pc[(f0 << 1) | m0], t), e01 = work_emit(s_pc[(f0 << 1) | m1], t);
      const float e10 = work_emit(s_pc[(f1 << 1) | m0], t), e11 = work_emit(s_pc[(f1 << 1) | m1], t);
      const float bb00 = b0 * e00, bb01 = b1 * e01, bb10 = b2 * e10, bb11 = b3 * e11;
      const float *tr = W.tran + 3 * (size_t)m;
      const float t0 = tr[0], t1 = tr[1], t2 = tr[2];
      const float n0 = bb00 * t0 + (bb01 + bb10) * t1 + bb11 * t2;
      const float n1 = bb01 * t0 + (bb00 + bb11) * t1 + bb10 * t2;
      const float n2 = bb10 * t0 + (bb00 + bb11) * t1 + bb01 * t2;
      const float n3 = bb11 * t0 + (bb01 + bb10) * t1 + bb00 * t2;
      const float sum = 1.0f / (n0 + n1 + n2 + n3);
      b0 = n0 * sum;
      b1 = n1 * sum;
      b2 = n2 * sum;
      b3 = n3 * sum;
      beta[(size_t)(m - 1) * N + i] = make_float4(b0, b1, b2, b3);
    }
  }

  // forward pass with sampling (impute.cpp:418-484), sites 0 .. M-1 in blocks
  uint64_t *ha = W.hnew + (size_t)(2 * i) * wn, *hb = ha + wn;
  float l0 = 0, l1 = 0, l2 = 0, l3 = 0;
  uint64_t wa = 0, wb = 0;
  U4 x{0, 0, 0, 0};
  const float S = a.pen;
  float4 bn[B];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    hw[q] = rows[q][0];
    hwn[q] = wn > 1 ? rows[q][1] : 0ull;
  }
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const unsigned m = (unsigned)k;
    qn[k] = m < M ? q2[(size_t)m * N + i] : make_float2(0, 0);
    bn[k] = m < M ? beta[(size_t)m * N + i] : make_float4(0, 0, 0, 0);
  }
  for (unsigned base = 0; base < M; base += B) {
    float2 qc[B];
    float4 bc[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      qc[k] = qn[k];
      bc[k] = bn[k];
      const unsigned m = base + B + (unsigned)k;
      qn[k] = m < M ? q2[(size_t)m * N + i] : make_float2(0, 0);
      bn[k] = m < M ? beta[(size_t)m * N + i] : make_float4(0, 0, 0, 0);
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const unsigned m = base + (unsigned)k;
      if (m >= M) break;
      if ((m & 63u) == 0 && m) {
        const unsigned wx = (m >> 6) + 1;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          hw[q] = hwn[q];
          hwn[q] = wx < wn ? rows[q][wx] : 0ull;
        }
      }
      const unsigned f0 = hbit64(hw[0], m), f1 = hbit64(hw[1], m), m0 = hbit64(hw[2], m),
                     m1 = hbit64(hw[3], m);
      const float4 r00 = s_pc[(f0 << 1) | m0], r01 = s_pc[(f0 << 1) | m1],
                   r10 = s_pc[(f1 << 1) | m0], r11 = s_pc[(f1 << 1) | m1];
      const float3 t = work_prob(qc[k]);
      const float e00 = work_emit(r00, t), e01 = work_emit(r01, t), e10 = work_emit(r10, t),
                  e11 = work_emit(r11, t);
      if (m) {
        const float *tr = W.tran + 3 * (size_t)m;
        const float t0 = tr[0], t1 = tr[1], t2 = tr[2];
        const float n0 = l0 * t0 + l1 * t1 + l2 * t1 + l3 * t2;
        const float n1 = l0 * t1 + l1 * t0 + l2 * t2 + l3 * t1;
        const float n2 = l0 * t1 + l1 * t2 + l2 * t0 + l3 * t1;
        const float n3 = l0 * t2 + l1 * t1 + l2 * t1 + l3 * t0;
        l0 = n0 * e00;
        l1 = n1 * e01;
        l2 = n2 * e10;
        l3 = n3 * e11;
      } else {
        l0 = 0.25f * e00;
        l1 = 0.25f * e01;
        l2 = 0.25f * e10;
        l3 = 0.25f * e11;
      }
      const float sum = 1.0f / (l0 + l1 + l2 + l3);
      l0 *= sum;
      l1 *= sum;
      l2 *= sum;
      l3 *= sum;
      const float4 b = bc[k];
      const float p00 = l0 * b.x, p01 = l1 * b.y, p10 = l2 * b.z, p11 = l3 * b.w;
      float c00 = t.x * (p00 * r00.x + p01 * r01.x + p10 * r10.x + p11 * r11.x);
      float c01 = t.y * (p00 * r00.y + p01 * r01.y + p10 * r10.y + p11 * r11.y);
      float c10 = t.y * (p00 * r00.z + p01 * r01.z + p10 * r10.z + p11 * r11.z);
      float c11 = t.z * (p00 * r00.w + p01 * r01.w + p10 * r10.w + p11 * r11.w);
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
      if ((m & 63u) == 63u || m + 1 == M) {
        ha[m >> 6] = wa;
        hb[m >> 6] = wb;
        wa = wb = 0;
      }
    }
  }
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
    const dim3 grid((d->num_samples + 255) / 256, d->num_windows), block(256);
    const auto *wt = reinterpret_cast<const glp_work_window *>(dev);
    const auto *wo = reinterpret_cast<const size_t *>(dev + tbl);
    float4 *ws = static_cast<float4 *>(d->workspace);
    if (d->pen == 1.0f)
      hipLaunchKernelGGL((k_hmm_work<false>), grid, block, 0, s, wt, wo, ws, a);
    else
      hipLaunchKernelGGL((k_hmm_work<true>), grid, block, 0, s, wt, wo, ws, a);
    rc = hip_check(hipGetLastError(), "k_hmm_work launch");
  }
  (void)hipFreeAsync(dev, s);
  return rc;
}

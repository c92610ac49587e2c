// glp_work.hip -- gfx950 kernel of the haplotype resampling step that
// follows each generation's MH update: Impute::hmm_work (impute.cpp:380-485)
// for every sample, and its part of the C ABI (include/glphase_amd.h).
//
// Arithmetic contract: every float operation of impute.cpp:380-485 and of the
// emission / genotype-probability builder of Insti::initialize
// (insti.cpp:1345-1378) is kept in the reference's order, compiled with
// -ffp-contract=off, and 1.0f / x is the IEEE division (or a reciprocal
// proven equal to it where it is used); powf(x, S) is the
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
// lies inside one 64-site allele word), their GL pairs and transitions, and
// the lane's two parent words of that 64-site word.
struct WorkBlock {
  float2 q[kWorkBlock];
  float t[3 * kWorkBlock];
  uint64_t wx, wy;
};

// Loads block blk's inputs.  Site indices are clamped into [0, hi] (the
// window), so every load goes to a valid address and none sits behind a
// branch: one wait then covers a whole block, a block ahead.  Sites of the
// block outside the window get a copy of a real site's inputs, which the
// recursions never use.  The parent words come first, since a block's
// recursion needs them before its sums.
__device__ __forceinline__ void load_block(WorkBlock &b, int blk, int hi, gptr_t<float> q2,
                                           gptr_t<float> tran, gptr_t<uint64_t> rx,
                                           gptr_t<uint64_t> ry, unsigned N, unsigned i) {
  const int top = hi / kWorkBlock;  // the window's last block
  const int cb = blk < 0 ? 0 : (blk > top ? top : blk);
  const unsigned word = (unsigned)(kWorkBlock * cb) >> 6;
  b.wx = rx[word];
  b.wy = ry[word];
#pragma unroll
  for (int k = 0; k < kWorkBlock; ++k) {
    const int m = kWorkBlock * cb + k;
    const int mc = m > hi ? hi : m;
    b.q[k] = load_f2(q2 + 2 * ((size_t)mc * N + i));
    b.t[3 * k] = tran[3 * mc];
    b.t[3 * k + 1] = tran[3 * mc + 1];
    b.t[3 * k + 2] = tran[3 * mc + 2];
  }
}

// quad_perm DPP moves within each group of 4 lanes (one sample's 4 states,
// lane s = state s = 2 x + y for parents f_x, m_y)
template <int CTRL>
__device__ __forceinline__ float qperm(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, true));
}
constexpr int kQSwap1 = 0xB1;  // [1,0,3,2]
constexpr int kQSwap2 = 0x4E;  // [2,3,0,1]
constexpr int kQOpp = 0x1B;    // [3,2,1,0]: state 3 - s
constexpr int kQ0 = 0x00, kQ1 = 0x55, kQ2 = 0xAA, kQ3 = 0xFF;  // broadcasts of lane 0..3

// ((x0 + x1) + x2) + x3 of the quad's four values, in every lane: the
// reference's order of a 4-state sum.
__device__ __forceinline__ float quad_sum(float x) {
  float s = qperm<kQ0>(x) + qperm<kQ1>(x);
  s = s + qperm<kQ2>(x);
  return s + qperm<kQ3>(x);
}

// One lane's view of a sample for the two recursions.
struct WorkLane {
  gptr_t<uint64_t> rx, ry;  // this state's parents f_x, m_y
  gptr_t<float> q2, tran;
  float *out;               // [M][N] x 4: beta (backward) or scaled l (forward)
  const float4 *s_pc;       // rows pc[s][0..3] in LDS
  unsigned M, N, i, st;
};

// 1.0f / x: the IEEE division, or (FAST) the 3-instruction reciprocal, which
// equals it for normal x up to 2^126 and is NaN below (and 0 above) -- a NaN
// or zero sum leaves a NaN or all-zero state that the caller detects.
template <bool FAST>
__device__ __forceinline__ float work_recip(float x) {
  return FAST ? recip_normal(x) : 1.0f / x;
}

// Backward step over block blk (sites 8 blk + 7 down to 8 blk, those in
// [1, M-1]) of impute.cpp:398-416: site m turns beta_m into beta_{m-1},
// written to the workspace.
template <bool FAST>
__device__ __forceinline__ void bwd_block(const WorkLane &L, const WorkBlock &cur, int blk,
                                          float &b) {
  constexpr int B = kWorkBlock;
  const unsigned M = L.M, N = L.N, i = L.i, st = L.st;
  float e[B];
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const unsigned m = (unsigned)(B * blk + k);
    const unsigned row = (hbit64(cur.wx, m) << 1) | hbit64(cur.wy, m);
    e[k] = work_emit(L.s_pc[row], work_prob(cur.q[k]));
  }
#pragma unroll
  for (int k = B - 1; k >= 0; --k) {
    const int mi = B * blk + k;
    if (mi > (int)M - 1 || mi < 1) continue;
    const float t0 = cur.t[3 * k], t1 = cur.t[3 * k + 1], t2 = cur.t[3 * k + 2];
    const float bb = b * e[k];
    // bb_s t0 + (pair-cross sum) t1 + bb_{3-s} t2; the cross sum is
    // bb01 + bb10 for s = 00, 11 and bb00 + bb11 for s = 01, 10
    const float cross = qperm<kQSwap1>(bb) + qperm<kQSwap2>(bb);
    const float n = bb * t0 + cross * t1 + qperm<kQOpp>(bb) * t2;
    const float sum = work_recip<FAST>(quad_sum(n));
    b = n * sum;
    L.out[4 * ((size_t)(mi - 1) * N + i) + st] = b;
  }
}

// Backward recursion (impute.cpp:398-416) over the window, blocks from the
// last to the first, each block's inputs loaded a block ahead into the
// other of two buffers (no copy between them, so no wait for the load just
// issued).  Returns whether the last beta is a proper one (its quad sum
// positive), which fails exactly when a FAST reciprocal met a subnormal,
// zero or huge sum.
template <bool FAST>
__device__ __forceinline__ bool work_backward(const WorkLane &L) {
  const unsigned M = L.M, N = L.N, i = L.i, st = L.st;
  const int hi = (int)M - 1, top_blk = hi / kWorkBlock;
  float b = 1;
  L.out[4 * ((size_t)(M - 1) * N + i) + st] = 1.0f;
  WorkBlock x, y;
  load_block(x, top_blk, hi, L.q2, L.tran, L.rx, L.ry, N, i);
  for (int blk = top_blk;;) {
    load_block(y, blk - 1, hi, L.q2, L.tran, L.rx, L.ry, N, i);
    bwd_block<FAST>(L, x, blk, b);
    if (--blk < 0) break;
    load_block(x, blk - 1, hi, L.q2, L.tran, L.rx, L.ry, N, i);
    bwd_block<FAST>(L, y, blk, b);
    if (--blk < 0) break;
  }
  return quad_sum(b) > 0.0f;
}

// Forward step over block blk (sites 8 blk .. 8 blk + 7, those below M) of
// impute.cpp:418-455; each site's slot of the forward table takes the
// scaled state probability l.
template <bool FAST>
__device__ __forceinline__ void fwd_block(const WorkLane &L, const WorkBlock &cur, int blk,
                                          float &l) {
  constexpr int B = kWorkBlock;
  const unsigned M = L.M, N = L.N, i = L.i, st = L.st;
  // forward coefficient of state j's value in this state's update:
  // t[popcount(st ^ j)] (impute.cpp:435-445)
  const bool own[4] = {st == 0, st == 1, st == 2, st == 3};
  const bool opp[4] = {st == 3, st == 2, st == 1, st == 0};
  float e[B];
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const unsigned m = (unsigned)(B * blk + k);
    const unsigned row = (hbit64(cur.wx, m) << 1) | hbit64(cur.wy, m);
    e[k] = work_emit(L.s_pc[row], work_prob(cur.q[k]));
  }
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const unsigned m = (unsigned)(B * blk + k);
    if (m >= M) break;
    if (m) {
      const float t0 = cur.t[3 * k], t1 = cur.t[3 * k + 1], t2 = cur.t[3 * k + 2];
      float c[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = own[j] ? t0 : (opp[j] ? t2 : t1);
      const float n = qperm<kQ0>(l) * c[0] + qperm<kQ1>(l) * c[1] + qperm<kQ2>(l) * c[2] +
                      qperm<kQ3>(l) * c[3];
      l = n * e[k];
    } else {
      l = 0.25f * e[k];
    }
    const float sum = work_recip<FAST>(quad_sum(l));
    l *= sum;
    L.out[4 * ((size_t)m * N + i) + st] = l;
  }
}

// Forward recursion (impute.cpp:418-455), sites 0 .. M-1, with the same two
// buffers as work_backward.  It needs no beta, so it runs beside the
// backward recursion.  Returns whether the last state is a proper one.
template <bool FAST>
__device__ __forceinline__ bool work_forward(const WorkLane &L) {
  const unsigned M = L.M, N = L.N, i = L.i;
  const int hi = (int)M - 1, top_blk = hi / kWorkBlock;
  float l = 0;
  WorkBlock x, y;
  load_block(x, 0, hi, L.q2, L.tran, L.rx, L.ry, N, i);
  for (int blk = 0;;) {
    load_block(y, blk + 1, hi, L.q2, L.tran, L.rx, L.ry, N, i);
    fwd_block<FAST>(L, x, blk, l);
    if (++blk > top_blk) break;
    load_block(x, blk + 1, hi, L.q2, L.tran, L.rx, L.ry, N, i);
    fwd_block<FAST>(L, y, blk, l);
    if (++blk > top_blk) break;
  }
  return quad_sum(l) > 0.0f;
}

// The two recursions of impute.cpp:398-470, side by side: one lane per
// (window, sample, state, recursion); even wavefronts run the backward
// recursion of 16 samples, odd ones the forward recursion of the same 16
// (each wavefront takes one code path).  Lane s of a quad holds state s's
// value, the cross-state terms come by quad DPP moves in the reference's
// operand order, and the four lanes compute each site's normaliser
// together.  The backward recursion writes beta to the first table of the
// window's workspace, the forward one the scaled l to the second ([M][N]
// float4 each, site-major: a quad stores one float4, a wavefront 16
// samples' contiguous float4s); the sampling step multiplies them into the
// posteriors l * beta, the reference's product.  Sites go in aligned blocks
// of 8 whose GL pairs, transitions and parent words are loaded a block
// ahead.  A recursion runs with the short reciprocal first; a wavefront in
// which some sample met a sum outside its exact range redoes it with the
// IEEE division, so every stored value is the division's.  Parent indices
// past the panel are clamped (the MH update never produces them).
__global__ void __launch_bounds__(256)
    k_hmm_work_pass(const glp_work_window *__restrict__ wins, const size_t *__restrict__ ws_off,
                    float4 *__restrict__ ws, WorkArgs a) {
  __shared__ float4 s_pc[4];  // rows pc[s][0..3]
  load_pc_rows(s_pc, a);
  const unsigned w = blockIdx.y;
  const unsigned g = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned wave = g >> 6, fwd = wave & 1u;
  const unsigned i = (wave >> 1) * 16u + ((g & 63u) >> 2), st = g & 3u;
  if (i >= a.N) return;  // whole quads leave together
  const glp_work_window W = wins[w];
  WorkLane L;
  L.M = W.num_sites;
  L.N = a.N;
  L.i = i;
  L.st = st;
  L.out = reinterpret_cast<float *>(ws + ws_off[w] + (fwd ? (size_t)L.M * a.N : 0));
  L.q2 = gmem(W.het_aa);
  L.tran = gmem(W.tran);
  L.s_pc = s_pc;
  {
    const unsigned wn = W.words_per_hap;
    unsigned hx = gmem(W.parents)[4 * (size_t)i + (st >> 1)];
    unsigned hy = gmem(W.parents)[4 * (size_t)i + 2 + (st & 1u)];
    hx = hx < a.H ? hx : a.H - 1;
    hy = hy < a.H ? hy : a.H - 1;
    L.rx = gmem(W.panel) + (size_t)hx * wn;
    L.ry = gmem(W.panel) + (size_t)hy * wn;
  }
  if (fwd) {
    if (__builtin_expect(__any(!work_forward<true>(L)), 0)) work_forward<false>(L);
  } else {
    if (__builtin_expect(__any(!work_backward<true>(L)), 0)) work_backward<false>(L);
  }
}

// The sampling of impute.cpp:457-484, one lane per (window, sample, 64-site
// word), sample fastest so a wavefront's reads of a site coalesce: from the
// posteriors p = l * beta (the two tables k_hmm_work_pass wrote) and the GL
// pair, the genotype
// weights c_g = (t_g * sum_s p_s pc[s][g])^S, one Philox word per site, the
// genotype choice, and the two new haplotype words.  POW: S != 1 (powf is
// the identity at S = 1).
template <bool POW>
__global__ void __launch_bounds__(256)
    k_hmm_work_sample(const glp_work_window *__restrict__ wins, const size_t *__restrict__ ws_off,
                      const float4 *__restrict__ ws, WorkArgs a) {
  __shared__ float4 s_pc[4];
  load_pc_rows(s_pc, a);
  const unsigned w = blockIdx.y;
  const glp_work_window W = wins[w];
  const unsigned M = W.num_sites, wn = W.words_per_hap, N = a.N;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * wn) return;
  const unsigned i = (unsigned)(t % N), wd = (unsigned)(t / N);
  const auto beta = gmem(reinterpret_cast<const float *>(ws + ws_off[w]));
  const auto lfwd = gmem(reinterpret_cast<const float *>(ws + ws_off[w] + (size_t)M * N));
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
    const float4 bq = load_f4(beta + 4 * ((size_t)m * N + i));
    const float4 lq = load_f4(lfwd + 4 * ((size_t)m * N + i));
    const float4 p = make_float4(lq.x * bq.x, lq.y * bq.y, lq.z * bq.z, lq.w * bq.w);
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
    b += 2 * (size_t)d->windows[w].num_sites * d->num_samples * sizeof(float4);
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
    o += 2 * (size_t)d->windows[w].num_sites * d->num_samples;
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
    // two wavefronts (backward, forward) per 16 samples
    const size_t waves = 2 * (((size_t)d->num_samples + 15) / 16);
    const dim3 grid((unsigned)((64 * waves + 255) / 256), d->num_windows), block(256);
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

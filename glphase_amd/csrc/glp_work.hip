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


// Component j (a compile-time index) of v.
__device__ __forceinline__ float f4c(const float4 &v, int j) {
  return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
}

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

// quad_perm DPP moves within each group of 4 lanes (one sample's 4 states,
// lane s = state s = 2 x + y for parents f_x, m_y).  Every lane of a
// quad_perm move has a source lane, so the move needs no old value and the
// compiler folds it into the consuming add or multiply.
template <int CTRL>
__device__ __forceinline__ float qperm(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, true));
}
constexpr int kQSwap1 = 0xB1;  // [1,0,3,2]
constexpr int kQSwap2 = 0x4E;  // [2,3,0,1]
constexpr int kQOpp = 0x1B;    // [3,2,1,0]: state 3 - s
constexpr int kQ0 = 0x00, kQ1 = 0x55, kQ2 = 0xAA, kQ3 = 0xFF;  // broadcasts of lane 0..3

// ((x0 + x1) + x2) + x3 of the quad's four values, in every lane: the
// reference's order of a 4-state sum.  The four broadcasts are separate
// moves of x (a DPP operation whose other operand a VALU has just written
// waits two wait states; moves of x wait once, for x, and the three sums
// are plain adds).  The two wait states x needs are taken in the same block.
__device__ __forceinline__ float quad_sum(float x) {
  float a, b, c, d;
  asm volatile(
      "s_nop 1\n\t"
      "v_mov_b32_dpp %0, %4 quad_perm:[0,0,0,0] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_mov_b32_dpp %1, %4 quad_perm:[1,1,1,1] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_mov_b32_dpp %2, %4 quad_perm:[2,2,2,2] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "v_mov_b32_dpp %3, %4 quad_perm:[3,3,3,3] row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "=&v"(a), "=&v"(b), "=&v"(c), "=&v"(d)
      : "v"(x));
  float s = a + b;
  s = s + c;
  return s + d;
}

// 1.0f / x: the IEEE division, or (FAST) the 3-instruction reciprocal, which
// equals it for normal x up to 2^126 and is NaN below (and 0 above) -- a NaN
// or zero sum leaves a NaN or all-zero state that the caller detects.
template <bool FAST>
__device__ __forceinline__ float work_recip(float x) {
  return FAST ? recip_normal(x) : 1.0f / x;
}

// ---- the emission table (k_work_emit) ----

// Insti::initialize's emissions (insti.cpp:1345-1378) for every (site,
// sample) of a window: the four pc rows against the genotype probabilities
// of the GL pair, one float4 per (site, sample), site-major.
__global__ void __launch_bounds__(256)
    k_work_emit(const float2 *__restrict__ het_aa, size_t n, float4 *__restrict__ emit,
                WorkArgs a) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float3 p = work_prob(load_f2(gmem(reinterpret_cast<const float *>(het_aa + t))));
  float e[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    e[r] = work_emit(make_float4(a.pc[4 * r], a.pc[4 * r + 1], a.pc[4 * r + 2], a.pc[4 * r + 3]),
                     p);
  store_f4(reinterpret_cast<float *>(emit + t), e[0], e[1], e[2], e[3]);
}

// ---- the two recursions (k_hmm_work_pass) ----
//
// Sites go in aligned blocks of 32 (half a 64-site panel word).  Each lane
// gathers its state's emission at every site of a block from the window's
// emission table -- the float of the row its two parental alleles select --
// a whole block ahead, into the other of two register buffers; the parents'
// allele bits that pick the rows arrive a block before that.  The window's
// transitions are staged through LDS in chunks of 256 sites,
// double-buffered with one workgroup barrier per chunk: the forward
// recursion reads, per site and state, the coefficient vector
// (t[popcount(s ^ j)], j = 0..3) in one LDS read, the backward one
// (t0, t1, t2).  The emission table and the output tables are reached
// through buffer descriptors, rebased per block, with the site's row offset
// in the scalar offset.
constexpr int kPassB = 32;                     // sites per block
constexpr int kPassChunk = 256;                // sites per staged chunk of transitions
constexpr int kPassBPC = kPassChunk / kPassB;  // blocks per chunk
constexpr int kPassWaves = 2;                  // wavefronts per workgroup, one recursion
constexpr int kPassThreads = 64 * kPassWaves;
constexpr int kPassStage = kPassChunk / kPassThreads;  // chunk sites staged per lane
static_assert(kPassChunk % kPassThreads == 0, "chunk sites split evenly over the lanes");
// the largest sample count whose per-block offsets stay below 2^31 bytes
constexpr unsigned kPassMaxSamples = 1u << 22;

typedef __amdgpu_buffer_rsrc_t buf_rsrc;

// Raw (untyped, unswizzled) buffer over device memory from byte address
// base, the same in every lane (read from the first, so the descriptor is
// scalar); word 3 selects 32-bit data, the record count covers every offset
// below 2^31.
__device__ __forceinline__ buf_rsrc raw_buffer(uintptr_t base) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
  return __builtin_amdgcn_make_buffer_rsrc(
      reinterpret_cast<void *>(((uintptr_t)hi << 32) | lo), (short)0, 0x7fffffff, 0x00020000);
}

// The output tables hold, per group of 4 sites (g = m / 4), per sample and
// per state, the 4 sites' values: float [G][N][4][4] with G = ceil(M / 4),
// so that a lane stores 4 sites at once and a wavefront's store covers 16
// samples' contiguous 1 KB.  Sites past M - 1 in the last group are padding.
__host__ __device__ inline size_t pass_groups(unsigned M) { return (M + 3u) / 4u; }
// float4 units of one output table
__host__ __device__ inline size_t pass_table_f4(unsigned M, unsigned N) {
  return 4 * pass_groups(M) * N;
}

// One lane's view of a sample for one recursion.
struct PassLane {
  uintptr_t emit, out;       // byte addresses: emission table [M][N] float4, output table
  uint32_t row, grow;        // bytes per emission site row (16 N) and output group (64 N)
  uint32_t e_off, out_off;   // the lane's offsets in a row: 16 i; 64 i + 16 st
  gptr_t<uint32_t> px, py;   // parents f_x, m_y: allele bits, 32 sites per word
  int M;
};

struct PassBlock {
  float e[kPassB];  // the lane's emission at the block's sites
};

// The 16 low bits of v moved to the even bits of a 32-bit word.
__device__ __forceinline__ uint32_t spread16(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  return (v | (v << 1)) & 0x55555555u;
}

// Issues the gathers of block blk, whose parents' allele bits are (x, y).
// FULL: all 32 sites lie in the window; otherwise site indices are clamped
// to M - 1 (with the row bits of the unclamped site: any row of a valid
// site is a valid address), so every gather reads inside the table.  The
// two bit words are interleaved once per block, so that a site's row
// 2 x_k + y_k is one bit-field read.
template <bool FULL>
__device__ __forceinline__ void pass_gather(PassBlock &b, const PassLane &L, int blk, uint32_t x,
                                            uint32_t y) {
  const int m0 = kPassB * blk;
  const buf_rsrc r = raw_buffer(L.emit + (uintptr_t)m0 * L.row);
  const int last = L.M - 1 - m0;
  const uint32_t z[2] = {(spread16(x) << 1) | spread16(y),
                         (spread16(x >> 16) << 1) | spread16(y >> 16)};
#pragma unroll
  for (int k = 0; k < kPassB; ++k) {
    const int kc = FULL ? k : (k < last ? k : last);
    const uint32_t off = L.e_off + (((z[k >> 4] >> (2 * (k & 15))) & 3u) << 2);
    b.e[k] = __uint_as_float(
        __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, kc * (int)L.row, 0));
  }
}

__device__ __forceinline__ void pass_gather_any(PassBlock &b, const PassLane &L, int blk,
                                                uint32_t x, uint32_t y) {
  if (kPassB * blk + kPassB <= L.M)
    pass_gather<true>(b, L, blk, x, y);
  else
    pass_gather<false>(b, L, blk, x, y);
}

__device__ __forceinline__ void pass_copy(PassBlock &dst, const PassBlock &src) {
#pragma unroll
  for (int k = 0; k < kPassB; ++k) dst.e[k] = src.e[k];
}

// Stores the lane's values of the 4 sites of group j of the descriptor's
// base (4 consecutive floats).
typedef uint32_t glp_u4s __attribute__((ext_vector_type(4)));
__device__ __forceinline__ glp_u4s pack4(const float (&o)[4]) {
  glp_u4s v;
  v.x = __float_as_uint(o[0]);
  v.y = __float_as_uint(o[1]);
  v.z = __float_as_uint(o[2]);
  v.w = __float_as_uint(o[3]);
  return v;
}

__device__ __forceinline__ void pass_store4(const float (&o)[4], const buf_rsrc &r,
                                            const PassLane &L, int j) {
  __builtin_amdgcn_raw_buffer_store_b128(pack4(o), r, (int)L.out_off, j * (int)L.grow, 0);
}

// The data registers of a store may not be overwritten before the store has
// completed (the memory counter waits for it), so the data of the groups
// stored last stay in two rotating register sets: set p (the group's
// parity) is refilled only two groups, 8 sites, after its store.
struct StoreRing {
  glp_u4s v[2];
};

template <int P>
__device__ __forceinline__ void pass_store4_ring(StoreRing &R, const float (&o)[4],
                                                 const buf_rsrc &r, const PassLane &L, int j) {
  asm volatile("" ::"v"(R.v[P]));  // the last store of set P keeps its registers until here
  R.v[P] = pack4(o);
  __builtin_amdgcn_raw_buffer_store_b128(R.v[P], r, (int)L.out_off, j * (int)L.grow, 0);
}

// Forward steps over block blk (impute.cpp:418-455); co: the block's first
// site's coefficient vector for this lane's state (site k at co[4 k]).  Each
// site's entry of the table takes the scaled l, stored per 4 sites from o.
// FULL: sites 1 .. M-1 only; otherwise site 0 starts the recursion, sites
// past M - 1 are skipped and the window's last group is stored whole.
template <bool FAST, bool FULL>
__device__ __forceinline__ void fwd_pass_block(const PassLane &L, const PassBlock &b, int blk,
                                               const float4 *co, float &l, float (&o)[4],
                                               StoreRing &R) {
  const int m0 = kPassB * blk;
  const buf_rsrc r = raw_buffer(L.out + (uintptr_t)(m0 / 4) * L.grow);
  float4 cn = co[0];  // each site's LDS read is issued a site ahead of its use
#pragma unroll
  for (int k = 0; k < kPassB; ++k) {
    const int m = m0 + k;
    if (!FULL && m >= L.M) break;
    const float4 c = cn;
    if (k + 1 < kPassB) cn = co[4 * (k + 1)];
    __builtin_amdgcn_sched_barrier(0);
    if (FULL || m) {
      float p0 = qperm<kQ0>(l) * c.x, p1 = qperm<kQ1>(l) * c.y, p2 = qperm<kQ2>(l) * c.z,
            p3 = qperm<kQ3>(l) * c.w;
      // the four products in four registers before the sums (reusing one for
      // the next product costs a wait state)
      asm volatile("" : "+v"(p0), "+v"(p1), "+v"(p2), "+v"(p3));
      const float n = p0 + p1 + p2 + p3;
      l = n * b.e[k];
    } else {
      l = 0.25f * b.e[k];
    }
    l *= work_recip<FAST>(quad_sum(l));
    o[k & 3] = l;
    if ((k & 3) == 3) {
      // group k >> 2 of a block of 8 groups: its parity is its global one
      if ((k >> 2) & 1)
        pass_store4_ring<1>(R, o, r, L, k >> 2);
      else
        pass_store4_ring<0>(R, o, r, L, k >> 2);
    }
  }
  if (!FULL && (L.M & 3) && L.M - 1 - m0 < kPassB) pass_store4(o, r, L, (L.M - 1 - m0) >> 2);
}

// Backward steps over block blk, sites 32 blk + 31 down to 32 blk
// (impute.cpp:398-416); tr: the block's first site's (t0, t1, t2).  Site m
// turns beta_m into beta_{m-1}, entry m - 1 of the table: slot (m - 1) & 3 of
// o, whose group is stored when its lowest entry is in.  FULL: every site in
// 1 .. M-1; otherwise the others are skipped.
template <bool FAST, bool FULL>
__device__ __forceinline__ void bwd_pass_block(const PassLane &L, const PassBlock &b, int blk,
                                               const float4 *tr, float &v, float (&o)[4],
                                               StoreRing &R) {
  const int m0 = kPassB * blk;
  // entry m0 - 1 + k sits in group m0 / 4 + (k - 1) / 4 (the one of k = 0
  // is stored from the block below)
  const buf_rsrc r = raw_buffer(L.out + (uintptr_t)(m0 / 4) * L.grow);
  float4 tn = tr[kPassB - 1];  // each site's LDS read is issued a site ahead of its use
#pragma unroll
  for (int k = kPassB - 1; k >= 0; --k) {
    const int m = m0 + k;
    const float4 t = tn;
    if (k > 0) tn = tr[k - 1];
    __builtin_amdgcn_sched_barrier(0);
    if (!FULL && (m > L.M - 1 || m < 1)) continue;
    const float bb = v * b.e[k];
    // bb_s t0 + (pair-cross sum) t1 + bb_{3-s} t2; the cross sum is
    // bb01 + bb10 for s = 00, 11 and bb00 + bb11 for s = 01, 10
    const float cross = qperm<kQSwap1>(bb) + qperm<kQSwap2>(bb);
    const float n = bb * t.x + cross * t.y + qperm<kQOpp>(bb) * t.z;
    v = n * work_recip<FAST>(quad_sum(n));
    o[(k + 3) & 3] = v;
    if ((k & 3) == 1) {
      if ((k >> 2) & 1)
        pass_store4_ring<1>(R, o, r, L, k >> 2);
      else
        pass_store4_ring<0>(R, o, r, L, k >> 2);
    }
  }
}

// Transitions of one chunk held in registers between their load and their
// LDS write: rows min(256 ch + s, M - 1) for the lane's sites s.
struct PassStage {
  float t[kPassStage][3];
};

__device__ __forceinline__ void stage_load(PassStage &S, gptr_t<float> tran, int ch, int M) {
#pragma unroll
  for (int r = 0; r < kPassStage; ++r) {
    const int m = kPassChunk * ch + (int)threadIdx.x + kPassThreads * r;
    const int mc = m < M - 1 ? m : M - 1;
    S.t[r][0] = tran[3 * mc];
    S.t[r][1] = tran[3 * mc + 1];
    S.t[r][2] = tran[3 * mc + 2];
  }
}

// Writes a chunk into LDS: forward, per site the four states' coefficient
// vectors (state s: t[popcount(s ^ j)] for j = 0..3, impute.cpp:435-445);
// backward, per site (t0, t1, t2, 0).
template <bool FWD>
__device__ __forceinline__ void stage_write(const PassStage &S, float4 *buf) {
#pragma unroll
  for (int r = 0; r < kPassStage; ++r) {
    const int s = (int)threadIdx.x + kPassThreads * r;
    float t0 = S.t[r][0], t1 = S.t[r][1], t2 = S.t[r][2];
    // the values are taken here, a chunk after their loads: no copy of them
    // may be scheduled next to the loads (it would wait for them there)
    asm volatile("" : "+v"(t0), "+v"(t1), "+v"(t2));
    if (FWD) {
      buf[4 * s] = make_float4(t0, t1, t1, t2);
      buf[4 * s + 1] = make_float4(t1, t0, t2, t1);
      buf[4 * s + 2] = make_float4(t1, t2, t0, t1);
      buf[4 * s + 3] = make_float4(t2, t1, t1, t0);
    } else {
      buf[s] = make_float4(t0, t1, t2, 0.0f);
    }
  }
}

// Workgroup barrier that orders LDS accesses only: global loads and stores
// in flight stay in flight (a full __syncthreads waits for all of them).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Moves the workgroup onto chunk ch: writes its staged transitions into LDS
// buffer ch & 1 (last read two chunks ago, before the previous barrier),
// waits for every wavefront, and loads chunk next (if in the window).
template <bool FWD>
__device__ __forceinline__ void stage_enter(PassStage &S, float4 *stage, gptr_t<float> tran,
                                            int ch, int next, int nch, int M) {
  stage_write<FWD>(S, stage + (FWD ? 4 : 1) * kPassChunk * (ch & 1));
  lds_barrier();
  if (next >= 0 && next < nch) stage_load(S, tran, next, M);
}

// Forward recursion (impute.cpp:418-455), sites 0 .. M-1.  Returns whether
// the last state is a proper one (its quad sum positive), which fails
// exactly when a FAST reciprocal met a subnormal, zero or huge sum.  Every
// wavefront of the workgroup runs the same chunk steps and barriers; idle
// ones (no sample) skip the loads and steps.  At each block the gathers of
// the next block are issued first, then the allele bits of the one after,
// then the block's steps.
template <bool FAST>
__device__ bool pass_forward(const PassLane &L, bool active, gptr_t<float> tran, float4 *stage,
                             int st) {
  const int M = L.M, top = (M - 1) / kPassB, nch = top / kPassBPC + 1;
  PassStage S;
  stage_load(S, tran, 0, M);
  stage_enter<true>(S, stage, tran, 0, 1, nch, M);
  auto co = [&](int blk) {
    const int ch = blk / kPassBPC;
    return stage + 4 * (kPassChunk * (ch & 1) + kPassB * (blk % kPassBPC)) + st;
  };
  auto enter = [&](int blk) {
    if (blk % kPassBPC == 0) {
      const int ch = blk / kPassBPC;
      stage_enter<true>(S, stage, tran, ch, ch + 1, nch, M);
    }
  };
  float l = 0, o[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  StoreRing R = {};
  PassBlock A, B;
  uint32_t bx = 0, by = 0;  // allele bits of the next block to gather
  if (active) {
    const uint32_t x0 = L.px[0], y0 = L.py[0];
    if (top >= 1) {
      bx = L.px[1];
      by = L.py[1];
    }
    pass_gather_any(A, L, 0, x0, y0);
  }
  if (top > 0) {
    if (active) {
      pass_gather_any(B, L, 1, bx, by);
      if (top >= 2) {
        bx = L.px[2];
        by = L.py[2];
      }
      fwd_pass_block<FAST, false>(L, A, 0, co(0), l, o, R);
    }
    int blk = 1;  // in B
    enter(blk);
    for (;;) {
      if (blk == top) {
        if (active) pass_copy(A, B);
        break;
      }
      if (active) {
        pass_gather_any(A, L, blk + 1, bx, by);
        if (blk + 2 <= top) {
          bx = L.px[blk + 2];
          by = L.py[blk + 2];
        }
        fwd_pass_block<FAST, true>(L, B, blk, co(blk), l, o, R);
      }
      enter(++blk);  // in A
      if (blk == top) break;
      if (active) {
        pass_gather_any(B, L, blk + 1, bx, by);
        if (blk + 2 <= top) {
          bx = L.px[blk + 2];
          by = L.py[blk + 2];
        }
        fwd_pass_block<FAST, true>(L, A, blk, co(blk), l, o, R);
      }
      enter(++blk);  // in B
    }
  }
  if (active) fwd_pass_block<FAST, false>(L, A, top, co(top), l, o, R);
  return !active || quad_sum(l) > 0.0f;
}

// Backward recursion (impute.cpp:398-416), sites M-1 down to 1, with the
// same buffers and chunk steps in descending order.
template <bool FAST>
__device__ bool pass_backward(const PassLane &L, bool active, gptr_t<float> tran, float4 *stage) {
  const int M = L.M, top = (M - 1) / kPassB, nch = top / kPassBPC + 1;
  PassStage S;
  stage_load(S, tran, nch - 1, M);
  stage_enter<false>(S, stage, tran, nch - 1, nch - 2, nch, M);
  auto tr = [&](int blk) {
    const int ch = blk / kPassBPC;
    return stage + kPassChunk * (ch & 1) + kPassB * (blk % kPassBPC);
  };
  auto enter = [&](int blk) {
    if (blk % kPassBPC == kPassBPC - 1) {
      const int ch = blk / kPassBPC;
      stage_enter<false>(S, stage, tran, ch, ch - 1, nch, M);
    }
  };
  // beta_{M-1} = 1: every slot starts at 1, so the group of entry M - 1 holds
  // it (and padding past it) when stored; stored at once if M - 1 is its
  // group's lowest entry
  float v = 1, o[4] = {1.0f, 1.0f, 1.0f, 1.0f};
  StoreRing R = {};
  PassBlock A, B;
  uint32_t bx = 0, by = 0;  // allele bits of the next block to gather
  if (active) {
    if (((M - 1) & 3) == 0) pass_store4(o, raw_buffer(L.out), L, (M - 1) >> 2);
    const uint32_t xt = L.px[top], yt = L.py[top];
    if (top >= 1) {
      bx = L.px[top - 1];
      by = L.py[top - 1];
    }
    pass_gather_any(A, L, top, xt, yt);
  }
  if (top > 0) {
    if (active) {
      pass_gather<true>(B, L, top - 1, bx, by);
      if (top >= 2) {
        bx = L.px[top - 2];
        by = L.py[top - 2];
      }
      bwd_pass_block<FAST, false>(L, A, top, tr(top), v, o, R);
    }
    int blk = top - 1;  // in B
    enter(blk);
    for (;;) {
      if (blk == 0) {
        if (active) pass_copy(A, B);
        break;
      }
      if (active) {
        pass_gather<true>(A, L, blk - 1, bx, by);
        if (blk >= 2) {
          bx = L.px[blk - 2];
          by = L.py[blk - 2];
        }
        bwd_pass_block<FAST, true>(L, B, blk, tr(blk), v, o, R);
      }
      enter(--blk);  // in A
      if (blk == 0) break;
      if (active) {
        pass_gather<true>(B, L, blk - 1, bx, by);
        if (blk >= 2) {
          bx = L.px[blk - 2];
          by = L.py[blk - 2];
        }
        bwd_pass_block<FAST, true>(L, A, blk, tr(blk), v, o, R);
      }
      enter(--blk);  // in B
    }
  }
  if (active) bwd_pass_block<FAST, false>(L, A, 0, tr(0), v, o, R);
  return !active || quad_sum(v) > 0.0f;
}

// The two recursions of impute.cpp:398-470, side by side: workgroups of
// even blockIdx.x run the backward recursion of 16 kPassWaves samples, odd
// ones the forward recursion of the same samples; one lane per (sample,
// state), lane s of a quad holding state s's value.  The cross-state terms
// come by quad DPP moves in the reference's operand order, and the four
// lanes compute each site's normaliser together.  The backward recursion
// writes beta to the first table of the window's workspace, the forward one
// the scaled l to the second (per group of 4 sites, sample and state the 4
// sites' values); the sampling step multiplies them into the posteriors
// l * beta, the reference's product.  A workgroup runs its recursion with
// the short reciprocal first; if some sample met a sum outside its exact
// range, the workgroup redoes it with the IEEE division, so every stored
// value is the division's.  Lanes past the last sample repeat the last
// sample's work and store its (identical) values.  Parent indices past the
// panel are clamped (the MH update never produces them).
__global__ void __launch_bounds__(kPassThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
    k_hmm_work_pass(const glp_work_window *__restrict__ wins, const size_t *__restrict__ ws_off,
                    float4 *__restrict__ ws, WorkArgs a) {
  __shared__ float4 s_stage[2 * 4 * kPassChunk];
  const unsigned w = blockIdx.y;
  const bool fwd = blockIdx.x & 1u;
  const unsigned lane = threadIdx.x & 63u;
  // the wavefront's first sample (uniform: the branches on it are too)
  const unsigned i0 = __builtin_amdgcn_readfirstlane(
      ((blockIdx.x >> 1) * kPassWaves + (threadIdx.x >> 6)) * 16u);
  const bool active = i0 < a.N;
  const unsigned i = min(i0 + (lane >> 2), a.N - 1), st = lane & 3u;
  const glp_work_window W = wins[w];
  PassLane L;
  L.M = (int)W.num_sites;
  L.emit = reinterpret_cast<uintptr_t>(W.emit);
  L.out = reinterpret_cast<uintptr_t>(ws + ws_off[w] +
                                     (fwd ? pass_table_f4(W.num_sites, a.N) : 0));
  L.row = 16u * a.N;
  L.grow = 64u * a.N;
  L.e_off = 16u * i;
  L.out_off = 64u * i + 16u * st;
  {
    const unsigned wn = W.words_per_hap;
    unsigned hx = gmem(W.parents)[4 * (size_t)i + (st >> 1)];
    unsigned hy = gmem(W.parents)[4 * (size_t)i + 2 + (st & 1u)];
    hx = hx < a.H ? hx : a.H - 1;
    hy = hy < a.H ? hy : a.H - 1;
    L.px = gmem(reinterpret_cast<const uint32_t *>(W.panel + (size_t)hx * wn));
    L.py = gmem(reinterpret_cast<const uint32_t *>(W.panel + (size_t)hy * wn));
  }
  const auto tran = gmem(W.tran);
  if (fwd) {
    const bool ok = pass_forward<true>(L, active, tran, s_stage, (int)st);
    if (__builtin_expect(__syncthreads_or(!ok), 0))
      pass_forward<false>(L, active, tran, s_stage, (int)st);
  } else {
    const bool ok = pass_backward<true>(L, active, tran, s_stage);
    if (__builtin_expect(__syncthreads_or(!ok), 0))
      pass_backward<false>(L, active, tran, s_stage);
  }
}

// The sampling of impute.cpp:457-484, one lane per (window, sample, 64-site
// word), sample fastest so a wavefront's reads coalesce: from the posteriors
// p = l * beta (the two tables k_hmm_work_pass wrote, read per group of 4
// sites: each state's 4 values in one load) and the GL pair, the genotype
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
  const auto lfwd = gmem(reinterpret_cast<const float *>(ws + ws_off[w] + pass_table_f4(M, N)));
  const auto q2 = gmem(W.het_aa);
  uint64_t hw[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    unsigned h = gmem(W.parents)[4 * (size_t)i + q];
    h = h < a.H ? h : a.H - 1;
    hw[q] = gmem(W.panel)[(size_t)h * wn + wd];
  }
  const float S = a.pen;
  const unsigned G = (unsigned)pass_groups(M);
  const unsigned g1 = 16u * wd + 16u < G ? 16u * wd + 16u : G;
  uint64_t wa = 0, wb = 0;
  for (unsigned gi = 16u * wd; gi < g1; ++gi) {
    // states 0..3 of the group's 4 sites, from each table
    const size_t rec = 16 * ((size_t)gi * N + i);
    float4 bs[4], ls[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      bs[s] = load_f4(beta + rec + 4 * s);
      ls[s] = load_f4(lfwd + rec + 4 * s);
    }
    const U4 x = draw(a.seed, kWorkTag | gi, i, w, a.gen);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
    const unsigned m = 4u * gi + (unsigned)j;
    if (m >= M) break;
    const float4 p = make_float4(f4c(ls[0], j) * f4c(bs[0], j), f4c(ls[1], j) * f4c(bs[1], j),
                                 f4c(ls[2], j) * f4c(bs[2], j), f4c(ls[3], j) * f4c(bs[3], j));
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
    const uint32_t xm = j == 0 ? x.x : j == 1 ? x.y : j == 2 ? x.z : x.w;
    const unsigned g = work_choose(c00, c01, c10, c11, xm);
    wa |= (uint64_t)(g >> 1) << (m & 63u);
    wb |= (uint64_t)(g & 1u) << (m & 63u);
    }
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
  if (d->num_samples > kPassMaxSamples)
    return set_error(GLP_ERR_ARG, "num_samples above 4,194,304");
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

// Workspace layout: per window, the beta table then the scaled-l table
// (float [ceil(M / 4)][N][4][4] each), windows in order; after them the
// emission tables ([M][N] float4) of the windows given none, in window order.
extern "C" size_t glp_work_workspace_bytes(const glp_work_desc *d) {
  if (!d || !d->windows) return 0;
  size_t b = 0;
  for (unsigned w = 0; w < d->num_windows; ++w) {
    const unsigned M = d->windows[w].num_sites;
    b += 2 * pass_table_f4(M, d->num_samples) * sizeof(float4);
    if (!d->windows[w].emit) b += (size_t)M * d->num_samples * sizeof(float4);
  }
  return b;
}

namespace {

WorkArgs work_args(unsigned N, unsigned H, const float *pc16) {
  WorkArgs a{};
  a.N = N;
  a.H = H;
  for (int k = 0; k < 16; ++k) a.pc[k] = pc16[k];
  return a;
}

void launch_emit(const float *het_aa, size_t cells, float *emit, const WorkArgs &a,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_work_emit, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const float2 *>(het_aa), cells,
                     reinterpret_cast<float4 *>(emit), a);
}

}  // namespace

extern "C" int glp_work_emit(const float *het_aa, unsigned num_sites, unsigned num_samples,
                             const float *pc16, float *emit, void *stream) {
  if (!het_aa || !pc16 || !emit) return set_error(GLP_ERR_ARG, "null pointer");
  if (num_sites == 0 || num_samples == 0) return set_error(GLP_ERR_ARG, "empty window");
  const size_t cells = (size_t)num_sites * num_samples;
  if ((cells + 255) / 256 > 0x7FFFFFFFu) return set_error(GLP_ERR_ARG, "window too large");
  launch_emit(het_aa, cells, emit, work_args(num_samples, 1, pc16),
              static_cast<hipStream_t>(stream));
  return hip_check(hipGetLastError(), "work emit launch");
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
  auto *hw = reinterpret_cast<glp_work_window *>(host);
  size_t *off = reinterpret_cast<size_t *>(host + tbl);
  float4 *ws = static_cast<float4 *>(d->workspace);
  WorkArgs a = work_args(d->num_samples, d->num_haps, d->pc16);
  a.pen = d->pen;
  a.seed = d->seed;
  a.gen = d->gen;
  size_t o = 0;
  for (unsigned w = 0; w < d->num_windows; ++w) {
    off[w] = o;
    o += 2 * pass_table_f4(d->windows[w].num_sites, d->num_samples);
  }
  // windows without an emission table get one built in the workspace
  for (unsigned w = 0; w < d->num_windows; ++w) {
    if (hw[w].emit) continue;
    const size_t cells = (size_t)hw[w].num_sites * d->num_samples;
    float *e = reinterpret_cast<float *>(ws + o);
    launch_emit(hw[w].het_aa, cells, e, a, s);
    hw[w].emit = e;
    o += cells;
  }
  // a copy from pageable memory has left the host buffer when it returns
  rc = hip_check(hipMemcpyAsync(dev, host, tbl + offb, hipMemcpyHostToDevice, s), "table copy");
  delete[] host;
  if (rc == GLP_OK) {
    unsigned max_wn = 0;
    for (unsigned w = 0; w < d->num_windows; ++w)
      max_wn = std::max(max_wn, d->windows[w].words_per_hap);
    // two workgroups (backward, forward) per 16 kPassWaves samples
    const unsigned per_group = 16u * kPassWaves;
    const dim3 grid(2u * ((d->num_samples + per_group - 1) / per_group), d->num_windows),
        block(kPassThreads);
    const size_t words = (size_t)d->num_samples * max_wn;
    const dim3 grid2((unsigned)((words + 255) / 256), d->num_windows);
    const auto *wt = reinterpret_cast<const glp_work_window *>(dev);
    const auto *wo = reinterpret_cast<const size_t *>(dev + tbl);
    hipLaunchKernelGGL(k_hmm_work_pass, grid, block, 0, s, wt, wo, ws, a);
    const dim3 block2(256);
    if (d->pen == 1.0f)
      hipLaunchKernelGGL((k_hmm_work_sample<false>), grid2, block2, 0, s, wt, wo, ws, a);
    else
      hipLaunchKernelGGL((k_hmm_work_sample<true>), grid2, block2, 0, s, wt, wo, ws, a);
    rc = hip_check(hipGetLastError(), "hmm_work launch");
  }
  (void)hipFreeAsync(dev, s);
  return rc;
}

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
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/glphase_amd.h"
#include "glp_common.hpp"
#include "glp_internal.hpp"

namespace glp {

#define GLP_TRY_K(expr)              \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != GLP_OK) return rc_;   \
  } while (0)

// ---------------------------------------------------------------------------
// correctly rounded logf
// ---------------------------------------------------------------------------
// (float)log((double)x) is the correctly rounded logf of x for every positive
// float except 20, whose log lies so close to a float rounding midpoint that
// a 2-ulp change of the double log flips the float (exhaustive search over
// all positive floats; values from an 80-digit decimal log).  They sit in a
// 32-slot table under the perfect hash (bits * 0xc2eba581) >> 27.
__constant__ uint32_t c_logf_hard[32][2] = {
    {0x4c5d65a5u, 0x418f034bu}, {0x3bf86ef0u, 0xc09c399eu},
    {0xffffffffu, 0x00000000u}, {0xffffffffu, 0x00000000u},
    {0xffffffffu, 0x00000000u}, {0x41178febu, 0x400fe5e7u},
    {0x66a8c860u, 0x42595e46u}, {0xffffffffu, 0x00000000u},
    {0xffffffffu, 0x00000000u}, {0x1f116ab8u, 0xc233b53eu},
    {0x4d604ebeu, 0x419a352cu}, {0x29fd22f8u, 0xc1ee8859u},
    {0xffffffffu, 0x00000000u}, {0x79e7ec37u, 0x42a1ffb7u},
    {0x1a8446cbu, 0xc24d0a45u}, {0xffffffffu, 0x00000000u},
    {0x4665a9a6u, 0x41198725u}, {0x464d5b2bu, 0x4117bcf1u},
    {0x3c413d3au, 0xc08e158fu}, {0x4e85f412u, 0x41a6b811u},
    {0xffffffffu, 0x00000000u}, {0x5ee8984eu, 0x422e4a21u},
    {0xffffffffu, 0x00000000u}, {0x65d890d3u, 0x4254d1f9u},
    {0x111c87f8u, 0xc28085dfu}, {0x2c4c24b7u, 0xc1d48710u},
    {0xffffffffu, 0x00000000u}, {0x0dc8bba4u, 0xc289bac4u},
    {0xffffffffu, 0x00000000u}, {0xffffffffu, 0x00000000u},
    {0x6f31a8ecu, 0x42845a89u}, {0x38dcbe38u, 0xc1128ba4u}};

__device__ __forceinline__ float cr_logf(float x) {
  const uint32_t b = __float_as_uint(x);
  const uint32_t slot = (b * 0xc2eba581u) >> 27;
  const float r = (float)log((double)x);
  return c_logf_hard[slot][0] == b ? __uint_as_float(c_logf_hard[slot][1]) : r;
}

// ---------------------------------------------------------------------------
// LDS tables
// ---------------------------------------------------------------------------
// LDS layout: the spread tables at address 0, the emission table after
// them, then per wave the tran slots (and, F32 only, the log rings).
// Spread table: for parent p and an allele byte b of 8 consecutive sites,
// the 8-byte entry at byte 2048p + 8b puts bit i of b at bit 4+p of byte i
// (sites 0-3 in .x, 4-7 in .y), so OR-ing the 4 parents' entries gives one
// (nib << 4) byte per site.  Its address is the byte shifted by 3 (one
// instruction that selects the byte and shifts it) with 2048p in the read's
// offset field.
// Emission table: the 16-byte entry at byte kSpreadBytes + (code << 8) |
// (nib << 4) holds (e00, e11, e01, e10) of a site with GL code `code` whose
// 4 parent alleles are nib = f0 | f1 << 1 | m0 << 2 | m1 << 3, e_xy =
// FillEmit(code)[(f_x << 1) | m_y] (hmmLike.cu:39-44,124-127).  The even
// lane of a chain's lane pair reads the first 8 bytes, the odd lane the last
// 8: it ORs 8 into every site's byte (its mask oddm).
constexpr unsigned kSpreadBytes = 4 * 256 * 8;
constexpr unsigned kTranFloats = 96;   // tran of one 32-site chunk
constexpr unsigned kTranSlot = 128;    // a wave's LDS slot for it (a float pair per lane; a wave has two)
constexpr unsigned kRingSlots = 32;    // per lane, F32 accumulation only
// Forward-backward form: a slot holds the forward half's chunk of tran at
// float 0 and the backward half's at float kTranBOff (a float4 per lane:
// lanes 0-31 stage the forward rows, lanes 32-63 the backward ones; the
// 16-float gap keeps lanes 24-31's unread tail writes clear of the backward
// rows and puts the two halves' row reads in different banks).
constexpr unsigned kTranBOff = 144;
constexpr unsigned kTranSlotFB = kTranBOff + 128;

template <bool FB>
struct TranSlot {
  static constexpr unsigned floats = FB ? kTranSlotFB : kTranSlot;
};

__device__ __forceinline__ void init_lds(unsigned char *smem, const float4 *__restrict__ g_lut,
                                         unsigned lut_entries) {
  float4 *lut = reinterpret_cast<float4 *>(smem + kSpreadBytes);
  for (unsigned q = threadIdx.x; q < lut_entries; q += blockDim.x) lut[q] = g_lut[q];
  uint2 *spread = reinterpret_cast<uint2 *>(smem);
  for (unsigned q = threadIdx.x; q < 1024u; q += blockDim.x) {
    const unsigned p = q >> 8, b = q & 255u;
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (unsigned i = 0; i < 4; ++i) {
      lo |= ((b >> i) & 1u) << (8 * i + 4 + p);
      hi |= ((b >> (i + 4)) & 1u) << (8 * i + 4 + p);
    }
    spread[q] = make_uint2(lo, hi);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// lane pairs and the exact reciprocal
// ---------------------------------------------------------------------------
// A chain's forward pass runs on two adjacent lanes: the even lane carries
// states 00 and 11, the odd lane 01 and 10.  The other lane's values arrive
// through DPP quad permutations folded into the consuming add or multiply,
// so every float operation of the serial recurrence happens once per lane
// (F32: in the reference's order, with the reference's operands).
constexpr int kDppSwap = 0xB1;  // quad_perm [1,0,3,2]: the partner lane
constexpr int kDppEven = 0xA0;  // quad_perm [0,0,2,2]: the pair's even lane
constexpr int kDppOdd = 0xF5;   // quad_perm [1,1,3,3]: the pair's odd lane

template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, true));
}

// ---------------------------------------------------------------------------
// forward pass
// ---------------------------------------------------------------------------
// State of one lane: its two probabilities, and the record of rescales
// (impute.cpp:367-372).
//   F32 (the reference's arithmetic): a site whose sum falls below norm
//   multiplies its probabilities by 1/sum (the factor of a site that does
//   not rescale is 1.0f, so a rescale is a select rather than a branch and
//   x * 1.0f == x keeps the others exact); the factors go to a per-lane LDS
//   ring and their correctly rounded logs are subtracted, oldest first, at
//   the end of each 32-site chunk -- the same float sum in the same order.
//   MIXED: the FP32 recurrence, with l * t0 and l' * t2 fused into the
//   partner term by two FMAs per state, forms no site sums.  The pass
//   starts from l = 2^98 (= 1/4 * 2^100) and after every 8th site (sites
//   7, 15, 23, ...) the pair scales its four probabilities by the power of
//   two 2^k that brings their sum into [2^100, 2^101), which rounds nothing;
//   the lane adds up the k (starting from 100) and the score is
//   log(last sum) - (sum of k) ln 2 in FP64.  A site lowers the sum by at
//   most the smallest emission (mu^2 for a confident genotype that no
//   parent carries: 2^-33 at 10,000 haplotypes, 2^-13 for one mismatch), so
//   eight sites from 2^100 stay far above the normal range unless nearly
//   every one of them is a double mismatch.  The lane keeps the largest k;
//   a pass whose sum fell below 2^-99 at a scaling (k >= 200) or at its end,
//   or whose last sum is not finite, is redone with the scaling after every
//   site (score_set), whose sums stay above 2^100 times the smallest
//   emission.
struct PairState {
  float la, lb;
  int nres;  // MIXED: sum of the scale exponents so far
  int kmax;  // MIXED: largest scale exponent so far
  float fscore;
  unsigned nq;
};

// MIXED: the scalings bring the pair's sum into [2^kScaleT, 2^(kScaleT+1));
// one by 2^k with k >= kTinyK means the sum had fallen below 2^-99 since
// the previous one
constexpr int kScaleT = 100;
constexpr int kTinyK = 200;

constexpr double kLn2 = 0.6931471805599453;

// One site (impute.cpp:354-372) for the lane's two states; e = this lane's
// two emissions.  ring: this lane's first ring slot (stride 64 floats).
template <bool MIXED, bool FAST>
__device__ __forceinline__ void pair_step(PairState &st, const float2 e, float t0, float t1,
                                          float t2, float nrm, float *ring) {
  const float la = st.la, lb = st.lb;
  const float pown = la + lb;                 // l00 + l11 (even) / l01 + l10 (odd)
  const float pt = dpp<kDppSwap>(pown) * t1;  // the partner's pair sum * t1
  if (MIXED) {
    // the same sums with the two products of each fused into FMAs
    st.la = __builtin_fmaf(lb, t2, __builtin_fmaf(la, t0, pt)) * e.x;  // b00 / b01
    st.lb = __builtin_fmaf(la, t2, __builtin_fmaf(lb, t0, pt)) * e.y;  // b11 / b10
    return;
  }
  const float ba = la * t0 + pt + lb * t2;    // b00 / b01
  const float bb = lb * t0 + pt + la * t2;    // b11 / b10
  const float na = ba * e.x, nb = bb * e.y;
  // l00 + l01: the partner's na plus the own one (IEEE addition commutes,
  // so both lanes hold the reference's first partial sum bit for bit)
  float s = dpp<kDppSwap>(na) + na;
  s = dpp<kDppOdd>(nb) + s;   // + l10
  s = dpp<kDppEven>(nb) + s;  // + l11
  const bool r = s < nrm;
  const float sc = r ? (FAST ? recip_normal(s) : recip_small(s)) : 1.0f;
  st.la = na * sc;
  st.lb = nb * sc;
  ring[64u * st.nq] = sc;
  st.nq += r ? 1u : 0u;
}

// MIXED, after every 8th site: scale the pair's probabilities by 2^k with
// (l00 + l11) + (l01 + l10) in [2^100, 2^101) after it (frexp exponent e,
// k = 101 - e; the scaling is exact while the result stays normal, which
// the guard checks).  Both lanes form the same sum (addition commutes), so
// the pair scales together.
template <bool TRACK>
__device__ __forceinline__ void mixed_scale(PairState &st) {
  const float p = st.la + st.lb;
  const float s = dpp<kDppSwap>(p) + p;
  const int k = (kScaleT + 1) - __builtin_amdgcn_frexp_expf(s);
  st.la = __builtin_amdgcn_ldexpf(st.la, k);
  st.lb = __builtin_amdgcn_ldexpf(st.lb, k);
  st.nres += k;
  if (TRACK) st.kmax = k > st.kmax ? k : st.kmax;
}

// The pair's l00 + l01 + l10 + l11 in the reference's order.
__device__ __forceinline__ float pair_sum(const PairState &st) {
  float s = dpp<kDppSwap>(st.la) + st.la;
  s = dpp<kDppOdd>(st.lb) + s;
  return dpp<kDppEven>(st.lb) + s;
}

// F32: subtract the logs of the ring's factors, oldest first.
__device__ __forceinline__ void flush_ring(PairState &st, const float *ring) {
  for (unsigned k = 0; __any(k < st.nq); ++k)
    if (k < st.nq) st.fscore -= cr_logf(ring[64u * k]);
  st.nq = 0;
}

// The kernels of this file declare no static LDS (checked on the host before
// each launch), so their dynamic LDS starts at LDS address 0 and a table
// offset is an LDS address: a read takes it as is, with constant parts in
// its offset field.
__device__ __forceinline__ uint2 lds_u2(uint32_t a) {
  const glp_u2v v = *reinterpret_cast<const __attribute__((address_space(3))) glp_u2v *>((uintptr_t)a);
  return make_uint2(v.x, v.y);
}

// The emission half-entry at byte a of the emission table.
__device__ __forceinline__ float2 lut_entry(uint32_t a) {
  const glp_f2v v = *reinterpret_cast<const __attribute__((address_space(3))) glp_f2v *>(
      (uintptr_t)(a + kSpreadBytes));
  return make_float2(v.x, v.y);
}

// Parent p's spread entry for allele byte g of word h.
__device__ __forceinline__ uint2 spread_entry(uint32_t h, int g, int p) {
  return lds_u2((((h >> (8 * g)) & 255u) << 3) + 2048u * p);
}

// Byte offset of site s's emission half-entry: byte 0 = (nib << 4) | 8 *
// (odd lane) from the spread word, byte 1 = the site's code.
template <int BITS>
__device__ __forceinline__ uint32_t lut_addr(uint32_t cw, uint32_t nbw, unsigned s) {
  if (BITS == 8) {
    // v_perm_b32: result byte 0 <- nbw byte (s & 3), byte 1 <- cw byte
    // (3 - (s & 3)) (codes are MSB-first), bytes 2, 3 <- 0
    const uint32_t sel = 0x0c0c0000u | ((7u - (s & 3u)) << 8) | (s & 3u);
    return __builtin_amdgcn_perm(cw, nbw, sel);
  } else {
    const uint32_t code = (cw >> (28u - 4u * (s & 7u))) & 15u;
    return (code << 8) | ((nbw >> (8u * (s & 3u))) & 0xF8u);
  }
}

// The sites of a chunk without the group pipeline (used for a window's
// last chunk when it is shorter than 32 sites).  h = the 4 parents' 32-bit
// allele words, cw = the chunk's code words, tr = the wave's LDS copy of the
// chunk's tran, n = sites in the chunk (TAIL: n < 32), nrm0 = the threshold
// of the chunk's first site (0 at the window's site 0, which only
// initialises, impute.cpp:339-347); one0: the chunk's first site emits 1
// (the identity step that opens the backward half of an odd window).
template <int BITS, bool MIXED, bool TAIL, bool FAST>
__device__ __forceinline__ void fwd_chunk(PairState &st, const uint32_t (&h)[4],
                                          const uint32_t (&cw)[BITS], const float *tr, unsigned n,
                                          uint32_t oddm, float norm, float nrm0, float *ring,
                                          bool one0 = false) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    if (TAIL && (unsigned)(8 * g) >= n) break;
    const uint2 s0 = spread_entry(h[0], g, 0), s1 = spread_entry(h[1], g, 1);
    const uint2 s2 = spread_entry(h[2], g, 2), s3 = spread_entry(h[3], g, 3);
    const uint32_t nb[2] = {s0.x | s1.x | s2.x | s3.x | oddm, s0.y | s1.y | s2.y | s3.y | oddm};
    constexpr int CPW = 32 / BITS;
    // emissions and transitions of the group, loaded ahead of the recurrence
    float2 e[8];
    float4 t4[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) t4[q] = reinterpret_cast<const float4 *>(tr)[6 * g + q];
    const float *tg = reinterpret_cast<const float *>(t4);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int s = 8 * g + k;
      if (TAIL && (unsigned)s >= n) break;
      const uint32_t a = lut_addr<BITS>(cw[s / CPW], nb[k >> 2], (unsigned)(BITS == 8 ? s : k));
      e[k] = lut_entry(a);
      if (s == 0 && one0) e[k] = make_float2(1.0f, 1.0f);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int s = 8 * g + k;
      if (TAIL && (unsigned)s >= n) break;
      pair_step<MIXED, FAST>(st, e[k], tg[3 * k], tg[3 * k + 1], tg[3 * k + 2], s == 0 ? nrm0 : norm,
                       ring);
      // MIXED: every 8th site (FAST), or every site (the redo)
      if (MIXED && (!FAST || k == 7)) mixed_scale<FAST>(st);
    }
  }
}

// LDS inputs of one 8-site group of a full chunk: the parents' spread
// entries, the sites' tran rows and the lane's emissions.
struct GroupLds {
  uint2 sp[4];
  float4 t4[6];
  float2 e[8];
};

// Group g's spread entries, from the chunk's allele words h.
__device__ __forceinline__ void group_load_sp(GroupLds &G, const uint32_t (&h)[4], int g) {
#pragma unroll
  for (int p = 0; p < 4; ++p) G.sp[p] = spread_entry(h[p], g, p);
}

// Group g's tran rows (from the chunk's LDS slot tr) and emissions, whose
// addresses need its spread entries.
template <int BITS>
__device__ __forceinline__ void group_load_et(GroupLds &G, const uint32_t (&cw)[BITS],
                                              const float *tr, int g, uint32_t oddm) {
  constexpr int CPW = 32 / BITS;
#pragma unroll
  for (int q = 0; q < 6; ++q) G.t4[q] = reinterpret_cast<const float4 *>(tr)[6 * g + q];
  const uint32_t nb[2] = {G.sp[0].x | G.sp[1].x | G.sp[2].x | G.sp[3].x | oddm,
                          G.sp[0].y | G.sp[1].y | G.sp[2].y | G.sp[3].y | oddm};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int s = 8 * g + k;
    const uint32_t a = lut_addr<BITS>(cw[s / CPW], nb[k >> 2], (unsigned)(BITS == 8 ? s : k));
    G.e[k] = lut_entry(a);
  }
}

// Sites k0 .. k0 + 3 of group g (k0 = 0 or 4), from its loaded inputs.
template <bool MIXED, bool FAST>
__device__ __forceinline__ void group_steps(PairState &st, const GroupLds &G, int g, int k0,
                                            float norm, float nrm0, float *ring) {
  const float *tg = reinterpret_cast<const float *>(G.t4);
#pragma unroll
  for (int k = k0; k < k0 + 4; ++k) {
    pair_step<MIXED, FAST>(st, G.e[k], tg[3 * k], tg[3 * k + 1], tg[3 * k + 2],
                           (g == 0 && k == 0) ? nrm0 : norm, ring);
    if (MIXED && !FAST) mixed_scale<false>(st);  // the redo: every site
  }
  if (MIXED && FAST && k0 == 4) mixed_scale<true>(st);  // after sites 7, 15, ...
}

// A full 32-site chunk as a software pipeline over its 4 groups, in two
// buffers A and B that the groups alternate between.  While group g runs,
// group g + 1's tran rows and emissions are read (a whole group ahead of
// their use) and, from its middle, group g + 2's spread entries (which the
// emission addresses need).  On entry A holds group 0's inputs and B group
// 1's spread entries; at exit the same holds for the next chunk (allele
// words hn, code words cn, tran in LDS slot trn, which this lane's part tpn
// reaches at trnw during group 2).  LDS reads past the window's end only
// read the tables.
template <int BITS, bool MIXED, bool FAST, class TV>
__device__ __forceinline__ void fwd_chunk_pipe(PairState &st, GroupLds &A, GroupLds &B,
                                               const uint32_t (&h)[4], const uint32_t (&cw)[BITS],
                                               const float *tr, const uint32_t (&hn)[4],
                                               const uint32_t (&cn)[BITS], const float *trn,
                                               float *trnw, TV tpn, uint32_t oddm, float norm,
                                               float nrm0, float *ring) {
  group_load_et<BITS>(B, cw, tr, 1, oddm);
  group_steps<MIXED, FAST>(st, A, 0, 0, norm, nrm0, ring);
  group_load_sp(A, h, 2);
  group_steps<MIXED, FAST>(st, A, 0, 4, norm, nrm0, ring);

  group_load_et<BITS>(A, cw, tr, 2, oddm);
  group_steps<MIXED, FAST>(st, B, 1, 0, norm, nrm0, ring);
  group_load_sp(B, h, 3);
  group_steps<MIXED, FAST>(st, B, 1, 4, norm, nrm0, ring);

  group_load_et<BITS>(B, cw, tr, 3, oddm);
  group_steps<MIXED, FAST>(st, A, 2, 0, norm, nrm0, ring);
  // every lane writes (the last lanes of each part into the slot's unread
  // tail), so that the write, and with it the tran load, is not under a branch
  *reinterpret_cast<TV *>(trnw) = tpn;
  group_load_sp(A, hn, 0);
  group_steps<MIXED, FAST>(st, A, 2, 4, norm, nrm0, ring);

  group_load_et<BITS>(A, cn, trn, 0, oddm);
  group_steps<MIXED, FAST>(st, B, 3, 0, norm, nrm0, ring);
  group_load_sp(B, hn, 1);
  group_steps<MIXED, FAST>(st, B, 3, 4, norm, nrm0, ring);
}

// The global inputs of a pass are read through raw buffer loads: the
// buffer descriptor (base, size) sits in scalar registers and each load
// adds a per-lane byte offset (a register fixed for the pass) to a scalar
// one (the chunk's), so no load needs a 64-bit address computed per lane.
// Raw buffers (stride 0) bounds-check the per-lane offset plus the
// instruction offset against the size, dword by dword: a dword at or past
// it reads 0 and touches no memory.  The panel and code buffers are
// declared 4 GB - 1 long and every offset into them stays inside the array
// (the chunk and word indices are clamped; the host refuses arrays of 4 GB
// or more); the tran buffer is declared its true size and relies on the
// check.
typedef __amdgpu_buffer_rsrc_t brsrc_t;
constexpr int kBufDword3 = 0x00020000;  // raw buffer, 32-bit data format

__device__ __forceinline__ brsrc_t buf_rsrc(const void *p, uint32_t bytes) {
  // the base must be wave-uniform: it is, but make it an SGPR for certain
  const uint64_t a = (uint64_t)(uintptr_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  void *u = (void *)(uintptr_t)(((uint64_t)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(u, 0, (int)bytes, kBufDword3);
}
// soff is wave-uniform in every call; readfirstlane states it, so that the
// compiler keeps it in an SGPR rather than looping over its distinct values
__device__ __forceinline__ uint32_t buf_u32(brsrc_t r, uint32_t voff, uint32_t soff) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(
      r, (int)voff, (int)__builtin_amdgcn_readfirstlane(soff), 0);
}

// Where one pass reads: the window's panel, codes and tran, the 4 parents'
// row offsets and the sample's code column (bytes, per lane); in the
// forward-backward form also this lane's byte offset of its part of a
// chunk's tran (the staging lanes' float4: forward rows from lanes 0-31,
// backward rows from lanes 32-63).
struct PassSrc {
  brsrc_t panel, codes, tran;
  uint32_t row[4];
  uint32_t col;     // 4 * sample
  uint32_t stride;  // 4 * number of samples: bytes between a sample's code words
  uint32_t tlane;
};

// The global inputs of chunk c: the 4 parents' allele words, the code
// words and this lane's part of the chunk's tran (a float pair: lanes 0-47
// carry a chunk's 96 floats; forward-backward form: a float4, lanes 0-23
// the forward half's 96 and lanes 32-55 the backward half's).  Every load
// runs for every c: past the window's end the panel and code offsets are
// clamped and the tran floats read 0 (or a later part's rows); those values
// are not used, except by reads of the LDS tables.
template <int BITS, bool FB>
struct ChunkIn {
  using TV = typename std::conditional<FB, float4, float2>::type;
  uint32_t h[4];
  uint32_t cw[BITS];
  TV tp;
};

template <int BITS, bool FB>
__device__ __forceinline__ void load_chunk(ChunkIn<BITS, FB> &in, unsigned c, unsigned nch,
                                           const PassSrc &src, unsigned nwords, unsigned lane) {
  const unsigned cc = c < nch ? c : nch - 1u;
  if constexpr (FB) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(
        src.tran, (int)src.tlane, (int)__builtin_amdgcn_readfirstlane(4u * kTranFloats * c), 0);
    in.tp = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                        __uint_as_float(v[3]));
  } else {
    // floats 96c + 2 lane and the next (tran has 3M floats; when 3M is odd
    // the last float stands alone and its partner reads 0)
    const uint32_t to = 4u * (kTranFloats * c + 2u * lane);
    in.tp = make_float2(__uint_as_float(buf_u32(src.tran, to, 0u)),
                        __uint_as_float(buf_u32(src.tran, to + 4u, 0u)));
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) in.h[p] = buf_u32(src.panel, src.row[p], 4u * cc);
  // word offsets step by stride from the chunk's first word, clamped to the
  // window's last word (one add and one min per word)
  const unsigned last = (nwords - 1u) * src.stride;
  unsigned off = cc * BITS * src.stride;
#pragma unroll
  for (int q = 0; q < BITS; ++q) {
    in.cw[q] = buf_u32(src.codes, src.col, off < last ? off : last);
    off += src.stride;
  }
}

// The forward-backward form of a pass (mixed precision).  The likelihood is
// l_{n-1}^T T_n beta_n for the forward variable l, the backward variable
// beta and the symmetric switch matrix T_n of tran row n.  The
// emission-weighted backward variable u_m = e_m o beta_m obeys u_{m-1} =
// e_{m-1} o (T_m u_m), which is the forward step itself.  So a lane quad
// runs both halves side by side with one instruction stream: lanes 0-1 the
// forward pair over sites 0 .. n-1 (n = ceil(M/2)), lanes 2-3 the backward
// pair over sites M-1 down to n, each n steps.  The backward half reads the
// window's data mirrored: its panel rows bit-reversed, its codes and tran
// rows in reverse order (k_fb_mirror lays them out after the forward
// half's), starting from u = 1 with an identity first row, and for odd M an
// identity first step (row (1, 0, 0), emission 1: one0).  Each half keeps
// the mixed pass's scaling and underflow guard.  The join runs in FP64: the
// backward pair forms g = T_n u_n (its row n, staged with its data), the
// quad exchanges halves, and every lane of it returns
// log((l00 g00 + l11 g11) + (l01 g01 + l10 g10)) - (kF + kB) ln 2.
struct FbInfo {
  bool back;  // this lane belongs to the backward pair
  bool one0;  // its first step is the identity (odd M)
  float t0, t1, t2;  // tran row n
};

template <bool FAST>
__device__ __forceinline__ double fb_join(const PairState &st, float last, const FbInfo &fb) {
  const int bad = (FAST && (st.kmax >= kTinyK || !(last >= 0x1p-99f && last <= 0x1.fffffep127f)))
                      ? 1 : 0;
  const double a = st.la, b = st.lb;
  const double partner = __shfl_xor(a + b, 1, 64);
  const double t0 = fb.t0, t1 = fb.t1, t2 = fb.t2;
  const double ga = (t0 * a + t1 * partner) + t2 * b;
  const double gb = (t0 * b + t1 * partner) + t2 * a;
  // the forward pair sends l, the backward pair g = T_n u
  const double ra = __shfl_xor(fb.back ? ga : a, 2, 64);
  const double rb = __shfl_xor(fb.back ? gb : b, 2, 64);
  const double d = fb.back ? ra * ga + rb * gb : a * ra + b * rb;
  const double tot = d + __shfl_xor(d, 1, 64);
  const int ks = st.nres + __shfl_xor(st.nres, 2, 64);
  if ((bad | __shfl_xor(bad, 2, 64)) != 0) return __builtin_nan("");
  return log(tot) - (double)ks * kLn2;
}

// Log-likelihood of one sample given 4 parents (father pair f0, f1, mother
// pair m0, m1: src.row), computed by a lane pair; both lanes return it
// (FB: by a lane quad over M = n steps per half, see fb_join; all four
// return it).  Every lane of the wavefront must call it together (M and
// tran wave-uniform): the lanes copy each 32-site chunk of tran into the
// wave's LDS buffer wtr, during the chunk before it.
template <int BITS, bool MIXED, bool FAST, bool FB>
__device__ __forceinline__ double forward_pass(const PassSrc &src, unsigned M,
                                               const unsigned char *smem, unsigned lut_entries,
                                               float *wtr, float *ring, float norm,
                                               const FbInfo &fb) {
  constexpr int CPW = 32 / BITS;
  constexpr unsigned SLOT = TranSlot<FB>::floats;
  using TV = typename ChunkIn<BITS, FB>::TV;
  const unsigned lane = threadIdx.x & 63u;
  // the odd lane reads the odd half of every emission entry
  const uint32_t oddm = (lane & 1u) ? 0x08080808u : 0u;
  PairState st;
  // MIXED starts scaled by 2^kScaleT (l = 1/4; the backward half u = 1)
  st.la = st.lb = FB ? (fb.back ? 0x1p100f : 0x1p98f) : (MIXED ? 0x1p98f : 0.25f);
  st.nres = MIXED ? kScaleT : 0;
  st.kmax = 0;
  st.fscore = 0.0f;
  st.nq = 0;
  const unsigned nch = (M + 31u) >> 5;
  const unsigned nwords = (M + CPW - 1) / CPW;
  // this lane's tran write (its staging part) and read (its half) offsets
  // within a slot
  const unsigned woff = FB ? (lane >= 32u ? kTranBOff : 0u) + 4u * (lane & 31u) : 2u * lane;
  const unsigned roff = FB && fb.back ? kTranBOff : 0u;
  // chunk ch's global inputs sit in buffer X (ch even) or Y (ch odd): the
  // loop body is written out twice, so that the next chunk's loads land in
  // the other buffer while this one is used, with no register copies
  ChunkIn<BITS, FB> X, Y;
  load_chunk<BITS, FB>(X, 0, nch, src, nwords, lane);
  if constexpr (!FB) {
    // site 0's row becomes (1, 0, 0), so that from l = 1/4 (F32) or 2^98
    // (MIXED) the generic step gives l_xy = l * e_xy exactly (the
    // forward-backward layout holds these rows already)
    if (lane == 0u) X.tp = make_float2(1.0f, 0.0f);
    if (lane == 1u) X.tp.x = 0.0f;
  }
  // chunk ch's tran sits in LDS slot ch & 1 of the wave's two; a full
  // chunk's last group writes the next chunk's slot
  *reinterpret_cast<TV *>(wtr + woff) = X.tp;
  GroupLds A, B;
  if (M >= 32u) {
    group_load_sp(A, X.h, 0);
    group_load_et<BITS>(A, X.cw, wtr + roff, 0, oddm);
    if (FB && fb.one0) A.e[0] = make_float2(1.0f, 1.0f);
    group_load_sp(B, X.h, 1);
  }
  auto chunk = [&](unsigned ch, const ChunkIn<BITS, FB> &cur, ChunkIn<BITS, FB> &nxt) {
    load_chunk<BITS, FB>(nxt, ch + 1u, nch, src, nwords, lane);
    // keep the loads here, a whole chunk ahead of their use (the scheduler
    // would otherwise sink the tran loads down to the LDS write that uses them)
    __builtin_amdgcn_sched_barrier(0);
    const unsigned base = ch << 5;
    const unsigned n = M - base < 32u ? M - base : 32u;
    const float nrm0 = ch == 0 ? 0.0f : norm;
    float *slot = wtr + SLOT * (ch & 1u), *slotn = wtr + SLOT * ((ch + 1u) & 1u);
    if (n == 32)
      fwd_chunk_pipe<BITS, MIXED, FAST, TV>(st, A, B, cur.h, cur.cw, slot + roff, nxt.h, nxt.cw,
                                            slotn + roff, slotn + woff, nxt.tp, oddm, norm, nrm0,
                                            ring);
    else  // the window's last chunk, shorter than 32 sites
      fwd_chunk<BITS, MIXED, true, FAST>(st, cur.h, cur.cw, slot + roff, n, oddm, norm, nrm0,
                                         ring, FB && fb.one0 && ch == 0u);
    if (!MIXED) flush_ring(st, ring);
  };
  for (unsigned ch = 0; ch < nch; ch += 2u) {
    chunk(ch, X, Y);
    if (ch + 1u == nch) break;
    chunk(ch + 1u, Y, X);
  }
  const float last = pair_sum(st);
  if constexpr (FB) {
    return fb_join<FAST>(st, last, fb);
  } else if (MIXED) {
    // a sum below 2^-99 (or 0), or one that is not finite: NaN asks
    // score_set for the redo
    if (FAST && (st.kmax >= kTinyK || !(last >= 0x1p-99f && last <= 0x1.fffffep127f)))
      return __builtin_nan("");
    return log((double)last) - (double)st.nres * kLn2;
  }
  return (double)(st.fscore + cr_logf(last));
}

// Sizes of a window's forward-backward layout (k_fb_mirror): n steps per
// half, half-rows of wh 64-bit words, nw code words per half column, nch
// chunks per half; tran holds the forward half's 96 nch floats, then the
// backward half's 96 nch + 128.
struct FbGeom {
  unsigned n, wh, nw, nch;
  __host__ __device__ FbGeom(unsigned M, unsigned bits) {
    n = (M + 1u) >> 1;
    wh = (n + 63u) >> 6;
    nw = (n * bits + 31u) >> 5;
    nch = (n + 31u) >> 5;
  }
  __host__ __device__ unsigned tran_floats() const { return 192u * nch + 128u; }
};

// LDS layout after the two tables: per wave, two slots of kTranSlot
// (FB: kTranSlotFB) floats of tran, then (F32 only) per wave a ring of
// kRingSlots x 64 floats.  FB: W's panel, packed codes and tran are the
// window's forward-backward layout (k_fb_mirror); H = its panel's rows.
template <int BITS, bool MIXED, bool FB = false>
__device__ __forceinline__ double score_set(const glp_window &W, const unsigned (&P)[4],
                                            unsigned sample, unsigned N, unsigned H,
                                            const unsigned char *smem, unsigned lut_entries,
                                            float norm) {
  constexpr unsigned SLOT = TranSlot<FB>::floats;
  const unsigned lane = threadIdx.x & 63u;
  const unsigned wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  float *tables_end = reinterpret_cast<float *>(const_cast<unsigned char *>(smem) +
                                                16u * lut_entries + kSpreadBytes);
  float *wtr = tables_end + 2u * SLOT * wave;
  float *ring = tables_end + 2u * SLOT * nwave + kRingSlots * 64u * wave + lane;
  PassSrc src;
  FbInfo fb{false, false, 1.0f, 0.0f, 0.0f};
  unsigned steps = W.num_sites;
  src.panel = buf_rsrc(W.panel, 0xFFFFFFFFu);
  src.codes = buf_rsrc(W.packed, 0xFFFFFFFFu);
  src.stride = 4u * N;
  if constexpr (FB) {
    const FbGeom g(W.num_sites, BITS);
    steps = g.n;
    fb.back = ((lane >> 1) & 1u) != 0u;
    fb.one0 = fb.back && (W.num_sites & 1u) != 0u;
    src.tran = buf_rsrc(W.tran, 4u * g.tran_floats());
    const uint32_t hb = fb.back ? H * g.wh : 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) src.row[q] = 8u * (P[q] * g.wh + hb);
    src.col = 4u * sample + (fb.back ? 4u * g.nw * N : 0u);
    src.tlane = 4u * ((lane >= 32u ? 96u * g.nch : 0u) + 4u * (lane & 31u));
    // the join's row n: the backward part's row n (tran row n of the window)
    const uint32_t ro = 4u * (96u * g.nch + 3u * g.n);
    fb.t0 = __uint_as_float(buf_u32(src.tran, 0u, ro));
    fb.t1 = __uint_as_float(buf_u32(src.tran, 0u, ro + 4u));
    fb.t2 = __uint_as_float(buf_u32(src.tran, 0u, ro + 8u));
  } else {
    src.tran = buf_rsrc(W.tran, 12u * W.num_sites);
#pragma unroll
    for (int q = 0; q < 4; ++q) src.row[q] = 8u * P[q] * W.words_per_hap;
    src.col = 4u * sample;
    src.tlane = 0u;
  }
  const double v =
      forward_pass<BITS, MIXED, true, FB>(src, steps, smem, lut_entries, wtr, ring, norm, fb);
  // F32: a subnormal (or zero) site sum made the short reciprocal NaN: the
  // wavefront redoes its passes with the exact one (recip_normal is exact
  // for normal sums, so a value that is not NaN is already the exact one).
  // MIXED: a sum fell below 2^-99 between two scalings, or the last one is
  // not finite: the redo scales after every site.  Each lane keeps its first value unless it was NaN.
  if (__builtin_expect(__any(v != v), 0)) {
    const double x =
        forward_pass<BITS, MIXED, false, FB>(src, steps, smem, lut_entries, wtr, ring, norm, fb);
    return v != v ? x : v;
  }
  return v;
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
  // kickstart (insti.cpp:1510-1513): the generation's first proposal from
  // the reference haplotypes only; Sampler::SampleHap(I, true)'s rejection
  // loop (sampler.cpp:12-27) is uniform over them
  if (a.kick && a.k0 + c == 0u && a.prop_mode != kPropHost) {
    const unsigned ref0 = 2u * a.N;
    if (a.prop_mode == kPropKList) {
      const unsigned *L = a.klist + ((size_t)w * a.N + i) * a.K;
      unsigned nref = 0;
      for (unsigned k = 0; k < a.K; ++k) nref += L[k] >= ref0 ? 1u : 0u;
      if (nref) {
        unsigned pick = uniform_int(x.y, nref), k = 0;
        for (;; ++k)
          if (L[k] >= ref0 && pick-- == 0u) break;
        cand = L[k];
      }
    } else {
      cand = ref0 + uniform_int(x.y, a.H - ref0);
    }
  }
}

// ---------------------------------------------------------------------------
// plain path: one lane pair (FB: lane quad) per (window, sample), one
// launch per MH iteration
// ---------------------------------------------------------------------------
template <int BITS, bool MIXED, bool FB>
__global__ void __launch_bounds__(256)
    k_plain(const glp_window *__restrict__ wins, KArgs a, unsigned step) {
  extern __shared__ __align__(16) unsigned char smem[];
  init_lds(smem, reinterpret_cast<const float4 *>(a.lut), a.lut_entries);
  constexpr unsigned UL = FB ? 4u : 2u;  // lanes per chain
  const unsigned w = blockIdx.y;
  const unsigned ig = (blockIdx.x * blockDim.x + threadIdx.x) / UL;
  // the forward pass runs on whole wavefronts: chains past N score a copy of
  // sample N-1 and write nothing; the first lane of a chain writes
  const bool valid = ig < a.N && (threadIdx.x & (UL - 1u)) == 0u;
  if (__ballot(ig < a.N) == 0ull) return;
  const unsigned i = ig < a.N ? ig : a.N - 1;
  const glp_window W = wins[w];
  const size_t row = (size_t)w * a.N + i;
  unsigned P[4];
  if (step == 0) {
    initial_parents(a, W, w, i, P);
    const double v = score_set<BITS, MIXED, FB>(W, P, i, a.N, a.H, smem, a.lut_entries, a.norm);
    if (!valid) return;
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
  P[0] = rp == 0 ? cand : P[0];
  P[1] = rp == 1 ? cand : P[1];
  P[2] = rp == 2 ? cand : P[2];
  P[3] = rp == 3 ? cand : P[3];
  const double prop = score_set<BITS, MIXED, FB>(W, P, i, a.N, a.H, smem, a.lut_entries, a.norm);
  if (!valid) return;
  const bool acc = mh_accept(prop, curr, (double)a.pen, ub);
  if (acc) {
    W.parents[4 * (size_t)i + rp] = cand;
    W.like[i] = prop;
  }
  if (a.trace_like) a.trace_like[row * (a.C + 1) + 1 + c] = prop;
  if (a.trace_acc) a.trace_acc[row * a.C + c] = acc ? 1 : 0;
}

// ---------------------------------------------------------------------------
// batched form: persistent work-queue kernel with speculative proposal windows
// ---------------------------------------------------------------------------
// Wavefronts walk the windows in `order` (longest first).  While on a window
// a wavefront keeps its data (sites, tran, panel, codes) wave-uniform and
// holds floor(32/Q) home slots (Q = qtab[window], 1..32).  A home slot owns
// one (window, sample) chain at a time; the chain's state is replicated over
// the slot's 2Q home lanes.  Idle home slots take the window's next samples
// from its task counter; when the window runs dry and every slot is idle the
// wavefront moves on.
//
// Each round the wavefront's 32 lane pairs are shared among its A active
// chains: in home-slot order, chain c gets a contiguous run of
// floor(32/A) pairs, one more for the first 32 mod A chains, so the chains
// left after their neighbours finish speculate deeper.  A chain's pairs
// score its next undecided proposals (the first pair of a fresh chain scores
// its initial parents), all from the chain's current parent set S.  Then
// the chain's home lanes replay the serial MH walk over those values in
// proposal order.  A value is the serial chain's value exactly when every
// slot of S changed in this walk so far equals the proposal's own replace
// slot; the walk stops at the first proposal for which that fails, and the
// next round re-scores from there.  The result equals the serial loop of
// insti.cpp:1505-1539 draw for draw; speculation only spends lane time.
// WAVES wavefronts per workgroup share one copy of the emission and spread
// tables: 4 (one workgroup per CU) or 8 (two wavefronts per SIMD where the
// LDS of the F32 ring or of fewer tables allows it; the compiler then keeps
// the kernel within 256 VGPRs).  With 8 and a.split, wavefronts 4-7 hold
// one home slot fewer than wavefronts 0-3: the chains spread over more
// workgroups, and where a SIMD runs one wavefront of each half, the one
// with fewer chains finishes its rounds first and leaves the SIMD's issue
// slots to the other.
// FB (the forward-backward form): a pass takes a lane quad instead of a
// pair, so the units below are quads and a wavefront holds 16 of them.
template <int BITS, bool MIXED, int WAVES, bool FB>
__global__ void __launch_bounds__(64 * WAVES)
    k_batched(const glp_window *__restrict__ wins, const unsigned *__restrict__ order,
              const unsigned *__restrict__ qtab, unsigned nwin, unsigned *__restrict__ wctr,
              unsigned long long *__restrict__ ctrs, KArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  init_lds(smem, reinterpret_cast<const float4 *>(a.lut), a.lut_entries);
  constexpr unsigned UL = FB ? 4u : 2u;  // lanes per pass unit
  constexpr unsigned NU = 64u / UL;      // units per wavefront
  const unsigned lane = threadIdx.x & 63u, pair = lane / UL;
  const bool even = (lane & (UL - 1u)) == 0u;
  const uint64_t lanes_below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const double pen = (double)a.pen;

  // wave-uniform window state; sbase = this lane's home slot's first lane,
  // home = the lane belongs to a home slot (pairs past the last whole slot
  // only score)
  int wo = -1;  // position in `order`
  unsigned bw = 0, Q = 1, sbase = lane & ~(UL - 1u);
  bool home = false;
  glp_window W{};
  // home-slot chain state, replicated over the slot's home lanes
  bool active = false, fresh = false;
  unsigned i = 0, kp = 0;
  unsigned S[4] = {0, 0, 0, 0};
  double cur = 0.0;
  unsigned long long n_eval = 0, n_useful = 0;
  // Round budget.  In every round each active chain decides at least one
  // proposal (the first one it scores from its current set is always
  // valid, and every chain holds at least one pair), so a task takes at
  // most C + 1 rounds, and a round with no active chain moves the wave to
  // its next window.  The wave's rounds therefore never exceed (C + 1) per
  // task it drew plus one per window and one to leave; a wave past that
  // budget flags the run as failed and exits instead of spinning.
  unsigned long long rounds = 0, budget = (unsigned long long)nwin + 2ull;

  for (;;) {
    if (++rounds > budget) {
      if (lane == 0) atomicOr(&ctrs[3], 1ull);
      break;
    }
    // 1. idle home slots take the next samples of the current window
    if (wo >= 0) {
      const bool want = home && lane == sbase && !active;
      const uint64_t wm = __ballot(want);
      if (wm) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(&wctr[bw], (unsigned)__popcll(wm));
        base = __builtin_amdgcn_readfirstlane(base);
        const unsigned rank = (unsigned)__popcll(wm & lanes_below);
        const bool took = want && base + rank < a.N;
        budget += (unsigned long long)__popcll(__ballot(took)) * ((unsigned long long)a.C + 1ull);
        unsigned t = took ? base + rank : 0xFFFFFFFFu;
        t = __shfl(t, sbase, 64);
        if (home && !active && t != 0xFFFFFFFFu) {
          i = t;
          initial_parents(a, W, bw, i, S);
          active = true;
          fresh = true;
          kp = 0;
        }
      }
    }
    // the active chains, one bit at each active home slot's first lane
    const uint64_t am = __ballot(home && lane == sbase && active);
    if (am == 0ull) {
      // 2. the window is done for this wavefront: move to the next one
      if (++wo >= (int)nwin) break;
      bw = __builtin_amdgcn_readfirstlane(order[wo]);
      Q = __builtin_amdgcn_readfirstlane(qtab[bw]);
      const unsigned slot = pair / Q;
      sbase = UL * Q * slot;
      // a.split: the second wavefront of each SIMD (wavefronts 4-7 of the
      // workgroup) holds one chain slot fewer than the first
      unsigned nslots = NU / Q;
      if (WAVES == 8 && a.split && threadIdx.x >= 256u && nslots > 1u) --nslots;
      home = slot < nslots;
      W = wins[bw];
      continue;
    }

    // 3. the pairs of the round: chain rank c (in home-slot order) holds the
    //    contiguous pairs [start_c, start_c + q_c), q_c = base + [c < extra]
    const unsigned A = (unsigned)__popcll(am);
    const unsigned base = NU / A, extra = NU % A, big = extra * (base + 1u);
    unsigned c, j;  // this pair's chain rank and place in the chain's run
    if (pair < big) {
      c = pair / (base + 1u);
      j = pair - c * (base + 1u);
    } else {
      c = extra + (pair - big) / base;
      j = pair - big - (c - extra) * base;
    }
    uint64_t m = am;  // the c-th set bit of am is the chain's home lane
    for (unsigned k = 0; k < c; ++k) m &= m - 1ull;
    const unsigned hl = (unsigned)__builtin_ctzll(m);
    const unsigned ci = __shfl(i, hl, 64), ckp = __shfl(kp, hl, 64);
    const bool cfresh = __shfl(fresh ? 1 : 0, hl, 64) != 0;
    unsigned P[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) P[q] = __shfl(S[q], hl, 64);

    // 4. this pair's job: the chain's initial set or one speculative proposal
    bool work = false;
    unsigned rp = 4, cand = 0;
    uint32_t ub = 0;
    {
      const int job = cfresh ? (int)j - 1 : (int)j;
      if (job < 0) {
        work = true;
      } else if (ckp + (unsigned)job < a.C) {
        work = true;
        proposal_draws(a, bw, ci, ckp + (unsigned)job, rp, cand, ub);
        P[0] = rp == 0 ? cand : P[0];
        P[1] = rp == 1 ? cand : P[1];
        P[2] = rp == 2 ? cand : P[2];
        P[3] = rp == 3 ? cand : P[3];
      }
    }
    // the whole wavefront runs the pass; pairs without work score their
    // chain's current set and the value is ignored
    double v = score_set<BITS, MIXED, FB>(W, P, ci, a.N, a.H, smem, a.lut_entries, a.norm);
    if (work && even) ++n_eval;

    // 5. serial MH walk over each active chain's values, replayed by every
    //    lane of its home slot (so the chain state stays replicated)
    const bool mine = home && active;
    const uint64_t below_base = sbase == 0u ? 0ull : (~0ull >> (64u - sbase));
    const unsigned myc = (unsigned)__popcll(am & below_base);  // this home slot's rank
    const unsigned qn = myc < extra ? base + 1u : base;
    const unsigned start = myc < extra ? myc * (base + 1u) : big + (myc - extra) * base;
    const unsigned qmax = base + (extra ? 1u : 0u);
    const int jstart = fresh ? 1 : 0;
    const double v0 = __shfl(v, (UL * start) & 63u, 64);
    if (mine && fresh) {
      cur = v0;
      if (lane == sbase && a.trace_like) a.trace_like[((size_t)bw * a.N + i) * (a.C + 1)] = v0;
    }
    unsigned changed = 0, decided = 0;
    bool stop = !mine;
    for (unsigned jj = 0; jj < qmax; ++jj) {
      // the reads run on every lane; only a chain's own run is used
      const unsigned src = (UL * (start + jj)) & 63u;
      const double vj = __shfl(v, src, 64);
      const unsigned rpj = __shfl(rp, src, 64);
      const unsigned cj = __shfl(cand, src, 64);
      const uint32_t uj = __shfl(ub, src, 64);
      if (!stop && (int)jj >= jstart) {
        const unsigned kk = kp + (jj - (unsigned)jstart);
        if (jj >= qn || kk >= a.C || (changed & ~(1u << rpj)) != 0u) {
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
          if (lane == sbase) {
            const size_t row = (size_t)bw * a.N + i;
            if (a.trace_like) a.trace_like[row * (a.C + 1) + 1 + kk] = vj;
            if (a.trace_acc) a.trace_acc[row * a.C + kk] = acc ? 1 : 0;
          }
        }
      }
    }
    if (mine) {
      kp += decided;
      fresh = false;
      if (lane == sbase) n_useful += decided + (jstart ? 1u : 0u);
      if (kp >= a.C) {
        if (lane == sbase) {
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

// Clears the statistics and the per-window task counters before a run (a
// kernel rather than a memset, so the run can be captured in a HIP graph).
__global__ void k_reset_ctrs(unsigned long long *ctrs, unsigned *wctr, unsigned nwin) {
  if (threadIdx.x < 4) ctrs[threadIdx.x] = 0ull;
  for (unsigned w = threadIdx.x; w < nwin; w += blockDim.x) wctr[w] = 0u;
}

// ---------------------------------------------------------------------------
// single evaluations: Impute::hmm_like batched
// ---------------------------------------------------------------------------
template <int BITS>
__global__ void __launch_bounds__(256)
    k_eval(const glp_window *__restrict__ wins, KArgs a,
           const unsigned *__restrict__ samples,
           const unsigned *__restrict__ parents, unsigned n, float *out) {
  extern __shared__ __align__(16) unsigned char smem[];
  init_lds(smem, reinterpret_cast<const float4 *>(a.lut), a.lut_entries);
  const unsigned eg = (blockIdx.x * blockDim.x + threadIdx.x) >> 1;
  if (__ballot(eg < n) == 0ull) return;
  const bool valid = eg < n && (threadIdx.x & 1u) == 0u;
  const unsigned e = eg < n ? eg : n - 1;
  const glp_window W = wins[0];
  const unsigned P[4] = {parents[4 * (size_t)e], parents[4 * (size_t)e + 1],
                         parents[4 * (size_t)e + 2], parents[4 * (size_t)e + 3]};
  const float v =
      (float)score_set<BITS, false>(W, P, samples[e], a.N, a.H, smem, a.lut_entries, a.norm);
  if (valid) out[e] = v;
}

// Every float with bit pattern in [lo, hi): recip_small bit-equal to the
// division below 2^-20; recip_normal bit-equal to it on [2^-126, 2^126] and
// NaN for subnormal x.
__global__ void k_selftest_recip(uint32_t lo, uint32_t hi, unsigned long long *bad,
                                 uint32_t *first) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t b = (uint64_t)lo + blockIdx.x * blockDim.x + threadIdx.x; b < hi; b += stride) {
    const float x = __uint_as_float((uint32_t)b);
    const uint32_t want = __float_as_uint(1.0f / x);
    const float fast = recip_normal(x);
    const bool ok_small = !(x < 0x1p-20f) || __float_as_uint(recip_small(x)) == want;
    const bool ok_fast = x < 0x1p-126f ? __builtin_isnan(fast)
                                       : (x > 0x1p126f || __float_as_uint(fast) == want);
    const bool ok = ok_small && ok_fast;
    if (!ok) {
      atomicAdd(bad, 1ull);
      atomicMin(first, (uint32_t)b);
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t lds_bytes(const KArgs &a, unsigned waves) {
  const size_t slot = a.precision == kPrecMixedFB ? kTranSlotFB : kTranSlot;
  size_t b = (size_t)a.lut_entries * 16 + kSpreadBytes + waves * 2 * slot * 4;
  if (a.precision == kPrecF32) b += waves * kRingSlots * 64 * 4;
  return b;
}

// Sets a kernel's dynamic LDS limit, after checking that it declares no
// static LDS: its emission-table reads take LDS addresses from 0.
static int allow_lds(const void *fn, size_t bytes) {
  hipFuncAttributes fa;
  GLP_TRY_K(hip_check(hipFuncGetAttributes(&fa, fn), "kernel attributes"));
  if (fa.sharedSizeBytes != 0) return set_error(GLP_ERR_HIP, "kernel declares static LDS");
  return hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes),
                   "dynamic LDS attribute");
}

template <int BITS>
static const void *plain_kernel_b(int precision) {
  if (precision == kPrecMixedFB) return (const void *)&k_plain<BITS, true, true>;
  if (precision == kPrecMixed) return (const void *)&k_plain<BITS, true, false>;
  return (const void *)&k_plain<BITS, false, false>;
}

const void *plain_kernel_fn(unsigned bits, int precision) {
  return bits == 8 ? plain_kernel_b<8>(precision) : plain_kernel_b<4>(precision);
}

// lanes of the plain path's grid: one pair per chain, a quad in the
// forward-backward form
unsigned plain_lanes(const KArgs &a) { return (a.precision == kPrecMixedFB ? 4u : 2u) * a.N; }

int launch_plain(const KArgs &a, const glp_window *d_wins, unsigned nwin,
                 unsigned step, hipStream_t stream) {
  const dim3 block(256), grid((plain_lanes(a) + 255) / 256, nwin);
  const size_t smem = lds_bytes(a, 4);
  const void *fn = plain_kernel_fn(a.bits, a.precision);
  GLP_TRY_K(allow_lds(fn, smem));
  KArgs aa = a;
  void *args[] = {(void *)&d_wins, (void *)&aa, (void *)&step};
  return hip_check(hipLaunchKernel(fn, grid, block, args, smem, stream), "k_plain launch");
}

const void *reset_kernel_fn() { return (const void *)&k_reset_ctrs; }

int prepare_kernel_lds(const void *fn, size_t bytes) { return allow_lds(fn, bytes); }

template <int BITS, int W>
static const void *batched_kernel_bw(int precision) {
  if (precision == kPrecMixedFB) return (const void *)&k_batched<BITS, true, W, true>;
  if (precision == kPrecMixed) return (const void *)&k_batched<BITS, true, W, false>;
  return (const void *)&k_batched<BITS, false, W, false>;
}

const void *batched_kernel_fn(unsigned bits, int precision, unsigned waves) {
  if (bits == 8)
    return waves == 8 ? batched_kernel_bw<8, 8>(precision) : batched_kernel_bw<8, 4>(precision);
  return waves == 8 ? batched_kernel_bw<4, 8>(precision) : batched_kernel_bw<4, 4>(precision);
}

int launch_batched(const KArgs &a, const glp_window *d_wins, const unsigned *d_order,
                   const unsigned *d_qtab, unsigned nwin, unsigned *d_wctr,
                   unsigned long long *d_ctrs, unsigned grid, unsigned waves,
                   hipStream_t stream) {
  if (waves != 4 && waves != 8) return set_error(GLP_ERR_ARG, "waves per workgroup must be 4 or 8");
  hipLaunchKernelGGL(k_reset_ctrs, dim3(1), dim3(256), 0, stream, d_ctrs, d_wctr, nwin);
  const void *fn = batched_kernel_fn(a.bits, a.precision, waves);
  void *args[] = {(void *)&d_wins, (void *)&d_order, (void *)&d_qtab, (void *)&nwin,
                  (void *)&d_wctr, (void *)&d_ctrs, (void *)&a};
  hipError_t err = hipLaunchKernel((void *)fn, dim3(grid), dim3(64 * waves), args,
                                   lds_bytes(a, waves), stream);
  return hip_check(err, "k_batched launch");
}

int batched_occupancy(unsigned bits, int precision, unsigned waves, size_t smem,
                      int *blocks_per_cu) {
  const void *fn = batched_kernel_fn(bits, precision, waves);
  GLP_TRY_K(allow_lds(fn, smem));
  return hip_check(
      hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, 64 * (int)waves, smem),
      "occupancy query");
}

// ---------------------------------------------------------------------------
// the forward-backward layout (k_fb_mirror)
// ---------------------------------------------------------------------------
// For a window of M sites (p = M mod 2, FbGeom g), built from the window's
// own panel ([H][wn] u64), packed codes ([nwords][N] u32, MSB-first) and
// tran ([M][3]):
//   panel [2][H][g.wh] u64: part 0 = the first g.wh words of each row;
//     part 1, bit j of row h = site M-1+p-j of row h (0 outside [0, M));
//   codes [2][g.nw][N] u32: part 0 = the first g.nw words of each column;
//     part 1, code slot j of column i (MSB-first) = the code of site
//     M-1+p-j (0 outside [0, M));
//   tran, forward rows r < 32 g.nch: (1, 0, 0) at r = 0, tran[r] for
//     r < M, else 0; then backward rows r (96 g.nch + 128 floats): (1, 0, 0)
//     for r <= p, tran[M+p-r] for r > p while M+p-r >= 1, else 0.
// Step j of the backward half then reads everything at index j.
__device__ __forceinline__ uint64_t sites64(const uint64_t *row, unsigned wn, unsigned M,
                                            long long lo) {
  // bits [lo, lo + 64) of the row (bit i = site lo + i), 0 outside [0, M)
  const long long w0 = lo >= 0 ? lo / 64 : -((63 - lo) / 64);
  const unsigned sh = (unsigned)(lo - 64 * w0);
  const uint64_t a = (w0 >= 0 && w0 < (long long)wn) ? row[w0] : 0ull;
  const uint64_t b = (w0 + 1 >= 0 && w0 + 1 < (long long)wn) ? row[w0 + 1] : 0ull;
  uint64_t x = sh ? (a >> sh) | (b << (64u - sh)) : a;
  const long long keep = (long long)M - lo;  // sites lo + i < M
  if (keep <= 0) return 0ull;
  if (keep < 64) x &= (1ull << keep) - 1ull;
  if (lo < 0) x &= ~0ull << (unsigned)(-lo < 64 ? -lo : 63);
  return x;
}

// Code word q of column i in the mirrored half: slot r (MSB-first) holds the
// code of site hi - r, hi = M-1+p - q*cpw, 0 outside [0, M).  The word's
// sites span at most two source words k0, k0 + 1; their 64 bits with the
// fields reversed put slot r at field d + r, so one shift extracts it.
template <int BITS>
__device__ __forceinline__ uint32_t mirror_code_word(const uint32_t *__restrict__ packed,
                                                     unsigned N, unsigned M, unsigned p,
                                                     unsigned q, unsigned i) {
  constexpr int CPW = 32 / BITS;
  const unsigned nsrc = (M * BITS + 31u) >> 5;  // code words per column
  const int hi = (int)(M - 1u + p) - (int)q * CPW;  // >= 1 for q < nw
  const int lo = hi - CPW + 1;                      // >= 2 - CPW
  const int k0 = (lo + CPW) / CPW - 1;              // floor(lo / CPW), >= -1
  const uint32_t a = (k0 >= 0 && (unsigned)k0 < nsrc) ? packed[(size_t)k0 * N + i] : 0u;
  const uint32_t b = (unsigned)(k0 + 1) < nsrc ? packed[(size_t)(k0 + 1) * N + i] : 0u;
  uint64_t x = __builtin_bswap64(((uint64_t)a << 32) | b);  // bytes reversed
  if (BITS == 4)
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
  // x's field f' holds source field 2 CPW - 1 - f'; slot r wants source
  // field hi - r - k0 CPW, i.e. field d + r with d = 2 CPW - 1 - hi + k0 CPW
  const unsigned d = (unsigned)(2 * CPW - 1 - hi + k0 * CPW);  // 1 .. CPW
  uint32_t w = (uint32_t)((x << (BITS * d)) >> 32);
  // slots r < r0 hold sites >= M, slots r > hi sites < 0
  const unsigned r0 = hi >= (int)M ? (unsigned)(hi - (int)M + 1) : 0u;
  if (r0) w &= 0xFFFFFFFFu >> (BITS * r0);
  if (hi < CPW - 1) w &= ~(0xFFFFFFFFu >> (BITS * (hi + 1)));
  return w;
}

// One work item per output word: panel words (both halves), code words
// (both halves), tran floats.  Every index fits 32 bits (fb_check bounds the
// panel and code arrays below 4 GiB).
template <int BITS>
__global__ void __launch_bounds__(256)
    k_fb_mirror(const FbWin *__restrict__ fw, unsigned N, unsigned H) {
  const FbWin f = fw[blockIdx.y];
  const unsigned M = f.M, p = M & 1u;
  const FbGeom g(M, BITS);
  const unsigned np = H * g.wh, nc = g.nw * N, ntf = 96u * g.nch;
  const unsigned ntb = 96u * g.nch + 128u;
  const unsigned total = np + nc + ntf + ntb;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += gridDim.x * blockDim.x) {
    if (t < np) {
      const unsigned h = t / g.wh, k = t - h * g.wh;
      const uint64_t *row = f.panel + (size_t)h * f.wn;
      f.cpanel[t] = row[k];
      const long long lo = (long long)M + p - 64ll * (k + 1);
      f.cpanel[np + t] = __builtin_bitreverse64(sites64(row, f.wn, M, lo));
    } else if (t < np + nc) {
      const unsigned u = t - np, q = u / N, i = u - q * N;
      f.ccodes[u] = f.packed[u];  // rows of N words in both arrays
      f.ccodes[nc + u] = mirror_code_word<BITS>(f.packed, N, M, p, q, i);
    } else {
      const unsigned u = t - np - nc;
      const bool back = u >= ntf;
      const unsigned fl = back ? u - ntf : u, r = fl / 3u, c = fl - 3u * r;
      float v = 0.0f;
      if (!back) {
        if (r == 0u) v = c == 0u ? 1.0f : 0.0f;
        else if (r < M) v = f.tran[3u * r + c];
      } else if (r <= p) {
        v = c == 0u ? 1.0f : 0.0f;
      } else if (M + p >= r + 1u) {
        v = f.tran[3u * (M + p - r) + c];
      }
      f.ctran[u] = v;
    }
  }
}

size_t fb_layout_bytes(unsigned M, unsigned bits, unsigned N, unsigned H) {
  const FbGeom g(M, bits);
  return 16u * (size_t)H * g.wh + 8u * (size_t)g.nw * N + 4u * (size_t)g.tran_floats();
}

int fb_check(unsigned M, unsigned bits, unsigned N, unsigned H) {
  const FbGeom g(M, bits);
  // buffer descriptors with 32-bit byte offsets over each array
  if (16ull * H * g.wh >= (1ull << 32) || 8ull * g.nw * N >= (1ull << 32))
    return set_error(GLP_ERR_ARG, "a window's forward-backward layout reaches 4 GiB");
  return GLP_OK;
}

int launch_fb_mirror(const FbWin *d_fw, unsigned nwin, unsigned max_items, unsigned N,
                     unsigned H, unsigned bits, hipStream_t stream) {
  unsigned gx = (max_items + 255u) / 256u;
  const unsigned cap = std::max(1u, 4096u / std::max(1u, nwin));
  gx = std::max(1u, std::min(gx, cap));
  if (bits == 8)
    hipLaunchKernelGGL(k_fb_mirror<8>, dim3(gx, nwin), dim3(256), 0, stream, d_fw, N, H);
  else
    hipLaunchKernelGGL(k_fb_mirror<4>, dim3(gx, nwin), dim3(256), 0, stream, d_fw, N, H);
  return hip_check(hipGetLastError(), "k_fb_mirror launch");
}

const void *fb_mirror_kernel_fn(unsigned bits) {
  return bits == 8 ? (const void *)&k_fb_mirror<8> : (const void *)&k_fb_mirror<4>;
}

unsigned fb_items(unsigned M, unsigned bits, unsigned N, unsigned H) {
  const FbGeom g(M, bits);
  return (unsigned)((size_t)H * g.wh + (size_t)g.nw * N + 192u * g.nch + 128u);
}

int launch_eval(const KArgs &a, const glp_window *d_wins, const unsigned *d_samples,
                const unsigned *d_parents, unsigned n, float *d_out, hipStream_t stream) {
  const dim3 block(256), grid((2 * n + 255) / 256);
  KArgs af = a;  // k_eval accumulates in F32 (Impute::hmm_like returns a float)
  af.precision = kPrecF32;
  const size_t smem = lds_bytes(af, 4);
  GLP_TRY_K(allow_lds(a.bits == 8 ? (const void *)&k_eval<8> : (const void *)&k_eval<4>, smem));
  if (a.bits == 8)
    hipLaunchKernelGGL((k_eval<8>), grid, block, smem, stream, d_wins, af, d_samples,
                       d_parents, n, d_out);
  else
    hipLaunchKernelGGL((k_eval<4>), grid, block, smem, stream, d_wins, af, d_samples,
                       d_parents, n, d_out);
  return hip_check(hipGetLastError(), "k_eval launch");
}

// ---------------------------------------------------------------------------
// the batched generation without allocations (glp_mh_run_noalloc)
// ---------------------------------------------------------------------------
// Every device buffer of a batched run beyond the caller's inputs and
// outputs -- statistics, per-window task counters, the longest-first window
// order, the lanes table and the window table -- sits in this file-scope
// array, one copy per device, held until the process ends.  Host code uses
// the per-device address from hipGetSymbolAddress (noalloc_scratch), never
// the array's name.  Layout (bytes): [0, 32) statistics (4 x u64),
// [256, ...) task counters, order, lanes (u32 each, kNoAllocMaxWindows
// apiece), then the window table (16-byte aligned).
constexpr unsigned kNoAllocMaxWindows = GLP_NOALLOC_MAX_WINDOWS;
constexpr size_t kNoAllocWctr = 256;
constexpr size_t kNoAllocOrder = kNoAllocWctr + 4 * (size_t)kNoAllocMaxWindows;
constexpr size_t kNoAllocQtab = kNoAllocOrder + 4 * (size_t)kNoAllocMaxWindows;
constexpr size_t kNoAllocWins = kNoAllocQtab + 4 * (size_t)kNoAllocMaxWindows;
constexpr size_t kNoAllocBytes = kNoAllocWins + sizeof(glp_window) * (size_t)kNoAllocMaxWindows;
static_assert(kNoAllocWins % 16 == 0, "window table alignment");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// order = the windows by site count, longest first, ties in index order (the
// plan's stable sort); qtab = Q for every window.  One thread per window.
__global__ void k_noalloc_order(const glp_window *__restrict__ wins, unsigned nwin, unsigned Q,
                                unsigned *__restrict__ order, unsigned *__restrict__ qtab) {
  const unsigned w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwin) return;
  const unsigned m = wins[w].num_sites;
  unsigned rank = 0;
  for (unsigned v = 0; v < nwin; ++v) {
    const unsigned mv = wins[v].num_sites;
    rank += (mv > m || (mv == m && v < w)) ? 1u : 0u;
  }
  order[rank] = w;
  qtab[w] = Q;
}

// The scratch array's address on device `dev` (the current device), looked
// up once per device.
constexpr int kNoAllocMaxDevices = 64;

static int noalloc_scratch(int dev, unsigned char **out) {
  static std::mutex mu;
  static unsigned char *addr[kNoAllocMaxDevices] = {};
  std::lock_guard<std::mutex> lock(mu);
  if (dev < 0 || dev >= kNoAllocMaxDevices) return set_error(GLP_ERR_ARG, "bad device");
  if (!addr[dev]) {
    void *p = nullptr;
    GLP_TRY_K(hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address"));
    addr[dev] = static_cast<unsigned char *>(p);
  }
  *out = addr[dev];
  return GLP_OK;
}

int run_noalloc(const KArgs &a, const glp_window *wins_h, unsigned nwin, unsigned Q,
                int num_cu, hipStream_t stream, uint64_t *stats_out) {
  if (nwin > kNoAllocMaxWindows)
    return set_error(GLP_ERR_ARG, "glp_mh_run_noalloc: more than GLP_NOALLOC_MAX_WINDOWS windows");
  int dev = 0;
  GLP_TRY_K(hip_check(hipGetDevice(&dev), "hipGetDevice"));
  unsigned char *base = nullptr;
  GLP_TRY_K(noalloc_scratch(dev, &base));
  auto *ctrs = reinterpret_cast<unsigned long long *>(base);
  auto *wctr = reinterpret_cast<unsigned *>(base + kNoAllocWctr);
  auto *order = reinterpret_cast<unsigned *>(base + kNoAllocOrder);
  auto *qtab = reinterpret_cast<unsigned *>(base + kNoAllocQtab);
  auto *wins = reinterpret_cast<glp_window *>(base + kNoAllocWins);
  if (a.precision == kPrecMixedFB)
    return set_error(GLP_ERR_ARG, "glp_mh_run_noalloc: the forward-backward form needs a plan "
                                  "(its window layouts are allocated)");
  const size_t smem = lds_bytes(a, 4);
  int bpc = 0;
  GLP_TRY_K(batched_occupancy(a.bits, a.precision, 4, smem, &bpc));
  if (bpc < 1) return set_error(GLP_ERR_ARG, "the batched kernel's workgroup does not fit a CU (LDS)");
  // the plan's grid: one wave per floor(32 / Q) chains, 4 waves a block
  const double waves = (double)nwin * std::ceil((double)a.N / (double)(32u / Q));
  const unsigned blocks = (unsigned)std::ceil(waves / 4.0);
  const unsigned grid = std::max(1u, std::min((unsigned)(bpc * num_cu), blocks));
  GLP_TRY_K(hip_check(hipMemcpyAsync(wins, wins_h, sizeof(glp_window) * nwin,
                                     hipMemcpyHostToDevice, stream), "window table copy"));
  hipLaunchKernelGGL(k_noalloc_order, dim3((nwin + 255) / 256), dim3(256), 0, stream, wins, nwin,
                     Q, order, qtab);
  GLP_TRY_K(hip_check(hipGetLastError(), "k_noalloc_order launch"));
  GLP_TRY_K(launch_batched(a, wins, order, qtab, nwin, wctr, ctrs, grid, 4, stream));
  if (stats_out) {
    unsigned long long h[4];
    GLP_TRY_K(hip_check(hipStreamSynchronize(stream), "synchronize"));
    GLP_TRY_K(hip_check(hipMemcpy(h, ctrs, sizeof(h), hipMemcpyDeviceToHost), "stats copy"));
    stats_out[0] = h[1];
    stats_out[1] = h[2];
    stats_out[2] = h[3];
    if (h[3]) return set_error(GLP_ERR_STATE, "batched kernel hit its round bound");
  }
  return GLP_OK;
}

}  // namespace glp

extern "C" int glp_selftest_recip(uint32_t lo_bits, uint32_t hi_bits, uint64_t *mismatches,
                                  uint32_t *first_bad) {
  using namespace glp;
  if (!mismatches || !first_bad) return set_error(GLP_ERR_ARG, "null output");
  if (hi_bits < lo_bits) return set_error(GLP_ERR_ARG, "hi_bits < lo_bits");
  unsigned long long *d = nullptr;
  GLP_TRY_K(hip_check(hipMalloc(&d, 16), "hipMalloc"));
  unsigned long long init[2] = {0ull, 0xFFFFFFFFull};
  int rc = hip_check(hipMemcpy(d, init, 16, hipMemcpyHostToDevice), "self-test copy");
  if (rc == GLP_OK) {
    hipLaunchKernelGGL(k_selftest_recip, dim3(4096), dim3(256), 0, nullptr, lo_bits, hi_bits, d,
                       reinterpret_cast<uint32_t *>(d + 1));
    rc = hip_check(hipGetLastError(), "self-test launch");
  }
  if (rc == GLP_OK) rc = hip_check(hipMemcpy(init, d, 16, hipMemcpyDeviceToHost), "self-test copy");
  (void)hipFree(d);
  if (rc != GLP_OK) return rc;
  *mismatches = init[0];
  *first_bad = (uint32_t)init[1];
  return GLP_OK;
}

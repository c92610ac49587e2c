// glp_knn.hip -- gfx950 kernels of the kNN candidate lists (KNN,
// kNN.cpp:8-225) and their part of the C ABI (include/glphase_amd.h).
//
// Three steps: per-site variant allele counts over the panel give the mask
// of kept sites (CalculateVarAfs / CutDownHaps); a masked, word-major copy of
// the panel feeds an all-pairs Hamming kernel on the matrix cores that
// writes 16-bit distances of a batch of query haplotypes (the samples' own)
// against the panel; one
// wavefront per query row then selects its neighbours from a histogram of
// the row's distances, an index-ordered compaction of the ties at the
// threshold and a bitonic sort by (distance, index), and one workgroup per
// sample interleaves A's and B's lists as ClusterHaps does (kNN.cpp:176-199).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../../include/glphase_amd.h"
#include "glp_common.hpp"
#include "glp_internal.hpp"

namespace glp {

constexpr unsigned kKnnBatch = 4096;     // query haplotypes per distance batch
constexpr unsigned kKnnCountChunk = 128;  // haplotypes per allele-count block
constexpr unsigned kKnnMaxHalf = 1024;   // ceil(K / 2) limit
constexpr unsigned kKnnBins = 2560;      // distance histogram bins per row
constexpr unsigned kSelVecs = 8;         // 16-byte reads per lane in flight per select step
constexpr unsigned kKnnKc = 8;           // 64-site words per staged chunk of the distance kernel
constexpr size_t kKnnDistLds = 2 * kKnnKc * 384 * 8;  // its two staging buffers (48 KB)
constexpr unsigned kKnnMaxWords = 1024;  // 65,536 sites

// Alt-allele counts of the 64 sites of word w over haplotypes [h0, h0 + kKnnCountChunk).
__global__ void __launch_bounds__(64)
    k_knn_count(const uint64_t *__restrict__ panel, unsigned H, unsigned wn,
                unsigned *__restrict__ counts) {
  const unsigned w = blockIdx.x, lane = threadIdx.x;
  const unsigned h0 = blockIdx.y * kKnnCountChunk;
  const unsigned h1 = h0 + kKnnCountChunk < H ? h0 + kKnnCountChunk : H;
  unsigned alt = 0;
  const auto p = gmem(panel);
#pragma unroll 8
  for (unsigned h = h0; h < h1; ++h) alt += (unsigned)(p[(size_t)h * wn + w] >> lane) & 1u;
  atomicAdd(&counts[64 * w + lane], alt);
}

// Kept-site mask from the counts: frequency alt / H as a double
// (kNN.cpp:248-250), folded to the minor allele frequency with using_maf.
__global__ void __launch_bounds__(64)
    k_knn_mask(const unsigned *__restrict__ counts, unsigned H, unsigned M, double lb, double ub,
               int maf, uint64_t *__restrict__ mask) {
  const unsigned w = blockIdx.x, lane = threadIdx.x, s = 64 * w + lane;
  double a = (double)counts[s] / H;
  if (maf) a = fabs(fabs(a - 0.5) - 0.5);
  const uint64_t b = __ballot(s < M && a >= lb && a <= ub);
  if (lane == 0) mask[w] = b;
}

// Word-major masked panel: pt[w][h] = panel[h][w] & mask[w].
__global__ void __launch_bounds__(256)
    k_knn_transpose(const uint64_t *__restrict__ panel, unsigned H, unsigned wn,
                    const uint64_t *__restrict__ mask, uint64_t *__restrict__ pt) {
  const unsigned h = blockIdx.x * 256u + threadIdx.x, w = blockIdx.y;
  if (h < H) pt[(size_t)w * H + h] = panel[(size_t)h * wn + w] & mask[w];
}

// Masked allele count of every haplotype: pc[h] = sum_w popcount(pt[w][h]).
__global__ void __launch_bounds__(256)
    k_knn_popc(const uint64_t *__restrict__ pt, unsigned H, unsigned wn, unsigned *__restrict__ pc) {
  const unsigned h = blockIdx.x * 256u + threadIdx.x;
  if (h >= H) return;
  const auto p = gmem(pt);
  unsigned c = 0;
  for (unsigned w = 0; w < wn; ++w) c += (unsigned)__popcll(p[(size_t)w * H + h]);
  pc[h] = c;
}

// The distances as a matrix product on the matrix cores: with a, b the 0/1
// vectors of two masked haplotypes, popcount(a ^ b) = |a| + |b| - 2 a.b, and
// the a.b of a tile of queries x haplotypes is an int8 product accumulated
// in int32 (v_mfma_i32_32x32x32_i8), exact.
//  - A block covers 128 queries x 256 haplotypes; each of its 4 wavefronts
//    owns 128 queries x 64 haplotypes (4 x 2 blocks of 32 x 32).
//  - The block stages kKnnKc 64-site words of its 384 rows at a time in LDS
//    with direct global-to-LDS loads (global_load_lds_dwordx4: 1 KB of the
//    word-major panel per wavefront load), double-buffered: chunk c + 1 is in
//    flight while chunk c's products run.
//  - For a 32-row block, lane l reads the word of row l % 32 and expands its
//    bits 16 (l / 32) .. + 15 and 32 + 16 (l / 32) .. + 15 into the two
//    16-byte operands of the word's two 32-site steps: query operands
//    through a 256-entry LDS table (byte -> 8 bytes of 0 / 1), haplotype
//    operands by multiplies (a nibble n spreads to bytes by
//    (n * 0x00204081) & 0x01010101), which splits the work between the LDS
//    and VALU pipes.
//  - Operand layout (lane l holds row l % 32, sites 16 (l / 32) .. + 15) and
//    result layout (lane l, register r: row 8 (r / 4) + 4 (l / 32) + r % 4,
//    column l % 32) are the instruction's.
// Rows past the batch or the panel read a valid row (clamped) and are not
// stored.
// Row stride of the distance block: H rounded up to 8 distances, so a row
// starts 16-byte aligned (the select kernel reads 8 distances at a time).
__host__ __device__ __forceinline__ size_t knn_row_stride(unsigned H) {
  return ((size_t)H + 7) & ~(size_t)7;
}

typedef int knn_v4i __attribute__((ext_vector_type(4)));
typedef int knn_v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void_t;
typedef const __attribute__((address_space(1))) void gbl_void_t;

__device__ __forceinline__ knn_v4i knn_expand_lut(const uint2 *lut, uint64_t w, unsigned off) {
  const unsigned x = (unsigned)(w >> off);
  const uint2 lo = lut[x & 255u], hi = lut[(x >> 8) & 255u];
  knn_v4i v;
  v.x = (int)lo.x;
  v.y = (int)lo.y;
  v.z = (int)hi.x;
  v.w = (int)hi.y;
  return v;
}

__device__ __forceinline__ knn_v4i knn_expand_mul(uint64_t w, unsigned off) {
  const unsigned x = (unsigned)(w >> off);
  knn_v4i v;
  v.x = (int)((__builtin_amdgcn_ubfe(x, 0, 4) * 0x00204081u) & 0x01010101u);
  v.y = (int)((__builtin_amdgcn_ubfe(x, 4, 4) * 0x00204081u) & 0x01010101u);
  v.z = (int)((__builtin_amdgcn_ubfe(x, 8, 4) * 0x00204081u) & 0x01010101u);
  v.w = (int)((__builtin_amdgcn_ubfe(x, 12, 4) * 0x00204081u) & 0x01010101u);
  return v;
}

// Stage chunk c (words c kKnnKc .. + kKnnKc - 1, clamped into the window) of
// the block's rows into buffer buf ([kKnnKc][384] words: query rows 0..127,
// then haplotype rows 0..255).  Wavefront v issues loads t = kPer v .. kPer
// v + kPer - 1: word t / 3, part t % 3 (queries, haplotypes 0..127, 128..255).
__device__ __forceinline__ void knn_stage(uint64_t *buf, const uint64_t *pt, unsigned H,
                                          unsigned wn, unsigned c, unsigned qrow, unsigned hrow) {
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  constexpr unsigned kPer = 3 * kKnnKc / 4;  // loads per wavefront
#pragma unroll
  for (unsigned m = 0; m < kPer; ++m) {
    const unsigned t = kPer * wave + m, wl = t / 3u, part = t % 3u;
    unsigned w = c * kKnnKc + wl;
    w = w < wn ? w : wn - 1;
    const unsigned base = part == 0 ? qrow : hrow + 128u * (part - 1u);
    unsigned r = base + 2u * lane;
    r = r < H - 2u ? r : H - 2u;  // an even pair of rows inside the panel
    __builtin_amdgcn_global_load_lds((gbl_void_t *)(pt + (size_t)w * H + r),
                                     (lds_void_t *)(buf + wl * 384u + 128u * part), 16, 0, 0);
  }
}

__global__ void __launch_bounds__(256, 2)
    k_knn_dist(const uint64_t *__restrict__ pt, const unsigned *__restrict__ pc, unsigned H,
               unsigned wn, unsigned q0, unsigned nq, uint16_t *__restrict__ d) {
  extern __shared__ __align__(16) uint64_t s_stage[];  // 2 x kKnnKc x 384 words
  __shared__ uint2 lut[256];
  __shared__ unsigned s_pa[128];  // |a| of the block's queries
  const unsigned qt = blockIdx.y * 128u;          // first query of the block
  const unsigned hb0 = blockIdx.x * 256u;         // first haplotype of the block
  const unsigned nc = (wn + kKnnKc - 1) / kKnnKc;  // chunks
  knn_stage(s_stage, pt, H, wn, 0, q0 + qt, hb0);
  if (threadIdx.x < 128u) {
    const unsigned q = qt + threadIdx.x;
    s_pa[threadIdx.x] = pc[q0 + (q < nq ? q : nq - 1)];
  }
  for (unsigned b = threadIdx.x; b < 256u; b += blockDim.x) {
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (unsigned i = 0; i < 4; ++i) {
      lo |= ((b >> i) & 1u) << (8 * i);
      hi |= ((b >> (i + 4)) & 1u) << (8 * i);
    }
    lut[b] = make_uint2(lo, hi);
  }
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const unsigned row = lane & 31u, half = lane >> 5, off = 16u * half;
  const unsigned ht = hb0 + wave * 64u;  // first haplotype of the wavefront
  unsigned pb[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const unsigned h = ht + 32u * j + row;
    pb[j] = pc[h < H ? h : H - 1];
  }
  knn_v16i acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0;
  for (unsigned c = 0; c < nc; ++c) {
    // chunk c has landed (this wavefront's loads), every wavefront is past
    // chunk c - 1 (so its buffer may take chunk c + 1)
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (c + 1 < nc)
      knn_stage(s_stage + ((c + 1) & 1u) * (kKnnKc * 384u), pt, H, wn, c + 1, q0 + qt, hb0);
    const uint64_t *buf = s_stage + (c & 1u) * (kKnnKc * 384u);
    const unsigned nw = wn - c * kKnnKc < kKnnKc ? wn - c * kKnnKc : kKnnKc;
#pragma unroll 2
    for (unsigned wl = 0; wl < nw; ++wl) {
      const uint64_t *wr = buf + wl * 384u;
      uint64_t wa[4], wb[2];
#pragma unroll
      for (int i = 0; i < 4; ++i) wa[i] = wr[32u * i + row];
#pragma unroll
      for (int j = 0; j < 2; ++j) wb[j] = wr[128u + 64u * wave + 32u * j + row];
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        knn_v4i fa[4], fb[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) fa[i] = knn_expand_lut(lut, wa[i], off + 32u * s);
#pragma unroll
        for (int j = 0; j < 2; ++j) fb[j] = knn_expand_mul(wb[j], off + 32u * s);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
      }
    }
  }
  // d = |a| + |b| - 2 a.b; lane l holds column l % 32 of each block
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const unsigned h = ht + 32u * j + row;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const unsigned qi = 32u * i + 8u * (r >> 2) + 4u * half + (r & 3), q = qt + qi;
        if (q < nq && h < H)
          d[(size_t)q * knn_row_stride(H) + h] =
              (uint16_t)(s_pa[qi] + pb[j] - 2u * (unsigned)acc[i][j][r]);
      }
  }
}

// Orders a wavefront's LDS accesses before and after this point (the
// wavefront's own lanes only; no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanes_below() {
  const unsigned lane = threadIdx.x & 63u;
  return lane ? ~0ull >> (64 - lane) : 0ull;
}

// The bin b of hist[0, nb) (nb a multiple of 64) holding the k-th smallest
// counted value (1 <= k <= total count): cum(< b) < k <= cum(<= b); *below
// = cum(< b).  Each lane sums nb / 64 consecutive bins, a lane scan finds the
// lane holding the k-th value, and that lane walks its bins.
__device__ unsigned wave_find_bin(const unsigned *hist, unsigned nb, unsigned k, unsigned *below) {
  const unsigned lane = threadIdx.x & 63u, per = nb / 64;
  const unsigned *mine = hist + lane * per;
  unsigned s = 0;
  for (unsigned i = 0; i < per; ++i) s += mine[i];
  unsigned incl = s;
  for (unsigned o = 1; o < 64; o <<= 1) {
    const unsigned t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  const unsigned excl = incl - s;
  const bool here = excl < k && k <= incl;
  const unsigned src = (unsigned)__builtin_ctzll(__ballot(here));
  unsigned bin = 0, cum = excl;
  if (here) {
    unsigned i = 0;
    while (cum + mine[i] < k) cum += mine[i++];
    bin = lane * per + i;
  }
  *below = (unsigned)__shfl((int)cum, (int)src, 64);
  return (unsigned)__shfl((int)bin, (int)src, 64);
}

// Sorts the k (distinct) keys of tmp[0, k) into out[0, k), where out =
// tmp + k: a bitonic sort of tmp[0, P), P = k rounded up to a power of two
// (at most 2k, so the padding lies in out), padded with the largest key;
// then the first k move to out.  The keys are distinct (their low word is
// the haplotype), so the order is the rank order by (distance, index).
__device__ void knn_sort_keys(uint64_t *tmp, unsigned k, uint64_t *out) {
  const unsigned lane = threadIdx.x & 63u;
  unsigned P = 1;
  while (P < k) P <<= 1;
  for (unsigned i = k + lane; i < P; i += 64) tmp[i] = ~0ull;
  wave_lds_sync();
  for (unsigned size = 2; size <= P; size <<= 1)
    for (unsigned stride = size >> 1; stride > 0; stride >>= 1) {
      for (unsigned i = lane; i < P / 2; i += 64) {
        const unsigned a = 2 * stride * (i / stride) + (i % stride), b = a + stride;
        const bool up = (a & size) == 0;
        const uint64_t ka = tmp[a], kb = tmp[b];
        if ((ka > kb) == up) {
          tmp[a] = kb;
          tmp[b] = ka;
        }
      }
      wave_lds_sync();
    }
  for (unsigned i = lane; i < k; i += 64) out[i] = tmp[i];
  wave_lds_sync();
}

// The k nearest haplotypes of row dr (H distances, padded to a multiple of
// 8; haplotypes 2u and 2u + 1 excluded) by (distance, index), as keys
// (d << 32 | h) sorted into out[0, k), over the whole row (the fallback of
// knn_select_wave).  One wavefront per row, no workgroup barrier:
//  1. a histogram of d >> shift (exact distances when shift = 0) in the
//     wavefront's own kKnnBins counters gives the threshold bin; with
//     shift > 0, a second histogram of the low bits inside that bin refines
//     it to the threshold distance t and the count of distances below t;
//  2. one pass in index order keeps every d < t and the first `need` with
//     d == t, placed by ballot prefix counts (so ties go by index);
//  3. a sort of the k keys (knn_sort_keys).
// A lane reads 8 distances (16 bytes) at a time and has kSelVecs such reads
// in flight per step of a pass: the passes are bound by the latency of the
// row's reads, not by their bytes.
__device__ __forceinline__ unsigned knn_elem(const uint4 &v, unsigned t) {
  const unsigned w = t < 2 ? v.x : t < 4 ? v.y : t < 6 ? v.z : v.w;
  return (t & 1u) ? w >> 16 : w & 0xFFFFu;
}

__device__ void knn_select_full(const uint16_t *__restrict__ dr, unsigned H, unsigned u,
                                unsigned k, unsigned shift, unsigned *hist, uint64_t *tmp,
                                uint64_t *out) {
  const unsigned lane = threadIdx.x & 63u;
  const auto row = gmem(reinterpret_cast<const uint32_t *>(dr));
  const unsigned nv = (H + 7u) >> 3;  // 8-distance vectors of the row
  // element e of vector p: 8 p + t, kept when inside the row and not one of
  // the sample's own haplotypes
  auto keep = [&](unsigned p, unsigned t) {
    const unsigned e = 8u * p + t;
    return e < H && (e >> 1) != u;
  };
  for (unsigned b = lane; b < kKnnBins; b += 64) hist[b] = 0;
  wave_lds_sync();
  for (unsigned base = 0; base < nv; base += 64 * kSelVecs) {
    uint4 v[kSelVecs];
#pragma unroll
    for (unsigned j = 0; j < kSelVecs; ++j) {
      const unsigned p = base + 64 * j + lane;
      v[j] = p < nv ? load_u4(row + 4 * p) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (unsigned j = 0; j < kSelVecs; ++j) {
      const unsigned p = base + 64 * j + lane;
      if (p < nv) {
#pragma unroll
        for (unsigned t = 0; t < 8; ++t)
          if (keep(p, t)) atomicAdd(&hist[knn_elem(v[j], t) >> shift], 1u);
      }
    }
  }
  wave_lds_sync();
  unsigned below;
  unsigned thr = wave_find_bin(hist, kKnnBins, k, &below);
  if (shift) {
    wave_lds_sync();
    hist[lane] = 0;
    wave_lds_sync();
    const unsigned lo_mask = (1u << shift) - 1;
    for (unsigned base = 0; base < nv; base += 64 * kSelVecs) {
      uint4 v[kSelVecs];
#pragma unroll
      for (unsigned j = 0; j < kSelVecs; ++j) {
        const unsigned p = base + 64 * j + lane;
        v[j] = p < nv ? load_u4(row + 4 * p) : make_uint4(0, 0, 0, 0);
      }
#pragma unroll
      for (unsigned j = 0; j < kSelVecs; ++j) {
        const unsigned p = base + 64 * j + lane;
        if (p < nv) {
#pragma unroll
          for (unsigned t = 0; t < 8; ++t) {
            const unsigned dv = knn_elem(v[j], t);
            if (keep(p, t) && (dv >> shift) == thr) atomicAdd(&hist[dv & lo_mask], 1u);
          }
        }
      }
    }
    wave_lds_sync();
    unsigned below_lo;
    const unsigned lo = wave_find_bin(hist, 64, k - below, &below_lo);
    thr = (thr << shift) | lo;
    below += below_lo;
  }
  const unsigned need = k - below;  // ties at thr to take, lowest indices first
  const uint64_t lt = lanes_below();
  unsigned nless = 0, nties = 0;
  // index order: 512-element steps ascending, lanes ascending, elements
  // 8 p .. 8 p + 7 of a lane in order
  for (unsigned base = 0; base < nv && (nless < below || nties < need);
       base += 64 * kSelVecs) {
    uint4 v[kSelVecs];
#pragma unroll
    for (unsigned j = 0; j < kSelVecs; ++j) {
      const unsigned p = base + 64 * j + lane;
      v[j] = p < nv ? load_u4(row + 4 * p) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (unsigned j = 0; j < kSelVecs; ++j) {
      const unsigned p = base + 64 * j + lane;
      unsigned ml = 0, me = 0;  // this lane's elements below / at the threshold
#pragma unroll
      for (unsigned t = 0; t < 8; ++t) {
        const unsigned dv = knn_elem(v[j], t);
        const bool ok = p < nv && keep(p, t);
        ml |= (ok && dv < thr) ? 1u << t : 0u;
        me |= (ok && dv == thr) ? 1u << t : 0u;
      }
      // elements of lower lanes come first
      unsigned cl = 0, ce = 0, tl = 0, te = 0;
#pragma unroll
      for (unsigned t = 0; t < 8; ++t) {
        const uint64_t bl = __ballot((ml >> t) & 1u), be = __ballot((me >> t) & 1u);
        cl += (unsigned)__popcll(bl & lt);
        ce += (unsigned)__popcll(be & lt);
        tl += (unsigned)__popcll(bl);
        te += (unsigned)__popcll(be);
      }
      unsigned pl = nless + cl, pe = nties + ce;
#pragma unroll
      for (unsigned t = 0; t < 8; ++t) {
        const uint64_t key = ((uint64_t)knn_elem(v[j], t) << 32) | (8u * p + t);
        if ((ml >> t) & 1u) tmp[pl++] = key;
        if ((me >> t) & 1u) {
          if (pe < need) tmp[below + pe] = key;
          ++pe;
        }
      }
      nless += tl;
      nties += te;
    }
  }
  wave_lds_sync();
  knn_sort_keys(tmp, k, out);
}

// Sum over the wavefront.
__device__ __forceinline__ unsigned wave_sum(unsigned x) {
  for (unsigned o = 32; o; o >>= 1) x += (unsigned)__shfl_xor((int)x, (int)o, 64);
  return x;
}

constexpr unsigned kKnnCand = kKnnBins / 2;  // candidate keys the histogram's LDS can hold

// The same selection, usually from one pass over the row:
//  1. a histogram of the distances of every 8th 8-distance vector (an
//     eighth of the row) gives a bound T with about twice k of the row's
//     distances at or below it (the sample's rank 2 ceil(k ns / nk) + 16);
//  2. one pass in index order compacts every kept distance <= T into the
//     histogram's memory (at most kKnnCand keys; a step without one costs a
//     ballot);
//  3. if there are at least k of them (and they fit), the k-th smallest
//     distance t comes from a binary search over [0, T] with wavefront
//     counts, then the keys with d < t and the first k - below with d == t
//     (still in index order) are sorted.  Every key the exact
//     selection keeps has d <= t <= T, so the result is the same;
// otherwise (too few or too many candidates) it runs knn_select_full.
__device__ void knn_select_wave(const uint16_t *__restrict__ dr, unsigned H, unsigned u,
                                unsigned k, unsigned shift, unsigned *hist, uint64_t *tmp,
                                uint64_t *out) {
  const unsigned lane = threadIdx.x & 63u;
  const auto row = gmem(reinterpret_cast<const uint32_t *>(dr));
  const unsigned nv = (H + 7u) >> 3;
  auto keep = [&](unsigned p, unsigned t) {
    const unsigned e = 8u * p + t;
    return e < H && (e >> 1) != u;
  };
  // 1. the sample
  for (unsigned b = lane; b < kKnnBins; b += 64) hist[b] = 0;
  wave_lds_sync();
  unsigned ns = 0;
  const unsigned nsv = (nv + 7u) >> 3;  // sampled vectors: p = 8 m
  for (unsigned base = 0; base < nsv; base += 64 * kSelVecs) {
    uint4 v[kSelVecs];
#pragma unroll
    for (unsigned j = 0; j < kSelVecs; ++j) {
      const unsigned m = base + 64 * j + lane;
      v[j] = m < nsv ? load_u4(row + 32 * m) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (unsigned j = 0; j < kSelVecs; ++j) {
      const unsigned m = base + 64 * j + lane;
      if (m < nsv) {
#pragma unroll
        for (unsigned t = 0; t < 8; ++t)
          if (keep(8u * m, t)) {
            atomicAdd(&hist[knn_elem(v[j], t) >> shift], 1u);
            ++ns;
          }
      }
    }
  }
  ns = wave_sum(ns);
  wave_lds_sync();
  const unsigned nk = H - 2;  // kept distances of the row
  const unsigned rs = 2u * (unsigned)(((uint64_t)k * ns + nk - 1) / nk) + 16u;
  if (rs > ns) {  // a sample too small to bound anything
    knn_select_full(dr, H, u, k, shift, hist, tmp, out);
    return;
  }
  unsigned below_s;
  const unsigned tb = wave_find_bin(hist, kKnnBins, rs, &below_s);
  const unsigned T = (tb << shift) | ((1u << shift) - 1u);
  wave_lds_sync();
  // 2. every kept d <= T, in index order
  uint64_t *cand = reinterpret_cast<uint64_t *>(hist);
  const uint64_t lt = lanes_below();
  unsigned cnt = 0;
  for (unsigned base = 0; base < nv; base += 64 * kSelVecs) {
    uint4 v[kSelVecs];
#pragma unroll
    for (unsigned j = 0; j < kSelVecs; ++j) {
      const unsigned p = base + 64 * j + lane;
      v[j] = p < nv ? load_u4(row + 4 * p) : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (unsigned j = 0; j < kSelVecs; ++j) {
      const unsigned p = base + 64 * j + lane;
      unsigned mc = 0;
#pragma unroll
      for (unsigned t = 0; t < 8; ++t)
        mc |= (p < nv && keep(p, t) && knn_elem(v[j], t) <= T) ? 1u << t : 0u;
      if (__ballot(mc != 0u) == 0ull) continue;  // wavefront-uniform
      unsigned cl = 0, tl = 0;
#pragma unroll
      for (unsigned t = 0; t < 8; ++t) {
        const uint64_t bt = __ballot((mc >> t) & 1u);
        cl += (unsigned)__popcll(bt & lt);
        tl += (unsigned)__popcll(bt);
      }
      unsigned pos = cnt + cl;
#pragma unroll
      for (unsigned t = 0; t < 8; ++t)
        if ((mc >> t) & 1u) {
          if (pos < kKnnCand) cand[pos] = ((uint64_t)knn_elem(v[j], t) << 32) | (8u * p + t);
          ++pos;
        }
      cnt += tl;
    }
  }
  wave_lds_sync();
  if (cnt < k || cnt > kKnnCand) {
    knn_select_full(dr, H, u, k, shift, hist, tmp, out);
    return;
  }
  // 3. the k-th smallest distance t among the candidates: the least t with
  // count(d <= t) >= k
  unsigned lo = 0, hi = T;
  while (lo < hi) {
    const unsigned mid = (lo + hi) >> 1;
    unsigned c = 0;
    for (unsigned i = lane; i < cnt; i += 64) c += (unsigned)(cand[i] >> 32) <= mid;
    if (wave_sum(c) >= k)
      hi = mid;
    else
      lo = mid + 1;
  }
  const unsigned thr = lo;
  unsigned below = 0;
  for (unsigned i = lane; i < cnt; i += 64) below += (unsigned)(cand[i] >> 32) < thr;
  below = wave_sum(below);
  const unsigned need = k - below;
  // keys with d < thr, then the first `need` with d == thr, in index order
  unsigned nless = 0, nties = 0;
  for (unsigned base = 0; base < cnt; base += 64) {
    const unsigned i = base + lane;
    const uint64_t key = i < cnt ? cand[i] : ~0ull;
    const unsigned dv = (unsigned)(key >> 32);
    const bool l = i < cnt && dv < thr, e = i < cnt && dv == thr;
    const uint64_t bl = __ballot(l), be = __ballot(e);
    const unsigned pl = nless + (unsigned)__popcll(bl & lt);
    const unsigned pe = nties + (unsigned)__popcll(be & lt);
    if (l) tmp[pl] = key;
    if (e && pe < need) tmp[below + pe] = key;
    nless += (unsigned)__popcll(bl);
    nties += (unsigned)__popcll(be);
  }
  wave_lds_sync();
  knn_sort_keys(tmp, k, out);
}

// One workgroup of two wavefronts per sample u in [u0, u0 + nu): wavefront 0
// selects from row 2(u - u0) of d (haplotype 2u, list A, kh nearest),
// wavefront 1 from the next row (haplotype 2u + 1, list B, kb nearest); then
// wavefront 0 interleaves A with B's nearest not already in A (kNN.cpp:176-199).
// Dynamic LDS: keys A (2 kh), B (2 kb), then 2 x kKnnBins counters.
__global__ void __launch_bounds__(128)
    k_knn_select(const uint16_t *__restrict__ d, unsigned H, unsigned u0, unsigned kh,
                 unsigned kb, unsigned kout, unsigned shift, unsigned *__restrict__ klist) {
  extern __shared__ uint64_t s_keys[];
  uint64_t *a_tmp = s_keys, *a_out = a_tmp + kh, *b_tmp = a_out + kh, *b_out = b_tmp + kb;
  unsigned *hist = reinterpret_cast<unsigned *>(b_out + kb);
  const unsigned wave = threadIdx.x >> 6, lane = threadIdx.x & 63u, u = u0 + blockIdx.x;
  const uint16_t *row = d + (size_t)(2 * blockIdx.x + wave) * knn_row_stride(H);
  if (wave == 0)
    knn_select_wave(row, H, u, kh, shift, hist, a_tmp, a_out);
  else
    knn_select_wave(row, H, u, kb, shift, hist + kKnnBins, b_tmp, b_out);
  __syncthreads();
  if (wave) return;
  unsigned *out = klist + (size_t)u * kout;
  for (unsigned j = lane; j < kh; j += 64) out[2 * j] = (unsigned)a_out[j];
  // A's haplotypes in an open-addressing set (the histograms' LDS, free
  // now): at most half full, so a probe takes a step or two
  unsigned *set = hist;
  unsigned ts = 2;
  while (ts < 2 * kh) ts <<= 1;  // <= 2 kKnnMaxHalf = 2048 <= kKnnBins
  const unsigned tmask = ts - 1, tshift = 32u - (unsigned)__builtin_ctz(ts);
  for (unsigned j = lane; j < ts; j += 64) set[j] = 0xFFFFFFFFu;
  wave_lds_sync();
  for (unsigned j = lane; j < kh; j += 64) {
    const unsigned idx = (unsigned)a_out[j];
    unsigned slot = (idx * 2654435761u) >> tshift;
    while (atomicCAS(&set[slot], 0xFFFFFFFFu, idx) != 0xFFFFFFFFu) slot = (slot + 1) & tmask;
  }
  wave_lds_sync();
  const uint64_t lt = lanes_below();
  unsigned placed = 0;
  for (unsigned base = 0; base < kb && placed < kh; base += 64) {
    const unsigned i = base + lane;
    bool fresh = false;
    unsigned idx = 0;
    if (i < kb) {
      idx = (unsigned)b_out[i];
      unsigned slot = (idx * 2654435761u) >> tshift, v;
      while ((v = set[slot]) != idx && v != 0xFFFFFFFFu) slot = (slot + 1) & tmask;
      fresh = v != idx;
    }
    const uint64_t bal = __ballot(fresh);
    const unsigned pos = placed + (unsigned)__popcll(bal & lt);
    if (fresh && pos < kh) out[2 * pos + 1] = idx;
    placed += (unsigned)__popcll(bal);
  }
}

}  // namespace glp

using namespace glp;

namespace {

struct KnnLayout {
  size_t mask, counts, pt, pc, dist, total;
};

KnnLayout knn_layout(const glp_knn_desc *d) {
  const size_t wn = (d->num_sites + 63) / 64, H = d->num_haps;
  const size_t batch = std::min<size_t>(kKnnBatch, 2 * (size_t)d->num_samples);
  KnnLayout L;
  L.mask = 0;
  L.counts = (wn * 8 + 255) & ~(size_t)255;
  L.pt = L.counts + ((64 * wn * 4 + 255) & ~(size_t)255);
  L.pc = L.pt + ((wn * H * 8 + 255) & ~(size_t)255);
  L.dist = L.pc + ((H * 4 + 255) & ~(size_t)255);
  L.total = L.dist + batch * knn_row_stride((unsigned)H) * 2;
  return L;
}

int check_knn(const glp_knn_desc *d) {
  if (!d) return set_error(GLP_ERR_ARG, "null descriptor");
  if (!d->panel || !d->klist) return set_error(GLP_ERR_ARG, "null panel or klist");
  if (d->num_haps < 4 || (d->num_haps & 1u)) return set_error(GLP_ERR_ARG, "num_haps must be even and >= 4");
  if (d->num_sites == 0 || (d->num_sites + 63) / 64 > kKnnMaxWords)
    return set_error(GLP_ERR_ARG, "num_sites must be in [1, 65536]");
  if (d->num_samples == 0 || d->num_samples > d->num_haps / 2)
    return set_error(GLP_ERR_ARG, "num_samples must be in [1, num_haps / 2]");
  if (d->K == 0 || d->K > 2 * kKnnMaxHalf) return set_error(GLP_ERR_ARG, "K must be in [1, 2048]");
  if (d->num_sites >= 65536u) return set_error(GLP_ERR_ARG, "distances are 16-bit: num_sites < 65536");
  if (!(d->freq_lb <= d->freq_ub)) return set_error(GLP_ERR_ARG, "freq_lb > freq_ub");
  return GLP_OK;
}

}  // namespace

extern "C" unsigned glp_knn_k_out(unsigned K, unsigned num_haps) {
  if (num_haps < 2) return 0;
  if (K > num_haps - 2) K = num_haps - 2;
  return 2 * ((K + 1) / 2);
}

extern "C" size_t glp_knn_workspace_bytes(const glp_knn_desc *d) {
  if (!d || d->num_sites == 0) return 0;
  return knn_layout(d).total;
}

extern "C" int glp_knn_build(const glp_knn_desc *d, void *stream) {
  int rc = check_knn(d);
  if (rc != GLP_OK) return rc;
  if (!d->workspace) return set_error(GLP_ERR_ARG, "null workspace");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned H = d->num_haps, M = d->num_sites, wn = (M + 63) / 64;
  const KnnLayout L = knn_layout(d);
  char *ws = static_cast<char *>(d->workspace);
  uint64_t *mask = reinterpret_cast<uint64_t *>(ws + L.mask);
  unsigned *counts = reinterpret_cast<unsigned *>(ws + L.counts);
  uint64_t *pt = reinterpret_cast<uint64_t *>(ws + L.pt);
  unsigned *pc = reinterpret_cast<unsigned *>(ws + L.pc);
  uint16_t *dist = reinterpret_cast<uint16_t *>(ws + L.dist);
  const unsigned kout = glp_knn_k_out(d->K, H), kh = kout / 2;
  const unsigned kb = std::min(2 * kh, H - 2);

  rc = hip_check(hipMemsetAsync(counts, 0, 64 * (size_t)wn * 4, s), "count reset");
  if (rc != GLP_OK) return rc;
  hipLaunchKernelGGL(k_knn_count, dim3(wn, (H + kKnnCountChunk - 1) / kKnnCountChunk), dim3(64), 0,
                     s, d->panel, H, wn, counts);
  hipLaunchKernelGGL(k_knn_mask, dim3(wn), dim3(64), 0, s, counts, H, M, d->freq_lb, d->freq_ub,
                     d->using_maf, mask);
  hipLaunchKernelGGL(k_knn_transpose, dim3((H + 255) / 256, wn), dim3(256), 0, s, d->panel, H, wn,
                     mask, pt);
  hipLaunchKernelGGL(k_knn_popc, dim3((H + 255) / 256), dim3(256), 0, s, pt, H, wn, pc);
  rc = hip_check(hipGetLastError(), "kNN setup launch");
  // distances are at most M: bins of 2^shift distances keep them below kKnnBins
  unsigned shift = 0;
  while ((M >> shift) >= kKnnBins) ++shift;
  const size_t lds = (size_t)(2 * kh + 2 * kb) * 8 + 2 * kKnnBins * 4;
  if (rc == GLP_OK)
    rc = hip_check(hipFuncSetAttribute((const void *)&k_knn_select,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                   "dynamic LDS attribute");
  if (rc == GLP_OK)
    rc = hip_check(hipFuncSetAttribute((const void *)&k_knn_dist,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)kKnnDistLds),
                   "dynamic LDS attribute");
  const unsigned per = kKnnBatch / 2;  // samples per batch
  for (unsigned u0 = 0; u0 < d->num_samples && rc == GLP_OK; u0 += per) {
    const unsigned nu = std::min(per, d->num_samples - u0), nq = 2 * nu;
    hipLaunchKernelGGL(k_knn_dist, dim3((H + 255) / 256, (nq + 127) / 128), dim3(256),
                       kKnnDistLds, s, pt, pc, H, wn, 2 * u0, nq, dist);
    hipLaunchKernelGGL(k_knn_select, dim3(nu), dim3(128), lds, s, dist, H, u0, kh, kb, kout,
                       shift, d->klist);
    rc = hip_check(hipGetLastError(), "kNN batch launch");
  }
  return rc;
}

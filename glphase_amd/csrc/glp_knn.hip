// glp_knn.hip -- gfx950 kernels of the kNN candidate lists (KNN,
// kNN.cpp:8-225) and their part of the C ABI (include/glphase_amd.h).
//
// Three steps: per-site variant allele counts over the panel give the mask
// of kept sites (CalculateVarAfs / CutDownHaps); a masked, word-major copy of
// the panel feeds an all-pairs Hamming kernel that writes 16-bit distances
// of a batch of query haplotypes (the samples' own) against the panel; one
// workgroup per sample then selects its neighbours by a two-pass 8-bit radix
// select, an index-ordered compaction of the ties at the threshold and a
// rank sort by (distance, index), and interleaves A's and B's lists as
// ClusterHaps does (kNN.cpp:176-199).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../../include/glphase_amd.h"
#include "glp_internal.hpp"

namespace glp {

constexpr unsigned kKnnBatch = 4096;     // query haplotypes per distance batch
constexpr unsigned kKnnCountChunk = 1024;  // haplotypes per allele-count block
constexpr unsigned kKnnMaxHalf = 1024;   // ceil(K / 2) limit
constexpr unsigned kKnnQpb = 8;          // queries per distance block
constexpr unsigned kKnnMaxWords = 1024;  // 65,536 sites

// Alt-allele counts of the 64 sites of word w over haplotypes [h0, h0 + 1024).
__global__ void __launch_bounds__(64)
    k_knn_count(const uint64_t *__restrict__ panel, unsigned H, unsigned wn,
                unsigned *__restrict__ counts) {
  const unsigned w = blockIdx.x, lane = threadIdx.x;
  const unsigned h0 = blockIdx.y * kKnnCountChunk;
  const unsigned h1 = h0 + kKnnCountChunk < H ? h0 + kKnnCountChunk : H;
  unsigned alt = 0;
  for (unsigned h = h0; h < h1; ++h) alt += (unsigned)(panel[(size_t)h * wn + w] >> lane) & 1u;
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

// Distances of queries q0 + blockIdx.y * 8 + q (q < 8, < q0 + nq) to every
// haplotype h: d[(query - q0) * H + h].
__global__ void __launch_bounds__(256)
    k_knn_dist(const uint64_t *__restrict__ pt, unsigned H, unsigned wn, unsigned q0,
               unsigned nq, uint16_t *__restrict__ d) {
  extern __shared__ uint64_t sq[];  // [wn][kKnnQpb]
  const unsigned qb = blockIdx.y * kKnnQpb;
  for (unsigned t = threadIdx.x; t < wn * kKnnQpb; t += blockDim.x) {
    const unsigned w = t / kKnnQpb, q = t % kKnnQpb;
    sq[t] = qb + q < nq ? pt[(size_t)w * H + q0 + qb + q] : 0ull;
  }
  __syncthreads();
  const unsigned h = blockIdx.x * 256u + threadIdx.x;
  if (h >= H) return;
  unsigned acc[kKnnQpb];
#pragma unroll
  for (unsigned q = 0; q < kKnnQpb; ++q) acc[q] = 0;
  for (unsigned w = 0; w < wn; ++w) {
    const uint64_t c = pt[(size_t)w * H + h];
#pragma unroll
    for (unsigned q = 0; q < kKnnQpb; ++q) acc[q] += (unsigned)__popcll(c ^ sq[w * kKnnQpb + q]);
  }
#pragma unroll
  for (unsigned q = 0; q < kKnnQpb; ++q)
    if (qb + q < nq) d[(size_t)(qb + q) * H + h] = (uint16_t)acc[q];
}

// Exclusive prefix sums of a 0/1 flag over the 256 threads of a block;
// returns the thread's prefix and sets *total.
__device__ __forceinline__ unsigned block_scan(bool f, unsigned *s_wave, unsigned *total) {
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t b = __ballot(f);
  const unsigned in_wave = (unsigned)__popcll(b & (lane ? (~0ull >> (64 - lane)) : 0ull));
  if (lane == 0) s_wave[wave] = (unsigned)__popcll(b);
  __syncthreads();
  unsigned before = 0, all = 0;
  for (unsigned v = 0; v < 4; ++v) {
    if (v < wave) before += s_wave[v];
    all += s_wave[v];
  }
  __syncthreads();
  *total = all;
  return before + in_wave;
}

// The k nearest haplotypes of row `dr` (length H; haplotypes 2u, 2u+1
// excluded) by (distance, index), as keys (d << 32 | idx) in out[0, k).
__device__ void knn_select(const uint16_t *__restrict__ dr, unsigned H, unsigned u, unsigned k,
                           uint64_t *out, uint64_t *tmp, unsigned *hist, unsigned *s_wave,
                           unsigned *s_misc) {
  const unsigned tid = threadIdx.x;
  // pass 1: high bytes
  for (unsigned b = tid; b < 256; b += 256) hist[b] = 0;
  __syncthreads();
  for (unsigned h = tid; h < H; h += 256)
    if ((h >> 1) != u) atomicAdd(&hist[dr[h] >> 8], 1u);
  __syncthreads();
  if (tid == 0) {
    unsigned cum = 0, b = 0;
    while (cum + hist[b] < k) cum += hist[b++];
    s_misc[0] = b;
    s_misc[1] = cum;
  }
  __syncthreads();
  const unsigned hb = s_misc[0], below = s_misc[1];
  __syncthreads();
  // pass 2: low bytes under the chosen high byte
  hist[tid] = 0;
  __syncthreads();
  for (unsigned h = tid; h < H; h += 256)
    if ((h >> 1) != u && (dr[h] >> 8) == hb) atomicAdd(&hist[dr[h] & 255u], 1u);
  __syncthreads();
  if (tid == 0) {
    unsigned cum = below, b = 0;
    while (cum + hist[b] < k) cum += hist[b++];
    s_misc[2] = (hb << 8) | b;  // threshold distance
    s_misc[3] = k - cum;        // ties at the threshold to take
  }
  __syncthreads();
  const unsigned t = s_misc[2], need = s_misc[3];
  // index-ordered compaction: every d < t, and the first `need` with d == t
  unsigned taken = 0, ties = 0;
  for (unsigned base = 0; base < H; base += 256) {
    const unsigned h = base + tid;
    const bool ok = h < H && (h >> 1) != u;
    const unsigned dv = ok ? dr[h] : 0xFFFFu;
    const bool eq = ok && dv == t;
    unsigned eq_total = 0;
    const unsigned eq_rank = block_scan(eq, s_wave, &eq_total);
    const bool take = (ok && dv < t) || (eq && ties + eq_rank < need);
    unsigned take_total = 0;
    const unsigned pos = block_scan(take, s_wave, &take_total);
    if (take) tmp[taken + pos] = ((uint64_t)dv << 32) | h;
    taken += take_total;
    ties += eq_total;
  }
  __syncthreads();
  // rank sort of the k distinct keys
  for (unsigned i = tid; i < k; i += 256) {
    const uint64_t key = tmp[i];
    unsigned r = 0;
    for (unsigned j = 0; j < k; ++j) r += tmp[j] < key;
    out[r] = key;
  }
  __syncthreads();
}

// One workgroup per sample u in [u0, u0 + nu): rows of d are the batch's
// queries from haplotype q0 = 2 * u0.
__global__ void __launch_bounds__(256)
    k_knn_select(const uint16_t *__restrict__ d, unsigned H, unsigned u0, unsigned kh,
                 unsigned kb, unsigned kout, unsigned *__restrict__ klist) {
  __shared__ uint64_t s_a[kKnnMaxHalf], s_b[2 * kKnnMaxHalf], s_tmp[2 * kKnnMaxHalf];
  __shared__ unsigned s_hist[256], s_wave[4], s_misc[4];
  const unsigned u = u0 + blockIdx.x;
  const uint16_t *da = d + (size_t)(2 * blockIdx.x) * H, *db = da + H;
  knn_select(da, H, u, kh, s_a, s_tmp, s_hist, s_wave, s_misc);
  knn_select(db, H, u, kb, s_b, s_tmp, s_hist, s_wave, s_misc);
  unsigned *row = klist + (size_t)u * kout;
  for (unsigned j = threadIdx.x; j < kh; j += 256) row[2 * j] = (unsigned)s_a[j];
  // B's nearest not already among A's, in B's order
  unsigned placed = 0;
  for (unsigned base = 0; base < kb && placed < kh; base += 256) {
    const unsigned i = base + threadIdx.x;
    bool fresh = false;
    unsigned idx = 0;
    if (i < kb) {
      idx = (unsigned)s_b[i];
      fresh = true;
      for (unsigned j = 0; j < kh && fresh; ++j) fresh = (unsigned)s_a[j] != idx;
    }
    unsigned total = 0;
    const unsigned pos = block_scan(fresh, s_wave, &total);
    if (fresh && placed + pos < kh) row[2 * (placed + pos) + 1] = idx;
    placed += total;
  }
}

}  // namespace glp

using namespace glp;

namespace {

struct KnnLayout {
  size_t mask, counts, pt, dist, total;
};

KnnLayout knn_layout(const glp_knn_desc *d) {
  const size_t wn = (d->num_sites + 63) / 64, H = d->num_haps;
  const size_t batch = std::min<size_t>(kKnnBatch, 2 * (size_t)d->num_samples);
  KnnLayout L;
  L.mask = 0;
  L.counts = (wn * 8 + 255) & ~(size_t)255;
  L.pt = L.counts + ((64 * wn * 4 + 255) & ~(size_t)255);
  L.dist = L.pt + wn * H * 8;
  L.total = L.dist + batch * H * 2;
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
  rc = hip_check(hipGetLastError(), "kNN setup launch");
  const size_t lds = (size_t)wn * kKnnQpb * 8;
  if (rc == GLP_OK)
    rc = hip_check(hipFuncSetAttribute((const void *)&k_knn_dist,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                   "dynamic LDS attribute");
  const unsigned per = kKnnBatch / 2;  // samples per batch
  for (unsigned u0 = 0; u0 < d->num_samples && rc == GLP_OK; u0 += per) {
    const unsigned nu = std::min(per, d->num_samples - u0), nq = 2 * nu;
    hipLaunchKernelGGL(k_knn_dist, dim3((H + 255) / 256, (nq + kKnnQpb - 1) / kKnnQpb), dim3(256),
                       lds, s, pt, H, wn, 2 * u0, nq, dist);
    hipLaunchKernelGGL(k_knn_select, dim3(nu), dim3(256), 0, s, dist, H, u0, kh, kb, kout,
                       d->klist);
    rc = hip_check(hipGetLastError(), "kNN batch launch");
  }
  return rc;
}

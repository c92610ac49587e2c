// glp_vq.hip -- the GL codebook's nearest-codeword search on gfx950
// (GLVQ::AssignPoints / FindGLCode, glVQ.cpp:149-165 and glVQ.hpp:23-34) and
// the device entry points of codebook training and code assignment
// (include/glphase_amd.h).
//
// The search is the O(points x codewords) part of training; the centroid and
// variance sums stay sequential on the host in point order (float and double,
// glVQ.cpp:104-117,149-163), so device-assisted training returns exactly the
// codebook of the host path.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <vector>

#include "../../include/glphase_amd.h"
#include "glp_host.hpp"
#include "glp_internal.hpp"

namespace glp {

// FindGLCode of (RR, Het): the first codeword of least
// sqrt((c.RR - RR)^2 + (c.Het - Het)^2), the differences taken in float and
// the rest in FP64, decided without most of the square roots.  A float
// difference squared is exact in FP64 (48 significant bits), so the sum of
// squares is one FMA: fma(dr, dr, dh * dh) rounds once, like dr*dr + dh*dh.
//  - the squares and the sum are thus the reference's values bit for bit;
//  - sqrt is monotone, so codeword c can only win when its squared sum s is
//    below the best one, bs;
//  - when s < round(bs * (1 - 2^-50)), the exact roots differ by a factor
//    below 1 - 1.7 * 2^-52 and the correctly rounded roots (relative error
//    <= 2^-53 each) are strictly ordered: c wins;
//  - only in the band between (a near tie) are the two rounded roots compared
//    as FindGLCode compares them, so equal roots keep the earlier codeword.
__device__ __forceinline__ double vq_dist2(float2 w, float RR, float Het) {
  const double dr = (double)(w.x - RR), dh = (double)(w.y - Het);
  return fma(dr, dr, dh * dh);
}

// One codeword of the search, near ties decided by the rounded roots.
__device__ __forceinline__ void vq_step(const float2 *cb, unsigned c, float RR, float Het,
                                        double &bs, double &thr, unsigned &best) {
  constexpr double kShrink = 1.0 - 0x1p-50;
  const double s = vq_dist2(cb[c], RR, Het);
  bool win = s < thr;
  if (!win && s < bs) {
    // near tie, rare: the empty asm keeps the compiler from computing both
    // square roots for every codeword and selecting afterwards
    asm volatile("");
    win = sqrt(s) < sqrt(bs);
  }
  if (win) {
    bs = s;
    thr = s * kShrink;
    best = c;
  }
}

// The codewords go in groups of 4 on a branch-free path that takes s < thr
// as the decision and only notes near ties; a group in which a lane of the
// wavefront met one is redone by vq_step from the state at its start.
__device__ __forceinline__ unsigned vq_nearest(const float2 *cb, unsigned n, float RR, float Het) {
  constexpr double kShrink = 1.0 - 0x1p-50;
  double bs = vq_dist2(cb[0], RR, Het), thr = bs * kShrink;
  unsigned best = 0, c = 1;
  for (; c + 4u <= n; c += 4u) {
    const double bs0 = bs;
    const unsigned best0 = best;
    bool tie = false;
#pragma unroll
    for (unsigned k = 0; k < 4u; ++k) {
      // thr of the current best, formed here rather than kept: one multiply
      // instead of a multiply and two selects
      const double s = vq_dist2(cb[c + k], RR, Het), th = bs * kShrink;
      const bool win = s < th;
      tie = tie || (!win && s < bs);
      bs = win ? s : bs;
      best = win ? c + k : best;
    }
    thr = bs * kShrink;
    if (__builtin_expect(__ballot(tie) != 0ull, 0)) {
      bs = bs0;
      thr = bs0 * kShrink;
      best = best0;
      for (unsigned k = 0; k < 4u; ++k) vq_step(cb, c + k, RR, Het, bs, thr, best);
    }
  }
  for (; c < n; ++c) vq_step(cb, c, RR, Het, bs, thr, best);
  return best;
}

// RAW: points are GL trios, normalised as glp_vq_codes does
// (RR, Het) / ((RR + Het) + AA); else (RR, Het) pairs.
template <bool RAW>
__global__ void __launch_bounds__(256)
    k_vq_assign(const float *__restrict__ pts, size_t n, const float2 *__restrict__ g_cb,
                unsigned ncode, uint8_t *__restrict__ out) {
  __shared__ float2 cb[256];
  for (unsigned c = threadIdx.x; c < ncode; c += blockDim.x) cb[c] = g_cb[c];
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
    float RR, Het;
    if (RAW) {
      RR = pts[3 * p];
      Het = pts[3 * p + 1];
      const float sum = RR + Het + pts[3 * p + 2];
      RR /= sum;
      Het /= sum;
    } else {
      RR = pts[2 * p];
      Het = pts[2 * p + 1];
    }
    out[p] = (uint8_t)vq_nearest(cb, ncode, RR, Het);
  }
}

// GLPack::GetPackedGLs / GLIdxs2uint32_t (glPack.cpp:32-192) on the device:
// the codes [n_samples][num_sites] (one byte each) of samples [s0, s0 + ns)
// become words [ceil(num_sites / per)][ns], per = 32 / bits codes a word,
// code j of a word at bit 32 - bits * (j + 1) (MSB first); a site past the
// last one takes the code of GL 0 (glPack.cpp:66-70).  A workgroup packs a
// tile of 64 samples x 256 sites: it reads the tile along the sample rows
// (a wavefront covers 64 consecutive bytes of one row) into LDS, then each
// lane packs one sample's words and a wavefront's stores cover 64
// consecutive samples of one word row.
constexpr unsigned kPackSites = 256, kPackSamples = 64;

// VEC: rows start on 16-byte boundaries (codes 16-byte aligned, num_sites a
// multiple of 16): the tile is read in 16-byte pieces, 16 lanes a row.
template <bool VEC>
__global__ void __launch_bounds__(256)
    k_pack_codes(const uint8_t *__restrict__ codes, unsigned num_sites, unsigned bits, unsigned s0,
                 unsigned ns, uint32_t *__restrict__ out) {
  // rows of 65 dwords: the pack phase's lanes (one row each) read distinct banks
  __shared__ __align__(16) uint32_t tile[kPackSamples][kPackSites / 4 + 1];
  uint8_t *tb = reinterpret_cast<uint8_t *>(&tile[0][0]);
  constexpr unsigned kRowBytes = 4 * (kPackSites / 4 + 1);
  const unsigned site0 = blockIdx.x * kPackSites, samp0 = blockIdx.y * kPackSamples;
  const unsigned t = threadIdx.x;
  const uint8_t pad = codes[0];
  if (VEC) {
    // piece q of the tile: row q / 16, bytes 16 * (q % 16) ..; the row's
    // pieces past num_sites (all or nothing: num_sites % 16 == 0) take pad
    const uint32_t pad4 = 0x01010101u * pad;
    for (unsigned q = t; q < kPackSamples * (kPackSites / 16); q += 256) {
      const unsigned k = q >> 4, site = site0 + 16 * (q & 15u), samp = samp0 + k;
      uint4 v = make_uint4(pad4, pad4, pad4, pad4);
      if (samp < ns && site < num_sites)
        v = *reinterpret_cast<const uint4 *>(codes + (size_t)(s0 + samp) * num_sites + site);
      uint32_t *d = &tile[k][4 * (q & 15u)];
      d[0] = v.x;
      d[1] = v.y;
      d[2] = v.z;
      d[3] = v.w;
    }
  } else {
    const unsigned site = site0 + t;
    for (unsigned k = 0; k < kPackSamples; ++k) {
      const unsigned samp = samp0 + k;
      uint8_t c = pad;
      if (samp < ns && site < num_sites) c = codes[(size_t)(s0 + samp) * num_sites + site];
      tb[k * kRowBytes + t] = c;
    }
  }
  __syncthreads();
  const unsigned per = 32u / bits, words = kPackSites / per;
  const unsigned nwords = (num_sites + per - 1u) / per;
  const unsigned lane = t & 63u, samp = samp0 + lane;
  if (samp >= ns) return;
  for (unsigned w = t >> 6; w < words; w += 4u) {
    const unsigned gw = site0 / per + w;
    if (gw >= nwords) break;
    // the word's per codes are per / 4 dwords of the row (byte j of the
    // run is code j), placed MSB first
    uint32_t word = 0;
    for (unsigned q = 0; q < per / 4u; ++q) {
      const uint32_t d = tile[lane][w * (per / 4u) + q];
#pragma unroll
      for (unsigned b = 0; b < 4u; ++b)
        word |= ((d >> (8u * b)) & 0xFFu) << (32u - bits * (4u * q + b + 1u));
    }
    out[(size_t)gw * ns + samp] = word;
  }
}

int launch_vq_assign(const float *pts_dev, size_t n, bool raw, const float *cb_dev,
                     unsigned ncode, uint8_t *out_dev, hipStream_t s) {
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  if (raw)
    hipLaunchKernelGGL((k_vq_assign<true>), dim3(blocks), dim3(256), 0, s, pts_dev, n,
                       reinterpret_cast<const float2 *>(cb_dev), ncode, out_dev);
  else
    hipLaunchKernelGGL((k_vq_assign<false>), dim3(blocks), dim3(256), 0, s, pts_dev, n,
                       reinterpret_cast<const float2 *>(cb_dev), ncode, out_dev);
  return hip_check(hipGetLastError(), "k_vq_assign launch");
}

}  // namespace glp

using namespace glp;

namespace {

// Device buffers of one training or coding call, released on every path.
struct VqBuffers {
  float *pts = nullptr, *cb = nullptr;
  uint8_t *lab = nullptr;
  ~VqBuffers() {
    if (pts) (void)hipFree(pts);
    if (cb) (void)hipFree(cb);
    if (lab) (void)hipFree(lab);
  }
};

struct HipFailure : std::exception {
  const char *what() const noexcept override { return "HIP call failed"; }
};

void must(int rc) {
  if (rc != GLP_OK) throw HipFailure();
}

}  // namespace

extern "C" int glp_vq_assign_device(const float *points_dev, size_t n, int raw,
                                    const float *codebook, unsigned n_codes, uint8_t *codes_dev,
                                    void *stream) {
  if (!points_dev || !codebook || !codes_dev || n_codes == 0 || n_codes > 256)
    return set_error(GLP_ERR_ARG, "bad argument");
  if (n == 0) return GLP_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  float *cb = nullptr;
  int rc = hip_check(hipMallocAsync((void **)&cb, 8 * (size_t)n_codes, s), "hipMallocAsync");
  if (rc != GLP_OK) return rc;
  rc = hip_check(hipMemcpyAsync(cb, codebook, 8 * (size_t)n_codes, hipMemcpyHostToDevice, s),
                 "codebook copy");
  if (rc == GLP_OK) rc = launch_vq_assign(points_dev, n, raw != 0, cb, n_codes, codes_dev, s);
  (void)hipFreeAsync(cb, s);
  return rc;
}

extern "C" int glp_vq_train_device(const float *gls, size_t n_gl, unsigned n_codes, uint64_t key,
                                   size_t training_size, float *codebook_out, unsigned *iters) {
  if (!gls || !codebook_out || n_gl == 0) return set_error(GLP_ERR_ARG, "null or empty input");
  try {
    VqBuffers b;
    size_t cap = 0;
    GLVQ::AssignFn fn = [&](const std::vector<std::pair<float, float>> &cb, const float *points,
                            size_t n, uint8_t *lab) {
      if (n > cap) {  // the training points are fixed for the whole run
        must(hip_check(hipMalloc((void **)&b.pts, 8 * n), "hipMalloc"));
        must(hip_check(hipMalloc((void **)&b.lab, n), "hipMalloc"));
        must(hip_check(hipMalloc((void **)&b.cb, 8 * 256), "hipMalloc"));
        must(hip_check(hipMemcpy(b.pts, points, 8 * n, hipMemcpyHostToDevice), "points copy"));
        cap = n;
      }
      std::vector<float> flat(2 * cb.size());
      for (size_t c = 0; c < cb.size(); ++c) {
        flat[2 * c] = cb[c].first;
        flat[2 * c + 1] = cb[c].second;
      }
      must(hip_check(hipMemcpy(b.cb, flat.data(), 4 * flat.size(), hipMemcpyHostToDevice),
                     "codebook copy"));
      must(launch_vq_assign(b.pts, n, false, b.cb, (unsigned)cb.size(), b.lab, nullptr));
      must(hip_check(hipMemcpy(lab, b.lab, n, hipMemcpyDeviceToHost), "labels copy"));
    };
    GLVQ vq(gls, n_gl, key, n_codes, training_size, fn);
    const auto cb = vq.GetCodeBook();
    for (size_t c = 0; c < cb.size(); ++c) {
      codebook_out[2 * c] = cb[c].first;
      codebook_out[2 * c + 1] = cb[c].second;
    }
    if (iters) *iters = vq.iterations();
    return GLP_OK;
  } catch (const HipFailure &) {
    return GLP_ERR_HIP;  // the message is already set
  } catch (const std::exception &e) {
    return set_error(GLP_ERR_ARG, e.what());
  }
}

extern "C" int glp_vq_codes_device(const float *gls, size_t n_gl, const float *codebook,
                                   unsigned n_codes, uint8_t *codes_out) {
  if (!gls || !codebook || !codes_out || n_codes == 0 || n_codes > 256)
    return set_error(GLP_ERR_ARG, "bad argument");
  if (n_gl == 0) return GLP_OK;
  try {
    VqBuffers b;
    must(hip_check(hipMalloc((void **)&b.pts, 12 * n_gl), "hipMalloc"));
    must(hip_check(hipMalloc((void **)&b.lab, n_gl), "hipMalloc"));
    must(hip_check(hipMalloc((void **)&b.cb, 8 * (size_t)n_codes), "hipMalloc"));
    must(hip_check(hipMemcpy(b.pts, gls, 12 * n_gl, hipMemcpyHostToDevice), "GL copy"));
    must(hip_check(hipMemcpy(b.cb, codebook, 8 * (size_t)n_codes, hipMemcpyHostToDevice),
                   "codebook copy"));
    must(launch_vq_assign(b.pts, n_gl, true, b.cb, n_codes, b.lab, nullptr));
    must(hip_check(hipMemcpy(codes_out, b.lab, n_gl, hipMemcpyDeviceToHost), "codes copy"));
    return GLP_OK;
  } catch (const HipFailure &) {
    return GLP_ERR_HIP;  // the message is already set
  }
}

extern "C" int glp_pack_codes_device(const uint8_t *codes_dev, unsigned n_samples,
                                     unsigned num_sites, unsigned bits, unsigned s0, unsigned s1,
                                     uint32_t *out_dev, void *stream) {
  if (!codes_dev || !out_dev || num_sites == 0 || s0 >= s1 || s1 > n_samples)
    return set_error(GLP_ERR_ARG, "bad argument");
  if (bits != 1 && bits != 2 && bits != 4 && bits != 8)
    return set_error(GLP_ERR_ARG, "bits must divide 8");
  const unsigned ns = s1 - s0;
  const dim3 grid((num_sites + kPackSites - 1) / kPackSites, (ns + kPackSamples - 1) / kPackSamples);
  const bool vec = (reinterpret_cast<uintptr_t>(codes_dev) & 15u) == 0 && num_sites % 16 == 0;
  if (vec)
    hipLaunchKernelGGL(k_pack_codes<true>, grid, dim3(256), 0, static_cast<hipStream_t>(stream),
                       codes_dev, num_sites, bits, s0, ns, out_dev);
  else
    hipLaunchKernelGGL(k_pack_codes<false>, grid, dim3(256), 0, static_cast<hipStream_t>(stream),
                       codes_dev, num_sites, bits, s0, ns, out_dev);
  return hip_check(hipGetLastError(), "k_pack_codes launch");
}

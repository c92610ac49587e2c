// glphase_amd_hmmlike.hpp -- the HMMLike class of GLPhase (hmmLike.hpp:39-101)
// with the reference's own signatures, running on libglphase_amd.so.
//
// This header replaces src/hmmLike.hpp in a GLPhase tree: Insti keeps
// constructing HMMLike with a gsl_rng& and calling UpdateSampler and
// RunHMMOnSamples (insti.cpp:1643-1644,1697-1698, 1439-1440), and
// hmmLike.cu / hmmLike.cpp leave the build.  It is header-only and expects
// the includer to have declared, as GLPhase's sources do:
//   gsl_rng and unsigned long gsl_rng_get(const gsl_rng *)   (gsl/gsl_rng.h)
//   GLPack with GetNumSamps, GetSampleStride, GetNextSampIdx, GetLastSampIdx,
//     GetNumSites, GetNumBitsPerGL, GetCodeBook, GetPackedGLs (glPack.hpp)
//   Sampler with unsigned SampleHap(unsigned excludeInd)    (sampler.hpp)
// Device errors throw std::runtime_error (the reference exits the process).
#pragma once

#include <stdint.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "glphase_amd.h"

namespace HMMLikeHelper {
// hmmLike.cpp:6-23: SoA [4][N] parent indices -> sample-major [N][4]
inline std::vector<unsigned> transposeHapIdxs(const std::vector<unsigned> &hapIdxs) {
  if (hapIdxs.size() % 4 != 0)
    throw std::runtime_error("transposeHapIdxs: size is not a multiple of 4");
  const size_t n = hapIdxs.size() / 4;
  std::vector<unsigned> out(hapIdxs.size());
  for (size_t s = 0; s < n; ++s)
    for (size_t q = 0; q < 4; ++q) out[q + 4 * s] = hapIdxs[s + q * n];
  return out;
}
}  // namespace HMMLikeHelper

class HMMLike {
 public:
  // hmmLike.hpp:75-78, hmmLike.cpp:32-65: device set-up, tran, mutation
  // matrix and codebook uploaded once; all GLs too when one call covers
  // every sample (sample stride == number of samples)
  HMMLike(const std::vector<uint64_t> &hapPanel, unsigned numHaps, GLPack &glPack,
          unsigned numCycles, const std::vector<float> &tran, const float (*mutationMat)[4][4],
          std::shared_ptr<Sampler> &sampler, gsl_rng &rng, int devID = 0)
      : m_inHapPanel(hapPanel), m_totalNumHaps(numHaps), m_glPack(glPack),
        m_totalNumSamps(glPack.GetNumSamps()), m_numCycles(numCycles), m_sampler(sampler),
        m_rng(rng) {
    if (m_numCycles == 0) throw std::runtime_error("HMMLike: numCycles must be > 0");
    const unsigned bits = m_glPack.GetNumBitsPerGL();
    const auto codeBook = m_glPack.GetCodeBook();
    if (codeBook.size() != (size_t(1) << bits))
      throw std::runtime_error("HMMLike: GLPack must use VQ with a 2^bits codebook");
    check(glp_create(devID, &m_ctx), "HMMLike device set-up");
    try {
      check(glp_set_tran(m_ctx, tran.data(), tran.size()), "CopyTranToDevice");
      float pc[16];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) pc[4 * i + j] = (*mutationMat)[i][j];
      check(glp_set_mutation_mat(m_ctx, pc), "CopyMutationMatToDevice");
      std::vector<float> cb;
      cb.reserve(2 * codeBook.size());
      for (const auto &p : codeBook) {
        cb.push_back(p.first);
        cb.push_back(p.second);
      }
      check(glp_set_codebook(m_ctx, cb.data(), (unsigned)codeBook.size()), "CopyCodeBookToDevice");
      if (m_totalNumSamps == m_glPack.GetSampleStride()) upload(m_glPack.GetPackedGLs());
    } catch (...) {
      glp_destroy(m_ctx);
      throw;
    }
  }
  ~HMMLike() { glp_destroy(m_ctx); }
  HMMLike(const HMMLike &) = delete;
  HMMLike &operator=(const HMMLike &) = delete;

  // hmmLike.hpp:85-91
  void UpdateSampler(std::shared_ptr<Sampler> &sampler) {
    if (sampler)
      m_sampler = sampler;
    else
      throw std::runtime_error("HMMLike::UpdateSampler was called with null pointer");
  }

  // hmmLike.cpp:69-111: the range of samples run and 4 parent indices per
  // sample (sample-major); the current panel is uploaded on every call, so
  // Insti's haps/hnew swap is seen
  std::vector<unsigned> RunHMMOnSamples(unsigned &firstSampIdx, unsigned &lastSampIdx) {
    const unsigned numRunSamps = m_glPack.GetSampleStride();
    if (numRunSamps == m_totalNumSamps) {
      firstSampIdx = 0;
      lastSampIdx = numRunSamps - 1;
    } else {
      firstSampIdx = m_glPack.GetNextSampIdx();
      upload(m_glPack.GetPackedGLs());
      lastSampIdx = m_glPack.GetLastSampIdx();
    }
    std::vector<unsigned> hapIdxs(4 * (size_t)numRunSamps);
    std::vector<unsigned> extraPropHaps((size_t)numRunSamps * m_numCycles);
    unsigned sampIdx = firstSampIdx;
    for (unsigned s = 0; s < numRunSamps; ++s, ++sampIdx) {
      for (unsigned q = 0; q < 4; ++q)
        hapIdxs[s + (size_t)q * numRunSamps] = m_sampler->SampleHap(sampIdx);
      for (unsigned c = 0; c < m_numCycles; ++c)
        extraPropHaps[(size_t)c * numRunSamps + s] = m_sampler->SampleHap(sampIdx);
    }
    const unsigned wn = (m_glPack.GetNumSites() + 63) / 64;
    const size_t words = std::min(m_inHapPanel.size(), (size_t)m_totalNumHaps * wn);
    check(glp_run_hmm_on_device(m_ctx, m_inHapPanel.data(), words, extraPropHaps.data(),
                                m_glPack.GetNumSites(), numRunSamps, m_numCycles,
                                hapIdxs.data(), gsl_rng_get(&m_rng)),
          "RunHMMOnDevice");
    return HMMLikeHelper::transposeHapIdxs(hapIdxs);
  }

  // The MH penalty of the accept test: the reference's kernel fixes 1; its
  // CPU loop anneals pen = min((g + 1) / B, 1)^2 (insti.cpp:1670-1671).
  void SetPen(float pen) { check(glp_set_pen(m_ctx, pen), "SetPen"); }

 private:
  static void check(int rc, const char *what) {
    if (rc != GLP_OK) throw std::runtime_error(std::string(what) + ": " + glp_last_error());
  }
  void upload(const std::vector<uint32_t> &packed) {
    check(glp_set_packed_gls(m_ctx, packed.data(), packed.size(), m_glPack.GetNumBitsPerGL()),
          "CopyPackedGLsToDevice");
  }

  const std::vector<uint64_t> &m_inHapPanel;
  const unsigned m_totalNumHaps;
  GLPack &m_glPack;
  const unsigned m_totalNumSamps;
  const unsigned m_numCycles;
  std::shared_ptr<Sampler> &m_sampler;
  gsl_rng &m_rng;
  glp_ctx *m_ctx = nullptr;
};

// glp_host.hpp -- C++ host classes of the drop-in: the GL codebook and
// packing (GLVQ, GLPack), proposal samplers, the model-parameter builders of
// Insti::initialize, and HMMLike, the object Insti::estimate drives
// (insti.cpp:1630-1698).  Their public interfaces follow the reference
// headers (glVQ.hpp, glPack.hpp, sampler.hpp, hmmLike.hpp); GSL's gsl_rng is
// replaced by glp::Rng, a Philox4x32-10 stream.
#pragma once

#include <stdint.h>

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/glphase_amd.h"

namespace glp {

// A sequential stream over Philox4x32-10 (key = seed, counter {n, stream}),
// standing in for gsl_rng (gsl_rng_get / gsl_rng_uniform /
// gsl_rng_uniform_int).
class Rng {
 public:
  explicit Rng(uint64_t seed = 0, uint32_t stream = 0) : m_seed(seed), m_stream(stream) {}
  void set(uint64_t seed) {
    m_seed = seed;
    m_n = 0;
    m_have = 0;
  }
  uint32_t get();                       // gsl_rng_get: 32 random bits
  double uniform() { return (double)get() * (1.0 / 4294967296.0); }  // [0, 1)
  uint32_t uniform_int(uint32_t n);     // [0, n)

 private:
  uint64_t m_seed;
  uint32_t m_stream, m_n = 0, m_have = 0;
  uint32_t m_buf[4] = {0, 0, 0, 0};
};

// ---- model parameters: insti.cpp:1254-1315 ---------------------------------
// mu = 1 / H_{hn-1}, then mu / (hn + mu) (insti.cpp:1254-1270).
float mutation_rate(unsigned hn);
// pc[i][j] = (1-mu)^2 on the diagonal, mu(1-mu) for one allele flip, mu^2 for
// two (insti.cpp:1307-1315); row-major 16 floats.
void mutation_matrix(float mu, float *pc16);
// tran[3m..3m+2] = ((1-r)^2, r(1-r), r^2) with r = min(r_morgans[m] *
// inflation, 1) (insti.cpp:1277-1300); tran[0..2] = 0.
void transition_matrix(const double *r_morgans, unsigned M, double inflation, float *tran);

// ---- GL vector quantisation: glVQ.hpp / glVQ.cpp ----------------------------
class GLVQ {
 public:
  // Nearest-codeword step of AssignPoints for the n training points of the
  // current codebook: lab[p] = FindGLCode(points[p]).  The default runs on
  // the host's threads; the device entry points pass one that runs on the
  // GPU (the other steps stay sequential on the host, which keeps the
  // reference's summation order).
  using AssignFn = std::function<void(const std::vector<std::pair<float, float>> &cb,
                                      const float *points, size_t n, uint8_t *lab)>;
  // BuildCodeBook (glVQ.cpp:26-146); training_size 0 = every GL (reference).
  GLVQ(const std::vector<float> &inGLs, Rng &rng, unsigned codeBookSize,
       size_t training_size = 0);
  // the same with the Philox key given directly
  GLVQ(const float *gls, size_t numGLs, uint64_t key, unsigned codeBookSize,
       size_t training_size = 0, AssignFn assign = nullptr);
  // FindGLCode (glVQ.cpp:156-165): first index of the smallest Euclidean
  // distance (compared in FP64 after sqrt).
  unsigned char FindGLCode(float RR, float Het) const;
  std::vector<std::pair<float, float>> GetCodeBook() const { return m_codeBook; }
  unsigned iterations() const { return m_iters; }

 private:
  std::vector<std::pair<float, float>> m_codeBook;
  unsigned m_iters = 0;
  void build(const float *gls, size_t numGLs, uint64_t key, unsigned codeBookSize,
             size_t training_size, const AssignFn &assign);
};

unsigned nearest(const std::vector<std::pair<float, float>> &cb, float RR, float Het);
unsigned trio_char(const float *gl);

// ---- GL packing: glPack.hpp / glPack.cpp ------------------------------------
struct GLPackInit {
  const std::vector<float> &inGLs;  // [numSamps][numSites][3]
  Rng &rng;
  unsigned numSamps = 0;
  unsigned sampleStride = 0;
  bool useVQ = false;
  unsigned numBitsPerGL = 8;
  size_t vqTrainingSize = 0;
  GLPackInit(const std::vector<float> &GLs, Rng &r) : inGLs(GLs), rng(r) {}
};

class GLPack {
 public:
  explicit GLPack(const GLPackInit &init);
  // GetPackedGLs (glPack.cpp:32-85): the next sampleStride samples, site-major
  // words, MSB-first codes; wraps to sample 0 after the last sample.
  std::vector<uint32_t> GetPackedGLs();
  unsigned GetSampleStride() const { return m_sampleStride; }
  unsigned GetNextSampIdx() const { return m_nextSampIdx; }
  unsigned GetLastSampIdx() const { return m_lastSampIdx; }
  unsigned GetNumSites() const { return m_numSites; }
  unsigned GetNumSamps() const { return m_numSamps; }
  unsigned GetNumBitsPerGL() const { return m_numBitsPerGL; }
  std::vector<std::pair<float, float>> GetCodeBook() const {
    return m_vq ? m_vq->GetCodeBook() : std::vector<std::pair<float, float>>();
  }
  // codes of every (sample, site), computed once (VQ mode)
  const std::vector<uint8_t> &codes() const { return m_codes; }

 private:
  const std::vector<float> &m_inGLs;
  unsigned m_numSamps, m_sampleStride, m_numSites;
  bool m_useVQ;
  unsigned m_numBitsPerGL, m_numGLsPerWord;
  unsigned m_nextSampIdx = 0, m_lastSampIdx = 0;
  std::unique_ptr<GLVQ> m_vq;
  std::vector<uint8_t> m_codes;
  uint32_t packWord(unsigned samp, unsigned site0) const;
};

// ---- proposal samplers: sampler.hpp ------------------------------------------
class Sampler {
 public:
  Sampler(Rng &rng, unsigned numSamples, unsigned numHaps);
  virtual ~Sampler() {}
  // Sampler::SampleHap(excludeInd, onlyFromRef) (sampler.cpp:12-27)
  unsigned SampleHap(unsigned excludeInd, bool onlyFromRef);
  virtual unsigned SampleHap(unsigned excludeInd) = 0;

 protected:
  Rng &m_rng;
  unsigned m_numSamples, m_numHaps;
};

// UnifSampler (sampler.cpp:29-39)
class UnifSampler : public Sampler {
 public:
  UnifSampler(Rng &rng, unsigned numSamples, unsigned numHaps)
      : Sampler(rng, numSamples, numHaps) {}
  using Sampler::SampleHap;
  unsigned SampleHap(unsigned excludeInd) override;
};

// KNN::SampleHap (kNN.cpp:74-90): uniform over a per-sample neighbour list
class KListSampler : public Sampler {
 public:
  KListSampler(Rng &rng, unsigned numSamples, unsigned numHaps,
               std::vector<std::vector<unsigned>> neighbors);
  using Sampler::SampleHap;
  unsigned SampleHap(unsigned excludeInd) override;

 private:
  std::vector<std::vector<unsigned>> m_neighbors;
};

// ---- HMMLike: hmmLike.hpp:39-101 -------------------------------------------
namespace HMMLikeHelper {
std::vector<unsigned> transposeHapIdxs(const std::vector<unsigned> &hapIdxs);
}

class HMMLike {
 public:
  // hmmLike.hpp:75-78 / hmmLike.cpp:32-65
  HMMLike(const std::vector<uint64_t> &hapPanel, unsigned numHaps, GLPack &glPack,
          unsigned numCycles, const std::vector<float> &tran,
          const float (*mutationMat)[4][4], std::shared_ptr<Sampler> &sampler, Rng &rng,
          int devID = 0);
  ~HMMLike();
  HMMLike(const HMMLike &) = delete;
  HMMLike &operator=(const HMMLike &) = delete;

  // hmmLike.hpp:85-91
  void UpdateSampler(std::shared_ptr<Sampler> &sampler) {
    if (sampler)
      m_sampler = sampler;
    else
      throw std::runtime_error("HMMLike::UpdateSampler was called with null pointer");
  }
  // hmmLike.cpp:69-111: 4 parents per sample, sample-major
  std::vector<unsigned> RunHMMOnSamples(unsigned &firstSampIdx, unsigned &lastSampIdx);
  // MH penalty (the reference kernel fixes 1, its CPU loop anneals)
  void SetPen(float pen);
  // Impute::hmm_like for many (sample, parents) at once (impute.cpp:327-375)
  std::vector<float> HmmLike(const std::vector<unsigned> &samples,
                             const std::vector<unsigned> &parents);

 private:
  const std::vector<uint64_t> &m_inHapPanel;
  const unsigned m_totalNumHaps;
  GLPack &m_glPack;
  const unsigned m_totalNumSamps;
  const unsigned m_numCycles;
  const std::vector<float> &m_tran;
  const float (*m_mutationMat)[4][4];
  std::shared_ptr<Sampler> &m_sampler;
  Rng &m_rng;
  glp_ctx *m_ctx = nullptr;
  unsigned m_firstSampIdx = 0;
};

}  // namespace glp

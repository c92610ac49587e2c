// glp_host.cpp -- C++ host classes of the drop-in (see glp_host.hpp) and the
// extern "C" helpers that expose them (include/glphase_amd.h, section 3).
#include "glp_host.hpp"

#include <algorithm>
#include <cassert>
#include <cfloat>
#include <climits>
#include <cmath>
#include <limits>
#include <string>
#include <thread>

#include "glp_common.hpp"

namespace glp {

// ---------------------------------------------------------------------------
// Rng
// ---------------------------------------------------------------------------
uint32_t Rng::get() {
  if (m_have == 0) {
    const U4 x = philox4x32_10(U4{m_n++, m_stream, 0u, 0u}, (uint32_t)m_seed,
                               (uint32_t)(m_seed >> 32));
    m_buf[0] = x.x;
    m_buf[1] = x.y;
    m_buf[2] = x.z;
    m_buf[3] = x.w;
    m_have = 4;
  }
  return m_buf[4 - m_have--];
}

uint32_t Rng::uniform_int(uint32_t n) { return glp::uniform_int(get(), n); }

// ---------------------------------------------------------------------------
// model parameters (insti.cpp:1254-1315)
// ---------------------------------------------------------------------------
float mutation_rate(unsigned hn) {
  float mu = 0;
  for (unsigned i = 1; i < hn; i++) mu += 1.0 / i;  // float += double, as fast mu
  mu = 1 / mu;
  return mu / (hn + mu);
}

void mutation_matrix(float mu, float *pc) {
  const float same = (1 - mu) * (1 - mu), one = mu * (1 - mu), both = mu * mu;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const int d = i ^ j;
      pc[4 * i + j] = d == 0 ? same : (d == 3 ? both : one);
    }
}

void transition_matrix(const double *r_morgans, unsigned M, double inflation, float *tran) {
  tran[0] = tran[1] = tran[2] = 0;
  for (unsigned m = M - 1; m; m--) {
    float r = (float)r_morgans[m];
    r *= inflation;
    r = r > 1 ? 1 : r;
    tran[m * 3] = (1 - r) * (1 - r);
    tran[m * 3 + 1] = r * (1 - r);
    tran[m * 3 + 2] = r * r;
  }
}

// ---------------------------------------------------------------------------
// GLVQ (glVQ.cpp).  GSL's draws become Philox draws under a 64-bit key taken
// from the caller's Rng, counter {n, tag}: tag 1 = initial random codewords,
// 2 = training subsample, 3 = re-seeding of empty clusters.
// ---------------------------------------------------------------------------
namespace {
U4 vq_draw(uint64_t key, uint32_t n, uint32_t tag) {
  return philox4x32_10(U4{n, tag, 0u, 0u}, (uint32_t)key, (uint32_t)(key >> 32));
}

// EuclidianDist (glVQ.hpp:23-34)
inline double vq_dist(float cr, float ch, float RR, float Het) {
  const double dr = (double)(cr - RR), dh = (double)(ch - Het);
  return std::sqrt(dr * dr + dh * dh);
}
}  // namespace

unsigned nearest(const std::vector<std::pair<float, float>> &cb, float RR, float Het) {
  unsigned best = 0;
  double bd = vq_dist(cb[0].first, cb[0].second, RR, Het);
  for (unsigned c = 1; c < cb.size(); ++c) {
    const double d = vq_dist(cb[c].first, cb[c].second, RR, Het);
    if (d < bd) {
      bd = d;
      best = c;
    }
  }
  return best;
}

template <class F>
static void parallel_for(size_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  if (nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const size_t step = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const size_t b = t * step, e = std::min(n, b + step);
    if (b < e) th.emplace_back(f, b, e);
  }
  for (auto &t : th) t.join();
}

GLVQ::GLVQ(const std::vector<float> &inGLs, Rng &rng, unsigned codeBookSize,
           size_t training_size) {
  if (inGLs.size() % 3 != 0 || inGLs.empty()) throw std::runtime_error("GLVQ: GLs must be trios");
  const uint64_t key = (uint64_t)rng.get() | ((uint64_t)rng.get() << 32);
  build(inGLs.data(), inGLs.size() / 3, key, codeBookSize, training_size, nullptr);
}

GLVQ::GLVQ(const float *gls, size_t numGLs, uint64_t key, unsigned codeBookSize,
           size_t training_size, AssignFn assign) {
  build(gls, numGLs, key, codeBookSize, training_size, assign);
}

void GLVQ::build(const float *inGLs, size_t numGLs, uint64_t key, unsigned codeBookSize,
                 size_t training_size, const AssignFn &device_assign) {
  if (numGLs == 0) throw std::runtime_error("GLVQ: no GLs");
  if (codeBookSize < 4) throw std::runtime_error("GLVQ: codebook size < 4");
  if (codeBookSize - 1 > std::numeric_limits<unsigned char>::max())
    throw std::runtime_error("codebook size is too large: " + std::to_string(codeBookSize));
  // glVQ.cpp:41-51
  m_codeBook.reserve(codeBookSize);
  m_codeBook.push_back(std::make_pair(0.0f, 1.0f));
  m_codeBook.push_back(std::make_pair(0.0f, 0.0f));
  m_codeBook.push_back(std::make_pair(1.0f, 0.0f));
  m_codeBook.push_back(std::make_pair((float)(1.0 / 3), (float)(1.0 / 3)));
  for (unsigned e = 4; e < codeBookSize; ++e) {
    const U4 x = vq_draw(key, e, 1);
    const float RR = (float)u01(x.x);
    const float Het = (float)(u01(x.y) * (1 - RR));
    m_codeBook.push_back(std::make_pair(RR, Het));
  }
  // glVQ.cpp:53-73: training points normalised as (AA + RR) + Het
  const size_t n = training_size && training_size < numGLs ? training_size : numGLs;
  std::vector<float> pts(2 * n);
  for (size_t it = 0; it < n; ++it) {
    size_t g = it;
    if (n < numGLs) g = (size_t)(((uint64_t)vq_draw(key, (uint32_t)it, 2).x * numGLs) >> 32);
    float RR = inGLs[3 * g], Het = inGLs[3 * g + 1];
    const float sum = inGLs[3 * g + 2] + RR + Het;
    RR /= sum;
    Het /= sum;
    pts[2 * it] = RR;
    pts[2 * it + 1] = Het;
  }
  std::vector<uint8_t> lab(n);
  std::vector<double> d2(n);
  // AssignPoints (glVQ.cpp:149-163); the variance is summed in point order
  auto assign = [&]() {
    if (device_assign) device_assign(m_codeBook, pts.data(), n, lab.data());
    parallel_for(n, [&](size_t b, size_t e) {
      for (size_t p = b; p < e; ++p) {
        const unsigned c = device_assign ? lab[p] : nearest(m_codeBook, pts[2 * p], pts[2 * p + 1]);
        lab[p] = (uint8_t)c;
        const double dr = (double)(pts[2 * p] - m_codeBook[c].first);
        const double dh = (double)(pts[2 * p + 1] - m_codeBook[c].second);
        d2[p] = dr * dr + dh * dh;
      }
    });
    double var = 0;
    for (size_t p = 0; p < n; ++p) var += d2[p];
    return var;
  };
  uint32_t reseed = 0;
  // UpdateCodeBook (glVQ.cpp:165-211): float sums in point order
  auto update = [&]() {
    std::vector<std::pair<float, float>> nc(m_codeBook.size(), std::make_pair(0.0f, 0.0f));
    std::vector<size_t> cnt(m_codeBook.size(), 0);
    for (size_t p = 0; p < n; ++p) {
      nc[lab[p]].first += pts[2 * p];
      nc[lab[p]].second += pts[2 * p + 1];
      ++cnt[lab[p]];
    }
    for (size_t c = 0; c < nc.size(); ++c) {
      if (cnt[c] > 0) {
        nc[c].first /= (float)cnt[c];
        nc[c].second /= (float)cnt[c];
      } else {
        const size_t p = (size_t)(((uint64_t)vq_draw(key, reseed++, 3).x * n) >> 32);
        nc[c] = std::make_pair(pts[2 * p], pts[2 * p + 1]);
      }
    }
    m_codeBook = nc;
  };
  double prev = std::numeric_limits<double>::max(), cur = assign();
  while (std::fabs(cur - prev) > cur / 100) {
    if (++m_iters > 1000)
      throw std::runtime_error("[GLVQ] Clustering failed to converge within 1000 iterations");
    prev = cur;
    update();
    cur = assign();
  }
}

unsigned char GLVQ::FindGLCode(float RR, float Het) const {
  return (unsigned char)nearest(m_codeBook, RR, Het);
}

// ---------------------------------------------------------------------------
// GLPack (glPack.cpp)
// ---------------------------------------------------------------------------
GLPack::GLPack(const GLPackInit &init)
    : m_inGLs(init.inGLs), m_numSamps(init.numSamps), m_sampleStride(init.sampleStride),
      m_numSites(init.numSamps ? (unsigned)(init.inGLs.size() / init.numSamps / 3) : 0),
      m_useVQ(init.useVQ), m_numBitsPerGL(init.numBitsPerGL),
      m_numGLsPerWord(init.numBitsPerGL ? 32u / init.numBitsPerGL : 0) {
  if (m_numSamps == 0 || m_sampleStride == 0 || m_sampleStride > m_numSamps)
    throw std::runtime_error("GLPack: bad sample count or stride");
  if (m_inGLs.size() % 3 != 0 || (m_inGLs.size() / 3) % m_numSamps != 0)
    throw std::runtime_error("GLPack: GLs are not numSamps x numSites trios");
  if (m_useVQ) {
    if (m_numBitsPerGL == 0 || m_numBitsPerGL > CHAR_BIT || 8 % m_numBitsPerGL != 0)
      throw std::runtime_error("GLPack: numBitsPerGL must divide 8");
    m_vq.reset(new GLVQ(m_inGLs, init.rng, 1u << m_numBitsPerGL, init.vqTrainingSize));
    // GLs2VQChar (glPack.cpp:195-219): normalise (RR + Het) + AA, FindGLCode
    const size_t ngl = m_inGLs.size() / 3;
    m_codes.resize(ngl);
    parallel_for(ngl, [&](size_t b, size_t e) {
      for (size_t g = b; g < e; ++g) {
        float RR = m_inGLs[3 * g], Het = m_inGLs[3 * g + 1];
        const float sum = RR + Het + m_inGLs[3 * g + 2];
        RR /= sum;
        Het /= sum;
        m_codes[g] = m_vq->FindGLCode(RR, Het);
      }
    });
  } else if (m_numBitsPerGL != 8) {
    throw std::runtime_error("GLPack: numBitsPerGL must be 8 without VQ");
  }
}

// GLTrio2Char / GL2HalfChar (glPack.cpp:221-243)
static unsigned half_char(float gl) {
  gl = std::max(gl - std::numeric_limits<float>::epsilon(), 0.0f);
  return (unsigned)std::floor(gl * 16u);
}
unsigned trio_char(const float *g) {
  const float sum = g[0] + g[1] + g[2];
  return (half_char(g[0] / sum) << 4) | half_char(g[1] / sum);
}

// GLIdxs2uint32_t (glPack.cpp:167-192); a site past the end uses GL index 0
uint32_t GLPack::packWord(unsigned samp, unsigned site0) const {
  uint32_t word = 0;
  for (unsigned j = 0; j < m_numGLsPerWord; ++j) {
    const unsigned m = site0 + j;
    const size_t g = m < m_numSites ? (size_t)samp * m_numSites + m : 0;
    if (m_useVQ)
      word |= (uint32_t)m_codes[g] << (32u - m_numBitsPerGL * (j + 1));
    else
      word ^= trio_char(&m_inGLs[3 * g]) << (24u - 8u * j);
  }
  return word;
}

std::vector<uint32_t> GLPack::GetPackedGLs() {
  m_lastSampIdx = m_nextSampIdx + m_sampleStride <= m_numSamps ? m_nextSampIdx + m_sampleStride - 1
                                                               : m_numSamps - 1;
  const unsigned ns = m_lastSampIdx - m_nextSampIdx + 1;
  const unsigned nw = (m_numSites + m_numGLsPerWord - 1) / m_numGLsPerWord;
  std::vector<uint32_t> out((size_t)nw * ns);
  for (unsigned w = 0; w < nw; ++w)
    for (unsigned s = 0; s < ns; ++s) out[(size_t)w * ns + s] = packWord(m_nextSampIdx + s, w * m_numGLsPerWord);
  m_nextSampIdx = m_lastSampIdx == m_numSamps - 1 ? 0 : m_nextSampIdx + m_sampleStride;
  return out;
}

// ---------------------------------------------------------------------------
// samplers (sampler.cpp, kNN.cpp)
// ---------------------------------------------------------------------------
Sampler::Sampler(Rng &rng, unsigned numSamples, unsigned numHaps)
    : m_rng(rng), m_numSamples(numSamples), m_numHaps(numHaps) {
  if (numSamples * 2 > numHaps) throw std::runtime_error("Sampler: fewer haplotypes than 2 x samples");
}

unsigned Sampler::SampleHap(unsigned excludeInd, bool onlyFromRef) {
  if (!onlyFromRef) return SampleHap(excludeInd);
  if (m_numHaps == 2 * m_numSamples) throw std::runtime_error("Sampler: no reference haplotypes");
  for (;;) {
    const unsigned h = SampleHap(excludeInd);
    if (h >= m_numSamples * 2) return h;
  }
}

unsigned UnifSampler::SampleHap(unsigned excludeInd) {
  if (excludeInd >= m_numSamples) throw std::runtime_error("UnifSampler: sample out of range");
  for (;;) {
    const unsigned h = m_rng.uniform_int(m_numHaps);
    if (h / 2 != excludeInd) return h;
  }
}

KListSampler::KListSampler(Rng &rng, unsigned numSamples, unsigned numHaps,
                           std::vector<std::vector<unsigned>> neighbors)
    : Sampler(rng, numSamples, numHaps), m_neighbors(std::move(neighbors)) {
  if (m_neighbors.size() != numSamples) throw std::runtime_error("KListSampler: one list per sample");
}

unsigned KListSampler::SampleHap(unsigned excludeInd) {
  const auto &l = m_neighbors.at(excludeInd);
  if (l.empty()) throw std::runtime_error("KListSampler: empty neighbour list");
  return l[m_rng.uniform_int((uint32_t)l.size())];
}

// ---------------------------------------------------------------------------
// HMMLike (hmmLike.cpp)
// ---------------------------------------------------------------------------
namespace HMMLikeHelper {
std::vector<unsigned> transposeHapIdxs(const std::vector<unsigned> &hapIdxs) {
  const unsigned hapsPerSamp = 4;
  if (hapIdxs.size() % hapsPerSamp != 0) throw std::runtime_error("transposeHapIdxs: size % 4 != 0");
  const size_t numSamps = hapIdxs.size() / hapsPerSamp;
  std::vector<unsigned> ret(hapIdxs.size());
  for (size_t s = 0; s < numSamps; ++s)
    for (unsigned h = 0; h < hapsPerSamp; ++h) ret[h + s * hapsPerSamp] = hapIdxs[s + h * numSamps];
  return ret;
}
}  // namespace HMMLikeHelper

static void throw_if(int rc, const char *what) {
  if (rc != GLP_OK) throw std::runtime_error(std::string(what) + ": " + glp_last_error());
}

HMMLike::HMMLike(const std::vector<uint64_t> &hapPanel, unsigned numHaps, GLPack &glPack,
                 unsigned numCycles, const std::vector<float> &tran,
                 const float (*mutationMat)[4][4], std::shared_ptr<Sampler> &sampler, Rng &rng,
                 int devID)
    : m_inHapPanel(hapPanel), m_totalNumHaps(numHaps), m_glPack(glPack),
      m_totalNumSamps(glPack.GetNumSamps()), m_numCycles(numCycles), m_tran(tran),
      m_mutationMat(mutationMat), m_sampler(sampler), m_rng(rng) {
  if (m_numCycles == 0) throw std::runtime_error("HMMLike: numCycles must be > 0");
  const unsigned bits = m_glPack.GetNumBitsPerGL();
  if (m_glPack.GetCodeBook().size() != (1u << bits))
    throw std::runtime_error("HMMLike: GLPack must use VQ with a 2^bits codebook");
  throw_if(glp_create(devID, &m_ctx), "HMMLike device setup");
  try {
    throw_if(glp_set_tran(m_ctx, m_tran.data(), m_tran.size()), "CopyTranToDevice");
    float pc[16];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) pc[4 * i + j] = (*m_mutationMat)[i][j];
    throw_if(glp_set_mutation_mat(m_ctx, pc), "CopyMutationMatToDevice");
    std::vector<float> cb;
    for (auto &p : m_glPack.GetCodeBook()) {
      cb.push_back(p.first);
      cb.push_back(p.second);
    }
    throw_if(glp_set_codebook(m_ctx, cb.data(), (unsigned)(cb.size() / 2)), "CopyCodeBookToDevice");
    if (m_totalNumSamps == m_glPack.GetSampleStride()) {
      const std::vector<uint32_t> packed = m_glPack.GetPackedGLs();
      throw_if(glp_set_packed_gls(m_ctx, packed.data(), packed.size(), bits), "CopyPackedGLsToDevice");
    }
  } catch (...) {
    glp_destroy(m_ctx);
    throw;
  }
}

HMMLike::~HMMLike() { glp_destroy(m_ctx); }

void HMMLike::SetPen(float pen) { throw_if(glp_set_pen(m_ctx, pen), "SetPen"); }

std::vector<unsigned> HMMLike::RunHMMOnSamples(unsigned &firstSampIdx, unsigned &lastSampIdx) {
  const unsigned numRunSamps = m_glPack.GetSampleStride();
  if (numRunSamps == m_totalNumSamps) {
    firstSampIdx = 0;
    lastSampIdx = numRunSamps - 1;
  } else {
    firstSampIdx = m_glPack.GetNextSampIdx();
    const std::vector<uint32_t> packed = m_glPack.GetPackedGLs();
    throw_if(glp_set_packed_gls(m_ctx, packed.data(), packed.size(), m_glPack.GetNumBitsPerGL()),
             "CopyPackedGLsToDevice");
    lastSampIdx = m_glPack.GetLastSampIdx();
  }
  const unsigned nrun = lastSampIdx - firstSampIdx + 1;
  // hmmLike.cpp:89-102: 4 initial parents and numCycles candidates per sample
  std::vector<unsigned> hapIdxs(4 * (size_t)nrun);
  std::vector<unsigned> extra((size_t)nrun * m_numCycles);
  unsigned sampIdx = firstSampIdx;
  for (unsigned s = 0; s < nrun; ++s, ++sampIdx) {
    for (unsigned q = 0; q < 4; ++q) hapIdxs[s + (size_t)q * nrun] = m_sampler->SampleHap(sampIdx);
    for (unsigned c = 0; c < m_numCycles; ++c) extra[(size_t)c * nrun + s] = m_sampler->SampleHap(sampIdx);
  }
  const unsigned wn = (m_glPack.GetNumSites() + 63) / 64;
  const size_t words = std::min(m_inHapPanel.size(), (size_t)m_totalNumHaps * wn);
  throw_if(glp_run_hmm_on_device(m_ctx, m_inHapPanel.data(), words, extra.data(),
                                 m_glPack.GetNumSites(), nrun, m_numCycles, hapIdxs.data(),
                                 m_rng.get()),
           "RunHMMOnDevice");
  m_firstSampIdx = firstSampIdx;
  return HMMLikeHelper::transposeHapIdxs(hapIdxs);
}

std::vector<float> HMMLike::HmmLike(const std::vector<unsigned> &samples,
                                    const std::vector<unsigned> &parents) {
  if (parents.size() != 4 * samples.size()) throw std::runtime_error("HmmLike: 4 parents per sample");
  std::vector<float> out(samples.size());
  const unsigned wn = (m_glPack.GetNumSites() + 63) / 64;
  const size_t words = std::min(m_inHapPanel.size(), (size_t)m_totalNumHaps * wn);
  throw_if(glp_hmm_like(m_ctx, m_inHapPanel.data(), words, m_glPack.GetNumSites(),
                        m_glPack.GetSampleStride(), (unsigned)samples.size(), samples.data(),
                        parents.data(), out.data()),
           "hmm_like");
  return out;
}

}  // namespace glp

// ---------------------------------------------------------------------------
// extern "C" helpers
// ---------------------------------------------------------------------------
namespace glp {
int set_error(int code, const char *msg);
}

extern "C" {

int glp_model_params(unsigned hn, unsigned num_sites, const double *r_morgans, double inflation,
                     float *tran_out, float *pc16_out, float *mu_out) {
  if (hn < 2 || num_sites == 0) return glp::set_error(GLP_ERR_ARG, "need hn >= 2 and num_sites > 0");
  const float mu = glp::mutation_rate(hn);
  if (mu_out) *mu_out = mu;
  if (pc16_out) glp::mutation_matrix(mu, pc16_out);
  if (tran_out) {
    if (!r_morgans) return glp::set_error(GLP_ERR_ARG, "tran needs r_morgans");
    glp::transition_matrix(r_morgans, num_sites, inflation, tran_out);
  }
  return GLP_OK;
}

int glp_vq_train(const float *gls, size_t n_gl, unsigned n_codes, uint64_t key,
                 size_t training_size, float *codebook_out, unsigned *iters) {
  if (!gls || !codebook_out || n_gl == 0) return glp::set_error(GLP_ERR_ARG, "null or empty input");
  try {
    glp::GLVQ vq(gls, n_gl, key, n_codes, training_size);
    const auto cb = vq.GetCodeBook();
    for (size_t c = 0; c < cb.size(); ++c) {
      codebook_out[2 * c] = cb[c].first;
      codebook_out[2 * c + 1] = cb[c].second;
    }
    if (iters) *iters = vq.iterations();
    return GLP_OK;
  } catch (const std::exception &e) {
    return glp::set_error(GLP_ERR_ARG, e.what());
  }
}

int glp_vq_codes(const float *gls, size_t n_gl, const float *codebook, unsigned n_codes,
                 uint8_t *codes_out) {
  if (!gls || !codebook || !codes_out || n_codes == 0 || n_codes > 256)
    return glp::set_error(GLP_ERR_ARG, "bad argument");
  std::vector<std::pair<float, float>> cb(n_codes);
  for (unsigned c = 0; c < n_codes; ++c) cb[c] = std::make_pair(codebook[2 * c], codebook[2 * c + 1]);
  glp::parallel_for(n_gl, [&](size_t b, size_t e) {
    for (size_t g = b; g < e; ++g) {
      float RR = gls[3 * g], Het = gls[3 * g + 1];
      const float sum = RR + Het + gls[3 * g + 2];
      RR /= sum;
      Het /= sum;
      codes_out[g] = (uint8_t)glp::nearest(cb, RR, Het);
    }
  });
  return GLP_OK;
}

int glp_pack_codes(const uint8_t *codes, unsigned n_samples, unsigned num_sites, unsigned bits,
                   unsigned s0, unsigned s1, uint32_t *out) {
  if (!codes || !out || num_sites == 0 || s0 >= s1 || s1 > n_samples)
    return glp::set_error(GLP_ERR_ARG, "bad argument");
  if (bits != 1 && bits != 2 && bits != 4 && bits != 8) return glp::set_error(GLP_ERR_ARG, "bits must divide 8");
  const unsigned per = 32 / bits, nw = (num_sites + per - 1) / per, ns = s1 - s0;
  glp::parallel_for(nw, [&](size_t b, size_t e) {
    for (size_t w = b; w < e; ++w)
      for (unsigned i = s0; i < s1; ++i) {
        uint32_t word = 0;
        for (unsigned j = 0; j < per; ++j) {
          const size_t m = w * per + j;
          const uint32_t c = m < num_sites ? codes[(size_t)i * num_sites + m] : codes[0];
          word |= c << (32u - bits * (j + 1));
        }
        out[w * ns + (i - s0)] = word;
      }
  });
  return GLP_OK;
}

int glp_pack_nonvq(const float *gls, unsigned n_samples, unsigned num_sites, unsigned s0,
                   unsigned s1, uint32_t *out) {
  if (!gls || !out || num_sites == 0 || s0 >= s1 || s1 > n_samples)
    return glp::set_error(GLP_ERR_ARG, "bad argument");
  const unsigned nw = (num_sites + 3) / 4, ns = s1 - s0;
  for (unsigned w = 0; w < nw; ++w)
    for (unsigned i = s0; i < s1; ++i) {
      uint32_t word = 0;
      for (unsigned j = 0; j < 4; ++j) {
        const unsigned m = w * 4 + j;
        const float *g = m < num_sites ? gls + 3 * ((size_t)i * num_sites + m) : gls;
        word ^= glp::trio_char(g) << (24u - 8u * j);
      }
      out[(size_t)w * ns + (i - s0)] = word;
    }
  return GLP_OK;
}

}  // extern "C"

// The reference-signature HMMLike (include/glphase_amd_hmmlike.hpp, with a
// gsl_rng&) on the planted-parent cases of cuda_tests.cpp:118-656.  GLPhase's
// gsl_rng, GLPack and Sampler are stood in for by test-only declarations over
// the library's C++ host classes, which have the same interfaces.
#include <cstdio>
#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../glphase_amd/csrc/glp_host.hpp"

// test-only: the two names of gsl/gsl_rng.h that the adapter uses
struct gsl_rng {
  glp::Rng r;
};
inline unsigned long gsl_rng_get(const gsl_rng *g) { return const_cast<gsl_rng *>(g)->r.get(); }
using GLPack = glp::GLPack;
using Sampler = glp::Sampler;

#include "glphase_amd_hmmlike.hpp"

static int g_fail = 0;
#define CHECK(c)                                                                  \
  do {                                                                            \
    if (!(c)) {                                                                   \
      std::fprintf(stderr, "  CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++g_fail;                                                                   \
    }                                                                             \
  } while (0)

#include "planted.hpp"

static void run(const char *name, const std::function<void()> &f) {
  const int before = g_fail;
  std::printf("[ RUN  ] %s\n", name);
  std::fflush(stdout);
  try {
    f();
  } catch (const std::exception &e) {
    std::fprintf(stderr, "  exception: %s\n", e.what());
    ++g_fail;
  }
  std::printf("[ %s ] %s\n", g_fail == before ? " OK " : "FAIL", name);
  std::fflush(stdout);
}

static void planted_two() {
  Planted P;
  glp::Rng rng(114);
  auto GLs = P.gls(2, rng);
  glp::GLPackInit init(GLs, rng);
  init.numSamps = 2;
  init.sampleStride = 2;
  init.useVQ = true;
  init.numBitsPerGL = 4;
  GLPack glPack(init);
  gsl_rng grng{glp::Rng(7)};
  std::shared_ptr<Sampler> sampler = std::make_shared<glp::UnifSampler>(rng, 6, 12);
  HMMLike hmm(P.hapPanel, 12, glPack, 400, P.tran, &P.mutMat, sampler, grng);
  unsigned first = 5, last = 5;
  auto h = hmm.RunHMMOnSamples(first, last);
  CHECK(first == 0 && last == 1 && h.size() == 8);
  check_pairs(h, 0, 5, 6, 10);
  check_pairs(h, 4, 7, 8, 11);
  std::shared_ptr<Sampler> null;
  bool threw = false;
  try {
    hmm.UpdateSampler(null);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  CHECK(threw);
}

static void planted_stride() {
  Planted P;
  glp::Rng rng(111);
  auto GLs = P.gls(4, rng);
  glp::GLPackInit init(GLs, rng);
  init.numSamps = 4;
  init.sampleStride = 2;
  init.useVQ = true;
  init.numBitsPerGL = 4;
  GLPack glPack(init);
  gsl_rng grng{glp::Rng(8)};
  std::shared_ptr<Sampler> sampler = std::make_shared<glp::UnifSampler>(rng, 6, 12);
  HMMLike hmm(P.hapPanel, 12, glPack, 400, P.tran, &P.mutMat, sampler, grng);
  hmm.SetPen(1.0f);
  unsigned first = 0, last = 0;
  auto h = hmm.RunHMMOnSamples(first, last);
  CHECK(first == 0 && last == 1 && h.size() == 8);
  check_pairs(h, 0, 5, 6, 10);
  check_pairs(h, 4, 7, 8, 11);
  h = hmm.RunHMMOnSamples(first, last);
  CHECK(first == 2 && last == 3 && h.size() == 8);
  check_pairs(h, 0, 5, 6, 11);
  check_pairs(h, 4, 7, 8, 10);
}

static void transpose() {
  std::vector<unsigned> soa{0, 1, 2, 10, 11, 12, 20, 21, 22, 30, 31, 32};
  CHECK(HMMLikeHelper::transposeHapIdxs(soa) ==
        (std::vector<unsigned>{0, 10, 20, 30, 1, 11, 21, 31, 2, 12, 22, 32}));
}

int main() {
  run("HMMLikeHelper.transposeHapIdxs (gsl adapter)", transpose);
  run("HMMLike.plantedParents (gsl adapter)", planted_two);
  run("HMMLike.plantedParentsStride (gsl adapter)", planted_stride);
  std::printf("%s (%d failed checks)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail ? 1 : 0;
}

// C++ drop-in tests: the reference's gtest cases for HMMLike, GLPack and the
// samplers (src/testing/cuda_tests.cpp, glPack_unittests.cpp) restated
// against the C++ classes of libglphase_amd.so.  `--cpu` runs the cases that
// need no GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../glphase_amd/csrc/glp_host.hpp"

using namespace glp;

static int g_fail = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "  CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++g_fail;                                                         \
    }                                                                   \
  } while (0)

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

// glPack_unittests.cpp:27-66: stride 3 over 5 samples, wrap to 0
static void pack_stride() {
  std::vector<float> GLs{0, 1.0 / 16, 15.0 / 16, 2.0 / 16, 3.0 / 16, 11.0 / 16,
                         4.0 / 16, 5.0 / 16, 7.0 / 16, 0, 0, 1, 0, 0, 1};
  Rng rng(1);
  GLPackInit init(GLs, rng);
  init.numSamps = 5;
  init.sampleStride = 3;
  GLPack glPack(init);
  for (int i = 0; i < 2; ++i) {
    CHECK(glPack.GetNextSampIdx() == 0);
    auto p = glPack.GetPackedGLs();
    CHECK(glPack.GetLastSampIdx() == 2);
    CHECK(p.size() == 3);
    CHECK(glPack.GetNextSampIdx() == 3);
    CHECK((p[0] >> 24) == 0);
    CHECK((p[1] >> 24) == 18);
    CHECK((p[2] >> 24) == 52);
    auto p2 = glPack.GetPackedGLs();
    CHECK(glPack.GetLastSampIdx() == 4);
    CHECK(p2.size() == 2);
  }
}

// glPack_unittests.cpp:220-250: VQ word count
static void pack_vq_size() {
  const unsigned numSamps = 5, numSites = 1024;
  std::vector<float> GLs(numSamps * numSites * 3, 0);
  for (unsigned s = 0; s < numSamps; ++s)
    for (unsigned i = 0; i < numSites; ++i) {
      GLs[3 * i + numSites * 3 * s] = 1;
      GLs[3 * i + numSites * 3 * s + 1] = (float)((i + s) % 7) / 10;
      GLs[3 * i + numSites * 3 * s + 2] = 0.5f;
    }
  Rng rng(111);
  GLPackInit init(GLs, rng);
  init.numSamps = numSamps;
  init.sampleStride = numSamps;
  init.useVQ = true;
  init.numBitsPerGL = 4;
  GLPack glPack(init);
  CHECK(glPack.GetCodeBook().size() == 16);
  for (int i = 0; i < 2; ++i) {
    CHECK(glPack.GetNextSampIdx() == 0);
    CHECK(glPack.GetPackedGLs().size() == numSamps * numSites * 4 / 32);
  }
}

static void transpose() {
  std::vector<unsigned> soa{0, 1, 2, 10, 11, 12, 20, 21, 22, 30, 31, 32};
  auto aos = HMMLikeHelper::transposeHapIdxs(soa);
  std::vector<unsigned> want{0, 10, 20, 30, 1, 11, 21, 31, 2, 12, 22, 32};
  CHECK(aos == want);
}

static void samplers() {
  Rng rng(3);
  UnifSampler s(rng, 6, 12);
  unsigned mx = 0;
  for (int i = 0; i < 500; ++i) {
    const unsigned h = s.SampleHap(0);
    CHECK(h >= 2 && h < 12);
    mx = h > mx ? h : mx;
  }
  CHECK(mx == 11);  // cuda_tests.cpp:304-310
  UnifSampler withRef(rng, 4, 12);
  for (int i = 0; i < 100; ++i) CHECK(withRef.SampleHap(1, true) >= 8);
}

#include "planted.hpp"

static void planted_two() {
  Planted P;
  Rng rng(114);
  auto GLs = P.gls(2, rng);
  GLPackInit init(GLs, rng);
  init.numSamps = 2;
  init.sampleStride = 2;
  init.useVQ = true;
  init.numBitsPerGL = 4;
  GLPack glPack(init);
  std::shared_ptr<Sampler> sampler = std::make_shared<UnifSampler>(rng, 6, 12);
  HMMLike hmm(P.hapPanel, 12, glPack, 400, P.tran, &P.mutMat, sampler, rng);
  unsigned first = 0, last = 0;
  auto h = hmm.RunHMMOnSamples(first, last);
  CHECK(first == 0 && last == 1 && h.size() == 8);
  check_pairs(h, 0, 5, 6, 10);
  check_pairs(h, 4, 7, 8, 11);
  // bad < good (cuda_tests.cpp:393-429)
  auto like = hmm.HmmLike({0, 0}, {2, 2, 2, 2, 5, 10, 6, 10});
  CHECK(like[1] > like[0]);
  CHECK(like[0] <= 0 && like[1] <= 0);
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
  // four samples packed two at a time (cuda_tests.cpp:500-653)
  Planted P;
  Rng rng(111);
  auto GLs = P.gls(4, rng);
  GLPackInit init(GLs, rng);
  init.numSamps = 4;
  init.sampleStride = 2;
  init.useVQ = true;
  init.numBitsPerGL = 4;
  GLPack glPack(init);
  std::shared_ptr<Sampler> sampler = std::make_shared<UnifSampler>(rng, 6, 12);
  HMMLike hmm(P.hapPanel, 12, glPack, 400, P.tran, &P.mutMat, sampler, rng);
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

int main(int argc, char **argv) {
  const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu") == 0;
  run("GLPack.packGLsOK", pack_stride);
  run("GLPack.packGLsVQmultisite", pack_vq_size);
  run("HMMLikeHelper.transposeHapIdxs", transpose);
  run("Sampler.uniformAndRef", samplers);
  if (!cpu_only) {
    run("HMMLike.plantedParents", planted_two);
    run("HMMLike.plantedParentsStride", planted_stride);
  }
  std::printf("%s (%d failed checks)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail ? 1 : 0;
}

// Planted-parent data of the reference's HMMLike tests (cuda_tests.cpp:118-656),
// shared by the C++ drop-in test programs.  Needs CHECK and glp::Rng.
#pragma once

struct Planted {
  static const unsigned numSites = 1024, numWords = 16, bigNumHaps = 12, wordSize = 64;
  std::vector<uint64_t> hapPanel;
  std::vector<float> tran;
  float mutMat[4][4];
  Planted() : hapPanel(numWords * bigNumHaps, 0), tran(numSites * 3, 0) {
    const uint64_t ones = ~(uint64_t)0;
    hapPanel.at(5 * numWords) = hapPanel.at(6 * numWords) = ones;
    hapPanel.at(5 * numWords + 5) = hapPanel.at(6 * numWords + 5) = ones;
    hapPanel.at(7 * numWords + 1) = hapPanel.at(8 * numWords + 1) = ones;
    hapPanel.at(7 * numWords + 7) = hapPanel.at(8 * numWords + 7) = ones;
    hapPanel.at(10 * numWords + 2) = ones;
    hapPanel.at(11 * numWords + 3) = ones;
    float mu = 0;
    for (unsigned i = 1; i < bigNumHaps; i++) mu += 1.0 / i;
    mu = 1 / mu;
    const float rho = 0.5 * mu * (numSites - 1) / (numSites * 1000.0) / 1;
    mu = mu / (bigNumHaps + mu);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        const int d = i ^ j;
        mutMat[i][j] = d == 0 ? (1 - mu) * (1 - mu) : (d == 3 ? mu * mu : mu * (1 - mu));
      }
    for (unsigned m = numSites - 1; m; m--) {
      const float rt = rho * 1000;
      const float r = rt / (rt + 4);
      tran[m * 3] = (1 - r) * (1 - r);
      tran[m * 3 + 1] = r * (1 - r);
      tran[m * 3 + 2] = r * r;
    }
  }
  std::vector<float> gls(unsigned numSamps, glp::Rng &rng) {
    std::vector<float> GLs(3 * numSites * numSamps);
    for (auto &g : GLs) g = (float)rng.uniform();
    const unsigned alt[4][3] = {{0, 5, 2}, {1, 7, 3}, {0, 3, 5}, {1, 2, 7}};
    for (unsigned s = 0; s < numSamps; ++s)
      for (unsigned k = 0; k < 3; ++k)
        for (unsigned i = 0; i < wordSize; ++i)
          GLs.at(3 * ((size_t)s * numSites + alt[s][k] * wordSize + i) + 2) = 100;
    return GLs;
  }
};

static void check_pairs(const std::vector<unsigned> &h, unsigned base, unsigned lo1, unsigned lo2,
                        unsigned other) {
  for (unsigned i = 0; i < 4; i += 2) {
    unsigned a = h[base + i], b = h[base + i + 1];
    if (a > b) std::swap(a, b);
    CHECK(b == other);
    CHECK(a >= lo1 && a <= lo2);
  }
}


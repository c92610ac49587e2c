/*
 * oracle.c -- CPU restatement of GLPhase's MH surrogate-parent update.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  Compiled by oracle/Makefile with -ffp-contract=off.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, from Salmon, Moraes, Dror & Shaw, "Parallel random numbers: */
/* as easy as 1, 2, 3", SC'11 (the Random123 family).  Replaces curand XORWOW */
/* (hmmLike.cu:220-226) and the GSL streams (insti.cpp:1506, MHSampler.hpp:92) */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += PHILOX_W0;
      k1 += PHILOX_W1;
    }
    uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
    uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void orc_draw(uint64_t seed, uint32_t k, uint32_t sample, uint32_t window,
              uint32_t gen, uint32_t out[4]) {
  const uint32_t ctr[4] = {k, sample, window, gen};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox(ctr, key, out);
}

uint32_t orc_uniform_int(uint32_t x, uint32_t n) {
  return (uint32_t)(((uint64_t)x * n) >> 32);
}

/* UnifSampler::SampleHap (sampler.cpp:29-39) draws uniformly over [0, H)
 * rejecting h/2 == sample; the same distribution in closed form. */
uint32_t orc_unif_excl(uint32_t x, uint32_t nhaps, uint32_t sample) {
  uint32_t r = orc_uniform_int(x, nhaps - 2);
  return r + (r >= 2 * sample ? 2u : 0u);
}

static double orc_u01(uint32_t x) { return (double)x * (1.0 / 4294967296.0); }

/* ------------------------------------------------------------------------ */
/* decode + emission                                                        */
/* ------------------------------------------------------------------------ */

/* impute.cpp:11-13 and hmmLike.cu:233 */
float orc_norm(void) { return powf(FLT_MIN, 2.0f / 3.0f); }

/* hmmLike.cu:12-28: code j of a word is MSB-first; GL2 = max(1-GL0-GL1, 0) */
void orc_unpack_gl(uint32_t word, const float *codebook, unsigned bits,
                   unsigned slot, float gl[3]) {
  const uint32_t mask = (1u << bits) - 1u;
  const uint32_t code = (word >> (32u - bits * slot - bits)) & mask;
  gl[0] = codebook[2 * code];
  gl[1] = codebook[2 * code + 1];
  float g2 = 1 - gl[0] - gl[1];
  gl[2] = g2 > 0.0f ? g2 : 0.0f;
}

/* hmmLike.cu:39-44 (FillEmit with mutationMat[i + 4j] = pc[i][j],
 * hmmLike.cu:341-344); the het GL is counted twice (insti.cpp:1376-1378). */
void orc_emit_lut(const float *codebook, unsigned ncode, const float *pc,
                  float *lut) {
  for (unsigned c = 0; c < ncode; ++c) {
    float gl[3];
    gl[0] = codebook[2 * c];
    gl[1] = codebook[2 * c + 1];
    float g2 = 1 - gl[0] - gl[1];
    gl[2] = g2 > 0.0f ? g2 : 0.0f;
    for (unsigned i = 0; i < 4; ++i)
      lut[4 * c + i] = pc[4 * i + 0] * gl[0] + pc[4 * i + 1] * gl[1] +
                       pc[4 * i + 2] * gl[1] + pc[4 * i + 3] * gl[2];
  }
}

/* hmmLike.cu:100-104: word (site / k) of sample i sits at [site/k][N] */
void orc_sample_emit(const uint32_t *packed, unsigned nsamp_packed,
                     unsigned sample, unsigned M, unsigned bits,
                     const float *lut, float *emit) {
  const unsigned k = 32u / bits;
  const uint32_t mask = (1u << bits) - 1u;
  for (unsigned m = 0; m < M; ++m) {
    uint32_t w = packed[(size_t)(m / k) * nsamp_packed + sample];
    unsigned slot = m % k;
    uint32_t code = (w >> (32u - bits * slot - bits)) & mask;
    memcpy(emit + 4 * (size_t)m, lut + 4 * (size_t)code, 4 * sizeof(float));
  }
}

/* ------------------------------------------------------------------------ */
/* forward pass                                                             */
/* ------------------------------------------------------------------------ */

/* impute.hpp:64-66 / hmmLike.cu:47-49: site s of a hap is bit s&63 of word s>>6 */
static inline unsigned hbit(const uint64_t *h, unsigned s) {
  return (unsigned)((h[s >> 6] >> (s & 63)) & 1u);
}

/* The reference calls logf (impute.cpp:349,373).  glibc's logf is not
 * correctly rounded (it differs by 1 ulp on ~0.03% of inputs), so the oracle
 * fixes logf as the correctly rounded natural log: (float)log((double)x),
 * except for the 20 positive floats whose log lies so close to a float
 * rounding midpoint that a 2-ulp change of the double log flips the float
 * (found by exhaustive search; values from an 80-digit decimal log). */
static const uint32_t orc_logf_hard[20][2] = {
    {0x0dc8bba4u, 0xc289bac4u}, {0x111c87f8u, 0xc28085dfu},
    {0x1a8446cbu, 0xc24d0a45u}, {0x1f116ab8u, 0xc233b53eu},
    {0x29fd22f8u, 0xc1ee8859u}, {0x2c4c24b7u, 0xc1d48710u},
    {0x38dcbe38u, 0xc1128ba4u}, {0x3bf86ef0u, 0xc09c399eu},
    {0x3c413d3au, 0xc08e158fu}, {0x41178febu, 0x400fe5e7u},
    {0x464d5b2bu, 0x4117bcf1u}, {0x4665a9a6u, 0x41198725u},
    {0x4c5d65a5u, 0x418f034bu}, {0x4d604ebeu, 0x419a352cu},
    {0x4e85f412u, 0x41a6b811u}, {0x5ee8984eu, 0x422e4a21u},
    {0x65d890d3u, 0x4254d1f9u}, {0x66a8c860u, 0x42595e46u},
    {0x6f31a8ecu, 0x42845a89u}, {0x79e7ec37u, 0x42a1ffb7u}};

#define ORC_LN2 0.6931471805599453

float orc_logf_cr(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  for (int i = 0; i < 20; ++i)
    if (orc_logf_hard[i][0] == b) {
      float r;
      memcpy(&r, &orc_logf_hard[i][1], 4);
      return r;
    }
  return (float)log((double)x);
}

/* Impute::hmm_like, impute.cpp:327-375, in its own operation order. */
float orc_hmm_like_f32(const uint64_t *haps, unsigned wn, const unsigned P[4],
                       const float *emit, const float *tran, unsigned M) {
  const uint64_t *f0 = haps + (size_t)P[0] * wn, *f1 = haps + (size_t)P[1] * wn,
                 *m0 = haps + (size_t)P[2] * wn, *m1 = haps + (size_t)P[3] * wn;
  const float norm = orc_norm();
  const float *e = emit, *t = tran;
  float sum, score = 0;
  float l00 = 0.25f * e[(hbit(f0, 0) << 1) | hbit(m0, 0)],
        l01 = 0.25f * e[(hbit(f0, 0) << 1) | hbit(m1, 0)];
  float l10 = 0.25f * e[(hbit(f1, 0) << 1) | hbit(m0, 0)],
        l11 = 0.25f * e[(hbit(f1, 0) << 1) | hbit(m1, 0)];
  float b00, b01, b10, b11;
  e += 4;
  t += 3;
  for (unsigned m = 1; m < M; m++, e += 4, t += 3) {
    b00 = l00 * t[0] + (l01 + l10) * t[1] + l11 * t[2];
    b01 = l01 * t[0] + (l00 + l11) * t[1] + l10 * t[2];
    b10 = l10 * t[0] + (l00 + l11) * t[1] + l01 * t[2];
    b11 = l11 * t[0] + (l01 + l10) * t[1] + l00 * t[2];
    l00 = b00 * e[(hbit(f0, m) << 1) | hbit(m0, m)];
    l01 = b01 * e[(hbit(f0, m) << 1) | hbit(m1, m)];
    l10 = b10 * e[(hbit(f1, m) << 1) | hbit(m0, m)];
    l11 = b11 * e[(hbit(f1, m) << 1) | hbit(m1, m)];
    if ((sum = l00 + l01 + l10 + l11) < norm) {
      sum = 1.0f / sum;
      score -= orc_logf_cr(sum);
      l00 *= sum;
      l01 *= sum;
      l10 *= sum;
      l11 *= sum;
    }
  }
  return score + orc_logf_cr(l00 + l01 + l10 + l11);
}

/* The reference's FP32 recurrence (impute.cpp:354-366), with each state's
 * products l * t0 and l' * t2 fused with the (pair sum) * t1 term into two
 * FMAs, and without the per-site rescale test (impute.cpp:367-372). The
 * pass starts from 1/4 * 2^100 and, after every 8th site (m = 7, 15, 23,
 * ...; every = 8) or after every site including site 0 (every = 1), the
 * four probabilities are scaled by 2^k, where e is the frexp exponent of
 * (l00 + l11) + (l01 + l10) and k = 101 - e brings that sum into
 * [2^100, 2^101); the scaling rounds nothing while the values stay normal.
 * The score is log(last sum) - (sum of k, from 100) ln 2 in FP64: the
 * reference's score -= log(factor) sum (impute.cpp:367-374) with every
 * factor exact and the sum of logs carried as an integer.  *tiny is set
 * when a scaling had k >= 200 (the sum had fallen below 2^-99) or the last
 * sum is below 2^-99 or not finite. */
double orc_hmm_like_mixed_pass(const uint64_t *haps, unsigned wn, const unsigned P[4],
                               const float *emit, const float *tran, unsigned M,
                               unsigned every, int *tiny) {
  const uint64_t *f0 = haps + (size_t)P[0] * wn, *f1 = haps + (size_t)P[1] * wn,
                 *m0 = haps + (size_t)P[2] * wn, *m1 = haps + (size_t)P[3] * wn;
  const float *e = emit, *t = tran;
  long ksum = 100;
  int ex;
  float pe, po;
  float l00 = 0x1p98f * e[(hbit(f0, 0) << 1) | hbit(m0, 0)],
        l01 = 0x1p98f * e[(hbit(f0, 0) << 1) | hbit(m1, 0)];
  float l10 = 0x1p98f * e[(hbit(f1, 0) << 1) | hbit(m0, 0)],
        l11 = 0x1p98f * e[(hbit(f1, 0) << 1) | hbit(m1, 0)];
  float b00, b01, b10, b11;
  for (unsigned m = 0; m < M; m++, e += 4, t += 3) {
    if (m) {
      pe = (l00 + l11) * t[1];
      po = (l01 + l10) * t[1];
      b00 = fmaf(l11, t[2], fmaf(l00, t[0], po));
      b01 = fmaf(l10, t[2], fmaf(l01, t[0], pe));
      b10 = fmaf(l01, t[2], fmaf(l10, t[0], pe));
      b11 = fmaf(l00, t[2], fmaf(l11, t[0], po));
      l00 = b00 * e[(hbit(f0, m) << 1) | hbit(m0, m)];
      l01 = b01 * e[(hbit(f0, m) << 1) | hbit(m1, m)];
      l10 = b10 * e[(hbit(f1, m) << 1) | hbit(m0, m)];
      l11 = b11 * e[(hbit(f1, m) << 1) | hbit(m1, m)];
    }
    if (every == 1 || (m & 7u) == 7u) {
      (void)frexpf((l00 + l11) + (l01 + l10), &ex);
      const int k = 101 - ex;
      if (k >= 200)
        *tiny = 1;
      l00 = ldexpf(l00, k);
      l01 = ldexpf(l01, k);
      l10 = ldexpf(l10, k);
      l11 = ldexpf(l11, k);
      ksum += k;
    }
  }
  const float last = l00 + l01 + l10 + l11;
  if (!(last >= 0x1p-99f && last <= FLT_MAX))
    *tiny = 1;
  return log((double)last) - (double)ksum * ORC_LN2;
}

/* Mixed precision: the pass scaled every 8th site; when its sum fell below
 * 2^-99 (a run of sites whose emissions are all near mu^2), its last sum is
 * not finite, or it is NaN, the pass scaled after every site. */
double orc_hmm_like_mixed(const uint64_t *haps, unsigned wn,
                          const unsigned P[4], const float *emit,
                          const float *tran, unsigned M) {
  int tiny = 0;
  const double v = orc_hmm_like_mixed_pass(haps, wn, P, emit, tran, M, 8, &tiny);
  if (!tiny && v == v)
    return v;
  return orc_hmm_like_mixed_pass(haps, wn, P, emit, tran, M, 1, &tiny);
}

/* Mixed precision, forward-backward form.  For any split, the likelihood of
 * the model of impute.cpp:327-375 is l_{n-1}^T T_n beta_n, where l is the
 * forward variable, beta_{m-1} = T_m (e_m o beta_m) the backward one
 * (beta_{M-1} = 1) and T_m the symmetric 4-state switch matrix of tran row m
 * (hmmLike.h:18-20).  The emission-weighted backward variable u_m = e_m o
 * beta_m follows u_{m-1} = e_{m-1} o (T_m u_m): the forward step's shape,
 * transition then emission.  So two halves of n = ceil(M/2) steps each run
 * the mixed pass's step (orc_hmm_like_mixed_pass): the forward half over
 * sites 0 .. n-1 from 1/4, the backward half over sites M-1 down to n from
 * 1 (step j at site M-1+p-j with tran row M+p-j, p = M mod 2; the first
 * step has tran row (1, 0, 0), and for odd M its emission is 1, an exact
 * identity step).  Each half starts scaled by 2^100 and scales by powers of
 * two after every 8th of its own steps (every = 8) or after every step
 * (every = 1), as the mixed pass does.  The halves join in FP64:
 *   g = T_n u_n   (tran row n, or the identity row when n = M, i.e. M = 1),
 *   L = (l00 g00 + l11 g11) + (l01 g01 + l10 g10),
 * and the score is log(L) - (kF + kB) ln 2.  *tiny is set when either half
 * had a scaling with k >= 200 or a last sum below 2^-99 or not finite. */
static void fb_half(const uint64_t *const par[4], const float *emit, const float *tran,
                    unsigned M, int backward, unsigned every, float x[4], long *ksum,
                    int *tiny) {
  const unsigned n = (M + 1) / 2, p = M & 1u;
  float l00, l01, l10, l11;
  l00 = l01 = l10 = l11 = backward ? 0x1p100f : 0x1p98f;
  long ks = 100;
  int ex;
  for (unsigned j = 0; j < n; ++j) {
    const unsigned s = backward ? M - 1 + p - j : j;
    float t0 = 1.0f, t1 = 0.0f, t2 = 0.0f;
    if (j > 0 && !(backward && j <= p)) {
      const float *t = tran + 3 * (size_t)(backward ? s + 1 : s);
      t0 = t[0];
      t1 = t[1];
      t2 = t[2];
    }
    float e00 = 1.0f, e01 = 1.0f, e10 = 1.0f, e11 = 1.0f;
    if (!(backward && p && j == 0)) {
      const float *e = emit + 4 * (size_t)s;
      e00 = e[(hbit(par[0], s) << 1) | hbit(par[2], s)];
      e01 = e[(hbit(par[0], s) << 1) | hbit(par[3], s)];
      e10 = e[(hbit(par[1], s) << 1) | hbit(par[2], s)];
      e11 = e[(hbit(par[1], s) << 1) | hbit(par[3], s)];
    }
    const float pe = (l00 + l11) * t1, po = (l01 + l10) * t1;
    const float b00 = fmaf(l11, t2, fmaf(l00, t0, po));
    const float b01 = fmaf(l10, t2, fmaf(l01, t0, pe));
    const float b10 = fmaf(l01, t2, fmaf(l10, t0, pe));
    const float b11 = fmaf(l00, t2, fmaf(l11, t0, po));
    l00 = b00 * e00;
    l01 = b01 * e01;
    l10 = b10 * e10;
    l11 = b11 * e11;
    if (every == 1 || (j & 7u) == 7u) {
      (void)frexpf((l00 + l11) + (l01 + l10), &ex);
      const int k = 101 - ex;
      if (k >= 200)
        *tiny = 1;
      l00 = ldexpf(l00, k);
      l01 = ldexpf(l01, k);
      l10 = ldexpf(l10, k);
      l11 = ldexpf(l11, k);
      ks += k;
    }
  }
  const float last = l00 + l01 + l10 + l11;
  if (!(last >= 0x1p-99f && last <= FLT_MAX))
    *tiny = 1;
  x[0] = l00;
  x[1] = l01;
  x[2] = l10;
  x[3] = l11;
  *ksum = ks;
}

double orc_hmm_like_fb_pass(const uint64_t *haps, unsigned wn, const unsigned P[4],
                            const float *emit, const float *tran, unsigned M,
                            unsigned every, int *tiny) {
  const uint64_t *const par[4] = {haps + (size_t)P[0] * wn, haps + (size_t)P[1] * wn,
                                  haps + (size_t)P[2] * wn, haps + (size_t)P[3] * wn};
  const unsigned n = (M + 1) / 2;
  float l[4], u[4];
  long kf = 0, kb = 0;
  fb_half(par, emit, tran, M, 0, every, l, &kf, tiny);
  fb_half(par, emit, tran, M, 1, every, u, &kb, tiny);
  double t0 = 1.0, t1 = 0.0, t2 = 0.0;
  if (n < M) {
    t0 = tran[3 * (size_t)n];
    t1 = tran[3 * (size_t)n + 1];
    t2 = tran[3 * (size_t)n + 2];
  }
  const double u00 = u[0], u01 = u[1], u10 = u[2], u11 = u[3];
  const double se = u00 + u11, so = u01 + u10;
  const double g00 = (t0 * u00 + t1 * so) + t2 * u11;
  const double g11 = (t0 * u11 + t1 * so) + t2 * u00;
  const double g01 = (t0 * u01 + t1 * se) + t2 * u10;
  const double g10 = (t0 * u10 + t1 * se) + t2 * u01;
  const double de = (double)l[0] * g00 + (double)l[3] * g11;
  const double dd = (double)l[1] * g01 + (double)l[2] * g10;
  return log(de + dd) - (double)(kf + kb) * ORC_LN2;
}

/* The forward-backward form: both halves scaled every 8th step; when either
 * flagged a sum below 2^-99 or a last sum not finite, or the result is NaN,
 * both scaled after every step. */
double orc_hmm_like_fb(const uint64_t *haps, unsigned wn, const unsigned P[4],
                       const float *emit, const float *tran, unsigned M) {
  int tiny = 0;
  const double v = orc_hmm_like_fb_pass(haps, wn, P, emit, tran, M, 8, &tiny);
  if (!tiny && v == v)
    return v;
  return orc_hmm_like_fb_pass(haps, wn, P, emit, tran, M, 1, &tiny);
}

/* FP64 forward oracle: the same model, normalised at every site. */
double orc_hmm_like_f64(const uint64_t *haps, unsigned wn, const unsigned P[4],
                        const float *emit, const float *tran, unsigned M) {
  const uint64_t *f0 = haps + (size_t)P[0] * wn, *f1 = haps + (size_t)P[1] * wn,
                 *m0 = haps + (size_t)P[2] * wn, *m1 = haps + (size_t)P[3] * wn;
  double l00 = 0.25 * emit[(hbit(f0, 0) << 1) | hbit(m0, 0)];
  double l01 = 0.25 * emit[(hbit(f0, 0) << 1) | hbit(m1, 0)];
  double l10 = 0.25 * emit[(hbit(f1, 0) << 1) | hbit(m0, 0)];
  double l11 = 0.25 * emit[(hbit(f1, 0) << 1) | hbit(m1, 0)];
  double s = l00 + l01 + l10 + l11, score = log(s);
  l00 /= s;
  l01 /= s;
  l10 /= s;
  l11 /= s;
  for (unsigned m = 1; m < M; m++) {
    const float *e = emit + 4 * (size_t)m, *t = tran + 3 * (size_t)m;
    double t0 = t[0], t1 = t[1], t2 = t[2];
    double b00 = l00 * t0 + (l01 + l10) * t1 + l11 * t2;
    double b01 = l01 * t0 + (l00 + l11) * t1 + l10 * t2;
    double b10 = l10 * t0 + (l00 + l11) * t1 + l01 * t2;
    double b11 = l11 * t0 + (l01 + l10) * t1 + l00 * t2;
    l00 = b00 * e[(hbit(f0, m) << 1) | hbit(m0, m)];
    l01 = b01 * e[(hbit(f0, m) << 1) | hbit(m1, m)];
    l10 = b10 * e[(hbit(f1, m) << 1) | hbit(m0, m)];
    l11 = b11 * e[(hbit(f1, m) << 1) | hbit(m1, m)];
    s = l00 + l01 + l10 + l11;
    score += log(s);
    l00 /= s;
    l01 /= s;
    l10 /= s;
    l11 /= s;
  }
  return score;
}

/* Sum over all 4^M hidden paths in FP64.  State xy = (father x, mother y);
 * a step keeps both (tran[0]), switches one (tran[1]) or both (tran[2]),
 * the model of hmmLike.h:18-20 and hmmLike.cu:111-122. */
double orc_hmm_like_brute(const uint64_t *haps, unsigned wn,
                          const unsigned P[4], const float *emit,
                          const float *tran, unsigned M) {
  if (M == 0 || M > 12)
    return NAN;
  const uint64_t *par[4];
  for (int i = 0; i < 4; ++i)
    par[i] = haps + (size_t)P[i] * wn;
  const uint64_t npaths = 1ull << (2 * M);
  double total = 0;
  for (uint64_t path = 0; path < npaths; ++path) {
    double p = 0.25;
    unsigned prev = 0;
    for (unsigned m = 0; m < M; ++m) {
      unsigned st = (unsigned)((path >> (2 * m)) & 3u);
      unsigned x = st >> 1, y = st & 1u;
      unsigned idx = (hbit(par[x], m) << 1) | hbit(par[2 + y], m);
      if (m) {
        unsigned px = prev >> 1, py = prev & 1u;
        unsigned nsw = (px != x) + (py != y);
        p *= (double)tran[3 * (size_t)m + nsw];
      }
      p *= (double)emit[4 * (size_t)m + idx];
      prev = st;
    }
    total += p;
  }
  return log(total);
}

/* ------------------------------------------------------------------------ */
/* MH loop: Insti::solve (insti.cpp:1478-1551) with MHSampler::SampleHap      */
/* (MHSampler.hpp:82-96) and fresh parents each generation (insti.cpp:1493).  */
/* Accept when prop >= curr, else when u < exp((prop - curr) * pen) in FP64.  */
/* ------------------------------------------------------------------------ */
static double like_of(const orc_window *w, const unsigned P[4],
                      const float *emit, int precision) {
  if (precision == ORC_PREC_MIXED)
    return orc_hmm_like_mixed(w->haps, w->wn, P, emit, w->tran, w->M);
  if (precision == ORC_PREC_MIXED_FB)
    return orc_hmm_like_fb(w->haps, w->wn, P, emit, w->tran, w->M);
  return (double)orc_hmm_like_f32(w->haps, w->wn, P, emit, w->tran, w->M);
}

static void mh_one(const orc_window *w, const orc_mh_args *a, unsigned i,
                   float *emit, unsigned *out_parents, double *out_curr,
                   double *trace_like, uint8_t *trace_acc) {
  unsigned P[4];
  uint32_t x[4];
  orc_sample_emit(w->packed, w->nsamp, i, w->M, w->bits, w->lut, emit);
  if (a->prop_mode == ORC_PROP_HOST) {
    for (int j = 0; j < 4; ++j)
      P[j] = a->init_parents[4 * (size_t)i + j];
  } else {
    orc_draw(a->seed, ORC_INIT_TAG, i, a->window_id, a->gen, x);
    for (int j = 0; j < 4; ++j)
      P[j] = a->prop_mode == ORC_PROP_KLIST
                 ? a->klist[(size_t)i * a->K + orc_uniform_int(x[j], a->K)]
                 : orc_unif_excl(x[j], w->nhaps, i);
  }
  double curr = like_of(w, P, emit, a->precision);
  if (trace_like)
    trace_like[(size_t)i * (a->ncycles + 1)] = curr;
  const double pen = (double)a->pen;
  for (unsigned n = 0; n < a->ncycles; ++n) {
    orc_draw(a->seed, a->k0 + n, i, a->window_id, a->gen, x);
    unsigned rp = x[0] & 3u, oh = P[rp];
    unsigned cand;
    if (a->prop_mode == ORC_PROP_HOST)
      cand = a->extra[(size_t)n * w->nsamp + i];
    else if (a->prop_mode == ORC_PROP_KLIST)
      cand = a->klist[(size_t)i * a->K + orc_uniform_int(x[1], a->K)];
    else
      cand = orc_unif_excl(x[1], w->nhaps, i);
    /* insti.cpp:1510-1513: with kickstarting, the first proposal of the
     * generation comes from the reference haplotypes [2 nsamp, nhaps) only;
     * Sampler::SampleHap(I, true) (sampler.cpp:12-27) rejects draws below
     * 2 nsamp, which leaves the draw uniform over the reference haplotypes
     * (UnifSampler) or over the K-list's entries among them (KNN; a list
     * with none keeps the usual draw: the reference would loop forever) */
    if (a->kickstart && a->k0 + n == 0 && a->prop_mode != ORC_PROP_HOST) {
      const unsigned ref0 = 2 * w->nsamp;
      if (a->prop_mode == ORC_PROP_KLIST) {
        const unsigned *L = a->klist + (size_t)i * a->K;
        unsigned nref = 0;
        for (unsigned k = 0; k < a->K; ++k)
          nref += L[k] >= ref0;
        if (nref) {
          unsigned pick = orc_uniform_int(x[1], nref);
          for (unsigned k = 0; k < a->K; ++k)
            if (L[k] >= ref0 && pick-- == 0) {
              cand = L[k];
              break;
            }
        }
      } else {
        cand = ref0 + orc_uniform_int(x[1], w->nhaps - ref0);
      }
    }
    P[rp] = cand;
    double prop = like_of(w, P, emit, a->precision);
    int acc = prop >= curr || orc_u01(x[2]) < exp((prop - curr) * pen);
    if (acc)
      curr = prop;
    else
      P[rp] = oh;
    if (trace_like)
      trace_like[(size_t)i * (a->ncycles + 1) + 1 + n] = prop;
    if (trace_acc)
      trace_acc[(size_t)i * a->ncycles + n] = (uint8_t)acc;
  }
  for (int j = 0; j < 4; ++j)
    out_parents[4 * (size_t)i + j] = P[j];
  if (out_curr)
    out_curr[i] = curr;
}

void orc_mh_window(const orc_window *w, const orc_mh_args *a, unsigned s0,
                   unsigned s1, unsigned *out_parents, double *out_curr,
                   double *trace_like, uint8_t *trace_acc) {
  float *emit = (float *)malloc(sizeof(float) * 4 * (size_t)w->M);
  for (unsigned i = s0; i < s1; ++i)
    mh_one(w, a, i, emit, out_parents, out_curr, trace_like, trace_acc);
  free(emit);
}

void orc_mh_window_omp(const orc_window *w, const orc_mh_args *a, unsigned s0,
                       unsigned s1, unsigned *out_parents, double *out_curr,
                       int nthreads) {
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
  {
    float *emit = (float *)malloc(sizeof(float) * 4 * (size_t)w->M);
#pragma omp for schedule(dynamic, 4)
    for (long i = (long)s0; i < (long)s1; ++i)
      mh_one(w, a, (unsigned)i, emit, out_parents, out_curr, NULL, NULL);
    free(emit);
  }
#else
  (void)nthreads;
  orc_mh_window(w, a, s0, s1, out_parents, out_curr, NULL, NULL);
#endif
}

/* ------------------------------------------------------------------------ */
/* GL vector quantisation: GLVQ (glVQ.cpp:26-165).  The GSL draws of the     */
/* reference are replaced by Philox draws under key `seed`, counter          */
/* {n, tag, 0, 0}: tag 1 = random initial codewords, 2 = training subsample,  */
/* 3 = re-seeding of empty clusters.                                         */
/* ------------------------------------------------------------------------ */

/* EuclidianDist (glVQ.hpp:23-34): sqrt(pow(float diff, 2) + ...) in FP64;
 * std::min_element keeps the first of equal distances. */
static inline double vq_dist(float cr, float ch, float RR, float Het) {
  double dr = (double)(cr - RR), dh = (double)(ch - Het);
  return sqrt(dr * dr + dh * dh);
}

unsigned orc_vq_find(const float *codebook, unsigned ncode, float RR,
                     float Het) {
  unsigned best = 0;
  double bd = vq_dist(codebook[0], codebook[1], RR, Het);
  for (unsigned c = 1; c < ncode; ++c) {
    double d = vq_dist(codebook[2 * c], codebook[2 * c + 1], RR, Het);
    if (d < bd) {
      bd = d;
      best = c;
    }
  }
  return best;
}

static void vq_draw(uint64_t seed, uint32_t n, uint32_t tag, uint32_t out[4]) {
  const uint32_t ctr[4] = {n, tag, 0, 0};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  orc_philox(ctr, key, out);
}

/* AssignPoints (glVQ.cpp:149-163) */
static double vq_assign(const float *pts, size_t n, const float *cb,
                        unsigned ncode, uint8_t *lab) {
  double var = 0;
  for (size_t p = 0; p < n; ++p) {
    float RR = pts[2 * p], Het = pts[2 * p + 1];
    unsigned c = orc_vq_find(cb, ncode, RR, Het);
    lab[p] = (uint8_t)c;
    double dr = (double)(RR - cb[2 * c]), dh = (double)(Het - cb[2 * c + 1]);
    var += dr * dr + dh * dh;
  }
  return var;
}

/* UpdateCodeBook (glVQ.cpp:165-211): float sums in point order */
static void vq_update(const float *pts, size_t n, const uint8_t *lab,
                      unsigned ncode, uint64_t seed, uint32_t *reseed,
                      float *cb) {
  float *s = (float *)calloc(2 * (size_t)ncode, sizeof(float));
  size_t *cnt = (size_t *)calloc(ncode, sizeof(size_t));
  for (size_t p = 0; p < n; ++p) {
    s[2 * lab[p]] += pts[2 * p];
    s[2 * lab[p] + 1] += pts[2 * p + 1];
    ++cnt[lab[p]];
  }
  for (unsigned c = 0; c < ncode; ++c) {
    if (cnt[c] > 0) {
      cb[2 * c] = s[2 * c] / (float)cnt[c];
      cb[2 * c + 1] = s[2 * c + 1] / (float)cnt[c];
    } else {
      uint32_t x[4];
      vq_draw(seed, (*reseed)++, 3, x);
      size_t p = (size_t)(((uint64_t)x[0] * n) >> 32);
      cb[2 * c] = pts[2 * p];
      cb[2 * c + 1] = pts[2 * p + 1];
    }
  }
  free(s);
  free(cnt);
}

/* BuildCodeBook (glVQ.cpp:26-146).  Returns 0, or -1 when Lloyd iterations
 * exceed 1000 (the reference throws). */
int orc_vq_train(const float *gls, size_t ngl, unsigned ncode, uint64_t seed,
                 size_t training_size, float *cb, unsigned *iters) {
  if (ncode < 4 || ncode > 256 || ngl == 0)
    return -2;
  cb[0] = 0;
  cb[1] = 1;
  cb[2] = 0;
  cb[3] = 0;
  cb[4] = 1;
  cb[5] = 0;
  cb[6] = (float)(1.0 / 3);
  cb[7] = (float)(1.0 / 3);
  for (unsigned c = 4; c < ncode; ++c) {
    uint32_t x[4];
    vq_draw(seed, c, 1, x);
    const float RR = (float)orc_u01(x[0]);
    const float Het = (float)(orc_u01(x[1]) * (1 - RR));
    cb[2 * c] = RR;
    cb[2 * c + 1] = Het;
  }
  size_t n = training_size && training_size < ngl ? training_size : ngl;
  float *pts = (float *)malloc(sizeof(float) * 2 * n);
  for (size_t it = 0; it < n; ++it) {
    size_t g = it;
    if (n < ngl) {
      uint32_t x[4];
      vq_draw(seed, (uint32_t)it, 2, x);
      g = (size_t)(((uint64_t)x[0] * ngl) >> 32);
    }
    float RR = gls[3 * g], Het = gls[3 * g + 1];
    float sum = gls[3 * g + 2] + RR + Het;
    RR /= sum;
    Het /= sum;
    pts[2 * it] = RR;
    pts[2 * it + 1] = Het;
  }
  uint8_t *lab = (uint8_t *)malloc(n);
  uint32_t reseed = 0;
  double prev = DBL_MAX, cur = vq_assign(pts, n, cb, ncode, lab);
  unsigned iter = 0;
  int rc = 0;
  while (fabs(cur - prev) > cur / 100) {
    if (++iter > 1000) {
      rc = -1;
      break;
    }
    prev = cur;
    vq_update(pts, n, lab, ncode, seed, &reseed, cb);
    cur = vq_assign(pts, n, cb, ncode, lab);
  }
  if (iters)
    *iters = iter;
  free(pts);
  free(lab);
  return rc;
}

/* GLPack::GLs2VQChar (glPack.cpp:195-219): normalise (sum = RR + Het + AA)
 * then FindGLCode */
void orc_vq_codes(const float *gls, size_t ngl, const float *codebook,
                  unsigned ncode, uint8_t *codes) {
  for (size_t g = 0; g < ngl; ++g) {
    float RR = gls[3 * g], Het = gls[3 * g + 1];
    float sum = RR + Het + gls[3 * g + 2];
    RR /= sum;
    Het /= sum;
    codes[g] = (uint8_t)orc_vq_find(codebook, ncode, RR, Het);
  }
}

/* GLPack::GetPackedGLs + GLIdxs2uint32_t (glPack.cpp:32-85,87-139): site-major,
 * sample-minor words, MSB-first codes; a site past the end takes GL index 0,
 * i.e. the code of sample 0, site 0. codes are [nsamp][M]. */
void orc_pack_codes(const uint8_t *codes, unsigned nsamp, unsigned M,
                    unsigned bits, unsigned s0, unsigned s1, uint32_t *out) {
  (void)nsamp;
  const unsigned k = 32u / bits;
  const unsigned nw = (M + k - 1) / k, ns = s1 - s0;
  for (unsigned w = 0; w < nw; ++w)
    for (unsigned i = s0; i < s1; ++i) {
      uint32_t word = 0;
      for (unsigned j = 0; j < k; ++j) {
        unsigned m = w * k + j;
        uint32_t c = m < M ? codes[(size_t)i * M + m] : codes[0];
        word |= c << (32u - bits * (j + 1));
      }
      out[(size_t)w * ns + (i - s0)] = word;
    }
}

/* GLPack::GLTrio2Char / GL2HalfChar (glPack.cpp:221-243): two 4-bit halves */
static unsigned half_char(float gl) {
  gl = gl - FLT_EPSILON;
  gl = gl > 0.0f ? gl : 0.0f;
  return (unsigned)floorf(gl * 16u);
}

static unsigned trio_char(const float *g) {
  float sum = g[0] + g[1] + g[2];
  return (half_char(g[0] / sum) << 4) | half_char(g[1] / sum);
}

void orc_pack_nonvq(const float *gls, unsigned nsamp, unsigned M, unsigned s0,
                    unsigned s1, uint32_t *out) {
  (void)nsamp;
  const unsigned nw = (M + 3) / 4, ns = s1 - s0;
  for (unsigned w = 0; w < nw; ++w)
    for (unsigned i = s0; i < s1; ++i) {
      uint32_t word = 0;
      for (unsigned j = 0; j < 4; ++j) {
        unsigned m = w * 4 + j;
        const float *g = m < M ? gls + 3 * ((size_t)i * M + m) : gls;
        word ^= trio_char(g) << (24u - 8u * j);
      }
      out[(size_t)w * ns + (i - s0)] = word;
    }
}

/* ------------------------------------------------------------------------ */
/* parameters: insti.cpp:1254-1315                                          */
/* ------------------------------------------------------------------------ */
float orc_mu(unsigned hn) {
  float mu = 0;
  for (unsigned i = 1; i < hn; i++)
    mu += 1.0 / i;
  mu = 1 / mu;
  return mu / (hn + mu);
}

void orc_pc(float mu, float *pc) {
  const float d = (1 - mu) * (1 - mu), o = mu * (1 - mu), a = mu * mu;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      int flips = ((i ^ j) & 1) + (((i ^ j) >> 1) & 1);
      pc[4 * i + j] = flips == 0 ? d : (flips == 1 ? o : a);
    }
}

/* insti.cpp:1277-1300 with r from a genetic map (Morgans, converted to the
 * float r), times the inflation factor (a double, insti.hpp:86), capped at
 * 1; tran[0..2] unused */
void orc_tran_from_morgans(const double *r_morgans, unsigned M,
                           double inflation, float *tran) {
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

/* insti.cpp:1253-1300 without a genetic map: rho from the SNP density of the
 * window's positions, r = (pos_m - pos_{m-1}) rho / (that + hn), times the
 * inflation factor (a double, insti.hpp:86), capped at 1; tran[0..2] unused */
void orc_tran_from_positions(const unsigned *pos, unsigned M, unsigned hn, double density,
                             double inflation, float *tran) {
  float mu = 0;
  for (unsigned i = 1; i < hn; i++)
    mu += 1.0 / i;
  mu = 1 / mu;
  const size_t firstPos = pos[0], lastPos = pos[M - 1];
  const float rho = lastPos == firstPos ? 0.0f
                                        : 0.5 * mu * (M - 1) / (lastPos - firstPos) / density;
  tran[0] = tran[1] = tran[2] = 0;
  for (unsigned m = M - 1; m; m--) {
    const float tmp = (pos[m] - pos[m - 1]) * rho;
    float r = tmp / (tmp + hn);
    r *= inflation;
    r = r > 1 ? 1 : r;
    tran[m * 3] = (1 - r) * (1 - r);
    tran[m * 3 + 1] = r * (1 - r);
    tran[m * 3 + 2] = r * r;
  }
}

/* ------------------------------------------------------------------------ */
/* Initial haplotypes: Insti::initialize (insti.cpp:1316-1365) for one window */
/* of het_aa [M][N][2] GL pairs; the three gsl_rng_uniform calls of (i, m)    */
/* become words 0, 1, 2 of Philox block {0xFFFFFFFE, i, m, window}.  haps is */
/* [2N][wn], zeroed here.                                                     */
/* ------------------------------------------------------------------------ */
static inline void set1(uint64_t *h, unsigned s) { h[s >> 6] |= 1ull << (s & 63); }

void orc_init_haps(const float *het_aa, unsigned N, unsigned M, uint64_t seed,
                   uint32_t window, uint64_t *haps) {
  const unsigned wn = (M + 63) / 64, hn = 2 * N;
  memset(haps, 0, sizeof(uint64_t) * (size_t)hn * wn);
  for (unsigned i = 0; i < N; i++) {
    uint64_t *ha = &haps[(size_t)i * 2 * wn], *hb = ha + wn;
    const float *p = &het_aa[i * 2];
    for (unsigned m = 0; m < M; m++, p += hn) {
      float t[3];
      const float r = 1 - p[0] - p[1];
      t[0] = r < 0.0f ? 0.0f : r; /* max(1 - p[0] - p[1], 0.0f) */
      t[1] = p[0] < 0.0f ? 0.0f : p[0];
      t[2] = p[1] < 0.0f ? 0.0f : p[1];
      uint32_t x[4];
      orc_draw(seed, 0xFFFFFFFEu, i, m, window, x);
      if (orc_u01(x[0]) < t[0])
        ;
      else if (orc_u01(x[1]) < t[1] / (t[1] + t[2])) {
        if (orc_u01(x[2]) < 0.5)
          set1(ha, m);
        else
          set1(hb, m);
      } else {
        set1(ha, m);
        set1(hb, m);
      }
    }
  }
}

/* Impute::replace (impute.cpp:564-597) for samples [0, N): hsum [2N][M]. */
void orc_replace(const uint64_t *haps, unsigned N, unsigned M, uint32_t *hsum) {
  const unsigned wn = (M + 63) / 64;
  for (unsigned I = 0; I < N; I++) {
    const uint64_t *oa = &haps[(size_t)I * 2 * wn], *ob = oa + wn;
    uint32_t *ha = &hsum[(size_t)I * 2 * M], *hb = ha + M;
    unsigned sis = 0, tra = 0;
    for (unsigned m = 0; m < M; m++) {
      if (hbit(oa, m)) {
        sis += ha[m];
        tra += hb[m];
      }
      if (hbit(ob, m)) {
        sis += hb[m];
        tra += ha[m];
      }
    }
    if (sis > tra)
      for (unsigned m = 0; m < M; m++) {
        ha[m] += hbit(oa, m);
        hb[m] += hbit(ob, m);
      }
    else
      for (unsigned m = 0; m < M; m++) {
        ha[m] += hbit(ob, m);
        hb[m] += hbit(oa, m);
      }
  }
}

/* ------------------------------------------------------------------------ */
/* hmm_work: Impute::hmm_work (impute.cpp:380-485).  The per-site uniform of   */
/* gsl_rng_uniform becomes Philox counter {0x80000000 | m>>2, sample, window, */
/* gen}, word m & 3.  Bits of ha/hb past M are left zero.                     */
/* ------------------------------------------------------------------------ */
/* powf as the correctly rounded float of the FP64 power (hmm_work's
 * annealing exponent, impute.cpp:457-469); libm powf differs between
 * libraries in the last bit. */
static float orc_powf(float x, float y) { return (float)pow((double)x, (double)y); }

static inline void set_bit(uint64_t *h, unsigned s, unsigned v) {
  if (v)
    h[s >> 6] |= (uint64_t)1 << (s & 63);
  else
    h[s >> 6] &= ~((uint64_t)1 << (s & 63));
}

/* Impute::hmm_work's weights (impute.cpp:380-470): the scaled backward
 * recursion, then the forward pass normalised at every site, the state
 * posteriors p_xy = l_xy * beta_xy, and per site the 4 genotype weights
 * c[m][g] = powf(prob_g * sum_xy p_xy * pc[s_xy][g], S), g = (ha << 1) | hb,
 * prob_g = prob[m][0], [1], [1], [2].  post (optional) receives the
 * unnormalised p_xy. */
void orc_hmm_work_weights(const uint64_t *haps, unsigned wn, const unsigned P[4],
                          const float *emit, const float *tran, const float *prob,
                          const float *pc, float S, unsigned M, float *c, float *post) {
  const uint64_t *f0 = haps + (size_t)P[0] * wn, *f1 = haps + (size_t)P[1] * wn,
                 *m0 = haps + (size_t)P[2] * wn, *m1 = haps + (size_t)P[3] * wn;
  float *beta = (float *)malloc(sizeof(float) * 4 * (size_t)M);
  const float *e = emit + 4 * (size_t)(M - 1), *t = tran + 3 * (size_t)(M - 1);
  float *b = beta + 4 * (size_t)(M - 1), sum;
  float l00 = 0, l01 = 0, l10 = 0, l11 = 0, b00, b01, b10, b11;
  b[0] = b[1] = b[2] = b[3] = 1;
  for (unsigned m = M - 1; m; m--, e -= 4, t -= 3) {
    b00 = b[0] * e[(hbit(f0, m) << 1) | hbit(m0, m)];
    b01 = b[1] * e[(hbit(f0, m) << 1) | hbit(m1, m)];
    b10 = b[2] * e[(hbit(f1, m) << 1) | hbit(m0, m)];
    b11 = b[3] * e[(hbit(f1, m) << 1) | hbit(m1, m)];
    b -= 4;
    b[0] = b00 * t[0] + (b01 + b10) * t[1] + b11 * t[2];
    b[1] = b01 * t[0] + (b00 + b11) * t[1] + b10 * t[2];
    b[2] = b10 * t[0] + (b00 + b11) * t[1] + b01 * t[2];
    b[3] = b11 * t[0] + (b01 + b10) * t[1] + b00 * t[2];
    sum = 1.0f / (b[0] + b[1] + b[2] + b[3]);
    b[0] *= sum;
    b[1] *= sum;
    b[2] *= sum;
    b[3] *= sum;
  }
  const float *p = prob;
  e = emit;
  t = tran;
  b = beta;
  for (unsigned m = 0; m < M; m++, e += 4, t += 3, p += 3, b += 4) {
    unsigned s00 = (hbit(f0, m) << 1) | hbit(m0, m);
    unsigned s01 = (hbit(f0, m) << 1) | hbit(m1, m);
    unsigned s10 = (hbit(f1, m) << 1) | hbit(m0, m);
    unsigned s11 = (hbit(f1, m) << 1) | hbit(m1, m);
    if (m) {
      b00 = l00 * t[0] + l01 * t[1] + l10 * t[1] + l11 * t[2],
      b01 = l00 * t[1] + l01 * t[0] + l10 * t[2] + l11 * t[1];
      b10 = l00 * t[1] + l01 * t[2] + l10 * t[0] + l11 * t[1],
      b11 = l00 * t[2] + l01 * t[1] + l10 * t[1] + l11 * t[0];
      l00 = b00 * e[s00];
      l01 = b01 * e[s01];
      l10 = b10 * e[s10];
      l11 = b11 * e[s11];
    } else {
      l00 = 0.25f * e[s00];
      l01 = 0.25f * e[s01];
      l10 = 0.25f * e[s10];
      l11 = 0.25f * e[s11];
    }
    sum = 1.0f / (l00 + l01 + l10 + l11);
    l00 *= sum;
    l01 *= sum;
    l10 *= sum;
    l11 *= sum;
    float p00 = l00 * b[0], p01 = l01 * b[1], p10 = l10 * b[2],
          p11 = l11 * b[3];
    if (post) {
      post[4 * (size_t)m + 0] = p00;
      post[4 * (size_t)m + 1] = p01;
      post[4 * (size_t)m + 2] = p10;
      post[4 * (size_t)m + 3] = p11;
    }
    float *cm = c + 4 * (size_t)m;
    cm[0] = orc_powf(p[0] * (p00 * pc[4 * s00 + 0] + p01 * pc[4 * s01 + 0] +
                             p10 * pc[4 * s10 + 0] + p11 * pc[4 * s11 + 0]),
                     S);
    cm[1] = orc_powf(p[1] * (p00 * pc[4 * s00 + 1] + p01 * pc[4 * s01 + 1] +
                             p10 * pc[4 * s10 + 1] + p11 * pc[4 * s11 + 1]),
                     S);
    cm[2] = orc_powf(p[1] * (p00 * pc[4 * s00 + 2] + p01 * pc[4 * s01 + 2] +
                             p10 * pc[4 * s10 + 2] + p11 * pc[4 * s11 + 2]),
                     S);
    cm[3] = orc_powf(p[2] * (p00 * pc[4 * s00 + 3] + p01 * pc[4 * s01 + 3] +
                             p10 * pc[4 * s10 + 3] + p11 * pc[4 * s11 + 3]),
                     S);
  }
  free(beta);
}

/* The sampling step of impute.cpp:472-484: per site one uniform u (word
 * m & 3 of the Philox block kWorkTag | (m >> 2), which replaces
 * gsl_rng_uniform), x = (float)(u * (c00 + c01 + c10 + c11)); the first g
 * with x < c00 + ... + c_g sets (ha, hb) = (g >> 1, g & 1). */
void orc_hmm_work_sample(const float *c, unsigned M, unsigned wn, uint64_t seed,
                         uint32_t sample, uint32_t window, uint32_t gen, uint64_t *ha,
                         uint64_t *hb) {
  uint32_t x[4] = {0, 0, 0, 0};
  for (unsigned w = 0; w < wn; ++w) ha[w] = hb[w] = 0;
  for (unsigned m = 0; m < M; ++m) {
    const float *cm = c + 4 * (size_t)m;
    if ((m & 3u) == 0) orc_draw(seed, 0x80000000u | (m >> 2), sample, window, gen, x);
    const float sum = (float)(orc_u01(x[m & 3u]) * (cm[0] + cm[1] + cm[2] + cm[3]));
    if (sum < cm[0]) {
      set_bit(ha, m, 0);
      set_bit(hb, m, 0);
    } else if (sum < cm[0] + cm[1]) {
      set_bit(ha, m, 0);
      set_bit(hb, m, 1);
    } else if (sum < cm[0] + cm[1] + cm[2]) {
      set_bit(ha, m, 1);
      set_bit(hb, m, 0);
    } else {
      set_bit(ha, m, 1);
      set_bit(hb, m, 1);
    }
  }
}

void orc_hmm_work(const uint64_t *haps, unsigned wn, const unsigned P[4],
                  const float *emit, const float *tran, const float *prob,
                  const float *pc, float S, unsigned M, uint64_t seed,
                  uint32_t sample, uint32_t window, uint32_t gen, uint64_t *ha,
                  uint64_t *hb) {
  float *c = (float *)malloc(sizeof(float) * 4 * (size_t)M);
  orc_hmm_work_weights(haps, wn, P, emit, tran, prob, pc, S, M, c, NULL);
  orc_hmm_work_sample(c, M, wn, seed, sample, window, gen, ha, hb);
  free(c);
}

/* Insti::initialize's per-site genotype probabilities and emissions
 * (insti.cpp:1345-1378, autosomal case): from the input pair (Het, AA),
 * t = (max(1 - Het - AA, 0), max(Het, 0), max(AA, 0)) and
 * e[j] = pc[j][0] t0 + pc[j][1] t1 + pc[j][2] t1 + pc[j][3] t2. */
void orc_prob_emit(const float *het_aa, unsigned M, const float *pc, float *prob, float *emit) {
  for (unsigned m = 0; m < M; ++m) {
    const float *q = het_aa + 2 * (size_t)m;
    float *t = prob + 3 * (size_t)m, *e = emit + 4 * (size_t)m;
    const float t0r = 1 - q[0] - q[1];
    t[0] = t0r > 0.0f ? t0r : 0.0f;
    t[1] = q[0] > 0.0f ? q[0] : 0.0f;
    t[2] = q[1] > 0.0f ? q[1] : 0.0f;
    for (unsigned j = 0; j < 4; ++j)
      e[j] = pc[4 * j + 0] * t[0] + pc[4 * j + 1] * t[1] + pc[4 * j + 2] * t[1] +
             pc[4 * j + 3] * t[2];
  }
}

/* ------------------------------------------------------------------------ */
/* kNN candidate lists (KNN::KNN / CalculateVarAfs / CutDownHaps /
 * ClusterHaps, kNN.cpp:8-225; Haplotype::HammingDist, haplotype.cpp:35-60).
 * Sites are kept when the variant allele frequency over all nhaps
 * haplotypes (as a double, alt count / nhaps; with using_maf the minor
 * allele frequency |(|af - 0.5|) - 0.5|) lies in [lb, ub].  For sample u
 * (haplotypes 2u, 2u+1) every other haplotype is ranked by its Hamming
 * distance over the kept sites; std::sort leaves the order of equal
 * distances unspecified, this restatement breaks ties by haplotype index.
 * k' = ceil(K / 2): out[2j] = hap A's j-th nearest, out[2j + 1] = the j-th
 * nearest of hap B not already listed.  K is first capped at nhaps - 2. */
typedef struct {
  unsigned d, idx;
} orc_dist;

static int orc_dist_cmp(const void *a, const void *b) {
  const orc_dist *x = (const orc_dist *)a, *y = (const orc_dist *)b;
  if (x->d != y->d) return x->d < y->d ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

unsigned orc_knn_k_out(unsigned K, unsigned nhaps) {
  if (K > nhaps - 2) K = nhaps - 2;
  return 2 * ((K + 1) / 2);
}

void orc_knn_mask(const uint64_t *haps, unsigned wn, unsigned nhaps, unsigned M, double lb,
                  double ub, int using_maf, uint64_t *mask, double *af) {
  for (unsigned w = 0; w < wn; ++w) mask[w] = 0;
  for (unsigned s = 0; s < M; ++s) {
    unsigned alt = 0;
    for (unsigned h = 0; h < nhaps; ++h) alt += hbit(haps + (size_t)h * wn, s);
    double a = (double)alt / nhaps;
    if (af) af[s] = a;
    if (using_maf) a = fabs(fabs(a - 0.5) - 0.5);
    if (a >= lb && a <= ub) mask[s >> 6] |= (uint64_t)1 << (s & 63);
  }
}

/* the list of sample u (haplotypes 2u, 2u + 1) into o[kout]; da, db hold
 * nhaps entries each */
static void knn_one(const uint64_t *haps, unsigned wn, unsigned nhaps, const uint64_t *mask,
                    unsigned u, unsigned kh, orc_dist *da, orc_dist *db, unsigned *o) {
  const uint64_t *a = haps + (size_t)(2 * u) * wn, *b = a + wn;
  unsigned n = 0;
  for (unsigned h = 0; h < nhaps; ++h) {
    if ((h >> 1) == u) continue;
    const uint64_t *c = haps + (size_t)h * wn;
    unsigned x = 0, y = 0;
    for (unsigned w = 0; w < wn; ++w) {
      x += (unsigned)__builtin_popcountll((a[w] ^ c[w]) & mask[w]);
      y += (unsigned)__builtin_popcountll((b[w] ^ c[w]) & mask[w]);
    }
    da[n].d = x;
    da[n].idx = h;
    db[n].d = y;
    db[n].idx = h;
    ++n;
  }
  qsort(da, n, sizeof(orc_dist), orc_dist_cmp);
  qsort(db, n, sizeof(orc_dist), orc_dist_cmp);
  for (unsigned j = 0; j < kh; ++j) o[2 * j] = da[j].idx;
  unsigned hb = 0;
  for (unsigned j = 0; j < kh; ++j) {
    for (;;) {
      const unsigned cand = db[hb++].idx;
      int seen = 0;
      for (unsigned t = 0; t < kh && !seen; ++t) seen = o[2 * t] == cand;
      for (unsigned t = 0; t < j && !seen; ++t) seen = o[2 * t + 1] == cand;
      if (!seen) {
        o[2 * j + 1] = cand;
        break;
      }
    }
  }
}

void orc_knn(const uint64_t *haps, unsigned wn, unsigned nhaps, unsigned M, unsigned nsamp,
             unsigned K, double lb, double ub, int using_maf, unsigned *out) {
  uint64_t *mask = (uint64_t *)malloc(sizeof(uint64_t) * wn);
  orc_knn_mask(haps, wn, nhaps, M, lb, ub, using_maf, mask, NULL);
  const unsigned kout = orc_knn_k_out(K, nhaps), kh = kout / 2;
  orc_dist *da = (orc_dist *)malloc(sizeof(orc_dist) * nhaps);
  orc_dist *db = (orc_dist *)malloc(sizeof(orc_dist) * nhaps);
  for (unsigned u = 0; u < nsamp; ++u)
    knn_one(haps, wn, nhaps, mask, u, kh, da, db, out + (size_t)u * kout);
  free(da);
  free(db);
  free(mask);
}

/* orc_knn for the samples rows[0..nrows) only (out[r] = the list of sample
 * rows[r]), over nthreads OpenMP threads */
void orc_knn_rows(const uint64_t *haps, unsigned wn, unsigned nhaps, unsigned M, unsigned K,
                  double lb, double ub, int using_maf, const unsigned *rows, unsigned nrows,
                  unsigned *out, int nthreads) {
  uint64_t *mask = (uint64_t *)malloc(sizeof(uint64_t) * wn);
  orc_knn_mask(haps, wn, nhaps, M, lb, ub, using_maf, mask, NULL);
  const unsigned kout = orc_knn_k_out(K, nhaps), kh = kout / 2;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
    orc_dist *da = (orc_dist *)malloc(sizeof(orc_dist) * nhaps);
    orc_dist *db = (orc_dist *)malloc(sizeof(orc_dist) * nhaps);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 8)
#endif
    for (long r = 0; r < (long)nrows; ++r)
      knn_one(haps, wn, nhaps, mask, rows[r], kh, da, db, out + (size_t)r * kout);
    free(da);
    free(db);
  }
  (void)nthreads;
  free(mask);
}

/*
 * oracle.h -- CPU restatement of GLPhase's MH surrogate-parent update.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library in
 * glphase_amd/ and the timed CPU baseline in bench.py.  Nothing in the product
 * path links, loads or calls it.
 *
 * Every function cites the reference source it restates (paths relative to
 * the GLPhase source tree).  Floating-point expressions keep the reference's
 * operation order and are compiled with -ffp-contract=off, so a float result
 * here is the IEEE value the reference's CPU build (g++ -std=c++11, x86-64
 * SSE, no FMA) computes for the same inputs.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_philox          Random123 known-answer vectors (SC'11 paper's KAT).
 *   orc_unpack_gl       known answers of src/testing/cuda_tests.cpp:66-96.
 *   orc_emit_lut        cuda_tests.cpp:98-116 (pc == 1 -> GL0+2*GL1+GL2) and
 *                       the transposed layout of cuda_tests.cpp:46-63.
 *   orc_pack_nonvq      known answers of glPack_unittests.cpp:27-117.
 *   orc_vq_find/_train  glPack_unittests.cpp:119-163,252-340 properties.
 *   orc_hmm_like_f32    closed forms (r=0.5, r=0, constant emission), 4^M
 *                       brute force in FP64 (M<=8), FP64 forward, swap
 *                       invariances.
 *   orc_hmm_like_fb     FP64 forward, 4^M brute force (odd and even M), the
 *                       mixed pass's closed forms, its forward half against
 *                       the mixed pass, the underflow guard's mu^2 runs.
 *   orc_hmm_like_f64    brute force (M<=8).
 *   orc_mh_window       MHSampler semantics of other_unittests.cpp:256-283,
 *                       planted-parent behaviour of cuda_tests.cpp:118-656.
 */
#ifndef GLPHASE_ORACLE_H
#define GLPHASE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 and the draw streams -------------------------------- */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#define ORC_INIT_TAG 0xFFFFFFFFu
/* out = Philox(ctr = {k, sample, window, gen}, key = seed) */
void orc_draw(uint64_t seed, uint32_t k, uint32_t sample, uint32_t window,
              uint32_t gen, uint32_t out[4]);
/* uniform integer in [0, n) from one 32-bit word: floor(x * n / 2^32) */
uint32_t orc_uniform_int(uint32_t x, uint32_t n);
/* uniform haplotype of [0, nhaps) excluding sample's own 2*s, 2*s+1 */
uint32_t orc_unif_excl(uint32_t x, uint32_t nhaps, uint32_t sample);

/* ---- GL decode and emission -------------------------------------------- */
float orc_norm(void);
/* correctly rounded logf (see oracle.c) */
float orc_logf_cr(float x);
void orc_unpack_gl(uint32_t word, const float *codebook, unsigned bits,
                   unsigned slot, float gl[3]);
/* lut[c*4+i] = FillEmit(codebook[c]); pc is row-major pc[i][j] */
void orc_emit_lut(const float *codebook, unsigned ncode, const float *pc,
                  float *lut);
/* per-site emission of one sample from the packed codes ([ceil(M/k)][N]) */
void orc_sample_emit(const uint32_t *packed, unsigned nsamp_packed,
                     unsigned sample, unsigned M, unsigned bits,
                     const float *lut, float *emit);

/* ---- forward pass -------------------------------------------------------- */
float orc_hmm_like_f32(const uint64_t *haps, unsigned wn, const unsigned P[4],
                       const float *emit, const float *tran, unsigned M);
double orc_hmm_like_mixed(const uint64_t *haps, unsigned wn,
                          const unsigned P[4], const float *emit,
                          const float *tran, unsigned M);
/* one mixed-precision pass, scaled every 8th site (every = 8) or after every
 * site (every = 1); *tiny is set when a sum fell below 2^-99 or the last sum
 * is not finite */
double orc_hmm_like_mixed_pass(const uint64_t *haps, unsigned wn, const unsigned P[4],
                               const float *emit, const float *tran, unsigned M,
                               unsigned every, int *tiny);
/* mixed precision, forward-backward form: two halves of ceil(M/2) steps
 * joined in FP64 at the middle (see oracle.c); the single pass as above */
double orc_hmm_like_fb(const uint64_t *haps, unsigned wn, const unsigned P[4],
                       const float *emit, const float *tran, unsigned M);
double orc_hmm_like_fb_pass(const uint64_t *haps, unsigned wn, const unsigned P[4],
                            const float *emit, const float *tran, unsigned M,
                            unsigned every, int *tiny);
double orc_hmm_like_f64(const uint64_t *haps, unsigned wn, const unsigned P[4],
                        const float *emit, const float *tran, unsigned M);
double orc_hmm_like_brute(const uint64_t *haps, unsigned wn,
                          const unsigned P[4], const float *emit,
                          const float *tran, unsigned M);

/* ---- MH chain over a range of samples ----------------------------------- */
enum { ORC_PROP_UNIFORM = 0, ORC_PROP_KLIST = 1, ORC_PROP_HOST = 2 };
enum { ORC_PREC_F32 = 0, ORC_PREC_MIXED = 1, ORC_PREC_MIXED_FB = 2 };

typedef struct {
  const uint64_t *haps;    /* [nhaps][wn] */
  unsigned wn, nhaps;
  const uint32_t *packed;  /* [ceil(M/k)][nsamp] codes */
  unsigned nsamp, M, bits;
  const float *lut;        /* [2^bits][4] */
  const float *tran;       /* [M][3] */
} orc_window;

typedef struct {
  unsigned ncycles;
  float pen;
  uint64_t seed;
  uint32_t window_id, gen, k0; /* proposal k uses counter k0 + k */
  int prop_mode, precision;
  const unsigned *klist;      /* [nsamp][K] for ORC_PROP_KLIST */
  unsigned K;
  const unsigned *init_parents; /* [nsamp][4] for ORC_PROP_HOST (else null) */
  const unsigned *extra;      /* [ncycles][nsamp] for ORC_PROP_HOST */
  int kickstart;              /* proposal k0 + n == 0 from the reference haps only */
} orc_mh_args;

/* Runs samples [s0, s1) sequentially.  out_parents [nsamp][4];
 * out_curr [nsamp]; trace_like [nsamp][ncycles+1] (initial, then each
 * proposal's value) and trace_acc [nsamp][ncycles] are optional. */
void orc_mh_window(const orc_window *w, const orc_mh_args *a, unsigned s0,
                   unsigned s1, unsigned *out_parents, double *out_curr,
                   double *trace_like, uint8_t *trace_acc);
/* the same, split over OpenMP threads (a generation is a Jacobi update) */
void orc_mh_window_omp(const orc_window *w, const orc_mh_args *a, unsigned s0,
                       unsigned s1, unsigned *out_parents, double *out_curr,
                       int nthreads);

/* ---- GL vector quantisation and packing --------------------------------- */
unsigned orc_vq_find(const float *codebook, unsigned ncode, float RR,
                     float Het);
int orc_vq_train(const float *gls, size_t ngl, unsigned ncode, uint64_t seed,
                 size_t training_size, float *codebook, unsigned *iters);
void orc_vq_codes(const float *gls, size_t ngl, const float *codebook,
                  unsigned ncode, uint8_t *codes);
void orc_pack_codes(const uint8_t *codes, unsigned nsamp, unsigned M,
                    unsigned bits, unsigned s0, unsigned s1, uint32_t *out);
void orc_pack_nonvq(const float *gls, unsigned nsamp, unsigned M,
                    unsigned s0, unsigned s1, uint32_t *out);

/* ---- model parameters ----------------------------------------------------- */
float orc_mu(unsigned hn);
void orc_pc(float mu, float *pc);
void orc_tran_from_morgans(const double *r_morgans, unsigned M,
                           double inflation, float *tran);
void orc_tran_from_positions(const unsigned *pos, unsigned M, unsigned hn, double density,
                             double inflation, float *tran);

/* ---- generation bookkeeping: initial haplotypes, allele tallies ---------- */
void orc_init_haps(const float *het_aa, unsigned N, unsigned M, uint64_t seed,
                   uint32_t window, uint64_t *haps);
void orc_replace(const uint64_t *haps, unsigned N, unsigned M, uint32_t *hsum);

/* ---- hmm_work: forward-backward haplotype sampling ----------------------- */
void orc_hmm_work_weights(const uint64_t *haps, unsigned wn, const unsigned P[4],
                          const float *emit, const float *tran, const float *prob,
                          const float *pc, float S, unsigned M, float *c, float *post);
void orc_hmm_work_sample(const float *c, unsigned M, unsigned wn, uint64_t seed,
                         uint32_t sample, uint32_t window, uint32_t gen, uint64_t *ha,
                         uint64_t *hb);
void orc_hmm_work(const uint64_t *haps, unsigned wn, const unsigned P[4],
                  const float *emit, const float *tran, const float *prob,
                  const float *pc, float S, unsigned M, uint64_t seed,
                  uint32_t sample, uint32_t window, uint32_t gen,
                  uint64_t *ha, uint64_t *hb);
/* per-site genotype probabilities and emissions from (Het, AA) GL pairs */
void orc_prob_emit(const float *het_aa, unsigned M, const float *pc, float *prob, float *emit);

/* ---- kNN candidate lists -------------------------------------------------- */
unsigned orc_knn_k_out(unsigned K, unsigned nhaps);
void orc_knn_mask(const uint64_t *haps, unsigned wn, unsigned nhaps, unsigned M, double lb,
                  double ub, int using_maf, uint64_t *mask, double *af);
void orc_knn(const uint64_t *haps, unsigned wn, unsigned nhaps, unsigned M, unsigned nsamp,
             unsigned K, double lb, double ub, int using_maf, unsigned *out);
void orc_knn_rows(const uint64_t *haps, unsigned wn, unsigned nhaps, unsigned M, unsigned K,
                  double lb, double ub, int using_maf, const unsigned *rows, unsigned nrows,
                  unsigned *out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

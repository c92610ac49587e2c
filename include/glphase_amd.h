/*
 * glphase_amd.h -- C ABI of the gfx950 GLPhase MH surrogate-parent engine.
 *
 * Plain pointers and sizes only.  Every call returns GLP_OK (0) or a
 * GLP_ERR_* code; glp_last_error() gives the calling thread's last message.
 * No call exits the process (the reference's device layer calls
 * exit(EXIT_FAILURE), hmmLike.cu:240,252,261,272,286,476,544).
 *
 * Two layers:
 *  1. The reference-compatible layer mirrors namespace HMMLikeCUDA
 *     (hmmLike.h:23-38): host buffers in, synchronous, one context per
 *     HMMLike object instead of process-global device state
 *     (hmmLike.cu:8-10,357-383).
 *  2. The device-resident engine takes device pointers and a hipStream_t
 *     (passed as void*), runs asynchronously and is what the benchmark times.
 *
 * Conventions shared with the reference:
 *  - haplotype h, site s is bit (s & 63) of word h*wn + (s >> 6), with
 *    wn = ceil(M / 64) (impute.hpp:53-66, insti.cpp:1204);
 *  - packed GL codes are site-major, sample-minor uint32 words, MSB-first,
 *    32/bits codes per word (glPack.cpp:112-192);
 *  - parent slots 0,1 are the father pair and 2,3 the mother pair
 *    (hmmLike.cu:61-63, impute.cpp:331-332);
 *  - tran is [M][3] = ((1-r)^2, r(1-r), r^2) per site, tran[0..2] unused
 *    (insti.cpp:1277-1300);
 *  - pc is the 4x4 mutation matrix, row-major pc[i][j] (insti.cpp:1307-1315).
 */
#ifndef GLPHASE_AMD_H
#define GLPHASE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GLP_OK = 0,
  GLP_ERR_ARG = 1,   /* invalid argument or size mismatch */
  GLP_ERR_HIP = 2,   /* HIP runtime error */
  GLP_ERR_STATE = 3, /* call made before the state it needs was set */
  GLP_ERR_NODEV = 4  /* no usable gfx950 device */
};

/* proposal modes (glp_mh_desc.prop_mode) */
enum { GLP_PROP_UNIFORM = 0, GLP_PROP_KLIST = 1, GLP_PROP_HOST = 2 };
/* Arithmetic of the forward pass (glp_mh_desc.precision).  GLP_PREC_F32: the reference's FP32
 * recurrence and FP32 log accumulation (bit-identical to Impute::hmm_like);
 * GLP_PREC_MIXED: the same FP32 recurrence, kept in range by exact power-of-two
 * scalings after every 8th site instead of the per-site 1/sum rescale, FP64
 * log-likelihood and FP64 MH ratio; GLP_PREC_MIXED_FB: the mixed arithmetic
 * evaluated as two halves side by side (the forward variable over sites
 * 0 .. ceil(M/2)-1 and the emission-weighted backward variable over the rest,
 * joined in FP64), which halves a pass's latency; it needs allocated window
 * layouts, so it runs through plans (glp_plan_*) only.  An FP16 emission table
 * was measured at up to 1.3e-4 relative error against the FP64 forward pass
 * (bound 1e-5) and is not provided. */
enum { GLP_PREC_F32 = 0, GLP_PREC_MIXED = 1, GLP_PREC_MIXED_FB = 2 };

const char *glp_last_error(void);
const char *glp_version(void);

typedef struct glp_ctx glp_ctx;

/* ===== 1. reference-compatible layer ===================================== */

/* HMMLikeCUDA::CheckDevice (hmmLike.h:24, hmmLike.cu:266-320): counts
 * devices and requires at least one gfx950 device. */
int glp_check_device(int *num_devices);

/* HMMLikeCUDA::SetDevice (hmmLike.h:25, hmmLike.cu:244-264) plus the state
 * the reference keeps in process globals; norm = powf(FLT_MIN, 2/3)
 * (hmmLike.cu:233). */
int glp_create(int dev_id, glp_ctx **out);

/* HMMLikeCUDA::Cleanup (hmmLike.h:37, hmmLike.cu:385-397); frees this
 * context only (no device reset). */
int glp_destroy(glp_ctx *ctx);

/* HMMLikeCUDA::CopyTranToDevice (hmmLike.h:26, hmmLike.cu:322-337).
 * n = 3 * num_sites floats. */
int glp_set_tran(glp_ctx *ctx, const float *tran, size_t n);

/* HMMLikeCUDA::CopyMutationMatToDevice (hmmLike.h:27, hmmLike.cu:339-355).
 * pc16 is row-major pc[i][j]. */
int glp_set_mutation_mat(glp_ctx *ctx, const float *pc16);

/* HMMLikeCUDA::CopyCodeBookToDevice (hmmLike.h:29-30, hmmLike.cu:369-383).
 * pairs = n_codes (RR, Het) float pairs; n_codes = 2^bits with bits 4 (the
 * reference GPU path's BITSPERCODE, globals.h:23) or 8.  The packing helpers
 * below also take 1- and 2-bit codes (glPack.cpp:26-27); the MH kernels do
 * not. */
int glp_set_codebook(glp_ctx *ctx, const float *pairs, unsigned n_codes);

/* HMMLikeCUDA::CopyPackedGLsToDevice (hmmLike.h:28, hmmLike.cu:357-366). */
int glp_set_packed_gls(glp_ctx *ctx, const uint32_t *packed, size_t n_words,
                       unsigned bits_per_code);

/* MH penalty for the accept test u < exp((prop - curr) * pen).  The
 * reference CUDA kernel fixes it at 1 (hmmLike.cu:159,196); its CPU MH loop
 * uses the annealing penalty (insti.cpp:1670-1671, MHSampler.hpp:92).
 * Default 1. */
int glp_set_pen(glp_ctx *ctx, float pen);

/* HMMLikeCUDA::RunHMMOnDevice (hmmLike.h:31-35, hmmLike.cu:399-572).
 * hap_idxs: in = initial parents, out = chosen parents, both [4][num_samples]
 * (slot-major, as the reference).  extra_prop_haps: [num_cycles][num_samples]
 * candidate per cycle.  The replace slot and the accept uniform of cycle c of
 * sample i are Philox draws keyed by seed, counter {c, i, 0, 0}. */
int glp_run_hmm_on_device(glp_ctx *ctx, const uint64_t *hap_panel,
                          size_t n_panel_words, const unsigned *extra_prop_haps,
                          unsigned num_sites, unsigned num_samples,
                          unsigned num_cycles, unsigned *hap_idxs,
                          unsigned long seed);

/* Impute::hmm_like / Insti::hmm_like (impute.hpp:80, impute.cpp:327-375,
 * insti.cpp:97), batched: out[e] = log-likelihood of sample samples[e] given
 * the 4 parents parents[4e..4e+3], with emissions decoded from the packed
 * codes through the codebook (hmmLike.cu:53-142). */
int glp_hmm_like(glp_ctx *ctx, const uint64_t *hap_panel, size_t n_panel_words,
                 unsigned num_sites, unsigned num_samples, unsigned n_evals,
                 const unsigned *samples, const unsigned *parents, float *out);

/* ===== 2. device-resident engine ========================================= */

/* One window (a run of sites scored together).  All pointers are device
 * pointers. */
typedef struct {
  unsigned num_sites;     /* M */
  unsigned words_per_hap; /* ceil(M / 64) */
  const uint64_t *panel;  /* [num_haps][words_per_hap] */
  const uint32_t *packed; /* [ceil(M * bits / 32)][num_samples] */
  const float *tran;      /* [M][3] */
  unsigned *parents;      /* [num_samples][4]; in (GLP_PROP_HOST) and out */
  double *like;           /* [num_samples] out: log-likelihood of parents */
} glp_window;

typedef struct {
  unsigned num_windows;
  const glp_window *windows; /* host array of num_windows */
  unsigned num_samples;      /* samples per window */
  unsigned num_haps;         /* haplotypes per panel (2N + reference) */
  unsigned bits_per_code;    /* 4 or 8 */
  const void *emit_lut;      /* device table built by glp_build_emit_lut */
  unsigned num_cycles;       /* MH proposals per sample (C) */
  unsigned k0;               /* proposal c uses Philox counter k0 + c */
  float pen;
  uint64_t seed;
  unsigned gen;
  int prop_mode;           /* GLP_PROP_* */
  const unsigned *klist;   /* GLP_PROP_KLIST: device [num_windows][N][K] */
  unsigned K;
  const unsigned *extra;   /* GLP_PROP_HOST: device [num_windows][C][N] */
  int precision;           /* GLP_PREC_* */
  double *trace_like;      /* optional device [num_windows][N][C+1] */
  uint8_t *trace_acc;      /* optional device [num_windows][N][C] */
  /* Insti::s_bKickStartFromRef (insti.cpp:1510-1513, main.cpp:273,430-433):
   * nonzero draws the candidate of proposal 0 of every generation (Philox
   * counter k0 + c == 0) from the reference haplotypes [2 * num_samples,
   * num_haps) only -- uniformly, or uniformly over the K-list's entries in
   * that range (GLP_PROP_KLIST; a list with none falls back to the usual
   * draw, where the reference's rejection loop would not end).  Needs
   * num_haps > 2 * num_samples; ignored under GLP_PROP_HOST. */
  int kickstart_from_ref;
} glp_mh_desc;

/* Bytes of the device emission table for codes of `bits` bits. */
size_t glp_emit_lut_bytes(unsigned bits, int precision);

/* Builds the emission table FillEmit(codebook[c]) (hmmLike.cu:12-44) for
 * every code and every combination of the 4 parents' alleles, and copies it
 * to lut_dev (device, glp_emit_lut_bytes bytes).  codebook and pc16 are host
 * arrays. */
int glp_build_emit_lut(const float *codebook, unsigned n_codes,
                       const float *pc16, int precision, void *lut_dev);

/* Plain path, step 0: draws 4 fresh parents per sample (insti.cpp:1493-1494;
 * kept from the input under GLP_PROP_HOST) and scores them. */
int glp_plain_init(const glp_mh_desc *desc, void *stream);

/* Plain path: one launch = one MH iteration = proposal c of every sample of
 * every window (findHapSet's cycle body, hmmLike.cu:185-207, with the CPU
 * accept rule of MHSampler.hpp:82-96). */
int glp_plain_step(const glp_mh_desc *desc, unsigned c, void *stream);

/* Batched form: a persistent work-queue kernel runs init + all C proposals of
 * every (window, sample) chain, windows longest first; each round,
 * lanes_per_chain forward passes (each on a pair of lanes; a quad under
 * GLP_PREC_MIXED_FB) evaluate consecutive proposals of one chain
 * speculatively from the same state, and a sequential walk keeps exactly the
 * serial MH result.  lanes_per_chain is 1..32 for every window (1..16 under
 * GLP_PREC_MIXED_FB: a wavefront holds floor(32 / lanes) chains, or
 * floor(16 / lanes); 0 = the deepest value whose wavefronts fit one per
 * SIMD, or under GLP_PREC_MIXED_FB the deepest whose chains fit the device
 * in 8-wavefront workgroups with split chain slots (glp_plan_set_split)
 * where that is deeper; see glp_plan_autotune).  Under GLP_PREC_MIXED_FB
 * the plan allocates
 * each window's forward-backward layout (about its panel and codes once
 * more) and every run rebuilds it from the windows' current data first. */
typedef struct glp_plan glp_plan;
int glp_plan_create(const glp_mh_desc *desc, unsigned lanes_per_chain,
                    glp_plan **out);
/* Runs one generation with the given gen / pen / seed (other fields from
 * create).  Stream-ordered; capturable into a HIP graph. */
int glp_plan_run(glp_plan *plan, unsigned gen, float pen, uint64_t seed,
                 void *stream);
/* Counters of the last completed run: forward passes evaluated (useful and
 * speculative) and rounds.  Synchronises the plan's stream. */
int glp_plan_stats(glp_plan *plan, uint64_t *evals, uint64_t *useful);
/* Lanes per chain of the longest window and grid size the plan launches
 * with. */
int glp_plan_lanes(glp_plan *plan, unsigned *lanes_per_chain, unsigned *grid);
/* Sets passes per chain round per window (1..32; 1..16 under GLP_PREC_MIXED_FB). */
int glp_plan_set_lanes(glp_plan *plan, const unsigned *lanes_per_window);
/* Wavefronts per workgroup of the batched kernel: 4 (the default; one
 * workgroup per CU with the 8-bit emission table) or 8 (the workgroup's 8
 * wavefronts share one copy of the emission and spread tables, two per
 * SIMD; refused with GLP_ERR_ARG where its LDS does not fit a CU).  The
 * autotuner tries both. */
int glp_plan_set_waves(glp_plan *plan, unsigned waves_per_workgroup);
int glp_plan_waves(glp_plan *plan, unsigned *waves_per_workgroup);
/* Uneven chain slots at 8 wavefronts per workgroup (split = 1): wavefronts
 * 4-7 of each workgroup hold one chain slot fewer than wavefronts 0-3
 * (floor(units / lanes) - 1, at least one), so a generation's chains spread
 * over more workgroups and, where a SIMD runs one wavefront of each half,
 * the one with fewer chains finishes first and leaves the SIMD's issue slots
 * to the other.  Results do not depend on it; ignored at 4 wavefronts.  The
 * autotuner tries both.  (No counterpart in hmmLike.h: the reference kernel
 * runs one chain per thread.) */
int glp_plan_set_split(glp_plan *plan, int split);
int glp_plan_split(glp_plan *plan, int *split);
/* Autotuner: times whole generations of the plan (every window, as a real
 * run shares the device) with HIP events on `stream`: first one
 * lanes-per-chain for every window from 1..8, 10, 12, 16, 32 (at most the
 * proposals + 1) at 4 and at 8 wavefronts per workgroup (at 8 with and
 * without the uneven chain slots of glp_plan_set_split), then, with several
 * window-size buckets (ceil(log2 sites)), each bucket's value in turn with
 * the others held and the workgroup size and split fixed; keeps the fastest.
 * A trial's time is the mean over generations gen .. gen + 3, each run from
 * the same saved state: window outputs (and inputs of GLP_PROP_HOST) are
 * restored after every run.  lanes_out (optional) receives the choice per window.
 * Synchronises `stream`. */
int glp_plan_autotune(glp_plan *plan, unsigned gen, float pen, uint64_t seed,
                      void *stream, unsigned *lanes_out);
/* Plain path over the plan's windows: one launch for the fresh parents, then
 * one launch per MH iteration (num_cycles launches), all stream-ordered. */
int glp_plan_run_plain(glp_plan *plan, unsigned gen, float pen, uint64_t seed,
                       void *stream);
int glp_plan_destroy(glp_plan *plan);

/* The same generations through a HIP graph: the plan builds the graph of one
 * generation on first use (batched: counter reset + persistent kernel;
 * plain: the C + 1 launches of glp_plan_run_plain), and each call rewrites
 * the nodes' (gen, pen, seed) and launches it on `stream`.  Results equal the
 * direct calls'.  Changing the plan's depths rebuilds the graph. */
int glp_plan_graph_run(glp_plan *plan, unsigned gen, float pen, uint64_t seed,
                       void *stream);
int glp_plan_graph_run_plain(glp_plan *plan, unsigned gen, float pen, uint64_t seed,
                             void *stream);

/* One batched generation that allocates nothing: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result as glp_plan_create(desc, lanes_per_chain) +
 * glp_plan_run(plan, desc->gen, desc->pen, desc->seed): the window table,
 * window order, lanes table, task counters and statistics live in a
 * 61,696-byte device array of the library (one per device, held until the
 * process ends) instead of plan buffers.  At most GLP_NOALLOC_MAX_WINDOWS
 * windows (larger descriptors are refused with GLP_ERR_ARG); the caller's
 * window buffers, emission table and traces are used as given.
 * Stream-ordered; with stats_out (host, 3 values: passes evaluated, useful
 * passes, round-bound flag) the call waits for `stream` and reports the
 * run's counters, returning GLP_ERR_STATE if the round bound was hit.
 * Every call on a device shares that device's array and leaves the last
 * call's contents in it: calls that overlap on one device (on different
 * streams or threads) race. */
#define GLP_NOALLOC_MAX_WINDOWS 1024
int glp_mh_run_noalloc(const glp_mh_desc *desc, unsigned lanes_per_chain, void *stream,
                       uint64_t *stats_out);

/* Self-test of the device reciprocals of the rescale and resampling steps
 * (no counterpart in the reference): compares them with the IEEE division
 * 1.0f / x for every float x with bit pattern in [lo_bits, hi_bits) on the
 * current device (the scaled one bit-equal below 2^-20; the short one
 * bit-equal on [2^-126, 2^126] and NaN for subnormal x, which sends the
 * work to an exact division).
 * mismatches receives the count, first_bad the lowest mismatching pattern
 * (0xFFFFFFFF if none).  Synchronous. */
int glp_selftest_recip(uint32_t lo_bits, uint32_t hi_bits, uint64_t *mismatches,
                       uint32_t *first_bad);

/* ----- haplotype resampling (Impute::hmm_work, impute.cpp:380-485) ------- */

/* One window of the resampling step that follows the MH update of a
 * generation (insti.cpp:1448-1452).  All pointers are device pointers. */
typedef struct {
  unsigned num_sites;
  unsigned words_per_hap;    /* ceil(num_sites / 64) */
  const uint64_t *panel;     /* [num_haps][wn] haplotypes the parents index */
  const unsigned *parents;   /* [N][4] chosen parents (f0, f1, m0, m1) */
  const float *tran;         /* [M][3] */
  const float *het_aa;       /* [M][N][2] input GL pairs (Het, AA), site-major */
  const float *emit;         /* [M][N][4] emission table of glp_work_emit (from the same
                                het_aa and pc16), or NULL: glp_hmm_work then builds it
                                in the workspace on every call */
  uint64_t *hnew;            /* out: [2N][wn], rows 2i, 2i+1 = sample i's */
} glp_work_window;

typedef struct {
  unsigned num_windows;
  const glp_work_window *windows; /* host array of num_windows */
  unsigned num_samples;           /* samples per window */
  unsigned num_haps;              /* rows of each panel */
  float pc16[16];                 /* mutation matrix pc[i][j] at 4 i + j */
  float pen;                      /* S of impute.cpp:457-469 */
  uint64_t seed;
  unsigned gen;
  void *workspace;                /* device, >= glp_work_workspace_bytes */
} glp_work_desc;

/* Device workspace the resampling needs: 32 B per sample and site of every
 * window (the scaled backward and forward probabilities, computed side by
 * side), plus 16 B per sample and site of every window whose emit is NULL. */
size_t glp_work_workspace_bytes(const glp_work_desc *desc);

/* Insti::initialize's emission table (insti.cpp:1345-1378) of one window,
 * site-major: emit[(m N + i) 4 + j] = pc[j][0] t0 + pc[j][1] t1 +
 * pc[j][2] t1 + pc[j][3] t2 with t = (max(1 - Het - AA, 0), max(Het, 0),
 * max(AA, 0)) of sample i's GL pair at site m, het_aa[m][i] (the reference's
 * per-sample emit arrays, [M][4] each, interleaved by site).  It depends on
 * the GLs and pc only, so a run builds it once and every generation's
 * glp_hmm_work reads it.  het_aa and emit are device pointers, pc16 a host
 * array (pc[i][j] at 4 i + j).  Stream-ordered. */
int glp_work_emit(const float *het_aa, unsigned num_sites, unsigned num_samples,
                  const float *pc16, float *emit, void *stream);

/* Impute::hmm_work for every sample of every window: scaled backward
 * recursion, forward pass with per-site posteriors, genotype weights
 * powf(prob_g * sum_xy p_xy pc[s_xy][g], pen) and one draw per site (word
 * m & 3 of Philox block 0x80000000 | (m >> 2) for (sample, window, gen),
 * replacing gsl_rng_uniform).  prob and emissions come from the GL pairs as
 * in Insti::initialize (insti.cpp:1345-1378).  Writes every hnew row in
 * full (bits past num_sites are 0).  Stream-ordered. */
int glp_hmm_work(const glp_work_desc *desc, void *stream);

/* ----- generation bookkeeping (Insti::initialize, Impute::replace) -------- */

/* Insti::initialize's random initial haplotypes (insti.cpp:1316-1365) of one
 * window: for sample i and site m, t = (max(1 - Het - AA, 0), max(Het, 0),
 * max(AA, 0)) from het_aa [M][N][2] (site-major GL pairs); the three
 * gsl_rng_uniform draws become words 0, 1, 2 of Philox block
 * {0xFFFFFFFE, i, m, window} under `seed` (doubles of [0, 1), 32 bits):
 * u0 < t0 gives 0|0; else u1 < t1 / (t1 + t2) a heterozygote, the
 * alternative allele on hap 2i when u2 < 0.5, else on hap 2i + 1; else 1|1.
 * Writes rows [0, 2N) of haps [.][ceil(M / 64)] in full (bits past M are 0).
 * The male chromosome-X branch is not supported.  Device pointers,
 * stream-ordered. */
int glp_init_haps(const float *het_aa, unsigned num_samples, unsigned num_sites, uint64_t seed,
                  unsigned window, uint64_t *haps, void *stream);

/* Impute::replace (impute.cpp:564-597) for every sample of one window: the
 * running tallies hsum [2N][M] (u32; rows 2i, 2i + 1 are sample i's ha, hb)
 * count haps rows 2i, 2i + 1 in the orientation that agrees more with them.
 * Call once per sampling generation after the haplotypes are updated; the
 * allele probabilities of the output are hsum / sampling generations
 * (Impute::result, impute.cpp:599-613).  Device pointers, stream-ordered. */
int glp_replace(const uint64_t *haps, unsigned num_samples, unsigned num_sites, uint32_t *hsum,
                void *stream);

/* ----- data formats (GLReader, Insti::save_vcf) -------------------------- */

typedef struct glp_stbin glp_stbin;

/* GLReader's SNPTools .bin path (glReader.cpp:25-90, 201-285): a gzip or
 * plain tab-separated file, header "chr pos allele name...", one line per
 * site with "Het AA" per sample.  Keeps the sites inside `region`
 * ("chrom:start-end", either bound may be empty; NULL or "" = all), at most
 * size_limit of them (0 = no limit; the rest are counted as not read), and
 * the samples named in the file keep_names_path (one name a line; NULL = all).
 * Errors as the reference: wrong column count, unsorted positions, a second
 * chromosome, a repeated site, an unparsable region. */
int glp_stbin_read(const char *path, const char *region, size_t size_limit,
                   const char *keep_names_path, glp_stbin **out);
int glp_stbin_dims(const glp_stbin *b, unsigned *num_samples, unsigned *num_sites,
                   size_t *num_not_read);
const char *glp_stbin_name(const glp_stbin *b, unsigned i);
int glp_stbin_site(const glp_stbin *b, unsigned m, const char **chrom, unsigned *pos,
                   const char **ref, const char **alt);
/* GLs site-major: drop_first = 0 gives [M][N][3] (max(0, 1 - Het - AA), Het,
 * AA) (gl_ret_t::STANDARD); drop_first = 1 gives [M][N][2] (Het, AA) / sum
 * (gl_ret_t::ST_DROP_FIRST, the layout Insti and glp_hmm_work use). */
int glp_stbin_gls(const glp_stbin *b, int drop_first, float *out);
void glp_stbin_free(glp_stbin *b);

/* Insti::initialize's transitions without a genetic map (insti.cpp:1253-1300):
 * mu from the harmonic number of hn haplotypes, rho = 0.5 mu (M - 1) /
 * (pos[M-1] - pos[0]) / density, r = d rho / (d rho + hn) for the distance d
 * to the previous site, times inflation, capped at 1.  tran_out [M][3]. */
int glp_tran_from_positions(const unsigned *pos, unsigned num_sites, unsigned hn,
                            double density, double inflation, float *tran_out);

/* Impute::result + Insti::save_vcf (impute.cpp:599-613, insti.cpp:2098-2167):
 * VCFv4.2 with GT:GP:APP per sample from the allele tallies hsum (host,
 * [2N][M], glp_replace) over sampling_gens generations; gzip when path ends
 * in ".gz".  The ##source line names this library. */
int glp_write_vcf(const char *path, unsigned num_samples, const char *const *names,
                  unsigned num_sites, const char *const *chrom, const unsigned *pos,
                  const char *const *ref, const char *const *alt, const uint32_t *hsum,
                  unsigned sampling_gens, const char *command_line);

/* ----- kNN candidate lists (KNN, kNN.cpp:8-225) --------------------------- */

/* Neighbour lists of samples 0 .. num_samples-1 (haplotypes 2u, 2u+1 of the
 * panel) over the whole panel, by Hamming distance on the sites whose
 * variant allele frequency over the panel (alt count / num_haps, a double;
 * the minor allele frequency with using_maf) lies in [freq_lb, freq_ub]
 * (KNN::CalculateVarAfs / CutDownHaps / ClusterHaps).  k' = ceil(K / 2), K
 * capped at num_haps - 2: row u = (A_0, B_0, A_1, B_1, ...), A_j the j-th
 * nearest haplotype of hap 2u, B_j the j-th nearest of hap 2u+1 not already
 * in the row; the sample's own haplotypes never appear.  Equal distances
 * rank by haplotype index (std::sort leaves them unspecified; below 17
 * candidates libstdc++ sorts stably, which is the same order). */
typedef struct {
  const uint64_t *panel;  /* device [num_haps][ceil(num_sites / 64)] */
  unsigned num_haps;      /* even */
  unsigned num_sites;
  unsigned num_samples;   /* <= num_haps / 2 */
  unsigned K;             /* 1 .. 2048 */
  double freq_lb, freq_ub;
  int using_maf;
  unsigned *klist;        /* out, device [num_samples][glp_knn_k_out(K, num_haps)] */
  void *workspace;        /* device, >= glp_knn_workspace_bytes */
} glp_knn_desc;

/* Row length of the lists: 2 * ceil(min(K, num_haps - 2) / 2). */
unsigned glp_knn_k_out(unsigned K, unsigned num_haps);
size_t glp_knn_workspace_bytes(const glp_knn_desc *desc);
/* Builds the lists (stream-ordered). */
int glp_knn_build(const glp_knn_desc *desc, void *stream);

/* ===== 3. host-side helpers of the drop-in ================================ */

/* Insti::initialize parameter builders (insti.cpp:1254-1315): mu from the
 * harmonic number of hn haplotypes, the 4x4 mutation matrix pc (row-major)
 * and tran[M][3] from per-site recombination distances in Morgans
 * (r_morgans[m] between sites m-1 and m; geneticMap.cpp:105-111) times
 * `inflation`, capped at 1.  Any output pointer may be NULL. */
int glp_model_params(unsigned hn, unsigned num_sites, const double *r_morgans,
                     double inflation, float *tran_out, float *pc16_out,
                     float *mu_out);

/* GLVQ::BuildCodeBook (glVQ.cpp:26-146) on n_gl GL trios: Lloyd iterations
 * from the 4 fixed codewords plus Philox-drawn ones (key), until the
 * variance changes by <= 1%.  training_size 0 trains on every GL, as the
 * reference; otherwise on a Philox-drawn subsample of that size. */
int glp_vq_train(const float *gls, size_t n_gl, unsigned n_codes, uint64_t key,
                 size_t training_size, float *codebook_out, unsigned *iters);

/* GLPack::GLs2VQChar + GLVQ::FindGLCode (glPack.cpp:195-219,
 * glVQ.cpp:156-165) for every GL trio. */
int glp_vq_codes(const float *gls, size_t n_gl, const float *codebook,
                 unsigned n_codes, uint8_t *codes_out);

/* glp_vq_train with the nearest-codeword search of every Lloyd iteration
 * (GLVQ::AssignPoints, glVQ.cpp:149-163) on the current device; the
 * centroid and variance sums stay sequential on the host in point order, so
 * codebook and iteration count equal glp_vq_train's.  Synchronous. */
int glp_vq_train_device(const float *gls, size_t n_gl, unsigned n_codes, uint64_t key,
                        size_t training_size, float *codebook_out, unsigned *iters);

/* glp_vq_codes on the current device (host buffers in and out). */
int glp_vq_codes_device(const float *gls, size_t n_gl, const float *codebook,
                        unsigned n_codes, uint8_t *codes_out);

/* The search kernel itself on device buffers: codes_dev[p] = nearest
 * codeword of point p, where points_dev holds GL trios (raw != 0,
 * normalised by (RR + Het) + AA as glp_vq_codes) or normalised (RR, Het)
 * pairs (raw == 0); codebook is a host array of n_codes (RR, Het) pairs.
 * Enqueued on `stream`. */
int glp_vq_assign_device(const float *points_dev, size_t n, int raw, const float *codebook,
                         unsigned n_codes, uint8_t *codes_dev, void *stream);

/* GLPack::GetPackedGLs (glPack.cpp:32-192) for samples [s0, s1) of codes
 * [n_samples][num_sites]; out = [ceil(num_sites*bits/32)][s1-s0]. */
int glp_pack_codes(const uint8_t *codes, unsigned n_samples,
                   unsigned num_sites, unsigned bits, unsigned s0, unsigned s1,
                   uint32_t *out);

/* glp_pack_codes on device buffers (codes_dev [n_samples][num_sites] bytes,
 * out_dev [ceil(num_sites*bits/32)][s1-s0] words), bits 1, 2, 4 or 8, the
 * same words; enqueued on `stream`. */
int glp_pack_codes_device(const uint8_t *codes_dev, unsigned n_samples,
                          unsigned num_sites, unsigned bits, unsigned s0, unsigned s1,
                          uint32_t *out_dev, void *stream);

/* The same without VQ: two 4-bit halves floor((GL - eps) * 16) per trio
 * (glPack.cpp:221-243). */
int glp_pack_nonvq(const float *gls, unsigned n_samples, unsigned num_sites,
                   unsigned s0, unsigned s1, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif

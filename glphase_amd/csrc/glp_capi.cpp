// glp_capi.cpp -- host side of the C ABI (include/glphase_amd.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/glphase_amd.h"
#include "glp_common.hpp"
#include "glp_internal.hpp"

namespace glp {

static thread_local std::string g_err;

int set_error(int code, const char *msg) {
  g_err = msg ? msg : "";
  return code;
}

int hip_check(hipError_t err, const char *what) {
  if (err == hipSuccess) return GLP_OK;
  g_err = std::string(what) + ": " + hipGetErrorString(err);
  return GLP_ERR_HIP;
}

#define GLP_TRY(expr)            \
  do {                           \
    int rc_ = (expr);            \
    if (rc_ != GLP_OK) return rc_; \
  } while (0)

// FillEmit (hmmLike.cu:39-44; same order as insti.cpp:1376-1378):
// E[c][i] = pc[i][0]*GL0 + pc[i][1]*GL1 + pc[i][2]*GL1 + pc[i][3]*GL2 with
// GL2 = max(1 - GL0 - GL1, 0) (hmmLike.cu:27).
void emit_table(const float *codebook, unsigned ncode, const float *pc, float *E) {
  for (unsigned c = 0; c < ncode; ++c) {
    const float g0 = codebook[2 * c], g1 = codebook[2 * c + 1];
    const float g2r = 1 - g0 - g1;
    const float g2 = g2r > 0.0f ? g2r : 0.0f;
    for (unsigned i = 0; i < 4; ++i)
      E[4 * c + i] = pc[4 * i + 0] * g0 + pc[4 * i + 1] * g1 + pc[4 * i + 2] * g1 +
                     pc[4 * i + 3] * g2;
  }
}

// (code, nib) -> (e00, e11, e01, e10), nib = f0 | f1<<1 | m0<<2 | m1<<3,
// e_xy = E[code][(f_x << 1) | m_y]: the states of the even lane of a lane
// pair first, then those of the odd lane.
void nibble_table(const float *E, unsigned ncode, float *out) {
  for (unsigned c = 0; c < ncode; ++c)
    for (unsigned nib = 0; nib < 16; ++nib) {
      const unsigned f0 = nib & 1u, f1 = (nib >> 1) & 1u, m0 = (nib >> 2) & 1u,
                     m1 = (nib >> 3) & 1u;
      float *o = out + 4 * (16 * (size_t)c + nib);
      o[0] = E[4 * c + ((f0 << 1) | m0)];
      o[1] = E[4 * c + ((f1 << 1) | m1)];
      o[2] = E[4 * c + ((f0 << 1) | m1)];
      o[3] = E[4 * c + ((f1 << 1) | m0)];
    }
}

static float norm_value() { return powf(FLT_MIN, 2.0f / 3.0f); }

static bool valid_bits(unsigned bits) { return bits == 4 || bits == 8; }

// grow-only device buffer
struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return GLP_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    GLP_TRY(hip_check(hipMalloc(&p, bytes ? bytes : 1), "hipMalloc"));
    n = bytes;
    return GLP_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

}  // namespace glp

using namespace glp;

// ---------------------------------------------------------------------------
// batched plan
// ---------------------------------------------------------------------------
struct glp_plan {
  int dev = 0, num_cu = 1, bpc = 1;  // bpc: resident workgroups per CU at wpb
  unsigned wpb = 4;                  // wavefronts per workgroup: 4 or 8
  int bpc_of[2] = {0, 0};            // resident workgroups per CU at 4 and 8 (0: does not fit)
  KArgs a{};
  unsigned nwin = 0, grid = 1;
  std::vector<glp_window> wins_h;
  std::vector<unsigned> order, q;  // windows longest first; passes per chain round
  DevBuf wins, order_d, qtab, wctr, ctrs;
  // forward-backward form: every window's layout (fbdata) and the mirror
  // kernel's jobs (fbjobs), rebuilt from the windows before every run
  DevBuf fbdata, fbjobs;
  unsigned fb_items_max = 0;
  // HIP graph of one generation: kind 1 = batched (reset + k_batched), kind 2
  // = plain (C + 1 launches), each after the layout build in the
  // forward-backward form; node arguments are rewritten for each run
  int graph_kind = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipGraphNode_t> nodes;
  size_t node0 = 0;  // index of the first rewritten node (k_batched, or the plain step 0)
};

static bool plan_fb(const glp_plan *p) { return p->a.precision == GLP_PREC_MIXED_FB; }

// Builds every window's forward-backward layout on stream s (no-op in the
// other forms).
static int plan_mirror(glp_plan *p, hipStream_t s) {
  if (!plan_fb(p)) return GLP_OK;
  return launch_fb_mirror(p->fbjobs.as<FbWin>(), p->nwin, p->fb_items_max, p->a.N, p->a.H,
                          p->a.bits, s);
}

static void plan_graph_reset(glp_plan *p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  p->nodes.clear();
  p->graph_kind = 0;
}

// Passes per chain round (pass units per chain) when none is given: the
// deepest speculation whose wavefronts still fit one per SIMD.  A chain's
// time is its rounds times one pass, and rounds fall with depth while the
// SIMDs are not shared; past one wavefront per SIMD two wavefronts split the
// issue slots and every pass slows down.
// A wavefront holds units(precision) pass units (32 lane pairs, or 16 lane
// quads in the forward-backward form), floor(units / Q) chain slots.
constexpr unsigned kMaxPasses = 32;
constexpr size_t kLdsPerCu = 160 * 1024;  // gfx950

static unsigned units(int precision) { return precision == GLP_PREC_MIXED_FB ? 16u : 32u; }

static double waves_for(unsigned nchains, unsigned Q, unsigned nu) {
  return std::ceil((double)nchains / (double)(nu / Q));
}

static unsigned auto_lanes(unsigned nchains, unsigned C, int num_cu, unsigned nu) {
  const double target = (double)num_cu * 4.0;  // one wave per SIMD
  unsigned best = 1;
  for (unsigned Q = 1; Q <= nu && Q <= C + 1; ++Q)
    if (waves_for(nchains, Q, nu) <= target) best = Q;
  return best;
}

// Chain slots of an 8-wavefront workgroup with the split, on a window of Q
// passes per chain round: wavefronts 0-3 hold floor(units / Q) each,
// wavefronts 4-7 one fewer (at least one).
static unsigned split_wg_slots(unsigned nu, unsigned Q) {
  const unsigned s = nu / Q;
  return 4 * s + 4 * (s > 1 ? s - 1 : s);
}

// The deepest Q whose chains, in 8-wavefront workgroups with the split, fit
// max_wgs resident workgroups at once (0: none does).
static unsigned auto_lanes_split(unsigned nchains, unsigned C, unsigned max_wgs, unsigned nu) {
  unsigned best = 0;
  for (unsigned Q = 1; Q <= nu && Q <= C + 1; ++Q)
    if (std::ceil((double)nchains / (double)split_wg_slots(nu, Q)) <= (double)max_wgs) best = Q;
  return best;
}

static unsigned plan_grid(const glp_plan *p, const std::vector<unsigned> &wins) {
  const bool split = p->wpb == 8 && p->a.split;
  double wgs = 0;
  for (unsigned w : wins)
    wgs += split ? (double)p->a.N / (double)split_wg_slots(units(p->a.precision), p->q[w])
                 : waves_for(p->a.N, p->q[w], units(p->a.precision)) / (double)p->wpb;
  const unsigned blocks = (unsigned)std::ceil(wgs);
  return std::max(1u, std::min((unsigned)(p->bpc * p->num_cu), blocks));
}

// Switches the plan to workgroups of `wpb` wavefronts (4 or 8); the
// occupancy at that size was measured at creation.
static int plan_use_waves(glp_plan *p, unsigned wpb) {
  if (wpb != 4 && wpb != 8) return set_error(GLP_ERR_ARG, "waves per workgroup must be 4 or 8");
  const int bpc = p->bpc_of[wpb == 8 ? 1 : 0];
  if (bpc < 1) return set_error(GLP_ERR_ARG, "workgroups of this many wavefronts do not fit a CU (LDS)");
  if (wpb != p->wpb) plan_graph_reset(p);  // the graph's node holds the kernel and block size
  p->wpb = wpb;
  p->bpc = bpc;
  return GLP_OK;
}

static int plan_upload_q(glp_plan *p) {
  return hip_check(hipMemcpy(p->qtab.p, p->q.data(), 4 * p->nwin, hipMemcpyHostToDevice),
                   "lanes table copy");
}

static int fill_args(const glp_mh_desc *d, KArgs &a) {
  if (!d || !d->windows || d->num_windows == 0)
    return set_error(GLP_ERR_ARG, "descriptor has no windows");
  if (!valid_bits(d->bits_per_code)) return set_error(GLP_ERR_ARG, "bits_per_code must be 4 or 8");
  if (!d->emit_lut) return set_error(GLP_ERR_STATE, "emit_lut not built");
  if (d->num_samples == 0) return set_error(GLP_ERR_ARG, "num_samples is 0");
  if (d->prop_mode == GLP_PROP_UNIFORM && d->num_haps < 2 * d->num_samples)
    return set_error(GLP_ERR_ARG, "num_haps < 2 * num_samples");
  if (d->prop_mode == GLP_PROP_KLIST && (!d->klist || d->K == 0))
    return set_error(GLP_ERR_ARG, "klist mode needs klist and K > 0");
  if (d->prop_mode == GLP_PROP_HOST && d->num_cycles > 0 && !d->extra)
    return set_error(GLP_ERR_ARG, "host proposal mode needs extra");
  if (d->prop_mode < 0 || d->prop_mode > 2) return set_error(GLP_ERR_ARG, "bad prop_mode");
  // main.cpp:430-433: kickstarting needs a reference panel
  if (d->kickstart_from_ref && d->prop_mode != GLP_PROP_HOST &&
      d->num_haps <= 2 * d->num_samples)
    return set_error(GLP_ERR_ARG, "kickstart_from_ref needs reference haplotypes beyond 2N");
  if (d->precision < 0 || d->precision > 2)
    return set_error(GLP_ERR_ARG,
                     "precision must be GLP_PREC_F32, GLP_PREC_MIXED or GLP_PREC_MIXED_FB");
  if (d->k0 + (uint64_t)d->num_cycles >= 0x80000000ull)
    return set_error(GLP_ERR_ARG, "proposal counters must stay below 2^31");
  for (unsigned w = 0; w < d->num_windows; ++w) {
    const glp_window &W = d->windows[w];
    if (W.num_sites == 0) return set_error(GLP_ERR_ARG, "window with 0 sites");
    if (W.words_per_hap != (W.num_sites + 63) / 64)
      return set_error(GLP_ERR_ARG, "words_per_hap must be ceil(num_sites / 64)");
    if (!W.panel || !W.packed || !W.tran || !W.parents || !W.like)
      return set_error(GLP_ERR_ARG, "window with a null pointer");
    // the kernels read a window's panel and packed codes through buffer
    // descriptors with 32-bit byte offsets
    const uint64_t code_words = ((uint64_t)W.num_sites * d->bits_per_code + 31) / 32;
    if (code_words * d->num_samples * 4 >= (1ull << 32))
      return set_error(GLP_ERR_ARG, "a window's packed codes reach 4 GiB");
    if ((uint64_t)d->num_haps * W.words_per_hap * 8 >= (1ull << 32))
      return set_error(GLP_ERR_ARG, "a window's panel reaches 4 GiB");
    if (d->precision == GLP_PREC_MIXED_FB)
      GLP_TRY(fb_check(W.num_sites, d->bits_per_code, d->num_samples, d->num_haps));
  }
  a.N = d->num_samples;
  a.H = d->num_haps;
  a.C = d->num_cycles;
  a.k0 = d->k0;
  a.pen = d->pen;
  a.norm = norm_value();
  a.seed = d->seed;
  a.gen = d->gen;
  a.prop_mode = d->prop_mode;
  a.klist = d->klist;
  a.K = d->K;
  a.extra = d->extra;
  a.trace_like = d->trace_like;
  a.trace_acc = d->trace_acc;
  a.lut = d->emit_lut;
  a.bits = d->bits_per_code;
  a.lut_entries = 16u << d->bits_per_code;
  a.precision = d->precision;
  a.kick = d->kickstart_from_ref ? 1 : 0;
  a.split = 0;
  return GLP_OK;
}

// Plain-path launches need the window table on the device; one short-lived
// copy per call keeps the calls stateless.
static int plain_launch(const glp_mh_desc *desc, unsigned step, void *stream) {
  KArgs a;
  GLP_TRY(fill_args(desc, a));
  if (a.precision == GLP_PREC_MIXED_FB)
    return set_error(GLP_ERR_ARG, "the forward-backward form runs through a plan "
                                  "(glp_plan_run_plain): its window layouts are allocated");
  hipStream_t s = static_cast<hipStream_t>(stream);
  glp_window *d_wins = nullptr;
  const size_t bytes = sizeof(glp_window) * desc->num_windows;
  GLP_TRY(hip_check(hipMallocAsync((void **)&d_wins, bytes, s), "hipMallocAsync"));
  int rc = hip_check(hipMemcpyAsync(d_wins, desc->windows, bytes, hipMemcpyHostToDevice, s),
                     "window table copy");
  if (rc == GLP_OK) rc = launch_plain(a, d_wins, desc->num_windows, step, s);
  (void)hipFreeAsync(d_wins, s);
  return rc;
}

extern "C" {

const char *glp_last_error(void) { return g_err.c_str(); }
const char *glp_version(void) { return "glphase_amd 0.1 (gfx950)"; }

size_t glp_emit_lut_bytes(unsigned bits, int precision) {
  (void)precision;
  if (!valid_bits(bits)) return 0;
  return (size_t)16 * 16 << bits;
}

int glp_build_emit_lut(const float *codebook, unsigned n_codes, const float *pc16,
                       int precision, void *lut_dev) {
  if (!codebook || !pc16 || !lut_dev) return set_error(GLP_ERR_ARG, "null argument");
  if (n_codes != 16 && n_codes != 256) return set_error(GLP_ERR_ARG, "n_codes must be 16 or 256");
  if (precision < 0 || precision > 2) return set_error(GLP_ERR_ARG, "bad precision");
  std::vector<float> E(4 * (size_t)n_codes), T(64 * (size_t)n_codes);
  emit_table(codebook, n_codes, pc16, E.data());
  nibble_table(E.data(), n_codes, T.data());
  return hip_check(hipMemcpy(lut_dev, T.data(), T.size() * sizeof(float), hipMemcpyHostToDevice),
                   "emit table copy");
}

int glp_plain_init(const glp_mh_desc *desc, void *stream) {
  return plain_launch(desc, 0, stream);
}

int glp_plain_step(const glp_mh_desc *desc, unsigned c, void *stream) {
  if (desc && c >= desc->num_cycles) return set_error(GLP_ERR_ARG, "cycle index >= num_cycles");
  return plain_launch(desc, c + 1, stream);
}

int glp_plan_create(const glp_mh_desc *desc, unsigned lanes_per_chain, glp_plan **out) {
  if (!out) return set_error(GLP_ERR_ARG, "null out");
  *out = nullptr;
  KArgs a;
  GLP_TRY(fill_args(desc, a));
  const unsigned nu = units(a.precision);
  if (lanes_per_chain > nu)
    return set_error(GLP_ERR_ARG, nu == 16u ? "lanes_per_chain must be 0 (automatic) or 1..16"
                                            : "lanes_per_chain must be 0 (automatic) or 1..32");
  glp_plan *p = new glp_plan();
  p->a = a;
  int rc = hip_check(hipGetDevice(&p->dev), "hipGetDevice");
  if (rc == GLP_OK)
    rc = hip_check(hipDeviceGetAttribute(&p->num_cu, hipDeviceAttributeMultiprocessorCount, p->dev),
                   "device attribute");
  // occupancy at both workgroup sizes; a size whose LDS exceeds the CU's is
  // not queried and cannot be chosen
  for (int k = 0; k < 2 && rc == GLP_OK; ++k) {
    const unsigned w = k ? 8u : 4u;
    if (lds_bytes(a, w) <= kLdsPerCu)
      rc = batched_occupancy(a.bits, a.precision, w, lds_bytes(a, w), &p->bpc_of[k]);
  }
  if (rc != GLP_OK) {
    delete p;
    return rc;
  }
  // a launch that cannot be resident is refused here rather than left to
  // fail (or to run with a grid sized from a bad occupancy figure)
  if (plan_use_waves(p, 4) != GLP_OK) {
    delete p;
    return set_error(GLP_ERR_ARG, "the batched kernel's workgroup does not fit a CU (LDS)");
  }
  p->nwin = desc->num_windows;
  p->wins_h.assign(desc->windows, desc->windows + p->nwin);
  p->order.resize(p->nwin);
  for (unsigned w = 0; w < p->nwin; ++w) p->order[w] = w;
  std::stable_sort(p->order.begin(), p->order.end(), [&](unsigned x, unsigned y) {
    return desc->windows[x].num_sites > desc->windows[y].num_sites;
  });
  unsigned Q = lanes_per_chain ? lanes_per_chain : auto_lanes(a.N * p->nwin, a.C, p->num_cu, nu);
  if (!lanes_per_chain && plan_fb(p) && p->bpc_of[1] >= 1) {
    // the forward-backward form: 8-wavefront workgroups with split chain
    // slots where they reach a deeper speculation that still fits the device
    // (at 5,000 chains: depth 5 on 250 workgroups against depth 3 at one
    // wavefront per SIMD)
    const unsigned Q8 =
        auto_lanes_split(a.N * p->nwin, a.C, (unsigned)(p->bpc_of[1] * p->num_cu), nu);
    if (Q8 > Q && plan_use_waves(p, 8) == GLP_OK) {
      p->a.split = 1;
      Q = Q8;
    }
  }
  p->q.assign(p->nwin, Q);
  p->grid = plan_grid(p, p->order);
  rc = p->wins.ensure(sizeof(glp_window) * p->nwin);
  if (rc == GLP_OK) rc = p->order_d.ensure(4 * p->nwin);
  if (rc == GLP_OK) rc = p->qtab.ensure(4 * p->nwin);
  if (rc == GLP_OK) rc = p->wctr.ensure(4 * p->nwin);
  if (rc == GLP_OK) rc = p->ctrs.ensure(4 * sizeof(unsigned long long));
  // the kernels read the windows through this table: in the forward-backward
  // form its panel, codes and tran point at the windows' layouts
  std::vector<glp_window> ktab(desc->windows, desc->windows + p->nwin);
  if (rc == GLP_OK && plan_fb(p)) {
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    std::vector<size_t> off(p->nwin);
    size_t total = 0;
    // per window: panel, codes and tran parts, each 256-byte aligned
    auto parts = [&](unsigned M, size_t *pb, size_t *cb) {
      const unsigned n = (M + 1) / 2, wh = (n + 63) / 64, nw = (n * a.bits + 31) / 32;
      *pb = up(16 * (size_t)a.H * wh);
      *cb = up(8 * (size_t)nw * a.N);
      return *pb + *cb + up(fb_layout_bytes(M, a.bits, a.N, a.H) - 16 * (size_t)a.H * wh -
                            8 * (size_t)nw * a.N);
    };
    for (unsigned w = 0; w < p->nwin; ++w) {
      size_t pb = 0, cb = 0;
      off[w] = total;
      total += parts(ktab[w].num_sites, &pb, &cb);
      p->fb_items_max = std::max(p->fb_items_max, fb_items(ktab[w].num_sites, a.bits, a.N, a.H));
    }
    rc = p->fbdata.ensure(total);
    if (rc == GLP_OK) rc = p->fbjobs.ensure(sizeof(FbWin) * p->nwin);
    std::vector<FbWin> jobs(p->nwin);
    for (unsigned w = 0; w < p->nwin && rc == GLP_OK; ++w) {
      const glp_window &W = desc->windows[w];
      const unsigned n = (W.num_sites + 1) / 2, wh = (n + 63) / 64;
      size_t pb = 0, cb = 0;
      (void)parts(W.num_sites, &pb, &cb);
      char *base = p->fbdata.as<char>() + off[w];
      FbWin &j = jobs[w];
      j.panel = W.panel;
      j.packed = W.packed;
      j.tran = W.tran;
      j.cpanel = reinterpret_cast<uint64_t *>(base);
      j.ccodes = reinterpret_cast<uint32_t *>(base + pb);
      j.ctran = reinterpret_cast<float *>(base + pb + cb);
      j.M = W.num_sites;
      j.wn = W.words_per_hap;
      ktab[w].panel = j.cpanel;
      ktab[w].packed = j.ccodes;
      ktab[w].tran = j.ctran;
      ktab[w].words_per_hap = wh;
    }
    if (rc == GLP_OK)
      rc = hip_check(hipMemcpy(p->fbjobs.p, jobs.data(), sizeof(FbWin) * p->nwin,
                               hipMemcpyHostToDevice), "layout job copy");
  }
  if (rc == GLP_OK)
    rc = hip_check(hipMemcpy(p->wins.p, ktab.data(), sizeof(glp_window) * p->nwin,
                             hipMemcpyHostToDevice), "window table copy");
  if (rc == GLP_OK)
    rc = hip_check(hipMemcpy(p->order_d.p, p->order.data(), 4 * p->nwin, hipMemcpyHostToDevice),
                   "window order copy");
  if (rc == GLP_OK) rc = plan_upload_q(p);
  if (rc != GLP_OK) {
    glp_plan_destroy(p);
    return rc;
  }
  *out = p;
  return GLP_OK;
}

int glp_mh_run_noalloc(const glp_mh_desc *desc, unsigned lanes_per_chain, void *stream,
                       uint64_t *stats_out) {
  if (desc && desc->num_windows > GLP_NOALLOC_MAX_WINDOWS)
    return set_error(GLP_ERR_ARG, "glp_mh_run_noalloc: more than GLP_NOALLOC_MAX_WINDOWS windows");
  KArgs a;
  GLP_TRY(fill_args(desc, a));
  if (lanes_per_chain > kMaxPasses)
    return set_error(GLP_ERR_ARG, "lanes_per_chain must be 0 (automatic) or 1..32");
  if (a.precision == GLP_PREC_MIXED_FB)
    return set_error(GLP_ERR_ARG, "glp_mh_run_noalloc: the forward-backward form needs a plan "
                                  "(its window layouts are allocated)");
  int dev = 0, num_cu = 1;
  GLP_TRY(hip_check(hipGetDevice(&dev), "hipGetDevice"));
  GLP_TRY(hip_check(hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev),
                    "device attribute"));
  if (lds_bytes(a, 4) > kLdsPerCu)
    return set_error(GLP_ERR_ARG, "the batched kernel's workgroup does not fit a CU (LDS)");
  const unsigned Q = lanes_per_chain ? lanes_per_chain
                                     : auto_lanes(a.N * desc->num_windows, a.C, num_cu, 32u);
  return run_noalloc(a, desc->windows, desc->num_windows, Q, num_cu,
                     static_cast<hipStream_t>(stream), stats_out);
}

int glp_plan_run(glp_plan *p, unsigned gen, float pen, uint64_t seed, void *stream) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  KArgs a = p->a;
  a.gen = gen;
  a.pen = pen;
  a.seed = seed;
  GLP_TRY(plan_mirror(p, static_cast<hipStream_t>(stream)));
  return launch_batched(a, p->wins.as<glp_window>(), p->order_d.as<unsigned>(),
                        p->qtab.as<unsigned>(), p->nwin, p->wctr.as<unsigned>(),
                        p->ctrs.as<unsigned long long>(), p->grid, p->wpb,
                        static_cast<hipStream_t>(stream));
}

int glp_plan_set_lanes(glp_plan *p, const unsigned *lanes_per_window) {
  if (!p || !lanes_per_window) return set_error(GLP_ERR_ARG, "null argument");
  const unsigned nu = units(p->a.precision);
  for (unsigned w = 0; w < p->nwin; ++w) {
    const unsigned Q = lanes_per_window[w];
    if (Q == 0 || Q > nu)
      return set_error(GLP_ERR_ARG, nu == 16u ? "lanes per chain must be 1..16"
                                              : "lanes per chain must be 1..32");
  }
  p->q.assign(lanes_per_window, lanes_per_window + p->nwin);
  plan_graph_reset(p);
  p->grid = plan_grid(p, p->order);
  return plan_upload_q(p);
}

int glp_plan_autotune(glp_plan *p, unsigned gen, float pen, uint64_t seed, void *stream,
                      unsigned *lanes_out) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  plan_graph_reset(p);
  hipStream_t s = static_cast<hipStream_t>(stream);
  KArgs a = p->a;
  a.gen = gen;
  a.pen = pen;
  a.seed = seed;
  // trials overwrite the windows' parents and log-likelihoods (inputs under
  // GLP_PROP_HOST): keep a copy and restore it after every trial
  const size_t pbytes = 16 * (size_t)a.N, lbytes = 8 * (size_t)a.N;
  DevBuf save;
  GLP_TRY(save.ensure((pbytes + lbytes) * p->nwin));
  auto copy_all = [&](bool to_save) -> int {
    for (unsigned w = 0; w < p->nwin; ++w) {
      char *base = save.as<char>() + (pbytes + lbytes) * w;
      const glp_window &W = p->wins_h[w];
      void *dp = to_save ? (void *)base : (void *)W.parents;
      const void *sp = to_save ? (const void *)W.parents : (const void *)base;
      GLP_TRY(hip_check(hipMemcpyAsync(dp, sp, pbytes, hipMemcpyDeviceToDevice, s), "save copy"));
      dp = to_save ? (void *)(base + pbytes) : (void *)W.like;
      sp = to_save ? (const void *)W.like : (const void *)(base + pbytes);
      GLP_TRY(hip_check(hipMemcpyAsync(dp, sp, lbytes, hipMemcpyDeviceToDevice, s), "save copy"));
    }
    return GLP_OK;
  };
  GLP_TRY(copy_all(true));
  // the windows' inputs do not change during the trials: one layout build
  GLP_TRY(plan_mirror(p, s));
  hipEvent_t e0, e1;
  GLP_TRY(hip_check(hipEventCreate(&e0), "event"));
  GLP_TRY(hip_check(hipEventCreate(&e1), "event"));
  // window-size buckets: ceil(log2(M))
  std::vector<std::vector<unsigned>> buckets;
  std::vector<int> bucket_of(64, -1);
  for (unsigned w : p->order) {
    unsigned lg = 0;
    while ((1u << lg) < p->wins_h[w].num_sites) ++lg;
    if (bucket_of[lg] < 0) {
      bucket_of[lg] = (int)buckets.size();
      buckets.emplace_back();
    }
    buckets[bucket_of[lg]].push_back(w);
  }
  // every trial runs the whole plan (all windows share the device, as in a
  // real run): first one depth for every window, then each bucket's depth in
  // turn with the others held (coordinate descent on the measured time)
  int rc = GLP_OK;
  // a trial's time is the mean over kTrialGens consecutive generations, each
  // from the saved state: every plan sees the same draws, and a plan's time
  // varies by several percent from one generation to the next (which chains
  // accept most), more than the plans near the best differ
  constexpr unsigned kTrialGens = 4;
  auto trial = [&](float *ms_out) -> int {
    GLP_TRY(plan_upload_q(p));
    const unsigned grid = plan_grid(p, p->order);
    float ms = 0;
    for (unsigned rep = 0; rep < kTrialGens; ++rep) {
      KArgs ar = a;
      ar.gen = gen + rep;
      GLP_TRY(hip_check(hipEventRecord(e0, s), "event record"));
      GLP_TRY(launch_batched(ar, p->wins.as<glp_window>(), p->order_d.as<unsigned>(),
                             p->qtab.as<unsigned>(), p->nwin, p->wctr.as<unsigned>(),
                             p->ctrs.as<unsigned long long>(), grid, p->wpb, s));
      GLP_TRY(hip_check(hipEventRecord(e1, s), "event record"));
      GLP_TRY(hip_check(hipEventSynchronize(e1), "event sync"));
      float t = 0;
      GLP_TRY(hip_check(hipEventElapsedTime(&t, e0, e1), "event time"));
      ms += t / kTrialGens;
      GLP_TRY(copy_all(false));
    }
    *ms_out = ms;
    return GLP_OK;
  };
  static const unsigned kTrials[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 16, 32};
  float best_ms = 1e30f;
  std::vector<unsigned> best_q = p->q;
  unsigned best_w = p->wpb;
  int best_split = p->a.split;
  // one depth for every window, at each workgroup size that fits, and at 8
  // wavefronts with and without the split
  for (unsigned w : {4u, 8u}) {
    if (rc != GLP_OK || p->bpc_of[w == 8 ? 1 : 0] < 1) continue;
    rc = plan_use_waves(p, w);
    for (unsigned Q : kTrials) {
      if (rc != GLP_OK || Q > a.C + 1 || Q > units(a.precision)) break;
      std::fill(p->q.begin(), p->q.end(), Q);
      for (int sp = 0; sp < (w == 8 ? 2 : 1); ++sp) {
        if (rc != GLP_OK || (sp && units(a.precision) / Q < 2)) break;
        p->a.split = a.split = sp;
        float ms = 0;
        rc = trial(&ms);
        if (rc == GLP_OK && ms < best_ms) {
          best_ms = ms;
          best_q = p->q;
          best_w = w;
          best_split = sp;
        }
      }
    }
  }
  p->a.split = a.split = best_split;
  if (rc == GLP_OK) rc = plan_use_waves(p, best_w);
  if (buckets.size() > 1) {
    for (auto &b : buckets) {
      for (unsigned Q : kTrials) {
        if (rc != GLP_OK || Q > a.C + 1 || Q > units(a.precision)) break;
        if (Q == best_q[b[0]]) continue;
        p->q = best_q;
        for (unsigned w : b) p->q[w] = Q;
        float ms = 0;
        rc = trial(&ms);
        if (rc == GLP_OK && ms < best_ms) {
          best_ms = ms;
          best_q = p->q;
        }
      }
    }
  }
  p->q = best_q;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == GLP_OK) rc = plan_upload_q(p);
  if (rc == GLP_OK) rc = hip_check(hipStreamSynchronize(s), "synchronize");
  p->grid = plan_grid(p, p->order);
  if (rc == GLP_OK && lanes_out)
    for (unsigned w = 0; w < p->nwin; ++w) lanes_out[w] = p->q[w];
  return rc;
}

int glp_plan_stats(glp_plan *p, uint64_t *evals, uint64_t *useful) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  unsigned long long h[4];
  GLP_TRY(hip_check(hipDeviceSynchronize(), "synchronize"));
  GLP_TRY(hip_check(hipMemcpy(h, p->ctrs.p, sizeof(h), hipMemcpyDeviceToHost), "stats copy"));
  if (evals) *evals = h[1];
  if (useful) *useful = h[2];
  if (h[3]) return set_error(GLP_ERR_STATE, "batched kernel hit its round bound");
  return GLP_OK;
}

int glp_plan_run_plain(glp_plan *p, unsigned gen, float pen, uint64_t seed, void *stream) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  KArgs a = p->a;
  a.gen = gen;
  a.pen = pen;
  a.seed = seed;
  hipStream_t s = static_cast<hipStream_t>(stream);
  GLP_TRY(plan_mirror(p, s));
  for (unsigned step = 0; step <= a.C; ++step)
    GLP_TRY(launch_plain(a, p->wins.as<glp_window>(), p->nwin, step, s));
  return GLP_OK;
}

}  // extern "C"

// Node parameters of the batched generation's kernel; `a` must outlive the
// call that consumes the parameters.
namespace {
struct BatchedArgs {
  const glp_window *wins;
  const unsigned *order, *qtab;
  unsigned nwin;
  unsigned *wctr;
  unsigned long long *ctrs;
  KArgs a;
  void *ptrs[7];
  void bind() {
    ptrs[0] = &wins;
    ptrs[1] = &order;
    ptrs[2] = &qtab;
    ptrs[3] = &nwin;
    ptrs[4] = &wctr;
    ptrs[5] = &ctrs;
    ptrs[6] = &a;
  }
};

struct PlainArgs {
  const glp_window *wins;
  KArgs a;
  unsigned step;
  void *ptrs[3];
  void bind() {
    ptrs[0] = &wins;
    ptrs[1] = &a;
    ptrs[2] = &step;
  }
};

hipKernelNodeParams node_params(const void *fn, dim3 grid, unsigned shmem, void **args,
                                unsigned threads = 256) {
  hipKernelNodeParams k{};
  k.func = const_cast<void *>(fn);
  k.gridDim = grid;
  k.blockDim = dim3(threads);
  k.sharedMemBytes = shmem;
  k.kernelParams = args;
  k.extra = nullptr;
  return k;
}
}  // namespace

static int plan_graph_build(glp_plan *p, int kind) {
  plan_graph_reset(p);
  const KArgs &a = p->a;
  const unsigned shmem = (unsigned)lds_bytes(a, kind == 1 ? p->wpb : 4u);
  GLP_TRY(hip_check(hipGraphCreate(&p->graph, 0), "hipGraphCreate"));
  int rc = GLP_OK;
  hipGraphNode_t prev = nullptr;
  auto add = [&](const hipKernelNodeParams &k) {
    hipGraphNode_t n = nullptr;
    if (rc == GLP_OK)
      rc = hip_check(hipGraphAddKernelNode(&n, p->graph, prev ? &prev : nullptr, prev ? 1 : 0, &k),
                     "graph node");
    if (rc == GLP_OK) {
      p->nodes.push_back(n);
      prev = n;
    }
  };
  if (plan_fb(p)) {
    // the windows' forward-backward layouts, from the windows' current data
    const FbWin *jobs = p->fbjobs.as<FbWin>();
    unsigned N = a.N, H = a.H;
    void *margs[] = {&jobs, &N, &H};
    unsigned gx = std::max(1u, std::min((p->fb_items_max + 255u) / 256u,
                                        std::max(1u, 4096u / std::max(1u, p->nwin))));
    hipKernelNodeParams mk{};
    mk.func = const_cast<void *>(fb_mirror_kernel_fn(a.bits));
    mk.gridDim = dim3(gx, p->nwin);
    mk.blockDim = dim3(256);
    mk.sharedMemBytes = 0;
    mk.kernelParams = margs;
    mk.extra = nullptr;
    add(mk);
  }
  if (kind == 1) {
    const void *fn = batched_kernel_fn(a.bits, a.precision, p->wpb);
    if (rc == GLP_OK) rc = prepare_kernel_lds(fn, shmem);
    unsigned long long *ctrs = p->ctrs.as<unsigned long long>();
    unsigned *wctr = p->wctr.as<unsigned>();
    unsigned nwin = p->nwin;
    void *rargs[] = {&ctrs, &wctr, &nwin};
    add(node_params(reset_kernel_fn(), dim3(1), 0, rargs));
    BatchedArgs ba{p->wins.as<glp_window>(), p->order_d.as<unsigned>(), p->qtab.as<unsigned>(),
                   p->nwin, wctr, ctrs, a, {}};
    ba.bind();
    p->node0 = p->nodes.size();
    add(node_params(fn, dim3(p->grid), shmem, ba.ptrs, 64 * p->wpb));
  } else {
    const void *fn = plain_kernel_fn(a.bits, a.precision);
    if (rc == GLP_OK) rc = prepare_kernel_lds(fn, shmem);
    p->node0 = p->nodes.size();
    for (unsigned step = 0; step <= a.C && rc == GLP_OK; ++step) {
      PlainArgs pa{p->wins.as<glp_window>(), a, step, {}};
      pa.bind();
      add(node_params(fn, dim3((plain_lanes(a) + 255) / 256, p->nwin), shmem, pa.ptrs));
    }
  }
  if (rc == GLP_OK) rc = hip_check(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0),
                                   "hipGraphInstantiate");
  if (rc != GLP_OK) {
    plan_graph_reset(p);
    return rc;
  }
  p->graph_kind = kind;
  return GLP_OK;
}

// Rewrites the graph's kernel arguments for (gen, pen, seed) and launches it.
static int plan_graph_run(glp_plan *p, int kind, unsigned gen, float pen, uint64_t seed,
                          hipStream_t s) {
  if (p->graph_kind != kind) GLP_TRY(plan_graph_build(p, kind));
  KArgs a = p->a;
  a.gen = gen;
  a.pen = pen;
  a.seed = seed;
  const unsigned shmem = (unsigned)lds_bytes(a, kind == 1 ? p->wpb : 4u);
  if (kind == 1) {
    BatchedArgs ba{p->wins.as<glp_window>(), p->order_d.as<unsigned>(), p->qtab.as<unsigned>(),
                   p->nwin, p->wctr.as<unsigned>(), p->ctrs.as<unsigned long long>(), a, {}};
    ba.bind();
    hipKernelNodeParams bk = node_params(batched_kernel_fn(a.bits, a.precision, p->wpb),
                                         dim3(p->grid), shmem, ba.ptrs, 64 * p->wpb);
    GLP_TRY(hip_check(hipGraphExecKernelNodeSetParams(p->exec, p->nodes[p->node0], &bk),
                      "graph update"));
  } else {
    const void *fn = plain_kernel_fn(a.bits, a.precision);
    for (unsigned step = 0; step <= a.C; ++step) {
      PlainArgs pa{p->wins.as<glp_window>(), a, step, {}};
      pa.bind();
      hipKernelNodeParams k =
          node_params(fn, dim3((plain_lanes(a) + 255) / 256, p->nwin), shmem, pa.ptrs);
      GLP_TRY(hip_check(hipGraphExecKernelNodeSetParams(p->exec, p->nodes[p->node0 + step], &k),
                        "graph update"));
    }
  }
  return hip_check(hipGraphLaunch(p->exec, s), "hipGraphLaunch");
}

extern "C" {

int glp_plan_graph_run(glp_plan *p, unsigned gen, float pen, uint64_t seed, void *stream) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  return plan_graph_run(p, 1, gen, pen, seed, static_cast<hipStream_t>(stream));
}

int glp_plan_graph_run_plain(glp_plan *p, unsigned gen, float pen, uint64_t seed, void *stream) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  return plan_graph_run(p, 2, gen, pen, seed, static_cast<hipStream_t>(stream));
}

int glp_plan_set_waves(glp_plan *p, unsigned waves_per_workgroup) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  GLP_TRY(plan_use_waves(p, waves_per_workgroup));
  p->grid = plan_grid(p, p->order);
  return GLP_OK;
}

int glp_plan_waves(glp_plan *p, unsigned *waves_per_workgroup) {
  if (!p || !waves_per_workgroup) return set_error(GLP_ERR_ARG, "null argument");
  *waves_per_workgroup = p->wpb;
  return GLP_OK;
}

int glp_plan_set_split(glp_plan *p, int split) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  const int v = split ? 1 : 0;
  if (v != p->a.split) plan_graph_reset(p);  // the graph's node holds the grid and arguments
  p->a.split = v;
  p->grid = plan_grid(p, p->order);
  return GLP_OK;
}

int glp_plan_split(glp_plan *p, int *split) {
  if (!p || !split) return set_error(GLP_ERR_ARG, "null argument");
  *split = p->a.split;
  return GLP_OK;
}

int glp_plan_lanes(glp_plan *p, unsigned *lanes, unsigned *grid) {
  if (!p) return set_error(GLP_ERR_ARG, "null plan");
  if (lanes) *lanes = p->q.empty() ? 0 : p->q[p->order[0]];
  if (grid) *grid = p->grid;
  return GLP_OK;
}

int glp_plan_destroy(glp_plan *p) {
  if (!p) return GLP_OK;
  plan_graph_reset(p);
  for (DevBuf *b : {&p->wins, &p->order_d, &p->qtab, &p->wctr, &p->ctrs, &p->fbdata, &p->fbjobs})
    b->release();
  delete p;
  return GLP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// reference-compatible layer
// ---------------------------------------------------------------------------
struct glp_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  std::vector<float> tran, pc, codebook;
  unsigned bits = 0;
  float pen = 1.0f;
  bool lut_dirty = true;
  DevBuf lut, tran_d, packed, panel, extra, parents, like, samples, evalp, out;
  size_t packed_words = 0;
};

static int ensure_lut(glp_ctx *c) {
  if (c->codebook.empty()) return set_error(GLP_ERR_STATE, "codebook not set");
  if (c->pc.size() != 16) return set_error(GLP_ERR_STATE, "mutation matrix not set");
  if (!c->lut_dirty) return GLP_OK;
  const unsigned ncode = (unsigned)(c->codebook.size() / 2);
  unsigned bits = 0;
  while ((1u << bits) < ncode) ++bits;
  if (bits != 4 && bits != 8) return set_error(GLP_ERR_ARG, "codebook size must be 16 or 256");
  GLP_TRY(c->lut.ensure(glp_emit_lut_bytes(bits, GLP_PREC_F32)));
  GLP_TRY(glp_build_emit_lut(c->codebook.data(), ncode, c->pc.data(), GLP_PREC_F32, c->lut.p));
  c->lut_dirty = false;
  return GLP_OK;
}

static int ctx_window(glp_ctx *c, const uint64_t *hap_panel, size_t n_panel_words,
                      unsigned num_sites, unsigned num_samples, glp_window &W) {
  if (!hap_panel || num_sites == 0 || num_samples == 0)
    return set_error(GLP_ERR_ARG, "null panel or zero sites/samples");
  if (c->tran.size() < 3 * (size_t)num_sites)
    return set_error(GLP_ERR_STATE, "transition matrix has fewer than 3 * num_sites entries");
  if (c->bits == 0) return set_error(GLP_ERR_STATE, "packed GLs not set");
  const unsigned per = 32 / c->bits;
  // hmmLike.cu:407-414
  const size_t expect = (size_t)num_samples * ((num_sites + per - 1) / per);
  if (c->packed_words != expect)
    return set_error(GLP_ERR_ARG, "packed GLs have the wrong size for num_sites x num_samples");
  const unsigned wn = (num_sites + 63) / 64;
  if (n_panel_words % wn != 0) return set_error(GLP_ERR_ARG, "panel size not a multiple of words per hap");
  // the kernels read the panel and codes with 32-bit byte offsets
  if ((uint64_t)n_panel_words * 8 >= (1ull << 32) || (uint64_t)expect * 4 >= (1ull << 32))
    return set_error(GLP_ERR_ARG, "panel or packed GLs reach 4 GiB");
  GLP_TRY(ensure_lut(c));
  if (c->codebook.size() != 2 * ((size_t)1 << c->bits))
    return set_error(GLP_ERR_STATE, "codebook size does not match bits_per_code of the packed GLs");
  GLP_TRY(c->panel.ensure(n_panel_words * 8));
  GLP_TRY(hip_check(hipMemcpyAsync(c->panel.p, hap_panel, n_panel_words * 8, hipMemcpyHostToDevice,
                                   c->stream), "panel copy"));
  GLP_TRY(c->parents.ensure((size_t)num_samples * 16));
  GLP_TRY(c->like.ensure((size_t)num_samples * 8));
  W.num_sites = num_sites;
  W.words_per_hap = wn;
  W.panel = c->panel.as<uint64_t>();
  W.packed = c->packed.as<uint32_t>();
  W.tran = c->tran_d.as<float>();
  W.parents = c->parents.as<unsigned>();
  W.like = c->like.as<double>();
  return GLP_OK;
}

extern "C" {

int glp_check_device(int *num_devices) {
  int n = 0;
  hipError_t err = hipGetDeviceCount(&n);
  if (num_devices) *num_devices = (err == hipSuccess) ? n : 0;
  if (err != hipSuccess || n == 0) return set_error(GLP_ERR_NODEV, "no HIP device");
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess &&
        std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
      return GLP_OK;
  }
  return set_error(GLP_ERR_NODEV, "no gfx950 device");
}

int glp_create(int dev_id, glp_ctx **out) {
  if (!out) return set_error(GLP_ERR_ARG, "null out");
  *out = nullptr;
  if (dev_id < 0) return set_error(GLP_ERR_ARG, "negative device id");
  int n = 0;
  GLP_TRY(glp_check_device(&n));
  if (dev_id >= n) return set_error(GLP_ERR_ARG, "device id out of range");
  hipDeviceProp_t prop;
  GLP_TRY(hip_check(hipGetDeviceProperties(&prop, dev_id), "device properties"));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(GLP_ERR_NODEV, "device is not gfx950");
  GLP_TRY(hip_check(hipSetDevice(dev_id), "hipSetDevice"));
  glp_ctx *c = new glp_ctx();
  c->dev = dev_id;
  int rc = hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "stream create");
  if (rc != GLP_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return GLP_OK;
}

int glp_destroy(glp_ctx *c) {
  if (!c) return GLP_OK;
  (void)hipSetDevice(c->dev);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (DevBuf *b : {&c->lut, &c->tran_d, &c->packed, &c->panel, &c->extra, &c->parents, &c->like,
                    &c->samples, &c->evalp, &c->out})
    b->release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GLP_OK;
}

int glp_set_tran(glp_ctx *c, const float *tran, size_t n) {
  if (!c || !tran || n == 0 || n % 3 != 0) return set_error(GLP_ERR_ARG, "tran must hold 3 floats per site");
  GLP_TRY(hip_check(hipSetDevice(c->dev), "hipSetDevice"));
  c->tran.assign(tran, tran + n);
  GLP_TRY(c->tran_d.ensure(n * sizeof(float)));
  return hip_check(hipMemcpy(c->tran_d.p, tran, n * sizeof(float), hipMemcpyHostToDevice), "tran copy");
}

int glp_set_mutation_mat(glp_ctx *c, const float *pc16) {
  if (!c || !pc16) return set_error(GLP_ERR_ARG, "null argument");
  c->pc.assign(pc16, pc16 + 16);
  c->lut_dirty = true;
  return GLP_OK;
}

int glp_set_codebook(glp_ctx *c, const float *pairs, unsigned n_codes) {
  if (!c || !pairs) return set_error(GLP_ERR_ARG, "null argument");
  if (n_codes != 16 && n_codes != 256) return set_error(GLP_ERR_ARG, "codebook must hold 16 or 256 codes");
  c->codebook.assign(pairs, pairs + 2 * (size_t)n_codes);
  c->lut_dirty = true;
  return GLP_OK;
}

int glp_set_packed_gls(glp_ctx *c, const uint32_t *packed, size_t n_words, unsigned bits) {
  if (!c || !packed || n_words == 0) return set_error(GLP_ERR_ARG, "null or empty packed GLs");
  if (!valid_bits(bits)) return set_error(GLP_ERR_ARG, "bits_per_code must be 4 or 8");
  GLP_TRY(hip_check(hipSetDevice(c->dev), "hipSetDevice"));
  GLP_TRY(c->packed.ensure(n_words * 4));
  c->packed_words = n_words;
  c->bits = bits;
  return hip_check(hipMemcpy(c->packed.p, packed, n_words * 4, hipMemcpyHostToDevice), "packed copy");
}

int glp_set_pen(glp_ctx *c, float pen) {
  if (!c) return set_error(GLP_ERR_ARG, "null ctx");
  if (!(pen >= 0.0f)) return set_error(GLP_ERR_ARG, "pen must be >= 0");
  c->pen = pen;
  return GLP_OK;
}

int glp_run_hmm_on_device(glp_ctx *c, const uint64_t *hap_panel, size_t n_panel_words,
                          const unsigned *extra_prop_haps, unsigned num_sites,
                          unsigned num_samples, unsigned num_cycles, unsigned *hap_idxs,
                          unsigned long seed) {
  if (!c || !hap_idxs || (num_cycles > 0 && !extra_prop_haps))
    return set_error(GLP_ERR_ARG, "null argument");
  GLP_TRY(hip_check(hipSetDevice(c->dev), "hipSetDevice"));
  glp_window W;
  GLP_TRY(ctx_window(c, hap_panel, n_panel_words, num_sites, num_samples, W));
  const unsigned wn = W.words_per_hap;
  const size_t nhaps = n_panel_words / wn;
  // parents: reference SoA [4][N] -> engine AoS [N][4]
  std::vector<unsigned> aos(4 * (size_t)num_samples);
  for (unsigned i = 0; i < num_samples; ++i)
    for (unsigned q = 0; q < 4; ++q) {
      const unsigned h = hap_idxs[(size_t)q * num_samples + i];
      if (h >= nhaps) return set_error(GLP_ERR_ARG, "initial parent index outside the panel");
      aos[4 * (size_t)i + q] = h;
    }
  for (size_t e = 0; e < (size_t)num_cycles * num_samples; ++e)
    if (extra_prop_haps[e] >= nhaps) return set_error(GLP_ERR_ARG, "proposal index outside the panel");
  GLP_TRY(hip_check(hipMemcpyAsync(W.parents, aos.data(), aos.size() * 4, hipMemcpyHostToDevice,
                                   c->stream), "parents copy"));
  if (num_cycles > 0) {
    GLP_TRY(c->extra.ensure((size_t)num_cycles * num_samples * 4));
    GLP_TRY(hip_check(hipMemcpyAsync(c->extra.p, extra_prop_haps, (size_t)num_cycles * num_samples * 4,
                                     hipMemcpyHostToDevice, c->stream), "proposal copy"));
  }
  glp_mh_desc d{};
  d.num_windows = 1;
  d.windows = &W;
  d.num_samples = num_samples;
  d.num_haps = (unsigned)nhaps;
  d.bits_per_code = c->bits;
  d.emit_lut = c->lut.p;
  d.num_cycles = num_cycles;
  d.pen = c->pen;
  d.seed = (uint64_t)seed;
  d.prop_mode = GLP_PROP_HOST;
  d.extra = c->extra.as<unsigned>();
  d.precision = GLP_PREC_F32;
  glp_plan *p = nullptr;
  GLP_TRY(glp_plan_create(&d, 0, &p));
  int rc = glp_plan_run(p, 0, c->pen, d.seed, c->stream);
  if (rc == GLP_OK)
    rc = hip_check(hipMemcpyAsync(aos.data(), W.parents, aos.size() * 4, hipMemcpyDeviceToHost,
                                  c->stream), "parents copy back");
  if (rc == GLP_OK) rc = hip_check(hipStreamSynchronize(c->stream), "synchronize");
  if (rc == GLP_OK) rc = glp_plan_stats(p, nullptr, nullptr);
  glp_plan_destroy(p);
  if (rc != GLP_OK) return rc;
  for (unsigned i = 0; i < num_samples; ++i)
    for (unsigned q = 0; q < 4; ++q) hap_idxs[(size_t)q * num_samples + i] = aos[4 * (size_t)i + q];
  return GLP_OK;
}

int glp_hmm_like(glp_ctx *c, const uint64_t *hap_panel, size_t n_panel_words, unsigned num_sites,
                 unsigned num_samples, unsigned n_evals, const unsigned *samples,
                 const unsigned *parents, float *out) {
  if (!c || !samples || !parents || !out) return set_error(GLP_ERR_ARG, "null argument");
  if (n_evals == 0) return GLP_OK;
  GLP_TRY(hip_check(hipSetDevice(c->dev), "hipSetDevice"));
  glp_window W;
  GLP_TRY(ctx_window(c, hap_panel, n_panel_words, num_sites, num_samples, W));
  const size_t nhaps = n_panel_words / W.words_per_hap;
  for (unsigned e = 0; e < n_evals; ++e) {
    if (samples[e] >= num_samples) return set_error(GLP_ERR_ARG, "sample index out of range");
    for (unsigned q = 0; q < 4; ++q)
      if (parents[4 * (size_t)e + q] >= nhaps) return set_error(GLP_ERR_ARG, "parent index outside the panel");
  }
  GLP_TRY(c->samples.ensure((size_t)n_evals * 4));
  GLP_TRY(c->evalp.ensure((size_t)n_evals * 16));
  GLP_TRY(c->out.ensure((size_t)n_evals * 4));
  GLP_TRY(hip_check(hipMemcpyAsync(c->samples.p, samples, (size_t)n_evals * 4, hipMemcpyHostToDevice,
                                   c->stream), "samples copy"));
  GLP_TRY(hip_check(hipMemcpyAsync(c->evalp.p, parents, (size_t)n_evals * 16, hipMemcpyHostToDevice,
                                   c->stream), "parents copy"));
  glp_window *d_win = nullptr;
  GLP_TRY(hip_check(hipMallocAsync((void **)&d_win, sizeof(glp_window), c->stream), "hipMallocAsync"));
  int rc = hip_check(hipMemcpyAsync(d_win, &W, sizeof(W), hipMemcpyHostToDevice, c->stream), "window copy");
  KArgs a{};
  a.N = num_samples;
  a.norm = norm_value();
  a.lut = c->lut.p;
  a.bits = c->bits;
  a.lut_entries = 16u << c->bits;
  if (rc == GLP_OK)
    rc = launch_eval(a, d_win, c->samples.as<unsigned>(), c->evalp.as<unsigned>(), n_evals,
                     c->out.as<float>(), c->stream);
  (void)hipFreeAsync(d_win, c->stream);
  if (rc == GLP_OK)
    rc = hip_check(hipMemcpyAsync(out, c->out.p, (size_t)n_evals * 4, hipMemcpyDeviceToHost, c->stream),
                   "result copy");
  if (rc == GLP_OK) rc = hip_check(hipStreamSynchronize(c->stream), "synchronize");
  return rc;
}

}  // extern "C"

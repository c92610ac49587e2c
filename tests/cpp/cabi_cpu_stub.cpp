// cabi_cpu_stub.cpp -- test-only CPU stand-in for the device half of the
// C ABI (include/glphase_amd.h, layer 1), built on the oracle.  Linking the
// C++ host classes (glp_host.cpp) and tests/cpp/test_hmmlike.cpp against it
// runs the whole HMMLike flow -- sampler draws, SoA/AoS transposes, strided
// GLPack uploads, MH chains -- on a machine without a GPU.  It is never part
// of libglphase_amd.so.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/glphase_amd.h"
#include "../../oracle/oracle.h"

namespace glp {
static std::string g_err;
int set_error(int code, const char *msg) {
  g_err = msg ? msg : "";
  return code;
}
}  // namespace glp

struct glp_ctx {
  std::vector<float> tran, pc, codebook;
  std::vector<uint32_t> packed;
  unsigned bits = 0;
  float pen = 1.0f;
};

extern "C" {

const char *glp_last_error(void) { return glp::g_err.c_str(); }

int glp_create(int dev_id, glp_ctx **out) {
  if (!out || dev_id < 0) return glp::set_error(GLP_ERR_ARG, "bad argument");
  *out = new glp_ctx();
  return GLP_OK;
}

int glp_destroy(glp_ctx *c) {
  delete c;
  return GLP_OK;
}

int glp_set_tran(glp_ctx *c, const float *tran, size_t n) {
  c->tran.assign(tran, tran + n);
  return GLP_OK;
}

int glp_set_mutation_mat(glp_ctx *c, const float *pc16) {
  c->pc.assign(pc16, pc16 + 16);
  return GLP_OK;
}

int glp_set_codebook(glp_ctx *c, const float *pairs, unsigned n_codes) {
  c->codebook.assign(pairs, pairs + 2 * (size_t)n_codes);
  return GLP_OK;
}

int glp_set_packed_gls(glp_ctx *c, const uint32_t *packed, size_t n_words, unsigned bits) {
  if (bits != 4 && bits != 8) return glp::set_error(GLP_ERR_ARG, "bits must be 4 or 8");
  c->packed.assign(packed, packed + n_words);
  c->bits = bits;
  return GLP_OK;
}

int glp_set_pen(glp_ctx *c, float pen) {
  c->pen = pen;
  return GLP_OK;
}

static int window(glp_ctx *c, const uint64_t *panel, size_t n_words, unsigned M, unsigned N,
                  std::vector<float> &lut, orc_window &w) {
  if (c->bits == 0 || c->tran.size() < 3 * (size_t)M || c->pc.size() != 16)
    return glp::set_error(GLP_ERR_STATE, "model not set");
  const unsigned per = 32 / c->bits;
  if (c->packed.size() != (size_t)N * ((M + per - 1) / per))
    return glp::set_error(GLP_ERR_ARG, "packed GLs have the wrong size");
  const unsigned ncode = 1u << c->bits;
  if (c->codebook.size() != 2 * (size_t)ncode) return glp::set_error(GLP_ERR_STATE, "codebook size");
  lut.resize(4 * (size_t)ncode);
  orc_emit_lut(c->codebook.data(), ncode, c->pc.data(), lut.data());
  w.wn = (M + 63) / 64;
  if (n_words % w.wn) return glp::set_error(GLP_ERR_ARG, "panel size");
  w.haps = panel;
  w.nhaps = (unsigned)(n_words / w.wn);
  w.packed = c->packed.data();
  w.nsamp = N;
  w.M = M;
  w.bits = c->bits;
  w.lut = lut.data();
  w.tran = c->tran.data();
  return GLP_OK;
}

int glp_run_hmm_on_device(glp_ctx *c, const uint64_t *hap_panel, size_t n_panel_words,
                          const unsigned *extra_prop_haps, unsigned num_sites,
                          unsigned num_samples, unsigned num_cycles, unsigned *hap_idxs,
                          unsigned long seed) {
  std::vector<float> lut;
  orc_window w;
  int rc = window(c, hap_panel, n_panel_words, num_sites, num_samples, lut, w);
  if (rc != GLP_OK) return rc;
  std::vector<unsigned> init(4 * (size_t)num_samples), out(4 * (size_t)num_samples);
  for (unsigned i = 0; i < num_samples; ++i)
    for (unsigned q = 0; q < 4; ++q) init[4 * (size_t)i + q] = hap_idxs[(size_t)q * num_samples + i];
  orc_mh_args a;
  std::memset(&a, 0, sizeof(a));
  a.ncycles = num_cycles;
  a.pen = c->pen;
  a.seed = (uint64_t)seed;
  a.prop_mode = ORC_PROP_HOST;
  a.precision = ORC_PREC_F32;
  a.init_parents = init.data();
  a.extra = extra_prop_haps;
  std::vector<double> curr(num_samples);
  orc_mh_window(&w, &a, 0, num_samples, out.data(), curr.data(), nullptr, nullptr);
  for (unsigned i = 0; i < num_samples; ++i)
    for (unsigned q = 0; q < 4; ++q) hap_idxs[(size_t)q * num_samples + i] = out[4 * (size_t)i + q];
  return GLP_OK;
}

int glp_hmm_like(glp_ctx *c, const uint64_t *hap_panel, size_t n_panel_words, unsigned num_sites,
                 unsigned num_samples, unsigned n, const unsigned *samples,
                 const unsigned *parents, float *out) {
  std::vector<float> lut;
  orc_window w;
  int rc = window(c, hap_panel, n_panel_words, num_sites, num_samples, lut, w);
  if (rc != GLP_OK) return rc;
  std::vector<float> emit(4 * (size_t)num_sites);
  for (unsigned e = 0; e < n; ++e) {
    orc_sample_emit(w.packed, num_samples, samples[e], num_sites, w.bits, w.lut, emit.data());
    out[e] = orc_hmm_like_f32(w.haps, w.wn, parents + 4 * (size_t)e, emit.data(), w.tran, num_sites);
  }
  return GLP_OK;
}

}  // extern "C"

// glp_internal.hpp -- declarations shared by the kernel file and the host
// side of the C ABI.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/glphase_amd.h"

namespace glp {

// Kernel arguments common to every launch (passed by value).
struct KArgs {
  unsigned N;   // samples per window
  unsigned H;   // haplotypes per panel
  unsigned C;   // MH proposals per sample
  unsigned k0;  // Philox counter of proposal 0
  float pen;
  float norm;   // powf(FLT_MIN, 2/3) (impute.cpp:11-13)
  uint64_t seed;
  unsigned gen;
  int prop_mode;
  const unsigned *klist;
  unsigned K;
  const unsigned *extra;
  double *trace_like;
  uint8_t *trace_acc;
  const void *lut;       // device emission table
  unsigned lut_entries;  // 16 * 2^bits
  unsigned bits;
  int precision;
  int kick;  // proposal 0 draws from the reference haplotypes [2N, H)
  int split;  // batched kernel at 8 wavefronts per workgroup: wavefronts 4-7 hold one chain slot fewer
};

int set_error(int code, const char *msg);
int hip_check(hipError_t err, const char *what);

int launch_plain(const KArgs &a, const glp_window *d_wins, unsigned nwin,
                 unsigned step, hipStream_t stream);
// LDS bytes of a workgroup of `waves` wavefronts (the plain path and the
// single evaluations use 4).
size_t lds_bytes(const KArgs &a, unsigned waves);
int launch_batched(const KArgs &a, const glp_window *d_wins, const unsigned *d_order,
                   const unsigned *d_qtab, unsigned nwin, unsigned *d_wctr,
                   unsigned long long *d_ctrs, unsigned grid, unsigned waves,
                   hipStream_t stream);
// Kernel entry points for HIP graph nodes, and the dynamic-LDS attribute they
// need before any launch.
const void *plain_kernel_fn(unsigned bits, int precision);
const void *batched_kernel_fn(unsigned bits, int precision, unsigned waves);
const void *reset_kernel_fn();
int prepare_kernel_lds(const void *fn, size_t bytes);
int batched_occupancy(unsigned bits, int precision, unsigned waves, size_t smem,
                      int *blocks_per_cu);
// lanes of the plain path's grid for one window (2 or 4 per chain)
unsigned plain_lanes(const KArgs &a);

// The forward-backward form (GLP_PREC_MIXED_FB) reads each window through a
// layout the library builds before every run (k_fb_mirror): the forward
// half's data and the backward half's mirrored copy.  One job per window:
// the window's own arrays and where its layout goes.
struct FbWin {
  const uint64_t *panel;   // [H][wn]
  const uint32_t *packed;  // [nwords][N]
  const float *tran;       // [M][3]
  uint64_t *cpanel;        // [2][H][wh]
  uint32_t *ccodes;        // [2][nw][N]
  float *ctran;            // forward rows, then backward rows
  unsigned M, wn;
};
size_t fb_layout_bytes(unsigned M, unsigned bits, unsigned N, unsigned H);
// GLP_OK, or an error when a layout array passes 32-bit byte offsets
int fb_check(unsigned M, unsigned bits, unsigned N, unsigned H);
// work items of one window's layout (threads of the mirror kernel)
unsigned fb_items(unsigned M, unsigned bits, unsigned N, unsigned H);
int launch_fb_mirror(const FbWin *d_fw, unsigned nwin, unsigned max_items, unsigned N,
                     unsigned H, unsigned bits, hipStream_t stream);
const void *fb_mirror_kernel_fn(unsigned bits);  // 4 or 8
// One batched generation with every device buffer beyond the caller's in
// the file-scope scratch array of glp_kernels.hip (glp_mh_run_noalloc).
int run_noalloc(const KArgs &a, const glp_window *wins_h, unsigned nwin, unsigned Q,
                int num_cu, hipStream_t stream, uint64_t *stats_out);
int launch_eval(const KArgs &a, const glp_window *d_wins,
                const unsigned *d_samples, const unsigned *d_parents,
                unsigned n, float *d_out, hipStream_t stream);

// Host emission tables (compiled with -ffp-contract=off, so the values equal
// the CPU reference's float arithmetic).
void emit_table(const float *codebook, unsigned ncode, const float *pc16,
                float *E);
void nibble_table(const float *E, unsigned ncode, float *out);

}  // namespace glp

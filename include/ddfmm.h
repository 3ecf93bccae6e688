/* C ABI of the MI355X directional-FMM evaluation pass.
 *
 * The reference exposes this path as C++ classes (no FFI): Wave3d setup/eval/check
 * (include/wave3d.hpp:219-265), Kernel3d::kernel (include/kernel3d.hpp:29-45) and
 * Mlib3d::setup (include/mlib3d.hpp:58-99).  Each entry point below states which
 * of those it replaces.  All functions return 0 on success and a nonzero code on
 * failure; ddfmm_last_error() then describes the failure (no exception crosses
 * the ABI).  Complex arrays are interleaved (re, im) doubles.
 */
#ifndef DDFMM_H
#define DDFMM_H

#include <stdint.h>

/* the library exports these entry points only (its C++ internals are hidden) */
#if defined(__GNUC__)
#define DDFMM_API __attribute__((visibility("default")))
#else
#define DDFMM_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ddfmm_ctx ddfmm_ctx;
typedef struct ddfmm_mlib ddfmm_mlib;

enum { DDFMM_OK = 0, DDFMM_EINVAL = 1, DDFMM_ESETUP = 2, DDFMM_EDEVICE = 3, DDFMM_ESTATE = 4,
       DDFMM_ENOMEM = 5 };
enum { DDFMM_KERNEL_HELM = 0, DDFMM_KERNEL_EXPR = 1 };   /* kernel3d.hpp:24-27 */
/* FP64: all FP64, check->equivalent solves applied in the reference's factored form.
   MIXED: FP32 vectors and translations (MFMA f32), FP64 solves as precomputed matrices.
   FP64_FUSED: FP64 everywhere with the MIXED mode's fused solve matrices.
   MIXED_SPLIT: MIXED with the translation GEMMs as split-BF16 products (every FP32
   operand the sum of three bf16 terms, products to third order on the bf16 matrix cores,
   FP32 accumulation). */
enum { DDFMM_FP64 = 0, DDFMM_MIXED = 1, DDFMM_FP64_FUSED = 2, DDFMM_MIXED_SPLIT = 3 };

/* Wave3d / Mlib3d / Kernel3d options (wave3d_setup.cpp:29-61, mlib3d.cpp:70-91,
   tt.cpp:103-119): K, ctr, ACCU, NPQ, ptsmax, maxlevel, cell count per edge numC
   (geomprtn.m(), wave3d.hpp:305), kernel type.  nrhs_max sizes the device buffers. */
typedef struct ddfmm_params {
    double K;
    double ctr[3];
    int ACCU, NPQ, ptsmax, maxlevel, numC, kernel;
    int nrhs_max;
    int precision;
} ddfmm_params;

typedef struct ddfmm_stats {
    int64_t npts, nleaf, nlf, nhf_out, nhf_in;
    int64_t near_interactions, leaf_op_interactions, s2m_interactions, x2chk_interactions;
    int64_t v_pairs, v_targets, v_sources, m2l_pairs, m2l_groups;
    double m2l_flops, gemm_flops;
    int64_t opmat_elems, device_bytes;
    int nlevels_lf, nlevels_hf, max_hf_width, nstages;
} ddfmm_stats;

DDFMM_API const char* ddfmm_last_error(void);

/* Create a context on HIP device `device`; device < 0 gives a host-only context that
   plans but never touches a GPU (tree/list inspection).  Replaces Wave3d(prefix)
   (wave3d.cpp:24-31). */
DDFMM_API int ddfmm_create(int device, ddfmm_ctx** out);
DDFMM_API int ddfmm_destroy(ddfmm_ctx* ctx);

/* Build the octree, interaction lists and direction sets on the host, read the
   operator library (ldfile/hdfile in the reference binary format), generate the
   translation matrices on the device and keep everything resident.  Replaces
   Mlib3d::setup (mlib3d.cpp:61-103) + Wave3d::setup (wave3d_setup.cpp:21-97).
   pos: n x 3 row-major doubles (point id = row). */
DDFMM_API int ddfmm_setup(ddfmm_ctx* ctx, const double* pos, int64_t n, const ddfmm_params* prm,
                const char* ldfile, const char* hdfile);

/* Operator library handle: the reference's Mlib3d read once (Mlib3d::setup,
   mlib3d.cpp:61-103, the ldname/hdname files in the reference binary format) and shared
   by every context set up from it (many geometries of one kernel in a serving process).
   ddfmm_setup itself reuses a library that a live handle holds for the same two files.
   ddfmm_mlib_widths lists the widths of the LF and HF tables (the keys of Mlib3d's
   w2ldmap / w2hdmap, mlib3d.hpp:74-75), up to cap of each. */
DDFMM_API int ddfmm_mlib_load(const char* ldfile, const char* hdfile, ddfmm_mlib** out);
DDFMM_API int ddfmm_mlib_free(ddfmm_mlib* mlib);
DDFMM_API int ddfmm_mlib_widths(const ddfmm_mlib* mlib, int cap, double* ld_widths, int* nld, double* hd_widths,
                      int* nhd);
/* ddfmm_setup with a shared operator library instead of the two file names. */
DDFMM_API int ddfmm_setup_mlib(ddfmm_ctx* ctx, const double* pos, int64_t n, const ddfmm_params* prm,
                     const ddfmm_mlib* mlib);

/* Potentials val = G * den for nrhs density vectors (host buffers, nrhs x n complex,
   point-id order).  Replaces Wave3d::eval (wave3d_eval.cpp:299-415) — every call is a
   fresh evaluation (the reference's eval is single-shot). */
DDFMM_API int ddfmm_eval(ddfmm_ctx* ctx, const double* den, double* val, int nrhs);

/* Same with device-resident buffers on `stream` (0 = the context's stream).
   Asynchronous with respect to the host.  Part of the work runs on a second stream of
   the context that is forked from and joined back into `stream` with events, so work
   queued on `stream` after this call sees the complete result.  Evaluations of one
   context share its work buffers: each one waits on the device for the end of the
   previous one, also when the two were issued on different streams. */
DDFMM_API int ddfmm_eval_device(ddfmm_ctx* ctx, const void* den_dev, void* val_dev, int nrhs, void* stream);

/* Graph replay: with on != 0, ddfmm_eval / ddfmm_eval_device capture the stage sequence
   of each right-hand-side count once into a HIP graph (both streams of the evaluation)
   and replay it on later calls; the caller's buffers reach the captured gather and
   scatter kernels through a device-side pointer pair set before each replay.  on = 0
   drops the graphs and issues the stages one by one again.  Timed evaluations
   (ddfmm_set_timing) always issue the stages one by one.  No counterpart in the
   reference (which has no device); results are bit-identical to the stage-by-stage
   path. */
DDFMM_API int ddfmm_set_graph(ddfmm_ctx* ctx, int on);

/* GEMM tile autotuning (no counterpart in the reference).  Every planned GEMM stage
   (a translation level: U2U, M2L, D2D, LF M2M/L2L, solves) is tiled with at most
   mbmax row blocks of 16 per tile (1..7; a group's rows spread evenly over its tiles).
   ddfmm_autotune times, for nrhs right-hand sides, every stage under each limit 1..7
   and under the planner's own tiling (limit 7, rows split further for small stages),
   one stage at a time on one stream (best of reps evaluations after a warm-up), keeps
   the fastest per stage and returns the per-stage limits in mbmax (count entries, up to
   cap) and the summed GEMM milliseconds before and after.  ddfmm_set_tiles applies a
   table (e.g. cached for the geometry); ddfmm_get_tiles reads the current table and
   stage names.  All three drop the captured graphs. */
DDFMM_API int ddfmm_autotune(ddfmm_ctx* ctx, int nrhs, int reps, int cap, int* mbmax, int* count,
                             double* ms_before, double* ms_after);
DDFMM_API int ddfmm_set_tiles(ddfmm_ctx* ctx, int count, const int* mbmax);
DDFMM_API int ddfmm_get_tiles(ddfmm_ctx* ctx, int cap, int* mbmax, char* names /* cap x 24 */, int* count);

/* HBM budget (the reference's LIMITED_MEMORY build, wave3d_eval.cpp:27-48 and 106-186,
   bounds its directional data level by level).  The downward directional check values
   always occupy one level at a time.  With a budget (bytes, applied at the next setup;
   0 = none), setup sizes the work buffers for the largest right-hand-side count `slab`
   (<= nrhs_max) that fits beside the tables, and ddfmm_eval / ddfmm_eval_device with
   more right-hand sides run in passes of `slab`; results are the same bits as in one
   pass.  Setup fails with DDFMM_ENOMEM, naming the minimum, when not even one
   right-hand side fits.  ddfmm_memory reports the device bytes of the tables (fixed),
   of one right-hand side's work buffers and staging (per_rhs), of the setup-only
   generator tables (setup) and the slab; also on host-only contexts, where it is the
   plan's requirement. */
DDFMM_API int ddfmm_set_hbm_budget(ddfmm_ctx* ctx, int64_t bytes);
DDFMM_API int ddfmm_memory(ddfmm_ctx* ctx, int64_t* fixed_bytes, int64_t* per_rhs_bytes, int64_t* setup_bytes,
                           int* slab);

/* Direct FP64 sum at targets (Wave3d::check's oracle sum, wave3d_check.cpp:66-76):
   out[r][i] = sum_j G(trg_i, src_j) den[r][j]; host buffers; points as rows. */
DDFMM_API int ddfmm_direct_sum(int kernel, const double* trg, int m, const double* src, int n,
                     const double* den, int nrhs, double* out);

/* The same direct sum on device buffers (trg m x 3, src n x 3, den nrhs x n complex,
   out nrhs x m complex), queued on `stream` (0 = the null stream), with no device
   allocation: the partial sums of its source ranges live in a static 32 MiB array of
   the library on each GPU.  No counterpart in the reference: for callers that cannot
   call hipMalloc (graph capture, allocation-free serving loops).  Results are
   bit-identical to ddfmm_direct_sum.  Every call on one GPU shares that GPU's array,
   so calls that overlap on one GPU race.  Sizes whose partial sums would exceed the
   array are rejected with DDFMM_EINVAL (the source split keeps them within it). */
DDFMM_API int ddfmm_direct_sum_noalloc(int kernel, const double* trg_dev, int m, const double* src_dev, int n,
                             const double* den_dev, int nrhs, double* out_dev, void* stream);

/* Dense kernel block (Kernel3d::kernel, kernel3d.cpp:26-107): out(i,j) = G(trg_i, src_j),
   m x n column-major complex; points as rows. */
DDFMM_API int ddfmm_kernel_matrix(int kernel, const double* trg, int m, const double* src, int n, double* out);

DDFMM_API int ddfmm_get_stats(ddfmm_ctx* ctx, ddfmm_stats* st);
/* Canonical text of every box's lists and direction sets (integer parity checks). */
DDFMM_API int ddfmm_dump_lists(ddfmm_ctx* ctx, char* buf, int64_t cap, int64_t* len);
/* Inspection: copy `count` complex entries of internal vector buffer `buf` (leaf-sorted
   densities/potentials, LF/HF equivalent densities and check values; numbering of
   ddfmm_amd/csrc/plan.hpp) for right-hand side `rhs`, starting at `off`. */
DDFMM_API int ddfmm_read_buffer(ddfmm_ctx* ctx, int buf, int rhs, int64_t off, int64_t count, double* out);
/* Per-stage timing: enable, then read, per stage of the last eval, its name, milliseconds
   (HIP events on the evaluation stream), algorithmic work per RHS (GEMM: useful complex
   flops 8*M*K*N; kernel apply: target-source interactions; FFT: spectra; V list: pairs)
   and kind (stage kind + 16 * arithmetic class, 0 translation / 1 FP64 solve).
   on = 2 issues every stage on the evaluation stream (no overlap between stages), so
   that each time is the stage's own. */
DDFMM_API int ddfmm_set_timing(ddfmm_ctx* ctx, int on);
DDFMM_API int ddfmm_stage_times(ddfmm_ctx* ctx, int cap, char* names /* cap x 24 */, double* ms, double* work,
                      int* kinds, int* count);
/* Per stage, the MFMA flops per RHS the GEMM kernel executes under its tiling (16-row
   blocks, 16-column waves, 4-deep k steps, present K blocks only; 0 for other stages):
   executed / useful work is the padding overhead.  No counterpart in the reference. */
DDFMM_API int ddfmm_stage_exec(ddfmm_ctx* ctx, int cap, double* exec, int* count);

#ifdef __cplusplus
}
#endif

#endif

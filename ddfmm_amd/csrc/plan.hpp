// Flattened evaluation plan: the octree, lists and operators of a Wave3d setup turned
// into device-resident descriptor tables.  The evaluation (wave3d_eval.cpp:299-415) is
// a fixed sequence of stages; every stage is one kernel launch over a contiguous range
// of descriptors.  Shared by the host planner and the HIP kernels, so POD only.
#pragma once

#include <cstdint>

namespace ddfmm {

// vector buffers ([rhs][...] complex, one rhs slab per right-hand side)
enum Buf : int {
    BUF_DEN = 0,    // leaf-sorted densities, N
    BUF_VAL,        // leaf-sorted potentials, N
    BUF_LF_UPEQ,    // LF upward equivalent densities, nLF x NL
    BUF_LF_CHK,     // LF upward check values, then downward check values
    BUF_LF_TMP,     // LF pinv middle vector
    BUF_LF_DNEQ,    // LF downward equivalent densities
    BUF_HF_UP,      // HF directional upward equivalent densities (out-slots)
    BUF_HF_DN,      // HF directional downward check values (in-slots)
    BUF_HF_A,       // HF per-level scratch
    BUF_HF_B,
    BUF_SPEC,       // LF V-list source spectra of one level, (2P)^3 each
    BUF_RED,        // HF M2L/L2L contributions of one level, one slot per (target, source)
    BUF_TSPEC,      // LF V-list target spectra of one level, (2P)^3 each, in V-target order
    BUF_NMIR,       // near-field mirror products, one slot per pair evaluated both ways (the
                    // slots into one leaf adjacent, in source-task order)
    BUF_COUNT
};
// setup-time GEMMs read and write the operator table itself (FP64), as buffer BUF_OPMAT
constexpr int BUF_OPMAT = BUF_COUNT;

// near-field pairs evaluated once for both directions: the evaluating leaf has at most this
// many points (one pass of four row tiles of 16: each wave holds every target)
constexpr int MIRROR_MAX_TARGETS = 64;

// kernel-apply: out(target_i) (+)= sum_src sum_j G(t_i, s_j) v_j
// point coordinates: pts[off + i] * scale + c
// moff >= 0 (near field, pairs of leaves evaluated once for both directions): the source
// is a leaf whose own task skips this task's leaf; the kernel also writes
// sum_i G(t_i, s_j) den(t_i) into BUF_NMIR[moff + j], which a reduction adds to the
// source leaf's potentials
struct KSrc {
    int64_t off;
    int64_t voff;
    int64_t moff = -1;
    double c[3];
    double scale;
    int count;
    int vbuf;
};
struct KTask {
    int64_t off;
    int64_t ooff;
    double c[3];
    double scale;
    int count;
    int obuf;
    int s0, s1;
    int mode;       // 0 store, 1 add
    int pad;
};

// grouped complex GEMM: C[:, ccol] (=, +=) diag(scale) * op(A) * B[:, bcol]
// B columns: contiguous (kb = 0, offset colB[bcol0 + n]) or blocked along K in blocks of kb
// rows, each block gathered from its own offset colB[bcol0 + n * K/kb + block], with the
// matching A column block at colA[acol0 + block] (used to concatenate the non-empty
// children of a box along K without padding).
struct GGroup {
    int64_t aoff;   // A offset in the operator table
    int64_t soff;   // row scale (length M) offset, or -1
    int M, K, lda, transA;
    int bbuf, cbuf;
    int ncols, col0;
    int mode;       // 0 store, 1 add
    int kb;         // K block size of the blocked gather (0 = contiguous)
    int bcol0;      // first B offset in colB
    int acol0;      // kb > 0: first A block offset in colA (block b at colA[acol0 + b])
};
struct GTile {
    int g;
    int m0, n0;     // first row, first column
    int mb;         // row blocks of 16 (at most GEMM_MB)
};
constexpr int GEMM_MB = 7;      // rows per tile <= 112
constexpr int GEMM_TN = 64;     // columns per tile (16 per wave)

// LF FFT-diagonal M2L (wave3d_eval.cpp:602-661)
struct FftSrc {
    int64_t in_off;     // LF_UPEQ slot
    int64_t spec_off;   // SPEC slot
    int grid;           // grid-index table offset
    int pad;
};
// One ST_VLIST stage covers the targets of one level: their spectra accumulate, per chunk
// and frequency slice, into BUF_TSPEC (target i of the stage at i * (2P)^3), then one
// inverse transform per target adds them into LF_CHK.
struct VTarget {
    int64_t out_off;    // LF_CHK slot
    int p0, p1;         // pair range
    int64_t tab_off;    // translation tensor table of this width (343 x (2P)^3), same for the stage
    int grid;
    int k0;             // first packed pair in vpk (a multiple of 8; the target's p1 - p0 packed
                        // pairs, then zero pairs up to a multiple of 8)
};
struct VPair {
    int64_t spec_off;
    int offidx;         // (a+3) + 7(b+3) + 49(c+3), (a, b, c) = target - source in box widths
    int pad;
};
// V-list chunk: Morton-consecutive targets [t0, t1) of one level whose V-list sources, the
// union vsrc[u0, u1) of at most VCH_SRC rows (row 0 a zero row, then source spectrum
// indices), are staged in LDS together.  The device reads each pair packed as
// local row | (offidx << 16); a zero entry (row 0) pads.
// The chunk's packed pairs are the contiguous range vpk[k0, k1) (its targets' lists in
// target order), at most VCH_PK words, staged in LDS with the sources.
struct VChunk {
    int t0, t1;
    int u0, u1;
    int k0, k1;
};
constexpr int VCH_SRC = 200;
constexpr int VCH_TRG = 128;
constexpr int VCH_PK = 4096;

// segmented reduction: out[out_off + m] (=|+=) sum_{j < cnt} in[red_off + j * M + m], in =
// the stage's buffer (BUF_RED: HF M2L/L2L contributions; BUF_NMIR: near-field mirror
// products); the contributions of one target are adjacent, in a fixed order: deterministic
// sums
struct RTarget {
    int64_t out_off;
    int64_t red_off;
    int cnt, M;
    int obuf, mode;     // mode 0 store, 1 add
};

// setup-time kernel matrix generation: out(i,j) = G((sgn*T_i + tshift) - (S_j + sshift))
struct KGen {
    int64_t toff, soff, out;   // offsets into the generator point table / operator table
    double tshift[3];
    double sshift[3];
    int m, n;
    int tneg, pad;
};

enum StageKind : int {
    ST_ZERO = 0, ST_GATHER, ST_SCATTER, ST_KAPPLY, ST_GEMM, ST_FFTFWD, ST_VLIST, ST_REDUCE,
    ST_FORK,        // the second stream waits for everything issued so far on the first (event buf)
    ST_JOIN         // the first stream waits for everything issued so far on the second (event buf)
};
// Two streams per evaluation: the caller's (0) carries the HF pipeline and the LF passes
// that depend on it; the second (1) runs, concurrently, the work that needs only the
// densities (near field) or only the LF upward pass (FFT, V list, X list).
constexpr int STAGE_EVENTS = 3;   // fork after gather, fork after the LF upward pass, join
// Arithmetic class of a stage.  In the mixed-precision mode translation stages
// (kernel-generated matrices: M2M/L2L, U2U/D2D child kernels, HF M2L, near field, V list)
// run in FP32 on FP32 vectors, while the ill-conditioned check->equivalent solves
// (LF pinv, HF E1/E2/E3 chains) keep FP64 matrices and FP64 accumulation.
enum StagePrec : int { PREC_TRANSLATION = 0, PREC_SOLVE = 1 };
struct Stage {
    int kind;
    int begin, end;     // descriptor range (tasks / tiles / sources / targets)
    int buf;            // ST_ZERO: buffer id; ST_FORK/ST_JOIN: event index; ST_VLIST: first chunk;
                        // ST_REDUCE: the contributions' buffer
    int64_t count;      // ST_ZERO: elements per rhs; ST_VLIST: chunks
    int prec;           // StagePrec
    int stream;         // 0: the evaluation (caller's) stream, 1: the concurrent second stream
    double work;        // per RHS: GEMM useful flops, KAPPLY interactions, FFT spectra, V-list pairs,
                        // reduction elements read
    int mb;             // ST_GEMM: row blocks of every tile in the range (one kernel instance);
                        // ST_KAPPLY: 1 when sources carry mirror slots (KSrc::moff)
    int lstage;         // ST_GEMM: logical stage (all tile heights of one planned GEMM stage), else -1
    double exec;        // ST_GEMM: MFMA flops per RHS the kernel executes (padding included)
    char name[24];
};

}  // namespace ddfmm

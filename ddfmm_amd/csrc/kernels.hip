// HIP kernels of the directional-FMM evaluation for gfx950 (CDNA4, wave64).
//
//   kapply_kernel  generated-kernel block apply: out(t_i) (+)= sum_j G(t_i, s_j) v_j with
//                  LDS-staged source tiles; serves the near field (U list), S2M, L2T,
//                  W/X lists (Kernel3d::kernel + zgemv, kernel3d.cpp:26-107,
//                  wave3d_eval.cpp:441-451, 550-560, 585-718)
//   gemm_kernel    grouped complex GEMM on MFMA (16x16x4, f64 or f32) over (box, rhs)
//                  columns with gather/scatter and store/add epilogues: every translation
//                  the reference applies as a per-box zgemv (vecmatop.cpp:102-125)
//   reduce_kernel  fixed-order sum of per-contribution slots: the HF M2L/L2L targets
//                  (the reference's zgemv accumulations into one vector) and the
//                  near-field mirror products of pairs evaluated once
//   fft_fwd_kernel / vlist_acc_kernel / vlist_inv_kernel
//                  FFT-diagonal LF M2L on the (2P)^3 grid
//                  (wave3d_eval.cpp:602-661)
//   kgen_kernel    setup-time translation-matrix generation (mlib3d.cpp:106-280 and the
//                  per-offset HF M2L matrices of wave3d_eval.cpp:876-928)
//
// Vectors are double2 (FP64 mode) or float2 (mixed mode).  In the mixed mode the
// kernel G is evaluated from FP64 coordinate differences with FP32 sqrt/rcp and the
// hardware sin/cos on the phase reduced to [0, 1) revolutions (k = 2*pi, so
// exp(ikr) = exp(2*pi*i*fract(r))).
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "device.hpp"

// No implicit FP contraction: a + b * c is fused only where the code says fma.  Kernels
// that serve different right-hand-side counts (1-RHS and multi-RHS instances) then round
// every operation the same way, which keeps each right-hand side's bits independent of
// the batch it is evaluated in.
#pragma clang fp contract(off)

namespace ddfmm {

namespace {

template <class V> struct Real;
template <> struct Real<double2> { using T = double; };
template <> struct Real<float2> { using T = float; };

template <class V>
__device__ __forceinline__ V mk(double x, double y) {
    V v;
    v.x = (typename Real<V>::T)x;
    v.y = (typename Real<V>::T)y;
    return v;
}

template <class V>
__device__ __forceinline__ void cfma(V& acc, V a, V b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}

// FP64 G(r), k = 2*pi: e^{ikr}/r (type 0) or e^{ikr} (type 1); r^2 < 1e-16 => r := 1.
__device__ __forceinline__ double2 green64(double dx, double dy, double dz, int ktype) {
    double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    r2 = r2 < 1e-16 ? 1.0 : r2;
    double r = sqrt(r2);
    double s, c;
    sincospi(2.0 * r, &s, &c);
    if (ktype == 0) {
        double ir = 1.0 / r;
        return make_double2(c * ir, s * ir);
    }
    return make_double2(c, s);
}

// FP32 G(r) from FP64 differences: r^2 in FP64 (full-rate FMAs, one conversion), 1/r from
// v_rsq_f32, r = r^2 / r, and v_sin_f32 / v_cos_f32 on the phase fract(r) in revolutions
// (three transcendental operations).
__device__ __forceinline__ float2 green32(double dx, double dy, double dz, int ktype) {
    float r2 = (float)fma(dz, dz, fma(dy, dy, dx * dx));
    r2 = r2 < 1e-16f ? 1.0f : r2;
    const float ir = __builtin_amdgcn_rsqf(r2);
    // phase in [0, 1) revolutions, the argument range of the hardware sin/cos: v_fract_f32
    // of the rounded r (its rounding, at most 2^-23 revolution for r < 4, is below that of
    // the rsq approximation in r)
    const float t = __builtin_amdgcn_fractf(r2 * ir);
    float s = __builtin_amdgcn_sinf(t), c = __builtin_amdgcn_cosf(t);
    if (ktype == 0) return make_float2(c * ir, s * ir);
    return make_float2(c, s);
}

template <class V>
__device__ __forceinline__ V green(double dx, double dy, double dz, int ktype);
template <>
__device__ __forceinline__ double2 green<double2>(double dx, double dy, double dz, int ktype) {
    return green64(dx, dy, dz, ktype);
}
template <>
__device__ __forceinline__ float2 green<float2>(double dx, double dy, double dz, int ktype) {
    return green32(dx, dy, dz, ktype);
}

template <class V>
__device__ __forceinline__ V* bufp(const BufTable& bt, int b) { return (V*)bt.p[b]; }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2): give
// XCD x the contiguous chunk [x*q + min(x, r), ...) of the n work items (q = n/8, r = n%8)
// so that neighbouring items (Morton-ordered boxes, tiles of one operator) share an L2.
__device__ __forceinline__ int xcd_item(int b, int n) {
    const int x = b & 7, k = b >> 3;
    const int q = n >> 3, r = n & 7;
    return x * q + min(x, r) + k;
}

// ---------------------------------------------------------------- kernel apply
constexpr int KA_THREADS = 256;

// One block per task.  A pass covers nt = min(count, 256) targets; when nt < 256 the
// 256 threads split into G = 256 / nt source groups (thread = target x group) that
// stride through each LDS source tile, and the partial sums are reduced through LDS.
template <class V, int RC>
__global__ __launch_bounds__(KA_THREADS) void kapply_kernel(
    const KTask* __restrict__ tasks, const KSrc* __restrict__ srcs,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    BufTable bt, int task0, int ntask, int nrhs, int ktype) {
    const int r0 = blockIdx.y * RC;
    const int nr = min(RC, nrhs - r0);
    __shared__ double sx[KA_THREADS], sy[KA_THREADS], sz[KA_THREADS];
    __shared__ V sv[RC][KA_THREADS];
    __shared__ V red[RC][KA_THREADS];
    // grid-stride over the tasks (a grid smaller than the task count keeps a launch on the
    // second stream to a few blocks per CU); gridDim.x is a multiple of 8 or >= ntask
    for (int bi = blockIdx.x; bi < ntask; bi += gridDim.x) {
    const KTask t = tasks[task0 + bi];   // tasks are ordered by cost, longest first
    for (int tb = 0; tb < t.count; tb += KA_THREADS) {
        const int nt = min(KA_THREADS, t.count - tb);
        const int G = KA_THREADS / nt;
        const int ti = threadIdx.x % nt, gi = threadIdx.x / nt;
        const bool act = gi < G;
        const int i = tb + ti;
        double tx = 0, ty = 0, tz = 0;
        if (act) {
            tx = px[t.off + i] * t.scale + t.c[0];
            ty = py[t.off + i] * t.scale + t.c[1];
            tz = pz[t.off + i] * t.scale + t.c[2];
        }
        V acc[RC];
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = mk<V>(0, 0);
        for (int si = t.s0; si < t.s1; ++si) {
            const KSrc s = srcs[si];
            const V* vb = bufp<V>(bt, s.vbuf) + s.voff;
            const int64_t vs = bt.stride[s.vbuf];
            for (int sb = 0; sb < s.count; sb += KA_THREADS) {
                const int n = min(KA_THREADS, s.count - sb);
                __syncthreads();
                if ((int)threadIdx.x < n) {
                    const int64_t j = s.off + sb + threadIdx.x;
                    sx[threadIdx.x] = px[j] * s.scale + s.c[0];
                    sy[threadIdx.x] = py[j] * s.scale + s.c[1];
                    sz[threadIdx.x] = pz[j] * s.scale + s.c[2];
#pragma unroll
                    for (int q = 0; q < RC; ++q)
                        sv[q][threadIdx.x] = q < nr ? vb[(int64_t)(r0 + q) * vs + sb + threadIdx.x] : mk<V>(0, 0);
                }
                __syncthreads();
                if (act) {
                    for (int j = gi; j < n; j += G) {
                        const V g = green<V>(tx - sx[j], ty - sy[j], tz - sz[j], ktype);
#pragma unroll
                        for (int q = 0; q < RC; ++q) cfma(acc[q], g, sv[q][j]);
                    }
                }
            }
        }
        if (G > 1) {                      // reduce the source groups of each target
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RC; ++q) red[q][threadIdx.x] = acc[q];
            __syncthreads();
            if (gi == 0) {
                for (int h = 1; h < G; ++h)
#pragma unroll
                    for (int q = 0; q < RC; ++q) {
                        const V o = red[q][ti + h * nt];
                        acc[q].x += o.x;
                        acc[q].y += o.y;
                    }
            }
        }
        if (gi == 0) {
            V* ob = bufp<V>(bt, t.obuf) + t.ooff + i;
            const int64_t os = bt.stride[t.obuf];
#pragma unroll
            for (int q = 0; q < RC; ++q) {
                if (q >= nr) break;
                V* o = ob + (int64_t)(r0 + q) * os;
                if (t.mode == 0) {
                    *o = acc[q];
                } else {
                    V v = *o;
                    v.x += acc[q].x;
                    v.y += acc[q].y;
                    *o = v;
                }
            }
        }
        __syncthreads();
    }
    }
}

// Kernel apply on the matrix cores (mixed precision): out_q(t) = sum_s G(t, s) v_q(s) for
// 8 * NB right-hand sides, per block of 16 targets x 4 sources as two real products on
// v_mfma_f32_16x16x4f32 whose 16 columns are n = (rhs q, re/im):
//   C += Re G . B1,   B1[s][(q, 0)] = Re v_q(s),   B1[s][(q, 1)] = Im v_q(s)
//   C += Im G . B2,   B2[s][(q, 0)] = -Im v_q(s),  B2[s][(q, 1)] = Re v_q(s)
// so that C[t][(q, 0)] = Re out_q(t) and C[t][(q, 1)] = Im out_q(t).  Lane i + 16 k holds
// A[i][k] and B[k][i]: it evaluates G for its own pair (target i, source k) on the VALU
// (FP64 differences, FP32 G, the formula of the VALU path) and feeds Re G and Im G to the
// two products unchanged; its B entries come from the density of source k for right-hand
// side i / 2.  A block owns one task; targets go in passes of 256 (16 row tiles of 16),
// the four waves split the row tiles (up to 4 each, every source step serving all of
// them) and, for up to 8 row tiles, the sources, with a fixed-order reduction through
// LDS.  Every column's k order is the same for any right-hand-side count, so each
// right-hand side keeps its bits in any batch.
constexpr int KM_ST = 256;    // sources per LDS tile

typedef float kacc4 __attribute__((ext_vector_type(4)));

// One wave's pass over the source groups gs = g0, g0 + gstep, ... < ng of an LDS tile for
// its NM row tiles.  The next group's coordinates and densities are read from LDS while
// the current group is evaluated.
template <int NM, int NB, int KT>
__device__ __forceinline__ void kmfma_steps(kacc4 (&acc)[4][NB], const double (&tx)[4], const double (&ty)[4],
                                            const double (&tz)[4], const double* sx, const double* sy,
                                            const double* sz, const float2 (*sv)[8 * NB + 1], int g0, int gstep,
                                            int ng, int g, int li) {
    const bool imcol = li & 1;
    if (g0 >= ng) return;
    int j = 4 * g0 + g;
    double xs = sx[j], ys = sy[j], zs = sz[j];
    float2 v[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) v[c] = sv[j][8 * c + (li >> 1)];
    for (int gs = g0; gs < ng; gs += gstep) {
        const int jn = gs + gstep < ng ? 4 * (gs + gstep) + g : j;
        const double xn = sx[jn], yn = sy[jn], zn = sz[jn];
        float2 vn[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) vn[c] = sv[jn][8 * c + (li >> 1)];
        float b1[NB], b2[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            b1[c] = imcol ? v[c].y : v[c].x;
            b2[c] = imcol ? v[c].x : -v[c].y;
        }
        float gr[NM], gi[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            const float2 gv = green32(tx[m] - xs, ty[m] - ys, tz[m] - zs, KT);
            gr[m] = gv.x;
            gi[m] = gv.y;
        }
#pragma unroll
        for (int m = 0; m < NM; ++m)
#pragma unroll
            for (int c = 0; c < NB; ++c) acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(gr[m], b1[c], acc[m][c], 0, 0, 0);
#pragma unroll
        for (int m = 0; m < NM; ++m)
#pragma unroll
            for (int c = 0; c < NB; ++c) acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(gi[m], b2[c], acc[m][c], 0, 0, 0);
        j = jn;
        xs = xn;
        ys = yn;
        zs = zn;
#pragma unroll
        for (int c = 0; c < NB; ++c) v[c] = vn[c];
    }
}

// Orders one wave's LDS writes before its later LDS reads (and those reads before later
// writes): LDS executes a wave's instructions in order, so only the compiler must not move
// the accesses across.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Near-field pairs evaluated once for both directions.  For a mirrored source leaf s of the
// task's leaf k (KSrc::moff >= 0) one wave takes 16-source tiles st = st0, st0 + 4, ...,
// each lane reading its four sources from memory (no block staging: the waves share no
// sources, so they run without barriers), and evaluates G over its row tiles m x the
// tile's 16 sources, four values per lane and row tile in the forward operand layout
// (lane i + 16 g: target i, sources 4 jj + g):
//   forward  acc[m] += G v_s                      (as kmfma_steps, groups 4 st + jj)
//   mirror   macc   += G^T den_k over the targets (K = targets 4 j + g of row tile m)
// The mirror's A operand is the transposed tile (lane i + 16 g: source i, targets 4 j + g),
// taken from per-wave LDS copies of two row tiles at a time.  G of all row tiles comes
// first and the products of different row tiles interleave (independent accumulators;
// the mirror sums even and odd row tiles apart and adds the two at the end).  The task has
// at most four row tiles and this wave holds them all (the waves split the sources), so
// macc is the tile's whole mirror: stored into the source's slot.
template <int NM, int NB, int KT>
__device__ __forceinline__ void kmfma_mirror(kacc4 (&acc)[4][NB], const double (&tx)[4], const double (&ty)[4],
                                             const double (&tz)[4], const double* px, const double* py,
                                             const double* pz, const float2* vb, int64_t vs,
                                             const float2 (*tv)[8 * NB + 1], float2* gt, int n, int st0, int g,
                                             int li, float* mslot, int64_t ms, int r0, int nr) {
    const bool imcol = li & 1;
    const int nst = (n + 15) >> 4;
    for (int st = st0; st < nst; st += 4) {
        float gr[NM][4], gi[NM][4];
        float b1[4][NB], b2[4][NB];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            // sources past n: coordinates 0 (a finite G) and zero densities
            const int j = 16 * st + 4 * jj + g;
            const bool in = j < n;
            const double xs = in ? px[j] : 0.0, ys = in ? py[j] : 0.0, zs = in ? pz[j] : 0.0;
#pragma unroll
            for (int c = 0; c < NB; ++c) {
                const int q = 8 * c + (li >> 1);
                const float2 v = in && q < nr ? vb[(int64_t)(r0 + q) * vs + j] : make_float2(0.f, 0.f);
                b1[jj][c] = imcol ? v.y : v.x;
                b2[jj][c] = imcol ? v.x : -v.y;
            }
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const float2 gv = green32(tx[m] - xs, ty[m] - ys, tz[m] - zs, KT);
                gr[m][jj] = gv.x;
                gi[m][jj] = gv.y;
            }
        }
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
            for (int m = 0; m < NM; ++m)
#pragma unroll
                for (int c = 0; c < NB; ++c) acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(gr[m][jj], b1[jj][c], acc[m][c], 0, 0, 0);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
            for (int m = 0; m < NM; ++m)
#pragma unroll
                for (int c = 0; c < NB; ++c) acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(gi[m][jj], b2[jj][c], acc[m][c], 0, 0, 0);
        kacc4 macc[2][NB];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < NB; ++c) macc[p][c] = kacc4{0, 0, 0, 0};
#pragma unroll
        for (int h = 0; h < (NM + 1) / 2; ++h) {
            constexpr int T = 16 * 17;       // one tile [source][target], rows padded to 17
            const int mh = NM - 2 * h < 2 ? NM - 2 * h : 2;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                if (p >= mh) break;
#pragma unroll
                for (int jj = 0; jj < 4; ++jj)
                    gt[p * T + (4 * jj + g) * 17 + li] = make_float2(gr[2 * h + p][jj], gi[2 * h + p][jj]);
            }
            wave_lds_order();
            float hr[2][4], hi[2][4], d1[2][4][NB], d2[2][4][NB];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                if (p >= mh) break;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float2 e = gt[p * T + li * 17 + 4 * j + g];
                    hr[p][j] = e.x;
                    hi[p][j] = e.y;
#pragma unroll
                    for (int c = 0; c < NB; ++c) {
                        const float2 w = tv[16 * (2 * h + p) + 4 * j + g][8 * c + (li >> 1)];
                        d1[p][j][c] = imcol ? w.y : w.x;
                        d2[p][j][c] = imcol ? w.x : -w.y;
                    }
                }
            }
            wave_lds_order();
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    if (p >= mh) break;
#pragma unroll
                    for (int c = 0; c < NB; ++c)
                        macc[p][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(hr[p][j], d1[p][j][c], macc[p][c], 0, 0, 0);
                }
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int p = 0; p < 2; ++p) {
                    if (p >= mh) break;
#pragma unroll
                    for (int c = 0; c < NB; ++c)
                        macc[p][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(hi[p][j], d2[p][j][c], macc[p][c], 0, 0, 0);
                }
        }
        // rows 4 g + v: sources 16 st + 4 g + v; column li: (rhs 8 c + li / 2, re/im)
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const int q = 8 * c + (li >> 1);
            if (q >= nr) continue;
            float* o = mslot + 2 * ((int64_t)(r0 + q) * ms) + (li & 1);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int jx = 16 * st + 4 * g + v;
                if (jx >= n) continue;
                o[2 * jx] = macc[0][c][v] + macc[1][c][v];
            }
        }
    }
}

// Waves per SIMD the register allocation targets: the 8-RHS instance at five (96 VGPRs, a
// few spilled: G on the vector ALU needs more waves in flight than the 111 VGPRs of a
// free allocation allow; config 3 near field 16.5 -> 15.3 ms), its mirror instance at
// three (168 VGPRs without spills); the 16-RHS instances are held by LDS to three and two
// blocks per CU and keep their free allocation.
template <int NB, bool MIR> struct KmfmaOcc { static constexpr int waves = NB == 1 ? (MIR ? 3 : 5) : 1; };

template <int NB, int KT, bool MIR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KmfmaOcc<NB, MIR>::waves))) void kapply_mfma_kernel(
    const KTask* __restrict__ tasks, const KSrc* __restrict__ srcs,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    BufTable bt, int task0, int ntask, int nrhs) {
    constexpr int NR = 8 * NB;                       // right-hand sides of the block
    __shared__ double sx[KM_ST], sy[KM_ST], sz[KM_ST];
    // densities, right-hand side fastest, rows padded by one entry (the staging stores of
    // consecutive sources then spread over the banks); after the last source tile, the
    // partial sums of waves that split the sources (4 waves x 4 row tiles x 64 lanes x 4 NB)
    __shared__ __align__(16) float2 sv[KM_ST][NR + 1];
    static_assert(sizeof(float2) * KM_ST * (NR + 1) >= sizeof(float) * 4 * 4 * 64 * 4 * NB, "reduction fits the tile");
    float* red = (float*)&sv[0][0];
    // mirror mode: the densities of a task that evaluates mirrored sources (at most 64
    // targets, one pass: every wave holds all its row tiles; rows past the task zero), the
    // mirror's operand, and one transposition tile per wave
    __shared__ __align__(16) float2 tv[MIR ? MIRROR_MAX_TARGETS : 1][NR + 1];
    __shared__ float2 gts[MIR ? 4 : 1][2 * 16 * 17];
    const int r0 = blockIdx.y * NR;
    const int nr = min(NR, nrhs - r0);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, li = lane & 15;
    for (int bi = blockIdx.x; bi < ntask; bi += gridDim.x) {
    const KTask t = tasks[task0 + bi];   // tasks are ordered by cost, longest first
    for (int tb = 0; tb < t.count; tb += 256) {
        const int nt = min(256, t.count - tb);
        const int nT = (nt + 15) >> 4;
        // up to 4 row tiles: every wave takes all of them and a quarter of the sources;
        // up to 8: two waves per row tile set; beyond: the waves split the row tiles
        const int wrt = nT > 8 ? 4 : nT > 4 ? 2 : 1, wsrc = 4 / wrt;
        const int myrt0 = wave % wrt, mysrc = wave / wrt;
        // this wave's row tiles: myrt0 + wrt * m < nT, m < nm
        const int nm = min(4, (nT - myrt0 + wrt - 1) / wrt);
        double tx[4], ty[4], tz[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int i = (myrt0 + wrt * m) * 16 + li;
            tx[m] = ty[m] = tz[m] = 0;
            if (m < nm && i < nt) {
                tx[m] = px[t.off + tb + i] * t.scale + t.c[0];
                ty[m] = py[t.off + tb + i] * t.scale + t.c[1];
                tz[m] = pz[t.off + tb + i] * t.scale + t.c[2];
            }
        }
        kacc4 acc[4][NB];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int c = 0; c < NB; ++c) acc[m][c] = kacc4{0, 0, 0, 0};
        if (MIR && t.count <= MIRROR_MAX_TARGETS) {    // block-uniform
            const float2* den = bufp<float2>(bt, BUF_DEN) + t.off;
            const int64_t ds = bt.stride[BUF_DEN];
            for (int e = tid; e < MIRROR_MAX_TARGETS * NR; e += 256) {
                const int i = e / NR, q = e % NR;
                tv[i][q] = (i < nt && q < nr) ? den[(int64_t)(r0 + q) * ds + i] : make_float2(0.f, 0.f);
            }
            __syncthreads();
        }
        for (int si = t.s0; si < t.s1; ++si) {
            const KSrc s = srcs[si];
            const float2* vb = bufp<float2>(bt, s.vbuf) + s.voff;
            const int64_t vs = bt.stride[s.vbuf];
            if (MIR && s.moff >= 0) {
                // the first tile rotates with the source, which spreads the tiles evenly
                // over the waves
                const int st0 = (wave + si) & 3;
                float* ms = (float*)(bufp<float2>(bt, BUF_NMIR) + s.moff);
                const int64_t mst = bt.stride[BUF_NMIR];
                float2* gt = gts[MIR ? wave : 0];
                const double *qx = px + s.off, *qy = py + s.off, *qz = pz + s.off;
                switch (nm) {      // wave-uniform; every wave holds the task's row tiles
                case 4: kmfma_mirror<4, NB, KT>(acc, tx, ty, tz, qx, qy, qz, vb, vs, tv, gt, s.count, st0, g, li, ms, mst, r0, nr); break;
                case 3: kmfma_mirror<3, NB, KT>(acc, tx, ty, tz, qx, qy, qz, vb, vs, tv, gt, s.count, st0, g, li, ms, mst, r0, nr); break;
                case 2: kmfma_mirror<2, NB, KT>(acc, tx, ty, tz, qx, qy, qz, vb, vs, tv, gt, s.count, st0, g, li, ms, mst, r0, nr); break;
                case 1: kmfma_mirror<1, NB, KT>(acc, tx, ty, tz, qx, qy, qz, vb, vs, tv, gt, s.count, st0, g, li, ms, mst, r0, nr); break;
                default: break;
                }
                continue;
            }
            for (int sb = 0; sb < s.count; sb += KM_ST) {
                const int n = min(KM_ST, s.count - sb);
                __syncthreads();
                {   // sources past n: coordinates 0 (a finite G) and zero densities
                    const int j = tid;
                    double x = 0, y = 0, z = 0;
                    if (j < n) {
                        x = px[s.off + sb + j] * s.scale + s.c[0];
                        y = py[s.off + sb + j] * s.scale + s.c[1];
                        z = pz[s.off + sb + j] * s.scale + s.c[2];
                    }
                    sx[j] = x;
                    sy[j] = y;
                    sz[j] = z;
#pragma unroll
                    for (int q = 0; q < NR; ++q)
                        sv[j][q] = (j < n && q < nr) ? vb[(int64_t)(r0 + q) * vs + sb + j] : make_float2(0.f, 0.f);
                }
                __syncthreads();
                // groups of 4 sources: this wave takes groups mysrc, mysrc + wsrc, ...; the
                // staged rows past n hold zero densities
                const int ng = (n + 3) >> 2;
                switch (nm) {      // wave-uniform
                case 4: kmfma_steps<4, NB, KT>(acc, tx, ty, tz, sx, sy, sz, sv, mysrc, wsrc, ng, g, li); break;
                case 3: kmfma_steps<3, NB, KT>(acc, tx, ty, tz, sx, sy, sz, sv, mysrc, wsrc, ng, g, li); break;
                case 2: kmfma_steps<2, NB, KT>(acc, tx, ty, tz, sx, sy, sz, sv, mysrc, wsrc, ng, g, li); break;
                case 1: kmfma_steps<1, NB, KT>(acc, tx, ty, tz, sx, sy, sz, sv, mysrc, wsrc, ng, g, li); break;
                default: break;    // a wave with no row tile in this pass
                }
            }
        }
        // waves that split the sources of one row tile: sum their partials (fixed order)
        if (wsrc > 1) {
            __syncthreads();
            auto slot = [&](int w, int m, int c, int v) { return ((w * 4 + m) * 64 + lane) * 4 * NB + 4 * c + v; };
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                if (m >= nm) break;
#pragma unroll
                for (int c = 0; c < NB; ++c)
#pragma unroll
                    for (int v = 0; v < 4; ++v) red[slot(wave, m, c, v)] = acc[m][c][v];
            }
            __syncthreads();
            if (mysrc == 0) {
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    if (m >= nm) break;
#pragma unroll
                    for (int c = 0; c < NB; ++c)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            float a = red[slot(wave, m, c, v)];
                            for (int h = 1; h < wsrc; ++h) a += red[slot(wave + h * wrt, m, c, v)];
                            acc[m][c][v] = a;
                        }
                }
            }
        }
        // D[row][col] of a 16x16 f32 MFMA: lane holds rows 4 * (lane / 16) + v, column lane % 16
        if (mysrc == 0) {
            float* ob = (float*)(bufp<float2>(bt, t.obuf) + t.ooff + tb);
            const int64_t os = bt.stride[t.obuf];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                if (m >= nm) break;
                const int rt = myrt0 + wrt * m;
#pragma unroll
                for (int c = 0; c < NB; ++c) {
                    const int q = 8 * c + (li >> 1);
                    if (q >= nr) continue;
                    float* oq = ob + 2 * ((int64_t)(r0 + q) * os) + (li & 1);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = rt * 16 + 4 * g + v;
                        if (i >= nt) continue;
                        if (t.mode == 0) oq[2 * i] = acc[m][c][v];
                        else oq[2 * i] += acc[m][c][v];
                    }
                }
            }
        }
        __syncthreads();
    }
    }
}

// ---------------------------------------------------------------- grouped complex GEMM (MFMA)
template <class R> struct Mfma;
// Operand layouts (both): A lane l holds A[l%16][l/16], B lane l holds B[l/16][l%16].
// Output register v of lane l holds D[row(l/16, v)][l%16]; the row mapping differs
// between the f32 and the f64 instruction (checked on gfx950).
template <> struct Mfma<float> {
    typedef float acc4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc4 op(float a, float b, acc4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};
template <> struct Mfma<double> {
    typedef double acc4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc4 op(double a, double b, acc4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};

constexpr int MAXKB = 8;           // at most 8 K blocks (children of a box)

// LDS layouts: complex values interleaved (one 64-bit (f32) or 128-bit (f64) access per
// element); every 16-lane (f32) or 8-lane (f64) group of a fragment read touches one
// contiguous 128-byte row segment or distinct banks:
//   A slab [k][m], row stride SA = 16 MB + 4 (f32) or + 2 (f64) complex: the 4 k rows
//   written by 4 neighbouring lanes in the transposed-A store land in distinct banks;
//   B slab [n][k], row stride SB = 17 (odd: 16 columns at one k in distinct banks).
template <class R> struct Lds;
template <> struct Lds<float> { static constexpr int SAPAD = 4; using C = float2; };
template <> struct Lds<double> { static constexpr int SAPAD = 2; using C = double2; };

// C[:, ccol] (=|+=) diag(scale) * op(A) * B[:, bcol] on MFMA 16x16x4 (f32 or f64).
// A tile covers MB row blocks of 16 (the group's rows split evenly over its tiles; tiles
// are launched per MB) and GEMM_TN = 64 columns, 16 per wave; complex products as four
// real MFMAs.  K runs over segments (the whole K, or the K blocks present in at least
// one of the tile's columns) in slabs of 16 with a register-prefetch pipeline: slab s+1
// is loaded while slab s runs on the matrix cores.  Padding stays within 16 rows, 4 k
// and 16 columns: the MFMAs stop at the last k step of a segment, and a wave whose
// columns are all empty, or all lack the current K block, skips the segment's MFMAs.
template <class TA, class V, class R, int MB, bool G3M = false>
__global__ __launch_bounds__(256) void gemm_kernel(
    const GGroup* __restrict__ groups, const GTile* __restrict__ tiles,
    const int64_t* __restrict__ colB, const int64_t* __restrict__ colC, const int64_t* __restrict__ colA,
    const TA* __restrict__ opm, BufTable bt, int tile0, int nr) {
    using M = Mfma<R>;
    using T = typename Real<V>::T;
    using RC = typename Lds<R>::C;
    constexpr int KS = 16, TM = 16 * MB, SA = TM + Lds<R>::SAPAD, SB = 17;
    constexpr int NA = TM * KS / 256, NB = GEMM_TN * KS / 256;   // elements per thread and slab
    // columns are (box, rhs) pairs, rhs fastest: column n is box n / nr of the group in
    // right-hand side n % nr (that rhs slab of the buffer).  A planned tile (GEMM_TN boxes)
    // becomes nr launched tiles of GEMM_TN such columns; those past the group's last
    // column exit at once.
    const int bx = xcd_item(blockIdx.x, gridDim.x);
    GTile tl = tiles[tile0 + bx / nr];
    const GGroup g = groups[tl.g];
    const int ncx = g.ncols * nr;
    tl.n0 = tl.n0 * nr + GEMM_TN * (bx % nr);
    if (tl.n0 >= ncx) return;
    const V* B = bufp<V>(bt, g.bbuf);
    V* C = bufp<V>(bt, g.cbuf);
    const int64_t sB = bt.stride[g.bbuf], sC = bt.stride[g.cbuf];
    // f32: two LDS slab buffers (one barrier per slab); f64 slabs are twice the size and
    // take twice the MFMA time per k, so one buffer and two barriers per slab
    constexpr int NBUF = sizeof(R) == 4 ? 2 : 1;
    __shared__ RC As[NBUF][KS * SA];
    __shared__ RC Bs[NBUF][GEMM_TN * SB];
    __shared__ int64_t bo[MAXKB][GEMM_TN];   // B offset of (block, column), -1 = absent
    __shared__ int64_t ao[MAXKB];            // A offset of each active segment
    __shared__ int used[MAXKB][4], act[MAXKB], wm[MAXKB], nact_s;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int kb = g.kb ? g.kb : g.K;                   // segment length
    const int nkb = g.kb ? g.K / g.kb : 1;

    // per-thread slab elements, fixed for the tile: A (row, k) and B (k, column)
    const int mrows = g.M - tl.m0;                      // rows >= mrows are zero
    const int64_t kstride = g.transA ? 1 : g.lda;       // A offset per k
    int64_t aoffs[NA];
    int akk[NA], alds[NA];
#pragma unroll
    for (int l = 0; l < NA; ++l) {
        const int e = tid + 256 * l;
        int mm, kk;
        if (g.transA) {              // global A[k + m*lda]: 4 consecutive k per row
            kk = (e & 3) | ((e / (4 * TM)) << 2);
            mm = (e >> 2) % TM;
        } else {                     // global A[m + k*lda]: consecutive rows
            mm = e % TM;
            kk = e / TM;
        }
        const int64_t m = tl.m0 + mm;
        aoffs[l] = g.transA ? kk + m * g.lda : m + kk * (int64_t)g.lda;
        akk[l] = mm < mrows ? kk : KS;                  // KS: never loaded
        alds[l] = kk * SA + mm;
    }
    int64_t bc[NB];                                     // column offsets of the current segment
    int bkk[NB], blds[NB];
#pragma unroll
    for (int l = 0; l < NB; ++l) {
        const int e = tid + 256 * l;
        bkk[l] = e % KS;
        blds[l] = (e / KS) * SB + (e % KS);
    }
    // this lane's output column, fetched early (the epilogue needs it last)
    const int ncol = tl.n0 + wave * 16 + li;
    const int64_t coff = ncol < ncx ? colC[g.col0 + ncol / nr] + (int64_t)(ncol % nr) * sC : -1;

    // segments.  Contiguous B (kb = 0): one segment, column offsets read straight into
    // registers, waves with columns run.  Blocked B: a table of the K blocks present in
    // the tile's columns, with per-wave masks.
    int nact = 1, wm0 = 0;
    if (g.kb == 0) {
        const int nv = ncx - tl.n0;
        wm0 = nv >= GEMM_TN ? 15 : (1 << ((nv + 15) >> 4)) - 1;
#pragma unroll
        for (int l = 0; l < NB; ++l) {
            const int n = tl.n0 + (tid + 256 * l) / KS;
            bc[l] = n < ncx ? colB[g.bcol0 + n / nr] + (int64_t)(n % nr) * sB : -1;
        }
    } else {
        if (tid < MAXKB * 4) used[tid >> 2][tid & 3] = 0;
        __syncthreads();
        for (int e = tid; e < nkb * GEMM_TN; e += 256) {
            const int h = e / GEMM_TN, j = e % GEMM_TN;
            const int n = tl.n0 + j;
            int64_t v = n < ncx ? colB[g.bcol0 + (int64_t)(n / nr) * nkb + h] : -1;
            if (v >= 0) v += (int64_t)(n % nr) * sB;
            bo[h][j] = v;
            if (v >= 0) used[h][j >> 4] = 1;
        }
        __syncthreads();
        if (tid == 0) {
            int c = 0;
            for (int h = 0; h < nkb; ++h) {
                const int w = used[h][0] | (used[h][1] << 1) | (used[h][2] << 2) | (used[h][3] << 3);
                if (w) {
                    act[c] = h;
                    wm[c] = w;
                    ++c;
                }
            }
            nact_s = c;
        }
        __syncthreads();
        nact = nact_s;
        if (tid < nact) ao[tid] = colA[g.acol0 + act[tid]];
        __syncthreads();
    }
    TA ra[NA];
    V rb[NB];
    auto load = [&](int i, int k0, bool newseg) {
        const TA* Ab = opm + (g.kb ? ao[i] : g.aoff) + k0 * kstride;
        const int kl = kb - k0;
        if (newseg) {
#pragma unroll
            for (int l = 0; l < NB; ++l) bc[l] = bo[act[i]][(tid + 256 * l) / KS];
        }
#pragma unroll
        for (int l = 0; l < NA; ++l) {
            TA a;
            a.x = 0;
            a.y = 0;
            if (akk[l] < kl) a = Ab[aoffs[l]];
            ra[l] = a;
        }
        const V* Bk = B + k0;
#pragma unroll
        for (int l = 0; l < NB; ++l) {
            V b = mk<V>(0, 0);
            if (bc[l] >= 0 && bkk[l] < kl) b = Bk[bc[l] + bkk[l]];
            rb[l] = b;
        }
    };
    auto store = [&](int q) {
#pragma unroll
        for (int l = 0; l < NA; ++l) {
            RC v;
            v.x = (R)ra[l].x;
            v.y = (R)ra[l].y;
            As[q][alds[l]] = v;
        }
#pragma unroll
        for (int l = 0; l < NB; ++l) {
            RC v;
            v.x = (R)rb[l].x;
            v.y = (R)rb[l].y;
            Bs[q][blds[l]] = v;
        }
    };

    // G3M (FP32 contexts of the e^{ikr}/r kernel): three real products per complex one,
    // summed over k apart: k1 = (Ar + Ai) Br, k2 = Ar (Bi - Br), k3 = Ai (Br + Bi); Re C =
    // k1 - k3 and Im C = k1 + k2 in the epilogue.  Where |Ar| and |Ai| (or |Br| and |Bi|)
    // differ widely, k1 and k3 are both about Ai Br and cancel: the e^{ikr} kernel's larger
    // equivalent densities then lose accuracy, so it keeps four products (Re C in cr, Im C
    // in ci), as FP64 does.
    static_assert(!G3M || sizeof(R) == 4, "three products in FP32 only");
    typename M::acc4 cr[MB], ci[MB], c3[G3M ? MB : 1];
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            cr[a][v] = 0;
            ci[a][v] = 0;
            if constexpr (G3M) c3[a][v] = 0;
        }

    const int bb = (wave * 16 + li) * SB + lk, aa = lk * SA + li;
    // MFMAs of slab (i, k0) from LDS buffer q: four sweeps over the row blocks per k step,
    // so consecutive MFMAs never accumulate into the same registers
    auto mfma_slab = [&](int i, int k0, int q) {
        if (!((((g.kb ? wm[i] : wm0)) >> wave) & 1)) return;
        const int nsteps = (min(KS, kb - k0) + 3) >> 2;
        for (int s = 0; s < nsteps; ++s) {
            const RC bv = Bs[q][bb + 4 * s];
            const R br = bv.x, bi = bv.y, mbi = -bi;
            R ar[MB], ai[MB];
#pragma unroll
            for (int a = 0; a < MB; ++a) {
                const RC av = As[q][aa + 4 * s * SA + 16 * a];
                ar[a] = av.x;
                ai[a] = av.y;
            }
            if constexpr (G3M) {   // cr = sum k1, ci = sum k2, c3 = sum k3
                const R bd = bi - br, bs = br + bi;
#pragma unroll
                for (int a = 0; a < MB; ++a) cr[a] = M::op(ar[a] + ai[a], br, cr[a]);
#pragma unroll
                for (int a = 0; a < MB; ++a) ci[a] = M::op(ar[a], bd, ci[a]);
#pragma unroll
                for (int a = 0; a < MB; ++a) c3[a] = M::op(ai[a], bs, c3[a]);
            } else {
#pragma unroll
                for (int a = 0; a < MB; ++a) cr[a] = M::op(ar[a], br, cr[a]);
#pragma unroll
                for (int a = 0; a < MB; ++a) ci[a] = M::op(ar[a], bi, ci[a]);
#pragma unroll
                for (int a = 0; a < MB; ++a) cr[a] = M::op(ai[a], mbi, cr[a]);
#pragma unroll
                for (int a = 0; a < MB; ++a) ci[a] = M::op(ai[a], br, ci[a]);
            }
        }
    };
    auto next = [&](int i, int k0, int& ni, int& nk0) {
        ni = i;
        nk0 = k0 + KS;
        if (nk0 >= kb) {
            ni = i + 1;
            nk0 = 0;
        }
    };
    if (nact > 0) {
        load(0, 0, g.kb != 0);
        store(0);
    }
    int i = 0, k0 = 0;
    if constexpr (NBUF == 2) {
        // slab s+1 waits in registers while slab s runs; it is written to the other buffer
        // after the MFMAs, then slab s+2 is fetched; one barrier per slab
        int ni, nk0;
        next(i, k0, ni, nk0);
        bool more = ni < nact;
        if (more) load(ni, nk0, ni != i);
        __syncthreads();
        int q = 0;
        while (i < nact) {
            mfma_slab(i, k0, q);
            int n2i = nact, n2k0 = 0;
            bool more2 = false;
            if (more) {
                store(q ^ 1);
                next(ni, nk0, n2i, n2k0);
                more2 = n2i < nact;
                if (more2) load(n2i, n2k0, n2i != ni);
            }
            __syncthreads();
            i = ni;
            k0 = nk0;
            ni = n2i;
            nk0 = n2k0;
            more = more2;
            q ^= 1;
        }
    } else {
        __syncthreads();
        while (i < nact) {
            int ni, nk0;
            next(i, k0, ni, nk0);
            const bool more = ni < nact;
            if (more) load(ni, nk0, ni != i);   // global loads in flight during the MFMAs below
            mfma_slab(i, k0, 0);
            __syncthreads();
            if (more) {
                store(0);
                __syncthreads();
            }
            i = ni;
            k0 = nk0;
        }
    }
    // D layout of a 16x16 MFMA tile: lane holds rows M::row(lane/16, v), column lane%16
    if (coff < 0) return;
    V* Cc = C + coff;
#pragma unroll
    for (int a = 0; a < MB; ++a) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = tl.m0 + 16 * a + M::row(lk, v);
            if (m >= g.M) continue;
            R re = cr[a][v], im = ci[a][v];
            if constexpr (G3M) {
                re = cr[a][v] - c3[a][v];
                im = cr[a][v] + ci[a][v];
            }
            if (g.soff >= 0) {
                const TA s = opm[g.soff + m];
                const R sr = (R)s.x, si = (R)s.y;
                const R t = re * sr - im * si;
                im = re * si + im * sr;
                re = t;
            }
            if (g.mode == 0) {
                Cc[m] = mk<V>(re, im);
            } else {
                V o = Cc[m];
                o.x += (T)re;
                o.y += (T)im;
                Cc[m] = o;
            }
        }
    }
}

template <class TA, class V, class R, bool G3M = false>
void launch_gemm(int mb, int n, int nrhs, hipStream_t st, const DevPlan& d, const TA* opm, int t0) {
    const dim3 grid(n * nrhs, 1);
#define DDFMM_GEMM_MB(X)                                                                                     \
    case X:                                                                                                  \
        gemm_kernel<TA, V, R, X, G3M><<<grid, 256, 0, st>>>(d.groups, d.tiles, d.colB, d.colC, d.colA, opm, \
                                                            d.bt, t0, nrhs);                                 \
        break;
    switch (mb) {
        DDFMM_GEMM_MB(1) DDFMM_GEMM_MB(2) DDFMM_GEMM_MB(3) DDFMM_GEMM_MB(4) DDFMM_GEMM_MB(5)
        DDFMM_GEMM_MB(6) DDFMM_GEMM_MB(7)
    }
#undef DDFMM_GEMM_MB
}

// ---------------------------------------------------------------- split-BF16 complex GEMM (MFMA)
// The translation GEMM of the mixed mode with every FP32 operand written as the sum of
// three bf16 terms (x = x1 + x2 + x3, exact for FP32 x) and the products of
// v_mfma_f32_16x16x32_bf16 kept down to the third order (x1y1, x1y2, x2y1, x1y3, x2y2,
// x3y1; the dropped terms are below 2^-32 |x||y|), accumulated in FP32: FP32 accuracy at
// 2.7x the FP32 matrix-core rate.  The operator table is split once at setup (Bf16s
// entries); the vectors are split when a slab is staged.  A complex product is one real
// product with K' = 32 per 16 complex k:
//   Re C = [Ar, -Ai] [Br; Bi],   Im C = [Ai, Ar] [Br; Bi],
// lane i + 16 g of a 16x16x32 block holding A[i][8g..8g+7] and B[8g..8g+7][i], so lanes
// g < 2 take the real parts of k-octet g and lanes g >= 2 the imaginary parts of octet
// g - 2.  Tiles, K segments, column gathers, wave skips and epilogue as in gemm_kernel.
struct Bf16s {
    unsigned short v[8];    // re1, im1, re2, im2, re3, im3, 0, 0
};
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short bf16_rne(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
// x = h1 + h2 + h3 (each bf16, round to nearest even; the remainders are exact in FP32)
__device__ __forceinline__ void split3(float x, unsigned short& h1, unsigned short& h2, unsigned short& h3) {
    h1 = bf16_rne(x);
    const float r1 = x - bf16_val(h1);
    h2 = bf16_rne(r1);
    h3 = bf16_rne(r1 - bf16_val(h2));
}

__global__ void to_bf16s_kernel(const double2* __restrict__ in, Bf16s* __restrict__ out, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        Bf16s e;
        split3((float)in[i].x, e.v[0], e.v[2], e.v[4]);
        split3((float)in[i].y, e.v[1], e.v[3], e.v[5]);
        e.v[6] = e.v[7] = 0;
        out[i] = e;
    }
}

union Oct {                   // 8 bf16 of one plane and k-octet
    uint4 u;
    unsigned short h[8];
    bf16x8 v;
};

template <int MB>
__global__ __launch_bounds__(256) void gemm_bf16s_kernel(
    const GGroup* __restrict__ groups, const GTile* __restrict__ tiles,
    const int64_t* __restrict__ colB, const int64_t* __restrict__ colC, const int64_t* __restrict__ colA,
    const Bf16s* __restrict__ opb, const float2* __restrict__ opm32, BufTable bt, int tile0, int nr) {
    typedef float acc4 __attribute__((ext_vector_type(4)));
    constexpr int KS = 16, TM = 16 * MB;
    const int bx = xcd_item(blockIdx.x, gridDim.x);
    GTile tl = tiles[tile0 + bx / nr];
    const GGroup g = groups[tl.g];
    const int ncx = g.ncols * nr;
    tl.n0 = tl.n0 * nr + GEMM_TN * (bx % nr);
    if (tl.n0 >= ncx) return;
    const float2* B = bufp<float2>(bt, g.bbuf);
    float2* C = bufp<float2>(bt, g.cbuf);
    const int64_t sB = bt.stride[g.bbuf], sC = bt.stride[g.cbuf];
    // [buffer][term][plane re/im][row or column][k-octet]: k-octet h of row r is stored
    // at h ^ ((r >> 3) & 1), so the 16 lanes of a fragment read (16 rows, one octet) and
    // the 16 lanes of a staging store cover all banks once
    __shared__ uint4 As[2][3][2][TM][2];
    __shared__ uint4 Bs[2][3][2][GEMM_TN][2];
    __shared__ int64_t bo[MAXKB][GEMM_TN];
    __shared__ int64_t ao[MAXKB];
    __shared__ int used[MAXKB][4], act[MAXKB], wm[MAXKB], nact_s;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int kb = g.kb ? g.kb : g.K;
    const int nkb = g.kb ? g.K / g.kb : 1;

    // staging units: A (row am, k-octet ah) for tid < 2 TM; B (column bn, k-octet bh) for
    // tid < 128
    const int mrows = g.M - tl.m0;
    const int64_t kstride = g.transA ? 1 : g.lda;
    const bool hasA = tid < 2 * TM, hasB = tid < 2 * GEMM_TN;
    const int am = tid % TM, ah = tid / TM;
    const int64_t aoff = g.transA ? (int64_t)(tl.m0 + am) * g.lda : (int64_t)(tl.m0 + am);
    const bool arow = hasA && am < mrows;
    const int bn = tid & (GEMM_TN - 1), bh = tid >> 6;
    int64_t bcol = -1;
    const int ncol = tl.n0 + wave * 16 + li;
    const int64_t coff = ncol < ncx ? colC[g.col0 + ncol / nr] + (int64_t)(ncol % nr) * sC : -1;

    int nact = 1, wm0 = 0;
    if (g.kb == 0) {
        const int nv = ncx - tl.n0;
        wm0 = nv >= GEMM_TN ? 15 : (1 << ((nv + 15) >> 4)) - 1;
        const int n = tl.n0 + bn;
        bcol = (hasB && n < ncx) ? colB[g.bcol0 + n / nr] + (int64_t)(n % nr) * sB : -1;
    } else {
        if (tid < MAXKB * 4) used[tid >> 2][tid & 3] = 0;
        __syncthreads();
        for (int e = tid; e < nkb * GEMM_TN; e += 256) {
            const int h = e / GEMM_TN, j = e % GEMM_TN;
            const int n = tl.n0 + j;
            int64_t v = n < ncx ? colB[g.bcol0 + (int64_t)(n / nr) * nkb + h] : -1;
            if (v >= 0) v += (int64_t)(n % nr) * sB;
            bo[h][j] = v;
            if (v >= 0) used[h][j >> 4] = 1;
        }
        __syncthreads();
        if (tid == 0) {
            int c = 0;
            for (int h = 0; h < nkb; ++h) {
                const int w = used[h][0] | (used[h][1] << 1) | (used[h][2] << 2) | (used[h][3] << 3);
                if (w) {
                    act[c] = h;
                    wm[c] = w;
                    ++c;
                }
            }
            nact_s = c;
        }
        __syncthreads();
        nact = nact_s;
        if (tid < nact) ao[tid] = colA[g.acol0 + act[tid]];
        __syncthreads();
    }
    Bf16s ra[8];
    float2 rb[8];
    auto load = [&](int i, int k0, bool newseg) {
        const int kl = kb - k0;
        if (newseg && hasB) bcol = bo[act[i]][bn];
        if (hasA) {
            const Bf16s* Ab = opb + (g.kb ? ao[i] : g.aoff) + k0 * kstride + aoff;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 8 * ah + j;
                Bf16s a = {};
                if (arow && k < kl) a = Ab[k * kstride];
                ra[j] = a;
            }
        }
        if (hasB) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int k = 8 * bh + j;
                float2 b = make_float2(0.f, 0.f);
                if (bcol >= 0 && k < kl) b = B[bcol + k0 + k];
                rb[j] = b;
            }
        }
    };
    auto store = [&](int q) {
        if (hasA) {
            const int sl = ah ^ ((am >> 3) & 1);
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    Oct o;
#pragma unroll
                    for (int j = 0; j < 8; ++j) o.h[j] = ra[j].v[2 * t + c];
                    As[q][t][c][am][sl] = o.u;
                }
        }
        if (hasB) {
            Oct o[3][2];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                split3(rb[j].x, o[0][0].h[j], o[1][0].h[j], o[2][0].h[j]);
                split3(rb[j].y, o[0][1].h[j], o[1][1].h[j], o[2][1].h[j]);
            }
            const int sl = bh ^ ((bn >> 3) & 1);
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int c = 0; c < 2; ++c) Bs[q][t][c][bn][sl] = o[t][c].u;
        }
    };

    acc4 cr[MB], ci[MB];
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            cr[a][v] = 0;
            ci[a][v] = 0;
        }
    // fragment addressing of this lane: octet lk & 1 (swizzled by row/column bit 3), plane
    // re for lanes lk < 2 and im for lk >= 2 (B and the Re-C operand of A), im/re for the
    // Im-C operand of A; the Re-C operand of lanes lk >= 2 is negated (-Ai)
    const int oct = (lk & 1) ^ ((li >> 3) & 1);
    const int px = lk >> 1, py = px ^ 1;
    const unsigned neg = lk >= 2 ? 0x80008000u : 0u;
    auto mfma_slab = [&](int i, int q) {
        if (!(((g.kb ? wm[i] : wm0) >> wave) & 1)) return;
        Oct bf[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) bf[t].u = Bs[q][t][px][wave * 16 + li][oct];
#pragma unroll
        for (int a = 0; a < MB; ++a) {
            Oct xr[3], yi[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                xr[t].u = As[q][t][px][16 * a + li][oct];
                xr[t].u.x ^= neg;
                xr[t].u.y ^= neg;
                xr[t].u.z ^= neg;
                xr[t].u.w ^= neg;
                yi[t].u = As[q][t][py][16 * a + li][oct];
            }
            // smallest terms first
            cr[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[2].v, bf[0].v, cr[a], 0, 0, 0);
            ci[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(yi[2].v, bf[0].v, ci[a], 0, 0, 0);
            cr[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[1].v, bf[1].v, cr[a], 0, 0, 0);
            ci[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(yi[1].v, bf[1].v, ci[a], 0, 0, 0);
            cr[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[0].v, bf[2].v, cr[a], 0, 0, 0);
            ci[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(yi[0].v, bf[2].v, ci[a], 0, 0, 0);
            cr[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[1].v, bf[0].v, cr[a], 0, 0, 0);
            ci[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(yi[1].v, bf[0].v, ci[a], 0, 0, 0);
            cr[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[0].v, bf[1].v, cr[a], 0, 0, 0);
            ci[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(yi[0].v, bf[1].v, ci[a], 0, 0, 0);
            cr[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xr[0].v, bf[0].v, cr[a], 0, 0, 0);
            ci[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(yi[0].v, bf[0].v, ci[a], 0, 0, 0);
        }
    };
    auto next = [&](int i, int k0, int& ni, int& nk0) {
        ni = i;
        nk0 = k0 + KS;
        if (nk0 >= kb) {
            ni = i + 1;
            nk0 = 0;
        }
    };
    if (nact > 0) {
        load(0, 0, g.kb != 0);
        store(0);
    }
    // slab s+1 waits in registers while slab s runs; one barrier per slab
    int i = 0, k0 = 0, ni, nk0;
    next(i, k0, ni, nk0);
    bool more = ni < nact;
    if (more) load(ni, nk0, ni != i);
    __syncthreads();
    int q = 0;
    while (i < nact) {
        mfma_slab(i, q);
        int n2i = nact, n2k0 = 0;
        bool more2 = false;
        if (more) {
            store(q ^ 1);
            next(ni, nk0, n2i, n2k0);
            more2 = n2i < nact;
            if (more2) load(n2i, n2k0, n2i != ni);
        }
        __syncthreads();
        i = ni;
        k0 = nk0;
        ni = n2i;
        nk0 = n2k0;
        more = more2;
        q ^= 1;
    }
    // D layout of the 16x16 output: lane holds rows 4 (lane / 16) + v, column lane % 16
    if (coff < 0) return;
    float2* Cc = C + coff;
#pragma unroll
    for (int a = 0; a < MB; ++a) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = tl.m0 + 16 * a + 4 * lk + v;
            if (m >= g.M) continue;
            float re = cr[a][v], im = ci[a][v];
            if (g.soff >= 0) {
                const float2 s = opm32[g.soff + m];
                const float t = re * s.x - im * s.y;
                im = re * s.y + im * s.x;
                re = t;
            }
            if (g.mode == 0) {
                Cc[m] = make_float2(re, im);
            } else {
                float2 o = Cc[m];
                o.x += re;
                o.y += im;
                Cc[m] = o;
            }
        }
    }
}

void launch_gemm_bf16s(int mb, int n, int nrhs, hipStream_t st, const DevPlan& d, int t0) {
    const dim3 grid(n * nrhs, 1);
#define DDFMM_GEMM_MB(X)                                                                                      \
    case X:                                                                                                   \
        gemm_bf16s_kernel<X><<<grid, 256, 0, st>>>(d.groups, d.tiles, d.colB, d.colC, d.colA,                 \
                                                   (const Bf16s*)d.opmatbf, d.opmat32, d.bt, t0, nrhs);       \
        break;
    switch (mb) {
        DDFMM_GEMM_MB(1) DDFMM_GEMM_MB(2) DDFMM_GEMM_MB(3) DDFMM_GEMM_MB(4) DDFMM_GEMM_MB(5)
        DDFMM_GEMM_MB(6) DDFMM_GEMM_MB(7)
    }
#undef DDFMM_GEMM_MB
}

// ---------------------------------------------------------------- pruned (2P)^3 DFT, P = 6
// The equivalent surface of a box sits on the 6^3 sub-cube [6, 12)^3 of the 12^3 grid
// (grid index = round((x + W/2)/step) + P), so the forward transform has non-zero input
// only there and the inverse transform is sampled only there: each 1-D pass runs on the
// lines that carry data and produces the outputs later passes use (18144 complex MACs
// per 3-D transform instead of 62208).  One thread per line, twiddles as constants.
struct Tw12 {
    static constexpr double c[12] = {1.0, 0.86602540378443865, 0.5, 0.0, -0.5, -0.86602540378443865,
                                     -1.0, -0.86602540378443865, -0.5, 0.0, 0.5, 0.86602540378443865};
    static constexpr double s[12] = {0.0, 0.5, 0.86602540378443865, 1.0, 0.86602540378443865, 0.5,
                                     0.0, -0.5, -0.86602540378443865, -1.0, -0.86602540378443865, -0.5};
};

// out[o * os] = sum_{q < NI} in[q * is] * exp(SGN * 2 pi i (q + QOFF)(o + OOFF) / 12), o < NO
template <class V, int NI, int NO, int QOFF, int OOFF, int SGN>
__device__ __forceinline__ void dft12_line(const V* in, int is, V* out, int os) {
    using T = typename Real<V>::T;
    V x[NI];
#pragma unroll
    for (int q = 0; q < NI; ++q) x[q] = in[q * is];
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        V acc = mk<V>(0, 0);
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            const int m = ((q + QOFF) * (o + OOFF)) % 12;
            const V w = mk<V>((T)Tw12::c[m], (T)(SGN * Tw12::s[m]));
            cfma(acc, x[q], w);
        }
        out[o * os] = acc;
    }
}

template <class V>
__global__ __launch_bounds__(256) void fft_fwd_kernel(const FftSrc* __restrict__ fs, const int* __restrict__ grid,
                                                      int NL, BufTable bt, int f0, int nf, int nrhs) {
    __shared__ V sub[216], A[432], B[864];
    const int rhs = blockIdx.y;
    const int t = threadIdx.x;
    for (int bi = blockIdx.x; bi < nf; bi += gridDim.x) {
    const FftSrc f = fs[f0 + xcd_item(bi, nf)];
    __syncthreads();
    const V* in = bufp<V>(bt, BUF_LF_UPEQ) + (int64_t)rhs * bt.stride[BUF_LF_UPEQ] + f.in_off;
    // spectra right-hand side fastest: frequency k of this source at (spec_off + k) * nrhs + rhs
    V* out = bufp<V>(bt, BUF_SPEC) + f.spec_off * nrhs + rhs;
    if (t < 216) sub[t] = mk<V>(0, 0);
    __syncthreads();
    for (int k = t; k < NL; k += blockDim.x) {
        const int g = grid[f.grid + k];
        sub[(g % 12 - 6) + 6 * ((g / 12) % 12 - 6) + 36 * (g / 144 - 6)] = in[k];
    }
    __syncthreads();
    if (t < 36)          // x lines (y', z'): 6 inputs -> 12 outputs, A[z'][y'][kx]
        dft12_line<V, 6, 12, 6, 0, -1>(sub + 6 * t, 1, A + 12 * t, 1);
    __syncthreads();
    if (t < 72)          // y lines (kx, z'): A -> B[z'][ky][kx]
        dft12_line<V, 6, 12, 6, 0, -1>(A + (t % 12) + 72 * (t / 12), 12, B + (t % 12) + 144 * (t / 12), 12);
    __syncthreads();
    if (t < 144)         // z lines (kx, ky): B -> spectrum [kz][ky][kx]
        dft12_line<V, 6, 12, 6, 0, -1>(B + t, 144, out + (int64_t)t * nrhs, 144 * nrhs);
    }
}

// Forward transforms of R right-hand sides of one source per block (nrhs > 1).  The
// scatter into the 6^3 sub-grids and the x and y passes run H right-hand sides at a time
// (H x 36 and H x 72 line transforms per pass over the block's threads) into each one's
// slab of B; the z pass then runs over (line, rhs) with the right-hand side fastest, so the
// spectra, stored right-hand side fastest, are written in runs of R consecutive values
// (64 bytes: whole write sectors).
template <class V, int R, int H>
__global__ __launch_bounds__(256) void fft_fwd_multi_kernel(const FftSrc* __restrict__ fs,
                                                            const int* __restrict__ grid, int NL, BufTable bt,
                                                            int f0, int nf, int nrhs) {
    __shared__ V sub[H][216], A[H][432], B[R][864];
    const int t = threadIdx.x;
    const int r0 = blockIdx.y * R, nr = min(R, nrhs - r0);
    const int64_t is = bt.stride[BUF_LF_UPEQ];
    for (int bi = blockIdx.x; bi < nf; bi += gridDim.x) {
        const FftSrc f = fs[f0 + xcd_item(bi, nf)];
        for (int h0 = 0; h0 < nr; h0 += H) {
            const int hn = min(H, nr - h0);
            const V* in = bufp<V>(bt, BUF_LF_UPEQ) + (int64_t)(r0 + h0) * is + f.in_off;
            __syncthreads();
            for (int e = t; e < H * 216; e += blockDim.x) sub[e / 216][e % 216] = mk<V>(0, 0);
            __syncthreads();
            for (int e = t; e < hn * NL; e += blockDim.x) {
                const int r = e / NL, k = e % NL;
                const int g = grid[f.grid + k];
                sub[r][(g % 12 - 6) + 6 * ((g / 12) % 12 - 6) + 36 * (g / 144 - 6)] = in[(int64_t)r * is + k];
            }
            __syncthreads();
            for (int e = t; e < hn * 36; e += blockDim.x) {    // x lines (y', z'): 6 inputs -> 12 outputs
                const int r = e / 36, l = e % 36;
                dft12_line<V, 6, 12, 6, 0, -1>(sub[r] + 6 * l, 1, A[r] + 12 * l, 1);
            }
            __syncthreads();
            for (int e = t; e < hn * 72; e += blockDim.x) {    // y lines (kx, z'): A -> B[rhs][z'][ky][kx]
                const int r = e / 72, l = e % 72;
                dft12_line<V, 6, 12, 6, 0, -1>(A[r] + (l % 12) + 72 * (l / 12), 12,
                                               B[h0 + r] + (l % 12) + 144 * (l / 12), 12);
            }
        }
        __syncthreads();
        V* out = bufp<V>(bt, BUF_SPEC) + f.spec_off * nrhs + r0;
        for (int q = t; q < 144 * nr; q += blockDim.x) {   // z lines (kx, ky) x rhs, rhs fastest
            const int line = q / nr, r = q % nr;
            dft12_line<V, 6, 12, 6, 0, -1>(B[r] + line, 144, out + (int64_t)line * nrhs + r, 144 * nrhs);
        }
    }
}

// V list, accumulation: T_t[f] = sum_{(t, s) in V} Tab[offidx(t, s)][f] * S_s[f] over the
// (2P)^3 frequencies.  A block owns one slice of VSlice<V>::F frequencies (64 B of one
// spectrum) of one chunk of Morton-consecutive targets.  It stages in LDS the slice of the
// level's translation tensor (343 offsets) and the slice of every source in the chunk's
// union (at most VCH_SRC), with independent loads only; the pair loop then reads both
// operands from LDS.  Items are numbered slice-major and given to the XCDs in contiguous
// ranges (xcd_item), so an XCD works through a few slices at a time and the source
// slices it stages come from its own L2.  Threads: F frequencies x 256 / F targets.
template <class V> struct VSlice {
    static constexpr int F = 64 / (int)sizeof(V);          // frequencies per slice
    // LDS row stride, unpadded: the two targets' 64-byte rows that a 16-lane read group
    // takes then sit in aligned bank halves (with a 72-byte pitch their bank windows nearly
    // always overlap), and the 1-RHS block's LDS (51 KB instead of 56) lets three blocks
    // share a CU instead of two; together 12% faster than the padded pitch
    static constexpr int SE = F;
};

constexpr int VL_THREADS = 512;

// Stage the chunk's packed pair words (vpk[k0, k1), a multiple of 8 words, 16-byte
// aligned) into LDS with int4 loads.
__device__ __forceinline__ void stage_pairs(const int* __restrict__ vpk, const VChunk& ch, int* pkl, int t) {
    const int4* src = (const int4*)(vpk + ch.k0);
    int4* dst = (int4*)pkl;
    const int n4 = (ch.k1 - ch.k0) >> 2;
    for (int e = t; e < n4; e += VL_THREADS) dst[e] = src[e];
}

// A block walks VL_SLICES consecutive frequency slices of one chunk: the chunk's packed
// pairs are staged once, and the tensor and source slices of the next frequency slice
// are loaded into registers while the current one is accumulated.
constexpr int VL_GROUPS = 108;    // slice groups per chunk: 1728 / (F * slices per group)

template <class V>
__global__ __launch_bounds__(VL_THREADS) void vlist_acc_kernel(const VTarget* __restrict__ vt,
                                                               const VChunk* __restrict__ vc,
                                                               const int* __restrict__ vsrc,
                                                               const int* __restrict__ vpk,
                                                               const V* __restrict__ opm, BufTable bt, int v0,
                                                               int c0, int nchunk) {
    constexpr int G3 = 1728, F = VSlice<V>::F, SE = VSlice<V>::SE, TPP = VL_THREADS / F;
    constexpr int NSL = G3 / F / VL_GROUPS;
    constexpr int NT = (343 * F + VL_THREADS - 1) / VL_THREADS, NS = (VCH_SRC * F + VL_THREADS - 1) / VL_THREADS;
    __shared__ V tab[343 * SE], src[VCH_SRC * SE];
    __shared__ __align__(16) int pkl[VCH_PK];
    const int t = threadIdx.x, fl = t % F;
    const int item = xcd_item(blockIdx.x, VL_GROUPS * nchunk);
    const int fg = (item / nchunk) * NSL * F;
    const VChunk ch = vc[c0 + item % nchunk];
    const V* T = opm + vt[v0].tab_off;
    const V* spec = bufp<V>(bt, BUF_SPEC);
    const int nu = (ch.u1 - ch.u0) * F;
    int64_t si[NS];                          // source row offsets, fixed for the block
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int e = t + VL_THREADS * k;
        si[k] = (e >= F && e < nu) ? (int64_t)vsrc[ch.u0 + e / F] * G3 + e % F : -1;
    }
    V a[NT], b[NS];
    auto fetch = [&](int f0) {
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int e = t + VL_THREADS * k;
            a[k] = e < 343 * F ? T[(int64_t)(e / F) * G3 + f0 + e % F] : mk<V>(0, 0);
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) b[k] = si[k] >= 0 ? spec[si[k] + f0] : mk<V>(0, 0);
    };
    fetch(fg);
    stage_pairs(vpk, ch, pkl, t);
    const V* srcl = src + fl;
    const V* tabl = tab + fl;
    for (int sl = 0; sl < NSL; ++sl) {
        const int f0 = fg + sl * F;
        __syncthreads();                     // the previous slice's reads of tab / src are done
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int e = t + VL_THREADS * k;
            if (e < 343 * F) tab[(e / F) * SE + e % F] = a[k];
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int e = t + VL_THREADS * k;
            if (e < nu) src[(e / F) * SE + e % F] = b[k];     // row 0 stays zero (si = -1)
        }
        __syncthreads();
        if (sl + 1 < NSL) fetch(f0 + F);     // in flight during the pair loop below
        V* out = bufp<V>(bt, BUF_TSPEC) + f0 + fl;
        for (int i = ch.t0 + t / F; i < ch.t1; i += TPP) {
            const VTarget tg = vt[i];
            V acc = mk<V>(0, 0);
            // the target's packed pairs (zero pairs pad to a multiple of 8; row 0 is zero),
            // read from LDS by every lane of its group (one address: a broadcast)
            const int n = (tg.p1 - tg.p0 + 7) & ~7;
            const int4* pk = (const int4*)(pkl + (tg.k0 - ch.k0));
            for (int j = 0; j < n; j += F) {
                V sv[F], tv[F];
#pragma unroll
                for (int u = 0; u < F; ++u) {
                    const int4 q = pk[(j + u) / 4];
                    const int w = (u % 4 == 0) ? q.x : (u % 4 == 1) ? q.y : (u % 4 == 2) ? q.z : q.w;
                    sv[u] = srcl[(w & 0xffff) * SE];
                    tv[u] = tabl[(w >> 16) * SE];
                }
#pragma unroll
                for (int u = 0; u < F; ++u) cfma(acc, sv[u], tv[u]);
            }
            out[(int64_t)(i - v0) * G3] = acc;
        }
    }
}

// V list, accumulation for several right-hand sides (nrhs > 1).  The source spectra are
// stored right-hand side fastest, SPEC[(row * (2P)^3 + f) * nrhs + r], so the values of
// one source for F frequencies and RC = R1 * R2 right-hand sides of a block are one
// 128-byte run per frequency.  A block owns (chunk, group of frequency slices of F
// frequencies, group of RC right-hand sides) and stages the slice of the translation
// tensor (343 offsets, shared by the group) and the union's source rows (128 bytes each).
// Eight lanes work on one target: lane l takes frequency l / R1 and the R2 right-hand
// sides (l % R1) * R2 + q, so the eight lanes read one whole source row per pair (all 32
// banks) and one tensor value serves R2 complex FMAs.  The accumulated target spectra go
// to TSPEC in the [rhs][target][(2P)^3] layout of the inverse transform.
template <class V> struct VMulti { static constexpr int R2 = 16 / (int)sizeof(V); };   // float2: 2, double2: 1

template <class V, int F, int R1>
__global__ __launch_bounds__(VL_THREADS) void vlist_acc_multi_kernel(const VTarget* __restrict__ vt,
                                                                     const VChunk* __restrict__ vc,
                                                                     const int* __restrict__ vsrc,
                                                                     const int* __restrict__ vpk,
                                                                     const V* __restrict__ opm, BufTable bt, int v0,
                                                                     int c0, int nchunk, int nrhs) {
    constexpr int G3 = 1728, R2 = VMulti<V>::R2, RC = R1 * R2, ROW = F * RC;
    constexpr int NSL = G3 / F / VL_GROUPS;
    constexpr int NT = (343 * F + VL_THREADS - 1) / VL_THREADS, NS = (VCH_SRC * ROW + VL_THREADS - 1) / VL_THREADS;
    static_assert(F * R1 == 8, "eight lanes per target");
    static_assert(ROW * (int)sizeof(V) == 128, "128-byte source rows");
    __shared__ V tab[343 * F];
    __shared__ __align__(16) V src[VCH_SRC * ROW];
    __shared__ __align__(16) int pkl[VCH_PK];
    const int t = threadIdx.x, l = t & 7;
    const int fl = l / R1, rq = (l % R1) * R2;
    const int item = xcd_item(blockIdx.x, VL_GROUPS * nchunk);
    const int fg = (item / nchunk) * NSL * F;
    const int rg = blockIdx.y * RC;
    const VChunk ch = vc[c0 + item % nchunk];
    const V* T = opm + vt[v0].tab_off;
    const V* spec = bufp<V>(bt, BUF_SPEC);
    const int nu = (ch.u1 - ch.u0) * ROW;
    int64_t si[NS];                          // element offsets at frequency 0, fixed for the block
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int e = t + VL_THREADS * k;
        const int u = e / ROW, w = e % ROW, f = w / RC, r = rg + w % RC;
        si[k] = (e >= ROW && e < nu && r < nrhs) ? ((int64_t)vsrc[ch.u0 + u] * G3 + f) * nrhs + r : -1;
    }
    V a[NT], b[NS];
    auto fetch = [&](int f0) {
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int e = t + VL_THREADS * k;
            a[k] = e < 343 * F ? T[(int64_t)(e / F) * G3 + f0 + e % F] : mk<V>(0, 0);
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) b[k] = si[k] >= 0 ? spec[si[k] + (int64_t)f0 * nrhs] : mk<V>(0, 0);
    };
    fetch(fg);
    stage_pairs(vpk, ch, pkl, t);
    const V* tabl = tab + fl;
    const V* srcl = src + l * R2;
    V* out = bufp<V>(bt, BUF_TSPEC);
    const int64_t os = bt.stride[BUF_TSPEC];
    for (int sl = 0; sl < NSL; ++sl) {
        const int f0 = fg + sl * F;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const int e = t + VL_THREADS * k;
            if (e < 343 * F) tab[e] = a[k];
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int e = t + VL_THREADS * k;
            if (e < nu) src[e] = b[k];       // row 0 and right-hand sides past nrhs are zero
        }
        __syncthreads();
        if (sl + 1 < NSL) fetch(f0 + F);     // in flight during the pair loop below
        for (int i = ch.t0 + t / 8; i < ch.t1; i += VL_THREADS / 8) {
            const VTarget tg = vt[i];
            V acc[R2];
#pragma unroll
            for (int q = 0; q < R2; ++q) acc[q] = mk<V>(0, 0);
            const int n = (tg.p1 - tg.p0 + 7) & ~7;
            const int4* pk = (const int4*)(pkl + (tg.k0 - ch.k0));
            for (int j = 0; j < n; j += 8) {
                V tv[8], sv[8][R2];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int4 q = pk[(j + u) / 4];
                    const int w = (u % 4 == 0) ? q.x : (u % 4 == 1) ? q.y : (u % 4 == 2) ? q.z : q.w;
                    tv[u] = tabl[(w >> 16) * F];
                    const V* sp = srcl + (w & 0xffff) * ROW;
                    if constexpr (R2 == 2) {
                        const float4 x = *(const float4*)sp;
                        sv[u][0] = mk<V>(x.x, x.y);
                        sv[u][R2 - 1] = mk<V>(x.z, x.w);
                    } else {
                        sv[u][0] = sp[0];
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int q = 0; q < R2; ++q) cfma(acc[q], sv[u][q], tv[u]);   // operand order of the 1-RHS kernel
            }
#pragma unroll
            for (int q = 0; q < R2; ++q) {
                const int r = rg + rq + q;
                if (r < nrhs) out[(int64_t)r * os + (int64_t)(i - v0) * G3 + f0 + fl] = acc[q];
            }
        }
    }
}

// V list, inverse: one block per target and group of R right-hand sides, the accumulated
// spectra back to the 6^3 sub-cube (pruned inverse DFT, 1/(2P)^3 scaling) and added into
// the target's check values.  Each pass runs the R right-hand sides' lines together
// (R x 144 x lines read straight from the spectra, R x 72 y lines, R x 36 z lines over the
// block's threads).  LDS: R x 1296 values (41 KB for R = 4 in FP32 or 2 in FP64).
template <class V, int R>
__global__ __launch_bounds__(256) void vlist_inv_kernel(const VTarget* __restrict__ vt, const int* __restrict__ grid,
                                                        int NL, BufTable bt, int v0, int nv, int nrhs) {
    constexpr int G3 = 1728;
    __shared__ V A[R][864], B[R][432];
    const int r0 = blockIdx.y * R, nr = min(R, nrhs - r0);
    const int t = threadIdx.x;
    const int64_t is = bt.stride[BUF_TSPEC], os = bt.stride[BUF_LF_CHK];
    for (int bi = blockIdx.x; bi < nv; bi += gridDim.x) {
        const int i = xcd_item(bi, nv);
        const VTarget tg = vt[v0 + i];
        const V* in = bufp<V>(bt, BUF_TSPEC) + (int64_t)r0 * is + (int64_t)i * G3;
        __syncthreads();
        for (int e = t; e < nr * 144; e += blockDim.x) {   // x lines (ky, kz): 12 inputs -> x' in [6, 12)
            const int r = e / 144, l = e % 144;
            dft12_line<V, 12, 6, 0, 6, 1>(in + (int64_t)r * is + 12 * l, 1, A[r] + 6 * l, 1);
        }
        __syncthreads();
        for (int e = t; e < nr * 72; e += blockDim.x) {    // y lines (x', kz): A -> B[kz][y'][x']
            const int r = e / 72, l = e % 72;
            dft12_line<V, 12, 6, 0, 6, 1>(A[r] + (l % 6) + 72 * (l / 6), 6, B[r] + (l % 6) + 36 * (l / 6), 6);
        }
        __syncthreads();
        for (int e = t; e < nr * 36; e += blockDim.x) {    // z lines (x', y'): B -> C[z'][y'][x'] (reuses A)
            const int r = e / 36, l = e % 36;
            dft12_line<V, 12, 6, 0, 6, 1>(B[r] + l, 36, A[r] + l, 36);
        }
        __syncthreads();
        V* out = bufp<V>(bt, BUF_LF_CHK) + (int64_t)r0 * os + tg.out_off;
        const typename Real<V>::T coef = (typename Real<V>::T)(1.0 / G3);
        for (int e = t; e < nr * NL; e += blockDim.x) {
            const int r = e / NL, k = e % NL;
            const int g = grid[tg.grid + k];
            const V v = A[r][(g % 12 - 6) + 6 * ((g / 12) % 12 - 6) + 36 * (g / 144 - 6)];
            V* o = out + (int64_t)r * os + k;
            V w = *o;
            w.x += v.x * coef;
            w.y += v.y * coef;
            *o = w;
        }
    }
}

// ---------------------------------------------------------------- setup-time generation
__global__ __launch_bounds__(256) void kgen_kernel(const KGen* __restrict__ jobs, const double* __restrict__ gp,
                                                   double2* __restrict__ opmat, int j0, int ktype) {
    const KGen k = jobs[j0 + blockIdx.x];
    const double sg = k.tneg ? -1.0 : 1.0;
    const int64_t mn = (int64_t)k.m * k.n;
    for (int64_t e = threadIdx.x; e < mn; e += blockDim.x) {
        const int i = (int)(e % k.m), j = (int)(e / k.m);
        const double* t = gp + 3 * (k.toff + i);
        const double* s = gp + 3 * (k.soff + j);
        double d[3];
        for (int q = 0; q < 3; ++q) d[q] = (sg * t[q] + k.tshift[q]) - (s[q] + k.sshift[q]);
        opmat[k.out + e] = green64(d[0], d[1], d[2], ktype);
    }
}

__global__ void to_f32_kernel(const double2* __restrict__ in, float2* __restrict__ out, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        out[e] = make_float2((float)in[e].x, (float)in[e].y);
}

template <class V>
__global__ void zero_kernel(V* p, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        p[e] = mk<V>(0, 0);
}

// The caller's densities are read in input order (coalesced: random 16-byte reads would
// fetch a whole line each) and written to their leaf-sorted positions.
template <class V>
__global__ void gather_kernel(const IoPtrs* __restrict__ io, V* __restrict__ out, const int* __restrict__ iperm,
                              int64_t N, int nrhs) {
    const double2* __restrict__ in = io->den;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * nrhs; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / N, j = e % N;
        const double2 v = in[e];
        out[r * N + iperm[j]] = mk<V>(v.x, v.y);
    }
}

template <class V>
__global__ void scatter_kernel(const V* __restrict__ in, const IoPtrs* __restrict__ io, const int* __restrict__ perm,
                               int64_t N, int nrhs) {
    double2* __restrict__ out = io->val;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * nrhs; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / N, i = e % N;
        const V v = in[e];
        out[r * N + perm[i]] = make_double2(v.x, v.y);
    }
}

// direct sum u(t_i) = sum_j G(t_i, s_j) f_j (Wave3d::check, wave3d_check.cpp:66-76).
// Block (target tile x, source range y, right-hand-side group z): 256 targets, one thread
// each, over the sources [y * chunk, (y + 1) * chunk) staged through LDS in tiles of 256;
// G is evaluated once per pair for the RC right-hand sides of the group.  With one source
// range the sums go straight to out, otherwise to part[y][rhs][target] for the fixed-order
// reduction below.
template <int RC>
__global__ __launch_bounds__(256) void direct_kernel(const double* __restrict__ trg, int M,
                                                     const double* __restrict__ src, int Ns,
                                                     const double2* __restrict__ den, int nrhs,
                                                     double2* __restrict__ out, int chunk, int ktype) {
    __shared__ double s3[256 * 3];
    __shared__ double2 sd[RC][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int r0 = blockIdx.z * RC, nr = min(RC, nrhs - r0);
    const bool ok = i < M;
    double tx = 0, ty = 0, tz = 0;
    if (ok) { tx = trg[3 * i]; ty = trg[3 * i + 1]; tz = trg[3 * i + 2]; }
    double2 acc[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q) acc[q] = make_double2(0, 0);
    const int j0 = blockIdx.y * chunk, j1 = min(Ns, j0 + chunk);
    for (int sb = j0; sb < j1; sb += 256) {
        const int n = min(256, j1 - sb);
        __syncthreads();
        if ((int)threadIdx.x < n) {
            const int j = sb + threadIdx.x;
            s3[3 * threadIdx.x] = src[3 * j];
            s3[3 * threadIdx.x + 1] = src[3 * j + 1];
            s3[3 * threadIdx.x + 2] = src[3 * j + 2];
#pragma unroll
            for (int q = 0; q < RC; ++q)
                sd[q][threadIdx.x] = q < nr ? den[(int64_t)(r0 + q) * Ns + j] : make_double2(0, 0);
        }
        __syncthreads();
        if (ok)
            for (int j = 0; j < n; ++j) {
                const double2 g = green64(tx - s3[3 * j], ty - s3[3 * j + 1], tz - s3[3 * j + 2], ktype);
#pragma unroll
                for (int q = 0; q < RC; ++q) cfma(acc[q], g, sd[q][j]);
            }
    }
    if (!ok) return;
    double2* o = out + (int64_t)blockIdx.y * nrhs * M + i;
#pragma unroll
    for (int q = 0; q < RC; ++q)
        if (q < nr) o[(int64_t)(r0 + q) * M] = acc[q];
}

// out[e] = sum_{y < S} part[y][e] in the order y = 0, 1, ... (deterministic)
__global__ void direct_reduce_kernel(const double2* __restrict__ part, int64_t n, int S, double2* __restrict__ out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double2 a = part[e];
        for (int y = 1; y < S; ++y) {
            const double2 b = part[(int64_t)y * n + e];
            a.x += b.x;
            a.y += b.y;
        }
        out[e] = a;
    }
}

// dense Kernel3d::kernel block: out(i,j) = G(trg_i, src_j), column-major M x N
template <class V>
__global__ void kmatrix_kernel(const double* __restrict__ trg, int M, const double* __restrict__ src, int Ns,
                               V* __restrict__ out, int ktype) {
    const int64_t mn = (int64_t)M * Ns;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % M), j = (int)(e / M);
        out[e] = green<V>(trg[3 * i] - src[3 * j], trg[3 * i + 1] - src[3 * j + 1], trg[3 * i + 2] - src[3 * j + 2],
                          ktype);
    }
}

// segmented reduction: target t gets the sum of its cnt adjacent contribution slots
// (M elements each) in a fixed order; one wave per target, four targets per block.  A lane
// owns rows lane + 64 r (r < R, R = 2 for M <= 128, 3 for M <= 192) and walks the slots
// once, all R loads of a slot in flight together; lanes past M read row M - 1 (in bounds)
// and store nothing.
template <class V, int R>
__device__ __forceinline__ void reduce_rows(const V* __restrict__ in, V* __restrict__ out, int M, int cnt, int mode,
                                            int m0, int lane) {
    using T = typename Real<V>::T;
    int idx[R];
    T sx[R], sy[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        idx[r] = min(m0 + lane + 64 * r, M - 1);
        sx[r] = 0;
        sy[r] = 0;
    }
#pragma unroll 4
    for (int j = 0; j < cnt; ++j) {
        const V* s = in + (int64_t)j * M;
        V v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = s[idx[r]];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sx[r] += v[r].x;
            sy[r] += v[r].y;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int m = m0 + lane + 64 * r;
        if (m < M) {
            if (mode) {
                const V o = out[m];
                sx[r] += o.x;
                sy[r] += o.y;
            }
            out[m] = mk<V>(sx[r], sy[r]);
        }
    }
}

template <class V>
__global__ __launch_bounds__(256) void reduce_kernel(const RTarget* __restrict__ rtrg, int n, BufTable bt, int t0,
                                                     const V* __restrict__ inb, int64_t ins) {
    const int ti = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (ti >= n) return;
    const RTarget r = rtrg[t0 + ti];
    const int rhs = blockIdx.y;
    const V* in = inb + (int64_t)rhs * ins + r.red_off;
    V* out = bufp<V>(bt, r.obuf) + (int64_t)rhs * bt.stride[r.obuf] + r.out_off;
    if (r.M <= 128) {
        reduce_rows<V, 2>(in, out, r.M, r.cnt, r.mode, 0, lane);
    } else {
        for (int m0 = 0; m0 < r.M; m0 += 192) reduce_rows<V, 3>(in, out, r.M, r.cnt, r.mode, m0, lane);
    }
}

// Grid caps of the launches on the second stream (multiples of 8: one XCD round), for
// running them as a background beside the GEMMs; very large = one block per item
constexpr int AUX_KAPPLY_BLOCKS = 1 << 30;
constexpr int AUX_BLOCKS = 1 << 30;

int grid_for(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 65536);
}

template <class V>
hipError_t launch_stage_t(const DevPlan& d, const Stage& s, int nrhs, hipStream_t st) {
    const int n = s.end - s.begin;
    switch (s.kind) {
    case ST_ZERO: {
        int64_t cnt = s.count * nrhs;   // rhs slabs are contiguous with stride = count
        if (cnt > 0) zero_kernel<V><<<grid_for(cnt, 256), 256, 0, st>>>((V*)d.bt.p[s.buf], cnt);
        break;
    }
    case ST_GATHER:
        gather_kernel<V><<<grid_for(d.N * nrhs, 256), 256, 0, st>>>(d.io, (V*)d.bt.p[BUF_DEN], d.iperm, d.N, nrhs);
        break;
    case ST_SCATTER:
        scatter_kernel<V><<<grid_for(d.N * nrhs, 256), 256, 0, st>>>((const V*)d.bt.p[BUF_VAL], d.io, d.perm, d.N,
                                                                     nrhs);
        break;
    case ST_KAPPLY: {
        const int gx = s.stream ? std::min(n, AUX_KAPPLY_BLOCKS) : n;
        if constexpr (sizeof(V) == 8) if (d.kmfma) {   // mixed precision: the products on the matrix cores
            const int nb = nrhs > 8 ? 2 : 1;
            const dim3 grid(gx, (nrhs + 8 * nb - 1) / (8 * nb));
#define DDFMM_KMFMA(NB, KT, MIR) \
    kapply_mfma_kernel<NB, KT, MIR><<<grid, 256, 0, st>>>(d.ktasks, d.ksrcs, d.px, d.py, d.pz, d.bt, s.begin, n, nrhs)
            const int sel = (nb == 2 ? 4 : 0) + (d.ktype ? 2 : 0) + (s.mb == 1 ? 1 : 0);
            switch (sel) {
            case 0: DDFMM_KMFMA(1, 0, false); break;
            case 1: DDFMM_KMFMA(1, 0, true); break;
            case 2: DDFMM_KMFMA(1, 1, false); break;
            case 3: DDFMM_KMFMA(1, 1, true); break;
            case 4: DDFMM_KMFMA(2, 0, false); break;
            case 5: DDFMM_KMFMA(2, 0, true); break;
            case 6: DDFMM_KMFMA(2, 1, false); break;
            case 7: DDFMM_KMFMA(2, 1, true); break;
            }
#undef DDFMM_KMFMA
            break;
        }
        if (s.mb == 1) return hipErrorInvalidValue;    // mirror slots need the matrix-core kernel
        if (nrhs == 1) {
            kapply_kernel<V, 1><<<dim3(gx, 1), KA_THREADS, 0, st>>>(d.ktasks, d.ksrcs, d.px, d.py, d.pz, d.bt,
                                                                    s.begin, n, nrhs, d.ktype);
        } else if (nrhs <= 4) {
            kapply_kernel<V, 4><<<dim3(gx, 1), KA_THREADS, 0, st>>>(d.ktasks, d.ksrcs, d.px, d.py, d.pz, d.bt,
                                                                    s.begin, n, nrhs, d.ktype);
        } else {    // G(t, s) evaluated once for 8 right-hand sides
            kapply_kernel<V, 8><<<dim3(gx, (nrhs + 7) / 8), KA_THREADS, 0, st>>>(d.ktasks, d.ksrcs, d.px, d.py, d.pz,
                                                                                d.bt, s.begin, n, nrhs, d.ktype);
        }
        break;
    }
    case ST_GEMM:
        if (s.mb < 1 || s.mb > GEMM_MB) return hipErrorInvalidValue;
        if (d.mixed && s.prec == PREC_SOLVE)
            launch_gemm<double2, float2, double>(s.mb, n, nrhs, st, d, d.opmat, s.begin);
        else if (d.mixed && d.opmatbf)
            launch_gemm_bf16s(s.mb, n, nrhs, st, d, s.begin);
        else if (d.mixed && d.ktype == 0)    // e^{ikr}/r: three real products per complex one
            launch_gemm<float2, float2, float, true>(s.mb, n, nrhs, st, d, d.opmat32, s.begin);
        else if (d.mixed)
            launch_gemm<float2, float2, float>(s.mb, n, nrhs, st, d, d.opmat32, s.begin);
        else
            launch_gemm<double2, double2, double>(s.mb, n, nrhs, st, d, d.opmat, s.begin);
        break;
    case ST_FFTFWD:
        if (d.P != 6) return hipErrorInvalidValue;
        if (nrhs == 1) {
            fft_fwd_kernel<V><<<dim3(s.stream ? std::min(n, AUX_BLOCKS) : n, 1), 256, 0, st>>>(d.fsrc, d.grid, d.NL,
                                                                                           d.bt, s.begin, n, 1);
        } else {    // 8 (FP32) or 4 (FP64) right-hand sides per block, two at a time: 66 or 76 KB
            // of LDS (an FP32 block still fits beside two translation-GEMM blocks)
            constexpr int R = sizeof(V) == 8 ? 8 : 4;
            fft_fwd_multi_kernel<V, R, 2><<<dim3(n, (nrhs + R - 1) / R), 256, 0, st>>>(d.fsrc, d.grid, d.NL, d.bt,
                                                                                 s.begin, n, nrhs);
        }
        break;
    case ST_VLIST: {
        if (d.P != 6) return hipErrorInvalidValue;
        if (n == 0) break;
        const int nchunk = (int)s.count;
        const V* opm = d.mixed ? (const V*)d.opmat32 : (const V*)d.opmat;
        if (nrhs == 1) {
            vlist_acc_kernel<V><<<dim3(VL_GROUPS * nchunk, 1), VL_THREADS, 0, st>>>(d.vtrg, d.vchunk, d.vsrc, d.vpk, opm,
                                                                              d.bt, s.begin, s.buf, nchunk);
        } else {
            // right-hand sides per block: 16 / 8 / 4 (FP32) or 8 / 4 / 2 (FP64) by nrhs
            constexpr int R2 = VMulti<V>::R2;
            if (nrhs >= 8 * R2) {
                vlist_acc_multi_kernel<V, 1, 8><<<dim3(VL_GROUPS * nchunk, (nrhs + 8 * R2 - 1) / (8 * R2)), VL_THREADS, 0, st>>>(
                    d.vtrg, d.vchunk, d.vsrc, d.vpk, opm, d.bt, s.begin, s.buf, nchunk, nrhs);
            } else if (nrhs >= 4 * R2) {
                vlist_acc_multi_kernel<V, 2, 4><<<dim3(VL_GROUPS * nchunk, (nrhs + 4 * R2 - 1) / (4 * R2)), VL_THREADS, 0, st>>>(
                    d.vtrg, d.vchunk, d.vsrc, d.vpk, opm, d.bt, s.begin, s.buf, nchunk, nrhs);
            } else {
                vlist_acc_multi_kernel<V, 4, 2><<<dim3(VL_GROUPS * nchunk, (nrhs + 2 * R2 - 1) / (2 * R2)), VL_THREADS, 0, st>>>(
                    d.vtrg, d.vchunk, d.vsrc, d.vpk, opm, d.bt, s.begin, s.buf, nchunk, nrhs);
            }
        }
        {
            constexpr int R = sizeof(V) == 8 ? 4 : 2;      // right-hand sides per block
            vlist_inv_kernel<V, R><<<dim3(s.stream ? std::min(n, AUX_BLOCKS) : n, (nrhs + R - 1) / R), 256, 0, st>>>(
                d.vtrg, d.grid, d.NL, d.bt, s.begin, n, nrhs);
        }
        break;
    }
    case ST_REDUCE:
        if (s.buf != BUF_RED && s.buf != BUF_NMIR) return hipErrorInvalidValue;
        reduce_kernel<V><<<dim3((n + 3) / 4, nrhs), 256, 0, st>>>(d.rtrg, n, d.bt, s.begin, (const V*)d.bt.p[s.buf],
                                                                  d.bt.stride[s.buf]);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void set_io_kernel(IoPtrs* io, const double2* den, double2* val) {
    io->den = den;
    io->val = val;
}

}  // namespace

// ---------------------------------------------------------------- launchers
hipError_t launch_set_io(IoPtrs* io, const double2* den, double2* val, hipStream_t st) {
    set_io_kernel<<<1, 1, 0, st>>>(io, den, val);
    return hipGetLastError();
}

hipError_t launch_stage(const DevPlan& d, const Stage& s, int nrhs, hipStream_t st) {
    return d.mixed ? launch_stage_t<float2>(d, s, nrhs, st) : launch_stage_t<double2>(d, s, nrhs, st);
}

hipError_t launch_setup_gemm(const DevPlan& d, const Stage& s, hipStream_t st) {
    if (s.kind != ST_GEMM || s.mb < 1 || s.mb > GEMM_MB) return hipErrorInvalidValue;
    launch_gemm<double2, double2, double>(s.mb, s.end - s.begin, 1, st, d, d.opmat, s.begin);
    return hipGetLastError();
}

hipError_t launch_kgen(const KGen* jobs, int njobs, const double* gp, double2* opmat, int ktype, hipStream_t st) {
    for (int j0 = 0; j0 < njobs; j0 += 65536) {
        int n = std::min(65536, njobs - j0);
        kgen_kernel<<<n, 256, 0, st>>>(jobs, gp, opmat, j0, ktype);
    }
    return hipGetLastError();
}

hipError_t launch_to_bf16s(const double2* in, void* out, int64_t n, hipStream_t st) {
    if (n > 0) to_bf16s_kernel<<<grid_for(n, 256), 256, 0, st>>>(in, (Bf16s*)out, n);
    return hipGetLastError();
}

hipError_t launch_to_f32(const double2* in, float2* out, int64_t n, hipStream_t st) {
    if (n > 0) to_f32_kernel<<<grid_for(n, 256), 256, 0, st>>>(in, out, n);
    return hipGetLastError();
}

DirectSplit direct_split(int M, int Ns, int nrhs) {
    DirectSplit d;
    d.rc = nrhs == 1 ? 1 : nrhs <= 4 ? 4 : 8;
    const int mb = (M + 255) / 256, rg = (nrhs + d.rc - 1) / d.rc;
    const int tiles = std::max(1, (Ns + 255) / 256);
    // about 1024 blocks: source ranges until the target tiles and groups alone fill that;
    // S * nrhs * M <= 1024 * 256 * 8 elements, so the partial sums fit in DIRECT_SCRATCH_BYTES
    int S = std::max(1, std::min(tiles, 1024 / std::max(1, mb * rg)));
    const int per = (tiles + S - 1) / S;
    d.chunk = per * 256;
    d.S = (std::max(Ns, 1) + d.chunk - 1) / d.chunk;
    d.part_elems = d.S > 1 ? (int64_t)d.S * nrhs * M : 0;
    return d;
}

hipError_t launch_direct(const double* trg, int M, const double* src, int Ns, const double2* den, int nrhs,
                         double2* out, double2* part, int ktype, hipStream_t st) {
    const DirectSplit d = direct_split(M, Ns, nrhs);
    if (d.S > 1 && !part) return hipErrorInvalidValue;
    double2* dst = d.S > 1 ? part : out;
    const dim3 grid((M + 255) / 256, d.S, (nrhs + d.rc - 1) / d.rc);
    if (d.rc == 1)
        direct_kernel<1><<<grid, 256, 0, st>>>(trg, M, src, Ns, den, nrhs, dst, d.chunk, ktype);
    else if (d.rc == 4)
        direct_kernel<4><<<grid, 256, 0, st>>>(trg, M, src, Ns, den, nrhs, dst, d.chunk, ktype);
    else
        direct_kernel<8><<<grid, 256, 0, st>>>(trg, M, src, Ns, den, nrhs, dst, d.chunk, ktype);
    if (d.S > 1) {
        const int64_t n = (int64_t)nrhs * M;
        direct_reduce_kernel<<<grid_for(n, 256), 256, 0, st>>>(part, n, d.S, out);
    }
    return hipGetLastError();
}

hipError_t launch_kmatrix(const double* trg, int M, const double* src, int Ns, void* out, int ktype, int f32,
                          hipStream_t st) {
    if (f32)
        kmatrix_kernel<float2><<<grid_for((int64_t)M * Ns, 256), 256, 0, st>>>(trg, M, src, Ns, (float2*)out, ktype);
    else
        kmatrix_kernel<double2><<<grid_for((int64_t)M * Ns, 256), 256, 0, st>>>(trg, M, src, Ns, (double2*)out, ktype);
    return hipGetLastError();
}

// ---------------------------------------------------------------- no-allocation direct sum
// Partial sums of the direct sum for callers that cannot allocate: one static array per
// GPU, its device address taken once per device with that device current.  Calls on one
// GPU share the array, so calls that overlap on one GPU race.
static __device__ __align__(256) double2 scratch[DIRECT_SCRATCH_BYTES / sizeof(double2)];

hipError_t direct_static_scratch(double2** out) {
    static std::mutex mu;
    static std::vector<double2*> ptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptr.size() <= dev) ptr.resize(dev + 1, nullptr);
    if (!ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptr[dev] = (double2*)p;
    }
    *out = ptr[dev];
    return hipSuccess;
}

}  // namespace ddfmm

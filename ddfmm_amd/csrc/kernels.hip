// HIP kernels of the directional-FMM evaluation for gfx950 (CDNA4, wave64).
//
//   kapply_kernel  generated-kernel block apply: out(t_i) += sum_j G(t_i, s_j) v_j with
//                  LDS-staged source tiles; serves the near field (U list), S2M, L2T,
//                  W/X lists (Kernel3d::kernel + zgemv, kernel3d.cpp:26-107,
//                  wave3d_eval.cpp:441-451, 550-560, 585-718)
//   gemm_kernel    grouped complex GEMM over (box, rhs) columns with gather/scatter and
//                  store/add/atomic epilogues: all translations that the reference does
//                  as per-box zgemv (vecmatop.cpp:102-125)
//   fft_fwd_kernel / vlist_kernel   FFT-diagonal LF M2L on the (2P)^3 grid
//                  (wave3d_eval.cpp:602-661)
//   kgen_kernel    setup-time translation-matrix generation (mlib3d.cpp:106-280 and the
//                  per-offset HF M2L matrices of wave3d_eval.cpp:876-928)
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace ddfmm {

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}

// G(r) for k = 2*pi: e^{ikr}/r (type 0) or e^{ikr} (type 1); r^2 < 1e-16 => r := 1.
__device__ __forceinline__ double2 green(double dx, double dy, double dz, int ktype) {
    double r2 = dx * dx + dy * dy + dz * dz;
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

constexpr int KA_THREADS = 128;

template <int RC>
__global__ __launch_bounds__(KA_THREADS) void kapply_kernel(
    const KTask* __restrict__ tasks, const KSrc* __restrict__ srcs,
    const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
    BufTable bt, int task0, int nrhs, int ktype) {
    const KTask t = tasks[task0 + blockIdx.x];
    const int r0 = blockIdx.y * RC;
    const int nr = min(RC, nrhs - r0);
    __shared__ double sx[KA_THREADS], sy[KA_THREADS], sz[KA_THREADS];
    __shared__ double2 sv[RC][KA_THREADS];
    for (int tb = 0; tb < t.count; tb += KA_THREADS) {
        const int i = tb + threadIdx.x;
        const bool ok = i < t.count;
        double tx = 0, ty = 0, tz = 0;
        if (ok) {
            tx = px[t.off + i] * t.scale + t.c[0];
            ty = py[t.off + i] * t.scale + t.c[1];
            tz = pz[t.off + i] * t.scale + t.c[2];
        }
        double2 acc[RC];
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = make_double2(0, 0);
        for (int si = t.s0; si < t.s1; ++si) {
            const KSrc s = srcs[si];
            const double2* vb = bt.p[s.vbuf] + s.voff;
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
                        sv[q][threadIdx.x] = q < nr ? vb[(int64_t)(r0 + q) * vs + sb + threadIdx.x]
                                                    : make_double2(0, 0);
                }
                __syncthreads();
                if (ok) {
                    for (int j = 0; j < n; ++j) {
                        double2 g = green(tx - sx[j], ty - sy[j], tz - sz[j], ktype);
#pragma unroll
                        for (int q = 0; q < RC; ++q) cfma(acc[q], g, sv[q][j]);
                    }
                }
            }
        }
        if (ok) {
            double2* ob = bt.p[t.obuf] + t.ooff + i;
            const int64_t os = bt.stride[t.obuf];
            for (int q = 0; q < nr; ++q) {
                double2* o = ob + (int64_t)(r0 + q) * os;
                if (t.mode == 0) {
                    *o = acc[q];
                } else {
                    double2 v = *o;
                    *o = make_double2(v.x + acc[q].x, v.y + acc[q].y);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- grouped GEMM
constexpr int GK = 16;

__global__ __launch_bounds__(256) void gemm_kernel(
    const GGroup* __restrict__ groups, const GTile* __restrict__ tiles,
    const int64_t* __restrict__ colB, const int64_t* __restrict__ colC,
    const double2* __restrict__ opmat, BufTable bt, int tile0) {
    const GTile tl = tiles[tile0 + blockIdx.x];
    const GGroup g = groups[tl.g];
    const int rhs = blockIdx.y;
    const int m0 = tl.mt * GEMM_TM, n0 = tl.nt * GEMM_TN;
    const double2* A = opmat + g.aoff;
    const double2* B = bt.p[g.bbuf] + (int64_t)rhs * bt.stride[g.bbuf];
    double2* C = bt.p[g.cbuf] + (int64_t)rhs * bt.stride[g.cbuf];
    __shared__ double2 As[GK][GEMM_TM + 1];
    __shared__ double2 Bs[GK][GEMM_TN + 1];
    __shared__ int64_t cb[GEMM_TN];
    const int tid = threadIdx.x;
    const int tx = tid % 16, ty = tid / 16;
    if (tid < GEMM_TN) cb[tid] = (n0 + tid < g.ncols) ? colB[g.col0 + n0 + tid] : -1;
    double2 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = make_double2(0, 0);
    for (int k0 = 0; k0 < g.K; k0 += GK) {
        __syncthreads();
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const int e = tid + 256 * l;
            int mm, kk;
            if (g.transA) { kk = e % GK; mm = e / GK; }
            else { mm = e % GEMM_TM; kk = e / GEMM_TM; }
            const int m = m0 + mm, k = k0 + kk;
            double2 v = make_double2(0, 0);
            if (m < g.M && k < g.K)
                v = g.transA ? A[k + (int64_t)m * g.lda] : A[m + (int64_t)k * g.lda];
            As[kk][mm] = v;
        }
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            const int e = tid + 256 * l;
            const int kk = e % GK, nn = e / GK;
            const int k = k0 + kk;
            const int64_t c = cb[nn];
            Bs[kk][nn] = (c >= 0 && k < g.K) ? B[c + k] : make_double2(0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; ++kk) {
            double2 a[4], b[2];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[kk][ty + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) cfma(acc[i][j], a[i], b[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = n0 + ty + 16 * j;
        if (n >= g.ncols) continue;
        double2* Cc = C + colC[g.col0 + n];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = m0 + tx + 16 * i;
            if (m >= g.M) continue;
            double2 v = acc[i][j];
            if (g.soff >= 0) v = cmul(v, opmat[g.soff + m]);
            if (g.mode == 0) {
                Cc[m] = v;
            } else if (g.mode == 1) {
                double2 o = Cc[m];
                Cc[m] = make_double2(o.x + v.x, o.y + v.y);
            } else {
                unsafeAtomicAdd(&Cc[m].x, v.x);
                unsafeAtomicAdd(&Cc[m].y, v.y);
            }
        }
    }
}

// ---------------------------------------------------------------- (2P)^3 DFT in LDS
template <int G>
struct Dft {
    // twiddles w_k = exp(-2 pi i k / G) (forward); the inverse uses their conjugates
    __device__ static void twiddles(double2* tw) {
        if ((int)threadIdx.x < G) {
            double s, c;
            sincospi(2.0 * threadIdx.x / G, &s, &c);
            tw[threadIdx.x] = make_double2(c, -s);
        }
    }
    // one pass of a length-G DFT along axis ax (0 fastest) from in to out
    __device__ static void pass(const double2* in, double2* out, const double2* tw, int ax, bool inverse) {
        constexpr int G3 = G * G * G;
        const int st = ax == 0 ? 1 : (ax == 1 ? G : G * G);
        const double sg = inverse ? -1.0 : 1.0;
        for (int e = threadIdx.x; e < G3; e += blockDim.x) {
            const int p = (e / st) % G;
            const int base = e - p * st;
            double2 acc = make_double2(0, 0);
#pragma unroll
            for (int q = 0; q < G; ++q) {
                const double2 w = tw[(p * q) % G];
                cfma(acc, in[base + q * st], make_double2(w.x, sg * w.y));
            }
            out[e] = acc;
        }
    }
};

template <int G>
__global__ __launch_bounds__(256) void fft_fwd_kernel(const FftSrc* __restrict__ fs, const int* __restrict__ grid,
                                                      int NL, BufTable bt, int f0) {
    constexpr int G3 = G * G * G;
    __shared__ double2 g0[G3], g1[G3];
    __shared__ double2 tw[G];
    Dft<G>::twiddles(tw);
    const FftSrc f = fs[f0 + blockIdx.x];
    const int rhs = blockIdx.y;
    const double2* in = bt.p[BUF_LF_UPEQ] + (int64_t)rhs * bt.stride[BUF_LF_UPEQ] + f.in_off;
    double2* out = bt.p[BUF_SPEC] + (int64_t)rhs * bt.stride[BUF_SPEC] + f.spec_off;
    for (int e = threadIdx.x; e < G3; e += blockDim.x) g0[e] = make_double2(0, 0);
    __syncthreads();
    for (int k = threadIdx.x; k < NL; k += blockDim.x) g0[grid[f.grid + k]] = in[k];
    __syncthreads();
    Dft<G>::pass(g0, g1, tw, 0, false);
    __syncthreads();
    Dft<G>::pass(g1, g0, tw, 1, false);
    __syncthreads();
    Dft<G>::pass(g0, g1, tw, 2, false);
    __syncthreads();
    for (int e = threadIdx.x; e < G3; e += blockDim.x) out[e] = g1[e];
}

template <int G>
__global__ __launch_bounds__(256) void vlist_kernel(const VTarget* __restrict__ vt, const VPair* __restrict__ vp,
                                                    const int* __restrict__ grid, const double2* __restrict__ opmat,
                                                    int NL, BufTable bt, int v0) {
    constexpr int G3 = G * G * G;
    constexpr int PER = (G3 + 255) / 256;
    __shared__ double2 g0[G3], g1[G3];
    __shared__ double2 tw[G];
    Dft<G>::twiddles(tw);
    const VTarget t = vt[v0 + blockIdx.x];
    const int rhs = blockIdx.y;
    const double2* spec = bt.p[BUF_SPEC] + (int64_t)rhs * bt.stride[BUF_SPEC];
    double2 acc[PER];
#pragma unroll
    for (int l = 0; l < PER; ++l) acc[l] = make_double2(0, 0);
    for (int p = t.p0; p < t.p1; ++p) {
        const VPair pr = vp[p];
        const double2* s = spec + pr.spec_off;
        const double2* T = opmat + t.tab_off + (int64_t)pr.offidx * G3;
#pragma unroll
        for (int l = 0; l < PER; ++l) {
            const int e = threadIdx.x + 256 * l;
            if (e < G3) cfma(acc[l], s[e], T[e]);
        }
    }
#pragma unroll
    for (int l = 0; l < PER; ++l) {
        const int e = threadIdx.x + 256 * l;
        if (e < G3) g0[e] = acc[l];
    }
    __syncthreads();
    Dft<G>::pass(g0, g1, tw, 0, true);
    __syncthreads();
    Dft<G>::pass(g1, g0, tw, 1, true);
    __syncthreads();
    Dft<G>::pass(g0, g1, tw, 2, true);
    __syncthreads();
    double2* out = bt.p[BUF_LF_CHK] + (int64_t)rhs * bt.stride[BUF_LF_CHK] + t.out_off;
    const double coef = 1.0 / G3;
    for (int k = threadIdx.x; k < NL; k += blockDim.x) {
        const double2 v = g1[grid[t.grid + k]];
        double2 o = out[k];
        out[k] = make_double2(o.x + v.x * coef, o.y + v.y * coef);
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
        opmat[k.out + e] = green(d[0], d[1], d[2], ktype);
    }
}

__global__ void zero_kernel(double2* p, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        p[e] = make_double2(0, 0);
}

__global__ void gather_kernel(const double2* __restrict__ in, double2* __restrict__ out, const int* __restrict__ perm,
                              int64_t N, int nrhs) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * nrhs; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / N, i = e % N;
        out[e] = in[r * N + perm[i]];
    }
}

__global__ void scatter_kernel(const double2* __restrict__ in, double2* __restrict__ out, const int* __restrict__ perm,
                               int64_t N, int nrhs) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * nrhs; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / N, i = e % N;
        out[r * N + perm[i]] = in[e];
    }
}

// direct sum u(t_i) = sum_j G(t_i, s_j) f_j (Wave3d::check, wave3d_check.cpp:66-76)
__global__ __launch_bounds__(256) void direct_kernel(const double* __restrict__ trg, int M, const double* __restrict__ src,
                                                     int Ns, const double2* __restrict__ den, int nrhs,
                                                     double2* __restrict__ out, int ktype) {
    __shared__ double s3[256 * 3];
    __shared__ double2 sd[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int rhs = blockIdx.y;
    const bool ok = i < M;
    double tx = 0, ty = 0, tz = 0;
    if (ok) { tx = trg[3 * i]; ty = trg[3 * i + 1]; tz = trg[3 * i + 2]; }
    double2 acc = make_double2(0, 0);
    for (int sb = 0; sb < Ns; sb += 256) {
        const int n = min(256, Ns - sb);
        __syncthreads();
        if ((int)threadIdx.x < n) {
            const int j = sb + threadIdx.x;
            s3[3 * threadIdx.x] = src[3 * j];
            s3[3 * threadIdx.x + 1] = src[3 * j + 1];
            s3[3 * threadIdx.x + 2] = src[3 * j + 2];
            sd[threadIdx.x] = den[(int64_t)rhs * Ns + j];
        }
        __syncthreads();
        if (ok)
            for (int j = 0; j < n; ++j)
                cfma(acc, green(tx - s3[3 * j], ty - s3[3 * j + 1], tz - s3[3 * j + 2], ktype), sd[j]);
    }
    if (ok) out[(int64_t)rhs * M + i] = acc;
}

// dense Kernel3d::kernel block: out(i,j) = G(trg_i, src_j), column-major M x N
__global__ void kmatrix_kernel(const double* __restrict__ trg, int M, const double* __restrict__ src, int Ns,
                               double2* __restrict__ out, int ktype) {
    const int64_t mn = (int64_t)M * Ns;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn; e += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % M), j = (int)(e / M);
        out[e] = green(trg[3 * i] - src[3 * j], trg[3 * i + 1] - src[3 * j + 1], trg[3 * i + 2] - src[3 * j + 2], ktype);
    }
}

int grid_for(int64_t n, int threads) {
    int64_t b = (n + threads - 1) / threads;
    return (int)std::min<int64_t>(std::max<int64_t>(b, 1), 65536);
}

}  // namespace

// ---------------------------------------------------------------- launchers
hipError_t launch_stage(const DevPlan& d, const Stage& s, int nrhs, hipStream_t st) {
    const int n = s.end - s.begin;
    switch (s.kind) {
    case ST_ZERO: {
        int64_t cnt = s.count * nrhs;   // rhs slabs are contiguous with stride = count
        if (cnt > 0) zero_kernel<<<grid_for(cnt, 256), 256, 0, st>>>(d.bt.p[s.buf], cnt);
        break;
    }
    case ST_GATHER:
        gather_kernel<<<grid_for(d.N * nrhs, 256), 256, 0, st>>>(d.den_io, d.bt.p[BUF_DEN], d.perm, d.N, nrhs);
        break;
    case ST_SCATTER:
        scatter_kernel<<<grid_for(d.N * nrhs, 256), 256, 0, st>>>(d.bt.p[BUF_VAL], d.val_io, d.perm, d.N, nrhs);
        break;
    case ST_KAPPLY:
        if (nrhs == 1) {
            kapply_kernel<1><<<dim3(n, 1), KA_THREADS, 0, st>>>(d.ktasks, d.ksrcs, d.px, d.py, d.pz, d.bt,
                                                                s.begin, nrhs, d.ktype);
        } else {
            kapply_kernel<4><<<dim3(n, (nrhs + 3) / 4), KA_THREADS, 0, st>>>(d.ktasks, d.ksrcs, d.px, d.py, d.pz,
                                                                            d.bt, s.begin, nrhs, d.ktype);
        }
        break;
    case ST_GEMM:
        gemm_kernel<<<dim3(n, nrhs), 256, 0, st>>>(d.groups, d.tiles, d.colB, d.colC, d.opmat, d.bt, s.begin);
        break;
    case ST_FFTFWD:
        if (d.P != 6) return hipErrorInvalidValue;
        fft_fwd_kernel<12><<<dim3(n, nrhs), 256, 0, st>>>(d.fsrc, d.grid, d.NL, d.bt, s.begin);
        break;
    case ST_VLIST:
        if (d.P != 6) return hipErrorInvalidValue;
        vlist_kernel<12><<<dim3(n, nrhs), 256, 0, st>>>(d.vtrg, d.vpair, d.grid, d.opmat, d.NL, d.bt, s.begin);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_kgen(const KGen* jobs, int njobs, const double* gp, double2* opmat, int ktype, hipStream_t st) {
    for (int j0 = 0; j0 < njobs; j0 += 65536) {
        int n = std::min(65536, njobs - j0);
        kgen_kernel<<<n, 256, 0, st>>>(jobs, gp, opmat, j0, ktype);
    }
    return hipGetLastError();
}

hipError_t launch_direct(const double* trg, int M, const double* src, int Ns, const double2* den, int nrhs,
                         double2* out, int ktype, hipStream_t st) {
    direct_kernel<<<dim3((M + 255) / 256, nrhs), 256, 0, st>>>(trg, M, src, Ns, den, nrhs, out, ktype);
    return hipGetLastError();
}

hipError_t launch_kmatrix(const double* trg, int M, const double* src, int Ns, double2* out, int ktype,
                          hipStream_t st) {
    kmatrix_kernel<<<grid_for((int64_t)M * Ns, 256), 256, 0, st>>>(trg, M, src, Ns, out, ktype);
    return hipGetLastError();
}

}  // namespace ddfmm

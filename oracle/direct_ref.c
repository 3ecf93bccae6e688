/* ORACLE (test infrastructure only): FP64 direct sum u(t_i) = sum_j G(t_i, s_j) f_j.
 *
 * Restates Wave3d::check's direct evaluation (wave3d_check.cpp:66-76) with the kernel of
 * Kernel3d::kernel (kernel3d.cpp:26-107): k = 2*pi, r^2 < 1e-16 => r := 1,
 * type 0: exp(ikr)/r, type 1: exp(ikr).  Pairwise sums are accumulated in
 * long double to keep the reference value well below the FMM's error.
 *
 * flags = kernel type | (threads << 8)   (threads 0 = OpenMP default)
 */
#include <math.h>
#include <complex.h>
#include <stddef.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void direct_sum(int nt, const double *trg, const double *src, int ns,
                const double *den /* nrhs x ns complex */, int nrhs,
                double *out /* nrhs x nt complex */, void *unused, int flags)
{
    (void)unused;
    int ktype = flags & 0xff;
    int threads = flags >> 8;
    const double K = 2.0 * M_PI;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int i = 0; i < nt; i++) {
        for (int r = 0; r < nrhs; r++) {
            long double accr = 0.0L, acci = 0.0L;
            const double *d = den + 2 * (size_t)r * ns;
            for (int j = 0; j < ns; j++) {
                double x = trg[3 * i] - src[3 * j];
                double y = trg[3 * i + 1] - src[3 * j + 1];
                double z = trg[3 * i + 2] - src[3 * j + 2];
                double r2 = x * x + y * y + z * z;
                if (r2 < 1e-16) r2 = 1.0;
                double rr = sqrt(r2);
                double s, c;
                sincos(K * rr, &s, &c);
                double gr = c, gi = s;
                if (ktype == 0) { gr /= rr; gi /= rr; }
                double fr = d[2 * j], fi = d[2 * j + 1];
                accr += (long double)(gr * fr - gi * fi);
                acci += (long double)(gr * fi + gi * fr);
            }
            out[2 * ((size_t)r * nt + i)] = (double)accr;
            out[2 * ((size_t)r * nt + i) + 1] = (double)acci;
        }
    }
}

/*
 * Oracle (test infrastructure only): plain-C restatement of the reference's
 * CPU coupling kernels.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the reference that
 * results are compared against, or as the timed CPU baseline -- never as part
 * of the product path.
 *
 * The loop nests follow the reference line by line so that this file can be
 * checked against it by reading:
 *   particle_interaction_scalar  yasfpy/functions/cpu_numba.py:21-95
 *   compute_scattering_cross_section  yasfpy/functions/cpu_numba.py:410-457
 * The OpenMP "parallel for" over output rows plays the role of numba.prange.
 *
 * Table layouts are the reference's (yasfpy/functions/misc.py:117-243), with
 * the target-particle axis restricted to rows [s1_lo, s1_lo + n_s1) so that a
 * bounded sample of a large matvec can be timed without O(N^2) tables:
 *   plm        [(2L+1)(2L+2)/2][n_s1][N]   double
 *   sph_h      [2L+1][n_s1][N]             double complex   (channel fixed)
 *   e_j_dm_phi [4L+1][n_s1][N]             double complex
 *   ab5        [nmax][nmax][2L+1]          double complex
 *   idx        [N*nmax][5]                 int64 (s, n, tau, l, m)
 * Complex arrays are interleaved (re, im) doubles.
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>

typedef double _Complex cplx;

/* (W x)[j1] for rows j1 of particles s1 in [s1_lo, s1_lo+n_s1); writes
 * wx[(s1-s1_lo)*nmax + n1]. Mirrors cpu_numba.py:21-95 including the full
 * p-range max(|dm|, |dl|+|dtau|) .. l1+l2 without zero skipping. */
void oracle_particle_interaction_scalar(int lmax, int particle_number,
                                        int s1_lo, int n_s1,
                                        const int64_t *idx, const double *x_ri,
                                        const double *ab5_ri,
                                        const double *plm, const double *sph_h_ri,
                                        const double *ephi_ri, double *wx_ri)
{
    const int nmax = 2 * lmax * (lmax + 2);
    const int P = 2 * lmax + 1;
    const int N = particle_number;
    const cplx *x = (const cplx *)x_ri;
    const cplx *ab5 = (const cplx *)ab5_ri;
    const cplx *sph_h = (const cplx *)sph_h_ri;
    const cplx *ephi = (const cplx *)ephi_ri;
    cplx *wx = (cplx *)wx_ri;
    const long plane = (long)n_s1 * N;

#pragma omp parallel for schedule(dynamic, 16)
    for (long r = 0; r < (long)n_s1 * nmax; ++r) {
        const long j1 = (long)s1_lo * nmax + r;
        const int s1 = (int)idx[j1 * 5 + 0];
        const int n1 = (int)idx[j1 * 5 + 1];
        const int tau1 = (int)idx[j1 * 5 + 2];
        const int l1 = (int)idx[j1 * 5 + 3];
        const int m1 = (int)idx[j1 * 5 + 4];
        const long row = (long)(s1 - s1_lo) * N;
        cplx acc = 0.0;
        for (int s2 = 0; s2 < N; ++s2) {
            if (s2 == s1)
                continue;
            const long base = (long)s2 * nmax;
            for (long j2 = base; j2 < base + nmax; ++j2) {
                const int n2 = (int)idx[j2 * 5 + 1];
                const int tau2 = (int)idx[j2 * 5 + 2];
                const int l2 = (int)idx[j2 * 5 + 3];
                const int m2 = (int)idx[j2 * 5 + 4];
                const int delta_tau = abs(tau1 - tau2);
                const int delta_l = abs(l1 - l2);
                const int delta_m = abs(m1 - m2);
                int p_start = delta_m;
                if (delta_l + delta_tau > p_start)
                    p_start = delta_l + delta_tau;
                const cplx phase_x =
                    ephi[(long)(m2 - m1 + 2 * lmax) * plane + row + s2] * x[j2];
                for (int p = p_start; p <= l1 + l2; ++p) {
                    acc += ab5[((long)n2 * nmax + n1) * P + p] *
                           plm[(long)(p * (p + 1) / 2 + delta_m) * plane + row + s2] *
                           phase_x * sph_h[(long)p * plane + row + s2];
                }
            }
        }
        wx[r] = acc;
    }
}

/* Sum_{j1,j2} conj(f[j1]) * phase * f[j2] * Sum_{p=|dm|}^{2L} ab5[n2,n1,p] *
 * plm[p(p+1)/2+|dm|] * sph_j[p] for one channel, including s1 == s2 terms.
 * Mirrors cpu_numba.py:410-457 (per-row partial sums reduced at the end).
 * Tables here cover all target particles (s1_lo = 0, n_s1 = N). */
void oracle_scattering_cross_section(int lmax, int particle_number,
                                     const int64_t *idx, const double *sfc_ri,
                                     const double *ab5_ri, const double *plm,
                                     const double *sph_j_ri, const double *ephi_ri,
                                     double *out_ri)
{
    const int nmax = 2 * lmax * (lmax + 2);
    const int P = 2 * lmax + 1;
    const int N = particle_number;
    const long jmax = (long)N * nmax;
    const cplx *f = (const cplx *)sfc_ri;
    const cplx *ab5 = (const cplx *)ab5_ri;
    const cplx *sph_j = (const cplx *)sph_j_ri;
    const cplx *ephi = (const cplx *)ephi_ri;
    const long plane = (long)N * N;
    cplx *partial = (cplx *)calloc((size_t)jmax, sizeof(cplx));

#pragma omp parallel for schedule(dynamic, 16)
    for (long j1 = 0; j1 < jmax; ++j1) {
        const int s1 = (int)idx[j1 * 5 + 0];
        const int n1 = (int)idx[j1 * 5 + 1];
        const int m1 = (int)idx[j1 * 5 + 4];
        cplx acc = 0.0;
        for (long j2 = 0; j2 < jmax; ++j2) {
            const int s2 = (int)idx[j2 * 5 + 0];
            const int n2 = (int)idx[j2 * 5 + 1];
            const int m2 = (int)idx[j2 * 5 + 4];
            const int delta_m = abs(m1 - m2);
            const long off = (long)s1 * N + s2;
            const cplx phase = ephi[(long)(m2 - m1 + 2 * lmax) * plane + off];
            cplx val = 0.0;
            for (int p = delta_m; p <= 2 * lmax; ++p)
                val += ab5[((long)n2 * nmax + n1) * P + p] *
                       plm[(long)(p * (p + 1) / 2 + delta_m) * plane + off] *
                       sph_j[(long)p * plane + off];
            val *= conj(f[j1]) * phase * f[j2];
            acc += val;
        }
        partial[j1] = acc;
    }
    cplx total = 0.0;
    for (long j1 = 0; j1 < jmax; ++j1)
        total += partial[j1];
    free(partial);
    out_ri[0] = creal(total);
    out_ri[1] = cimag(total);
}

#ifdef _OPENMP
#include <omp.h>
#endif

/* Set / query the OpenMP thread count used by the loops above. */
int oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0)
        omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

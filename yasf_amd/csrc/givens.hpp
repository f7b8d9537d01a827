// Givens rotations of the GMRES least-squares problem on the device: the
// update of one Hessenberg column (SciPy's inner loop,
// scipy/sparse/linalg/_isolve/iterative.py:755-795, with LAPACK zlartg's
// rotation convention) and the back substitution at a cycle's end.
#pragma once

#include <hip/hip_runtime.h>

namespace yasfhip {

__device__ __forceinline__ double cabs_dev(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }

// zlartg: [c s; -conj(s) c] [f; g] = [r; 0] with c real.
__device__ inline void lartg_dev(double2 f, double2 g, double &c, double2 &s, double2 &r) {
    if (g.x == 0.0 && g.y == 0.0) {
        c = 1.0;
        s = make_double2(0.0, 0.0);
        r = f;
        return;
    }
    if (f.x == 0.0 && f.y == 0.0) {
        c = 0.0;
        const double ag = cabs_dev(g);
        s = make_double2(g.x / ag, -g.y / ag);
        r = make_double2(ag, 0.0);
        return;
    }
    const double af = cabs_dev(f), ag = cabs_dev(g);
    const double d = sqrt(af * af + ag * ag);
    c = af / d;
    const double2 sg = make_double2(f.x / af, f.y / af);  // f / |f|
    // s = sg conj(g) / d
    s = make_double2((sg.x * g.x + sg.y * g.y) / d, (sg.y * g.x - sg.x * g.y) / d);
    r = make_double2(sg.x * d, sg.y * d);
}

// Hessenberg column j (col[0..j+1], col[j+1] the new subdiagonal entry):
// applies the rotations of columns 0..j-1, forms this column's rotation
// (kept in gc[j], gs[j]), zeroes col[j+1] and rotates the right-hand side g.
// Returns the new residual estimate |g[j+1]|.
__device__ inline double givens_column(double2 *col, int j, double *gc, double2 *gs, double2 *g) {
    for (int k = 0; k < j; ++k) {
        const double c = gc[k];
        const double2 sn = gs[k];
        const double2 n0 = col[k], n1 = col[k + 1];
        // col[k] = c n0 + sn n1 ; col[k+1] = -conj(sn) n0 + c n1
        col[k] = make_double2(c * n0.x + sn.x * n1.x - sn.y * n1.y, c * n0.y + sn.x * n1.y + sn.y * n1.x);
        col[k + 1] = make_double2(-(sn.x * n0.x + sn.y * n0.y) + c * n1.x, -(sn.x * n0.y - sn.y * n0.x) + c * n1.y);
    }
    double c;
    double2 sn, mag;
    lartg_dev(col[j], col[j + 1], c, sn, mag);
    gc[j] = c;
    gs[j] = sn;
    col[j] = mag;
    col[j + 1] = make_double2(0.0, 0.0);
    const double2 gj = g[j];
    // tmp = -conj(sn) g[j]
    const double2 tmp = make_double2(-(sn.x * gj.x + sn.y * gj.y), -(sn.x * gj.y - sn.y * gj.x));
    g[j] = make_double2(c * gj.x, c * gj.y);
    g[j + 1] = tmp;
    return cabs_dev(tmp);
}

// y = H^-1 g over the m leading columns (upper triangular after the
// rotations; column k starts at H + k * ld); y_k = 0 where H[k][k] = 0, as
// SciPy zeroes the last coefficient of a singular column.  One thread.
__device__ inline void backsolve_dev(const double2 *__restrict__ H, const double2 *__restrict__ g, int m, int ld,
                                     double2 *__restrict__ y) {
    for (int k = m - 1; k >= 0; --k) {
        double2 v = g[k];
        for (int i = k + 1; i < m; ++i) {
            const double2 hik = H[static_cast<long>(i) * ld + k];
            const double2 yi = y[i];
            v.x -= hik.x * yi.x - hik.y * yi.y;
            v.y -= hik.x * yi.y + hik.y * yi.x;
        }
        const double2 d = H[static_cast<long>(k) * ld + k];
        const double dd = d.x * d.x + d.y * d.y;
        y[k] = dd > 0.0 ? make_double2((v.x * d.x + v.y * d.y) / dd, (v.y * d.x - v.x * d.y) / dd)
                        : make_double2(0.0, 0.0);
    }
}

}  // namespace yasfhip

/*
 * ref_wrap.cpp -- TEST INFRASTRUCTURE ONLY: a C interface to the reference's own
 * Boost-free numerical islands, compiled from their sources where they lie
 * (oracle/ref_build.sh -> oracle/_ref/libref_islands.so) to pin the oracle and the
 * host pre-interpolation against the reference code itself.  Nothing in the
 * product links or loads this.
 *
 *   ALGLIB 3.15.0 spline1dgriddiffcubic (alglib-3.15.0/cpp/src/interpolation.cpp)
 *   ALGLIB 3.15.0 spline2dbuildbicubicv / spline2dcalc / spline2ddiff (the 2D maps)
 *   libtricubic 1.0 tricubic_get_coeff / tricubic_eval
 *     (tricubic-1.0/src/libtricubic/libtricubic.cpp)
 *
 * ref_preinterpol follows the call sequence of TabField3::CalcDerivs / PreInterpol
 * (src/field_3d.cpp:336-414): spline derivatives along x, y, z, then xy, xz, yz,
 * xyz, corner values scaled by the cell sizes, tricubic_get_coeff per cell.
 */
#include <vector>
#include "interpolation.h"
#include "tricubic.h"

namespace {
/* derivative along axis `dim` of a [nx][ny][nz] table (C order) */
void calc_derivs(const std::vector<double> g[3], const int n[3], const std::vector<double> &tab, int dim,
                 std::vector<double> &out) {
    out.assign(tab.size(), 0.0);
    alglib::real_1d_array x, y, d;
    x.setcontent(n[dim], g[dim].data());
    y.setlength(n[dim]);
    d.setlength(n[dim]);
    int d1 = (dim + 1) % 3, d2 = (dim + 2) % 3;
    int idx[3];
    for (idx[d1] = 0; idx[d1] < n[d1]; ++idx[d1])
        for (idx[d2] = 0; idx[d2] < n[d2]; ++idx[d2]) {
            for (idx[dim] = 0; idx[dim] < n[dim]; ++idx[dim])
                y[idx[dim]] = tab[((size_t)idx[0] * n[1] + idx[1]) * n[2] + idx[2]];
            alglib::spline1dgriddiffcubic(x, y, d);
            for (idx[dim] = 0; idx[dim] < n[dim]; ++idx[dim])
                out[((size_t)idx[0] * n[1] + idx[1]) * n[2] + idx[2]] = d[idx[dim]];
        }
}
} // namespace

extern "C" {

/* alglib::spline1dgriddiffcubic(x, y, d) with default (parabolic) end conditions */
int ref_spline_diff(int n, const double *x, const double *y, double *d) {
    alglib::real_1d_array X, Y, D;
    X.setcontent(n, x);
    Y.setcontent(n, y);
    alglib::spline1dgriddiffcubic(X, Y, D);
    for (int i = 0; i < n; ++i) d[i] = D[i];
    return 0;
}

/* libtricubic tricubic_get_coeff from the 8 x 8 corner data (f, fx, fy, fz, fxy, fxz, fyz, fxyz) */
int ref_tricubic_coeff(const double *corner64, double *a) {
    double c[8][8];
    for (int k = 0; k < 64; ++k) c[k / 8][k % 8] = corner64[k];
    tricubic_get_coeff(a, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
    return 0;
}

/* libtricubic tricubic_eval (value, or a partial derivative of order dx, dy, dz) */
double ref_tricubic_eval(const double *a, double x, double y, double z, int dx, int dy, int dz) {
    double c[64];
    for (int k = 0; k < 64; ++k) c[k] = a[k];
    if (dx == 0 && dy == 0 && dz == 0) return tricubic_eval(c, x, y, z);
    return tricubic_eval(c, x, y, z, dx, dy, dz);
}

/* TabField3::PreInterpol on one component: coef [(nx-1)(ny-1)(nz-1)][64] */
int ref_preinterpol(int nx, int ny, int nz, const double *xg, const double *yg, const double *zg,
                    const double *values, double *coef) {
    const int n[3] = {nx, ny, nz};
    std::vector<double> g[3] = {std::vector<double>(xg, xg + nx), std::vector<double>(yg, yg + ny),
                                std::vector<double>(zg, zg + nz)};
    std::vector<double> tab(values, values + (size_t)nx * ny * nz), dx, dy, dz, dxy, dxz, dyz, dxyz;
    calc_derivs(g, n, tab, 0, dx);
    calc_derivs(g, n, tab, 1, dy);
    calc_derivs(g, n, tab, 2, dz);
    calc_derivs(g, n, dx, 1, dxy);
    calc_derivs(g, n, dx, 2, dxz);
    calc_derivs(g, n, dy, 2, dyz);
    calc_derivs(g, n, dxy, 2, dxyz);
    for (int ix = 0; ix < nx - 1; ++ix)
        for (int iy = 0; iy < ny - 1; ++iy)
            for (int iz = 0; iz < nz - 1; ++iz) {
                double cx = g[0][ix + 1] - g[0][ix], cy = g[1][iy + 1] - g[1][iy], cz = g[2][iz + 1] - g[2][iz];
                double c[8][8];
                for (int k = 0; k < 8; ++k) {
                    size_t p = ((size_t)(ix + (k & 1)) * ny + (iy + ((k >> 1) & 1))) * nz + (iz + ((k >> 2) & 1));
                    c[0][k] = tab[p];
                    c[1][k] = dx[p] * cx;
                    c[2][k] = dy[p] * cy;
                    c[3][k] = dz[p] * cz;
                    c[4][k] = dxy[p] * cx * cy;
                    c[5][k] = dxz[p] * cx * cz;
                    c[6][k] = dyz[p] * cy * cz;
                    c[7][k] = dxyz[p] * cx * cy * cz;
                }
                double *a = coef + (((size_t)ix * (ny - 1) + iy) * (nz - 1) + iz) * 64;
                tricubic_get_coeff(a, c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7]);
            }
    return 0;
}

} // extern "C"

extern "C" {

/*
 * TabField's splines (src/field_2d.cpp:235-285, 289-337): alglib::spline2dbuildbicubicv
 * over the radial grid x[m] and axial grid y[n] of values[n][m], then at the points
 * (px, py): spline2dcalc (diff == 0; F only) or spline2ddiff (diff == 1; F, dF/dx, dF/dy).
 */
int ref_spline2d(int m, int n, const double *x, const double *y, const double *values, int np, const double *px,
                 const double *py, int diff, double *F, double *Fx, double *Fy) {
    alglib::real_1d_array X, Y, V;
    X.setcontent(m, x);
    Y.setcontent(n, y);
    V.setcontent((alglib::ae_int_t)m * n, values);
    alglib::spline2dinterpolant c;
    alglib::spline2dbuildbicubicv(X, m, Y, n, V, 1, c);
    for (int i = 0; i < np; ++i) {
        if (diff) {
            double fxy;
            alglib::spline2ddiff(c, px[i], py[i], F[i], Fx[i], Fy[i], fxy);
        } else {
            F[i] = alglib::spline2dcalc(c, px[i], py[i]);
        }
    }
    return 0;
}

}

/*
 * pt_precompute.hip -- TabField3::PreInterpol (src/field_3d.cpp:336-414) on the GPU.
 *
 * Seven spline-derivative passes over the sample table (CalcDerivs: d/dx, d/dy, d/dz
 * of F, then d/dy and d/dz of dF/dx, d/dz of dF/dy and of d2F/dxdy), each one thread
 * per grid line solving ALGLIB's tridiagonal system for parabolically terminated
 * cubic splines (spline1dgriddiffcubic), then one thread per cell forming the 64
 * corner data scaled by the cell sizes and multiplying them by libtricubic's 64 x 64
 * integer matrix.  Every operation follows the host routine pt_tricubic_precompute in
 * the same order (no contraction, all 64 products of each sum, zeros included), so
 * the coefficients are bit-identical to it.
 *
 * Work per cell: 4096 multiply-adds in FP64 and 8 x 8 B of gathered derivative data
 * per corner, 512 B of coefficients written per component (256 B more for the FP32
 * copy): the coefficient kernel is bound by FP64 issue rate, the line passes by HBM.
 */
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pt_device.h"

namespace ptp {

/* derivatives along `axis` of tab [nx][ny][nz] for every grid line; scr: [2][len][nlines] */
__global__ void __launch_bounds__(256) spline_lines_kernel(const double *__restrict__ g, int nx, int ny, int nz, int axis,
                                                           const double *__restrict__ tab, double *__restrict__ out,
                                                           double *__restrict__ scr) {
    const int n[3] = {nx, ny, nz};
    const long long stride[3] = {(long long)ny * nz, nz, 1};
    const int len = n[axis];
    /* lines are numbered so that consecutive threads touch consecutive z where z is not the axis */
    const int da = axis == 0 ? 1 : 0, db = axis == 2 ? 1 : 2;
    const long long nlines = (long long)n[da] * n[db];
    const long long line = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nlines) return;
    const long long ia = line / n[db], ib = line % n[db];
    const long long base = ia * stride[da] + ib * stride[db];
    const long long st = stride[axis];
    const double *y = tab + base;
    double *d = out + base;
    if (len == 2) {
        const double d0 = (y[st] - y[0]) / (g[1] - g[0]);
        d[0] = d0;
        d[st] = d0;
        return;
    }
    double *a2s = scr, *bs = scr + (size_t)len * nlines;
    /* row 0: a1 = 0, a2 = 1, a3 = 1, b = 2 (y1 - y0) / (x1 - x0) */
    double a2p = 1.0, a3p = 1.0;
    double yk = y[0], yk1 = y[st];
    double bp = 2 * (yk1 - yk) / (g[1] - g[0]);
    a2s[line] = a2p;
    bs[line] = bp;
    for (int k = 1; k <= len - 2; ++k) {
        const double ykm = yk;
        yk = yk1;
        yk1 = y[(long long)(k + 1) * st];
        const double a1 = g[k + 1] - g[k];
        const double a2 = 2 * (g[k + 1] - g[k - 1]);
        const double a3 = g[k] - g[k - 1];
        const double b = 3 * (yk - ykm) / (g[k] - g[k - 1]) * (g[k + 1] - g[k]) +
                         3 * (yk1 - yk) / (g[k + 1] - g[k]) * (g[k] - g[k - 1]);
        const double t = a1 / a2p;
        a2p = a2 - t * a3p;
        bp = b - t * bp;
        a3p = a3;
        a2s[(size_t)k * nlines + line] = a2p;
        bs[(size_t)k * nlines + line] = bp;
    }
    {   /* row n-1: a1 = 1, a2 = 1, a3 = 0, b = 2 (y[n-1] - y[n-2]) / (x[n-1] - x[n-2]) */
        const double b = 2 * (yk1 - yk) / (g[len - 1] - g[len - 2]);
        const double t = 1.0 / a2p;
        a2p = 1.0 - t * a3p;
        bp = b - t * bp;
    }
    double dk = bp / a2p;
    d[(long long)(len - 1) * st] = dk;
    for (int k = len - 2; k >= 0; --k) {
        const double a3 = k == 0 ? 1.0 : g[k] - g[k - 1];
        dk = (bs[(size_t)k * nlines + line] - a3 * dk) / a2s[(size_t)k * nlines + line];
        d[(long long)k * st] = dk;
    }
}

/* one thread per cell: corner data (F, Fx, Fy, Fz, Fxy, Fxz, Fyz, Fxyz at the 8 corners,
   derivatives scaled by the cell sizes) times the 64 x 64 matrix; der[8] -> tables of
   np values.  Output in libtricubic order a[i + 4j + 16k] (device_order 0) or in the
   tracker's order a[16i + 4j + k] (device_order 1), at coef[(cell * ncomp + comp) * 64] */
__global__ void __launch_bounds__(256) tricubic_coef_kernel(const double *__restrict__ xg, const double *__restrict__ yg,
                                                            const double *__restrict__ zg, int nx, int ny, int nz,
                                                            const double *const *__restrict__ der,
                                                            const double *__restrict__ A, int ncomp, int comp,
                                                            int device_order, double *__restrict__ coef) {
    __shared__ double As[64 * 64];
    for (int k = threadIdx.x; k < 64 * 64; k += blockDim.x) As[k] = A[k];
    __syncthreads();
    const long long ncell = (long long)(nx - 1) * (ny - 1) * (nz - 1);
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    const int iz = (int)(cell % (nz - 1));
    const int iy = (int)((cell / (nz - 1)) % (ny - 1));
    const int ix = (int)(cell / ((long long)(nz - 1) * (ny - 1)));
    const double cellx = xg[ix + 1] - xg[ix];
    const double celly = yg[iy + 1] - yg[iy];
    const double cellz = zg[iz + 1] - zg[iz];
    double xv[64];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const long long p = ((long long)(ix + (c & 1)) * ny + (iy + ((c >> 1) & 1))) * nz + (iz + ((c >> 2) & 1));
        xv[c] = der[0][p];
        xv[8 + c] = der[1][p] * cellx;
        xv[16 + c] = der[2][p] * celly;
        xv[24 + c] = der[3][p] * cellz;
        xv[32 + c] = der[4][p] * cellx * celly;
        xv[40 + c] = der[5][p] * cellx * cellz;
        xv[48 + c] = der[6][p] * celly * cellz;
        xv[56 + c] = der[7][p] * cellx * celly * cellz;
    }
    double *a = coef + ((size_t)cell * ncomp + comp) * 64;
#pragma unroll 1
    for (int i = 0; i < 64; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < 64; ++j) s += As[64 * i + j] * xv[j];
        const int o = device_order ? 16 * (i & 3) + 4 * ((i >> 2) & 3) + (i >> 4) : i;
        a[o] = s;
    }
}

/*
 * The mixed-precision table of a map from its FP64 table in the tracker's order: one thread
 * per cell and component writes its 4 doubles (the constant and linear coefficients, slots
 * 0, 1, 4, 16: 1, z, y, x) and 64 floats (all coefficients, those four zeroed) into the
 * cell's record (ptd::mix_cell_bytes: the doubles of all components first, then the floats).  The
 * FP32 remainder's error is bounded by 2^-20 sum_{i+j+k>=2} |a_ijk| for the value and by
 * 2^-20 sum |a_ijk| (i/hx + j/hy + k/hz) for the gradient (coefficients rounded to FP32 and
 * an FMA Horner over [0,1]^3, with room for the rounding of the cell coordinates); where the
 * gradient bound exceeds tol [field unit / m] or the value bound tol * min(hx, hy, hz), the
 * block is marked by a NaN constant term and the tracker reads the FP64 block instead.
 */
__global__ void __launch_bounds__(256) mixed_table_kernel(const double *__restrict__ xg, const double *__restrict__ yg,
                                                          const double *__restrict__ zg, int nx, int ny, int nz,
                                                          int ncomp, const double *__restrict__ coef, double tol,
                                                          unsigned char *__restrict__ mix, int *__restrict__ n_fp64) {
    const long long ncell = (long long)(nx - 1) * (ny - 1) * (nz - 1);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * ncomp) return;
    const long long cell = t / ncomp;
    const int iz = (int)(cell % (nz - 1));
    const int iy = (int)((cell / (nz - 1)) % (ny - 1));
    const int ix = (int)(cell / ((long long)(nz - 1) * (ny - 1)));
    const double hx = xg[ix + 1] - xg[ix], hy = yg[iy + 1] - yg[iy], hz = zg[iz + 1] - zg[iz];
    const double *a = coef + (size_t)t * 64;
    double vb = 0.0, gb = 0.0;
    for (int o = 0; o < 64; ++o) {
        const int i = o >> 4, j = (o >> 2) & 3, k = o & 3;
        if (i + j + k < 2) continue;
        const double m = fabs(a[o]);
        vb += m;
        gb += m * ((double)i / hx + (double)j / hy + (double)k / hz);
    }
    const double u = 1.0 / 1048576.0;   /* 2^-20 */
    const double hmin = fmin(hx, fmin(hy, hz));
    const bool keep64 = !(gb * u <= tol) || !(vb * u <= tol * hmin);
    const int comp = (int)(t - cell * ncomp);
    unsigned char *rec = mix + (size_t)cell * ptd::mix_cell_bytes(ncomp);
    double *lin = reinterpret_cast<double *>(rec + (size_t)comp * ptd::MIX_LIN_BYTES);
    float *rem = reinterpret_cast<float *>(rec + ptd::mix_lin_bytes(ncomp) + (size_t)comp * ptd::MIX_REM_BYTES);
    lin[0] = keep64 ? __builtin_nan("") : a[0];
    lin[1] = a[1];
    lin[2] = a[4];
    lin[3] = a[16];
    for (int o = 0; o < 64; ++o) rem[o] = (o == 0 || o == 1 || o == 4 || o == 16) ? 0.0f : (float)a[o];
    if (keep64) atomicAdd(n_fp64, 1);
}

} // namespace ptp

/* the seven derivative tables of one component (der[1..7]; der[0] = the samples) */
hipError_t pt_launch_tricubic_derivs(const double *const g[3], int nx, int ny, int nz, double *const der[8],
                                     double *scratch, hipStream_t stream) {
    const int n[3] = {nx, ny, nz};
    /* (source table, axis, output): dx, dy, dz, dxy, dxz, dyz, dxyz as CalcDerivs is called */
    const int plan[7][3] = {{0, 0, 1}, {0, 1, 2}, {0, 2, 3}, {1, 1, 4}, {1, 2, 5}, {2, 2, 6}, {4, 2, 7}};
    for (int s = 0; s < 7; ++s) {
        const int axis = plan[s][1];
        const long long nlines = (long long)nx * ny * nz / n[axis];
        const unsigned blocks = (unsigned)((nlines + 255) / 256);
        hipLaunchKernelGGL(ptp::spline_lines_kernel, dim3(blocks), dim3(256), 0, stream, g[axis], nx, ny, nz, axis,
                           der[plan[s][0]], der[plan[s][2]], scratch);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t pt_launch_tricubic_coef(const double *xg, const double *yg, const double *zg, int nx, int ny, int nz,
                                   const double *const *der_dev, const double *A, int ncomp, int comp, int device_order,
                                   double *coef, hipStream_t stream) {
    const long long ncell = (long long)(nx - 1) * (ny - 1) * (nz - 1);
    if (ncell <= 0) return hipSuccess;
    hipLaunchKernelGGL(ptp::tricubic_coef_kernel, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, stream, xg, yg,
                       zg, nx, ny, nz, der_dev, A, ncomp, comp, device_order, coef);
    return hipGetLastError();
}

hipError_t pt_launch_mixed_table(const double *xg, const double *yg, const double *zg, int nx, int ny, int nz,
                                 int ncomp, const double *coef, double tol, unsigned char *mix, int *n_fp64,
                                 hipStream_t stream) {
    const long long nblk = (long long)(nx - 1) * (ny - 1) * (nz - 1) * ncomp;
    if (nblk <= 0) return hipSuccess;
    hipLaunchKernelGGL(ptp::mixed_table_kernel, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, stream, xg, yg, zg,
                       nx, ny, nz, ncomp, coef, tol, mix, n_fp64);
    return hipGetLastError();
}

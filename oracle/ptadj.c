/*
 * ptadj.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the differentiable RK4
 * path (fixed-step RK4 of the non-relativistic neutron equation of motion,
 * src/particle.cpp:105-135, in B = sum_k s_k B_k plus gravity, specular box walls
 * as fixed maps after each step) and of its discrete adjoint, with every stage
 * stored (no checkpoints).  The GPU kernels (pt_adjoint.hip) must agree with it;
 * central finite differences pin both.
 *
 * Coefficient blocks here use libtricubic's order a[i + 4j + 16k] (x^i y^j z^k).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nmaps, nx, ny, nz;
    const double *g[3];
    const double *coef[8];   /* [(nx-1)(ny-1)(nz-1)][3][64], libtricubic order */
    double s[8];
    double lo[3], hi[3], dt, coupling, grav;
} adj_model;

typedef struct { double f, fx, fy, fz, fxx, fxy, fxz, fyy, fyz, fzz; } d2_t;

static void tric_d2(const double *a, double x, double y, double z, d2_t *o) {
    double q[4], qy[4], qz[4], qyy[4], qyz[4], qzz[4];
    for (int i = 0; i < 4; ++i) {
        double r = 0, ry = 0, rz = 0, ryy = 0, ryz = 0, rzz = 0;
        for (int j = 3; j >= 0; --j) {
            double c0 = a[i + 4 * j], c1 = a[i + 4 * j + 16], c2 = a[i + 4 * j + 32], c3 = a[i + 4 * j + 48];
            double p = ((c3 * z + c2) * z + c1) * z + c0;
            double pz = (3.0 * c3 * z + 2.0 * c2) * z + c1;
            double pzz = 6.0 * c3 * z + 2.0 * c2;
            ryy = ryy * y + 2.0 * ry;
            ry = ry * y + r;
            r = r * y + p;
            ryz = ryz * y + rz;
            rz = rz * y + pz;
            rzz = rzz * y + pzz;
        }
        q[i] = r; qy[i] = ry; qz[i] = rz; qyy[i] = ryy; qyz[i] = ryz; qzz[i] = rzz;
    }
    double F = 0, Fx = 0, Fxx = 0, Fy = 0, Fxy = 0, Fz = 0, Fxz = 0, Fyy = 0, Fyz = 0, Fzz = 0;
    for (int i = 3; i >= 0; --i) {
        Fxx = Fxx * x + 2.0 * Fx;
        Fx = Fx * x + F;
        F = F * x + q[i];
        Fxy = Fxy * x + Fy;
        Fy = Fy * x + qy[i];
        Fxz = Fxz * x + Fz;
        Fz = Fz * x + qz[i];
        Fyy = Fyy * x + qyy[i];
        Fyz = Fyz * x + qyz[i];
        Fzz = Fzz * x + qzz[i];
    }
    o->f = F; o->fx = Fx; o->fy = Fy; o->fz = Fz;
    o->fxx = Fxx; o->fxy = Fxy; o->fxz = Fxz; o->fyy = Fyy; o->fyz = Fyz; o->fzz = Fzz;
}

/* std::upper_bound cell search; 0 outside the grid */
static int locate(const adj_model *M, const double *p, long *cell, double r[3], double inv[3]) {
    const int n[3] = {M->nx, M->ny, M->nz};
    int idx[3];
    for (int a = 0; a < 3; ++a) {
        const double *g = M->g[a];
        int lo = 0, hi = n[a];
        while (lo < hi) { int mid = (lo + hi) / 2; if (g[mid] <= p[a]) lo = mid + 1; else hi = mid; }
        int k = lo - 1;
        if (k < 0 || k > n[a] - 2) return 0;
        idx[a] = k;
        inv[a] = 1.0 / (g[k + 1] - g[k]);
        r[a] = (p[a] - g[k]) * inv[a];
    }
    *cell = ((long)idx[0] * (M->ny - 1) + idx[1]) * (M->nz - 1) + idx[2];
    return 1;
}

typedef struct { double B[3], G[3][3], H[3][6]; int in; long cell; double r[3], inv[3]; } field2;

static void field_total(const adj_model *M, const double *p, field2 *F) {
    memset(F, 0, sizeof(*F));
    F->in = locate(M, p, &F->cell, F->r, F->inv);
    if (!F->in) return;
    double ix = F->inv[0], iy = F->inv[1], iz = F->inv[2];
    for (int k = 0; k < M->nmaps; ++k) {
        double sk = M->s[k];
        for (int c = 0; c < 3; ++c) {
            d2_t d;
            tric_d2(M->coef[k] + ((size_t)F->cell * 3 + c) * 64, F->r[0], F->r[1], F->r[2], &d);
            F->B[c] += sk * d.f;
            F->G[c][0] += sk * (d.fx * ix);
            F->G[c][1] += sk * (d.fy * iy);
            F->G[c][2] += sk * (d.fz * iz);
            F->H[c][0] += sk * (d.fxx * ix * ix);
            F->H[c][1] += sk * (d.fxy * ix * iy);
            F->H[c][2] += sk * (d.fxz * ix * iz);
            F->H[c][3] += sk * (d.fyy * iy * iy);
            F->H[c][4] += sk * (d.fyz * iy * iz);
            F->H[c][5] += sk * (d.fzz * iz * iz);
        }
    }
}

static void grad_absB(const field2 *F, double *absB, double g[3]) {
    *absB = sqrt(F->B[0] * F->B[0] + F->B[1] * F->B[1] + F->B[2] * F->B[2]);
    for (int j = 0; j < 3; ++j)
        g[j] = *absB > 0 ? (F->B[0] * F->G[0][j] + F->B[1] * F->G[1][j] + F->B[2] * F->G[2][j]) / *absB : 0.0;
}

static void rhs(const adj_model *M, const double *y, double *f) {
    field2 F;
    field_total(M, y, &F);
    double a, g[3];
    grad_absB(&F, &a, g);
    f[0] = y[3]; f[1] = y[4]; f[2] = y[5];
    f[3] = M->coupling * g[0];
    f[4] = M->coupling * g[1];
    f[5] = M->coupling * g[2] - M->grav;
}

static void rk4(const adj_model *M, const double *y, double *yn, double *k1, double *k2, double *k3) {
    double k4[6], u[6], h = M->dt, h2 = 0.5 * M->dt;
    rhs(M, y, k1);
    for (int q = 0; q < 6; ++q) u[q] = y[q] + h2 * k1[q];
    rhs(M, u, k2);
    for (int q = 0; q < 6; ++q) u[q] = y[q] + h2 * k2[q];
    rhs(M, u, k3);
    for (int q = 0; q < 6; ++q) u[q] = y[q] + h * k3[q];
    rhs(M, u, k4);
    for (int q = 0; q < 6; ++q) yn[q] = y[q] + (h / 6.0) * (k1[q] + 2.0 * k2[q] + 2.0 * k3[q] + k4[q]);
}

static int reflect(const adj_model *M, double *y) {
    int flags = 0;
    for (int a = 0; a < 3; ++a) {
        if (y[a] < M->lo[a]) { y[a] = 2.0 * M->lo[a] - y[a]; y[3 + a] = -y[3 + a]; flags |= 1 << a; }
        else if (y[a] > M->hi[a]) { y[a] = 2.0 * M->hi[a] - y[a]; y[3 + a] = -y[3 + a]; flags |= 1 << (a + 3); }
    }
    return flags;
}

static void vjp(const adj_model *M, const double *u, const double *w, double *ub, double *sb) {
    field2 F;
    field_total(M, u, &F);
    double absB, g[3];
    grad_absB(&F, &absB, g);
    ub[3] = w[0]; ub[4] = w[1]; ub[5] = w[2];
    ub[0] = ub[1] = ub[2] = 0;
    if (!F.in || !(absB > 0)) return;
    double wp[3] = {M->coupling * w[3], M->coupling * w[4], M->coupling * w[5]};
    double inv = 1.0 / absB;
    static const int hidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int j = 0; j < 3; ++j) {
        double acc = 0;
        for (int k = 0; k < 3; ++k) {
            double h = -g[j] * g[k];
            for (int c = 0; c < 3; ++c) h += F.G[c][j] * F.G[c][k] + F.B[c] * F.H[c][hidx[j][k]];
            acc += h * inv * wp[k];
        }
        ub[j] = acc;
    }
    double b[3] = {F.B[0] * inv, F.B[1] * inv, F.B[2] * inv};
    double wg = wp[0] * g[0] + wp[1] * g[1] + wp[2] * g[2];
    double alpha[3], beta[3][3];
    for (int c = 0; c < 3; ++c) {
        alpha[c] = (wp[0] * F.G[c][0] + wp[1] * F.G[c][1] + wp[2] * F.G[c][2] - wg * b[c]) * inv;
        for (int j = 0; j < 3; ++j) beta[c][j] = wp[j] * b[c];
    }
    for (int k = 0; k < M->nmaps; ++k) {
        double acc = 0;
        for (int c = 0; c < 3; ++c) {
            d2_t d;
            tric_d2(M->coef[k] + ((size_t)F.cell * 3 + c) * 64, F.r[0], F.r[1], F.r[2], &d);
            acc += alpha[c] * d.f + beta[c][0] * (d.fx * F.inv[0]) + beta[c][1] * (d.fy * F.inv[1]) +
                   beta[c][2] * (d.fz * F.inv[2]);
        }
        sb[k] += acc;
    }
}

/* forward of n particles (AoS y[n][6]) */
int ptadj_forward(const adj_model *M, int n, const double *y0, double *yT, int nsteps) {
    for (int i = 0; i < n; ++i) {
        double y[6], yn[6], k1[6], k2[6], k3[6];
        memcpy(y, y0 + 6 * (size_t)i, sizeof(y));
        for (int s = 0; s < nsteps; ++s) {
            rk4(M, y, yn, k1, k2, k3);
            reflect(M, yn);
            memcpy(y, yn, sizeof(y));
        }
        memcpy(yT + 6 * (size_t)i, y, sizeof(y));
    }
    return 0;
}

/* adjoint: lam [n][6] = dJ/dyT on entry, dJ/dy0 on exit; sbar [n][nmaps] */
int ptadj_adjoint(const adj_model *M, int n, const double *y0, double *lam, double *sbar, int nsteps) {
    double *st = (double *)malloc(sizeof(double) * 24 * (size_t)(nsteps > 0 ? nsteps : 1));
    int *fl = (int *)malloc(sizeof(int) * (size_t)(nsteps > 0 ? nsteps : 1));
    if (!st || !fl) { free(st); free(fl); return -1; }
    const double h = M->dt, h2 = 0.5 * M->dt;
    for (int i = 0; i < n; ++i) {
        double y[6], yn[6];
        memcpy(y, y0 + 6 * (size_t)i, sizeof(y));
        for (int s = 0; s < nsteps; ++s) {
            double *rec = st + 24 * (size_t)s;
            rk4(M, y, yn, rec + 6, rec + 12, rec + 18);
            memcpy(rec, y, sizeof(y));
            fl[s] = reflect(M, yn);
            memcpy(y, yn, sizeof(y));
        }
        double l[6], sb[8] = {0};
        memcpy(l, lam + 6 * (size_t)i, sizeof(l));
        for (int s = nsteps - 1; s >= 0; --s) {
            for (int a = 0; a < 3; ++a)
                if (fl[s] & ((1 << a) | (1 << (a + 3)))) { l[a] = -l[a]; l[3 + a] = -l[3 + a]; }
            const double *yj = st + 24 * (size_t)s, *k1 = yj + 6, *k2 = yj + 12, *k3 = yj + 18;
            double a1[6], a2[6], a3[6], a4[6], ub[6], ly[6], u[6];
            for (int q = 0; q < 6; ++q) {
                a4[q] = (h / 6.0) * l[q];
                a3[q] = (h / 6.0) * 2.0 * l[q];
                a2[q] = (h / 6.0) * 2.0 * l[q];
                a1[q] = (h / 6.0) * l[q];
                ly[q] = l[q];
            }
            for (int q = 0; q < 6; ++q) u[q] = yj[q] + h * k3[q];
            vjp(M, u, a4, ub, sb);
            for (int q = 0; q < 6; ++q) { ly[q] += ub[q]; a3[q] += h * ub[q]; }
            for (int q = 0; q < 6; ++q) u[q] = yj[q] + h2 * k2[q];
            vjp(M, u, a3, ub, sb);
            for (int q = 0; q < 6; ++q) { ly[q] += ub[q]; a2[q] += h2 * ub[q]; }
            for (int q = 0; q < 6; ++q) u[q] = yj[q] + h2 * k1[q];
            vjp(M, u, a2, ub, sb);
            for (int q = 0; q < 6; ++q) { ly[q] += ub[q]; a1[q] += h2 * ub[q]; }
            vjp(M, yj, a1, ub, sb);
            for (int q = 0; q < 6; ++q) l[q] = ly[q] + ub[q];
        }
        memcpy(lam + 6 * (size_t)i, l, sizeof(l));
        for (int k = 0; k < M->nmaps; ++k) sbar[(size_t)i * M->nmaps + k] = sb[k];
    }
    free(st);
    free(fl);
    return 0;
}

/*
 * pt_adjoint.hip -- differentiable tracking of the smooth inter-collision motion:
 * fixed-step RK4 for neutrons in B = sum_k s_k B_k (tricubic maps on one grid) plus
 * gravity, specular box walls applied as fixed maps after each step, and its
 * discrete adjoint with respect to the initial state and the map scales s_k.
 *
 * The reference integrates with adaptive dopri5 and has no derivatives; this path
 * follows its neutron equation of motion (TParticle::EquationOfMotion,
 * src/particle.cpp:105-135, non-relativistic limit: dv/dt = -g z + pol mu grad|B| / m)
 * and its tricubic interpolation (TabField3::Interpolate, src/field_3d.cpp:487-513)
 * with second derivatives added for the adjoint.
 *
 * Forward: one thread per particle, state checkpointed every K steps.
 * Adjoint: per checkpoint segment (last to first) the forward steps are recomputed
 * from the checkpoint, the stage data of each step kept in HBM, then the steps are
 * reversed: lambda <- (d y_{j+1} / d y_j)^T lambda and sbar += (d y_{j+1} / d s)^T lambda.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "pt_adjoint.h"

namespace pta {

#define PTA __device__ __forceinline__

/* ---- tricubic block evaluation with first and second derivatives --------- */
/* block layout a[16 i + 4 j + k] = coefficient of x^i y^j z^k (see pt_host.cpp) */
struct D2 { double f, fx, fy, fz, fxx, fxy, fxz, fyy, fyz, fzz; };

/* one row j (coefficients a0..a3 of x^i y^j z^0..3) of the z- and y-Horner sums of an
   x-power line, with their y- and z-derivatives */
PTA void d2_row(double a0, double a1, double a2, double a3, double y, double z, double &r, double &ry, double &rz,
                double &ryy, double &ryz, double &rzz) {
    /* z-Horner: p, p', p'' */
    const double p = ((a3 * z + a2) * z + a1) * z + a0;
    const double pz = (3.0 * a3 * z + 2.0 * a2) * z + a1;
    const double pzz = 6.0 * a3 * z + 2.0 * a2;
    /* y-Horner with y-derivatives (second derivative via the product rule) */
    ryy = ryy * y + 2.0 * ry;
    ry = ry * y + r;
    r = r * y + p;
    ryz = ryz * y + rz;
    rz = rz * y + pz;
    rzz = rzz * y + pzz;
}

/* x-Horner of every quantity over the four lines' sums, with x-derivatives */
PTA void d2_finish(const double q[4], const double qy[4], const double qz[4], const double qyy[4], const double qyz[4],
                   const double qzz[4], double x, bool hess, D2 &o) {
    double F = 0, Fx = 0, Fxx = 0, Fy = 0, Fxy = 0, Fz = 0, Fxz = 0, Fyy = 0, Fyz = 0, Fzz = 0;
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        Fxx = Fxx * x + 2.0 * Fx;
        Fx = Fx * x + F;
        F = F * x + q[i];
        Fxy = Fxy * x + Fy;
        Fy = Fy * x + qy[i];
        Fxz = Fxz * x + Fz;
        Fz = Fz * x + qz[i];
        if (hess) {
            Fyy = Fyy * x + qyy[i];
            Fyz = Fyz * x + qyz[i];
            Fzz = Fzz * x + qzz[i];
        }
    }
    o.f = F; o.fx = Fx; o.fy = Fy; o.fz = Fz;
    o.fxx = Fxx; o.fxy = Fxy; o.fxz = Fxz; o.fyy = Fyy; o.fyz = Fyz; o.fzz = Fzz;
}

PTA void tric_d2(const double *__restrict__ a, double x, double y, double z, bool hess, D2 &o) {
    /* q*: y-Horner results per power of x */
    double q[4], qy[4], qz[4], qyy[4], qyz[4], qzz[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double r = 0, ry = 0, rz = 0, ryy = 0, ryz = 0, rzz = 0;
#pragma unroll
        for (int j = 3; j >= 0; --j) {
            const double *c = a + 16 * i + 4 * j;
            d2_row(c[0], c[1], c[2], c[3], y, z, r, ry, rz, ryy, ryz, rzz);
        }
        q[i] = r; qy[i] = ry; qz[i] = rz; qyy[i] = ryy; qyz[i] = ryz; qzz[i] = rzz;
    }
    d2_finish(q, qy, qz, qyy, qyz, qzz, x, hess, o);
}

/*
 * Cooperative gathers, as the tracking kernels' trial steps do them: the active lanes of a
 * wave (n, ranks r) load each other's coefficients in units of one half x-power line (64 B,
 * two rows of four coefficients) of every lane's block, the lane of rank r loading, for
 * k = 0..3, chunk v & 3 of the half line of rank v >> 2 (v = k n + r), and the chunks reach
 * their owners through a 4 KB LDS slice per wave (chunk q of rank p at 1024 q + 16 (p ^ 2q)).
 * Every lane then runs the rows of its own block in tric_d2's order, so the results are
 * bit-identical.  All maps share the grid, so a lane's block offset is the same in each.
 */
constexpr int ACOOP_WAVES = 4;   /* waves of a 256-thread block */
__shared__ double2 acoop_lds[ACOOP_WAVES][4][64];

struct CoopIdx { uint32_t src[4], wb[4]; };

PTA uint32_t active_rank() {
    uint32_t r;
    asm volatile("v_mbcnt_lo_u32_b32 %0, exec_lo, 0\n\tv_mbcnt_hi_u32_b32 %0, exec_hi, %0" : "=v"(r));
    return r;
}

#define ACOOP_FENCE() asm volatile("" ::: "memory")

/* this lane's four (rank, chunk) load indices (16-byte units within a map) and LDS write offsets */
PTA void coop_setup(long cell, CoopIdx &ci) {
    const uint32_t n = (uint32_t)__builtin_popcountll(__builtin_amdgcn_read_exec());
    const uint32_t rank = active_rank();
    char *S = reinterpret_cast<char *>(acoop_lds[threadIdx.x >> 6]);
    reinterpret_cast<uint32_t *>(S)[rank] = (uint32_t)cell * 96u;
    ACOOP_FENCE();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t v = (uint32_t)k * n + rank, o = v >> 2, c = v & 3;
        ci.src[k] = reinterpret_cast<const uint32_t *>(S)[o] + c;
        ci.wb[k] = (c << 10) | ((o ^ (2 * c)) << 4);
    }
    ACOOP_FENCE();
}

struct AChunks { double2 a, b, c, d; };
PTA AChunks acoop_load(const double2 *__restrict__ t, const uint32_t *src, uint32_t off) {
    AChunks u;
    u.a = t[src[0] + off]; u.b = t[src[1] + off]; u.c = t[src[2] + off]; u.d = t[src[3] + off];
    return u;
}

/* tric_d2 of the three components of every active lane's cell in one map, gathered
   cooperatively: unit u = 8 c + 2 i + (1 - h) (component c, x-power i, half h) at 16-byte
   offset 32 c + 8 i + 4 h of the lane's cell; two units' loads in flight */
PTA void tric_d2_map_coop(const double *__restrict__ coef, const CoopIdx &ci, double x, double y, double z, bool hess,
                          D2 out[3]) {
    const double2 *t2 = reinterpret_cast<const double2 *>(coef);
    char *S = reinterpret_cast<char *>(acoop_lds[threadIdx.x >> 6]);
    auto off = [](int u) -> uint32_t { return (uint32_t)(32 * (u / 8) + 8 * ((u % 8) / 2) + 4 * (1 - (u & 1))); };
    AChunks buf[3];
    buf[0] = acoop_load(t2, ci.src, off(0));
    buf[1] = acoop_load(t2, ci.src, off(1));
    double q[4], qy[4], qz[4], qyy[4], qyz[4], qzz[4];
    double r = 0, ry = 0, rz = 0, ryy = 0, ryz = 0, rzz = 0;
#pragma unroll
    for (int u = 0; u < 24; ++u) {
        const int c = u / 8, i = (u % 8) / 2, h = 1 - (u & 1);
        if (u + 2 < 24) buf[(u + 2) % 3] = acoop_load(t2, ci.src, off(u + 2));
        const AChunks &vc = buf[u % 3];
        *reinterpret_cast<double2 *>(S + ci.wb[0]) = vc.a;
        *reinterpret_cast<double2 *>(S + ci.wb[1]) = vc.b;
        *reinterpret_cast<double2 *>(S + ci.wb[2]) = vc.c;
        *reinterpret_cast<double2 *>(S + ci.wb[3]) = vc.d;
        ACOOP_FENCE();
        const uint32_t rb = active_rank() << 4;
        double2 cc[4];
#pragma unroll
        for (int qq = 0; qq < 4; ++qq)
            cc[qq] = *reinterpret_cast<const double2 *>(S + ((rb ^ (uint32_t)(32 * qq)) + 1024u * qq));
        ACOOP_FENCE();
        if (h == 1) { r = 0; ry = 0; rz = 0; ryy = 0; ryz = 0; rzz = 0; }
        /* this half's rows: 2h + 1 (chunks 2, 3) then 2h (chunks 0, 1) */
        d2_row(cc[2].x, cc[2].y, cc[3].x, cc[3].y, y, z, r, ry, rz, ryy, ryz, rzz);
        d2_row(cc[0].x, cc[0].y, cc[1].x, cc[1].y, y, z, r, ry, rz, ryy, ryz, rzz);
        if (h == 0) {
            q[i] = r; qy[i] = ry; qz[i] = rz; qyy[i] = ryy; qyz[i] = ryz; qzz[i] = rzz;
            if (i == 3) d2_finish(q, qy, qz, qyy, qyz, qzz, x, hess, out[c]);
        }
    }
}

/* the map set's table is addressable in 16-byte units with 32-bit indices */
PTA bool coop_fits(const MapSet &M) {
    return (uint64_t)(M.nx - 1) * (uint64_t)(M.ny - 1) * (uint64_t)(M.nz - 1) * 96ull < (1ull << 32);
}

/* cell search (std::upper_bound semantics of TabField3::Interpolate) */
PTA bool locate(const MapSet &M, const double *p, long &cell, double r[3], double inv[3]) {
    const int n[3] = {M.nx, M.ny, M.nz};
    int idx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double *g = M.g[a];
        double gs = (p[a] - M.x0[a]) * M.inv_h[a];
        int k;
        if (!(gs >= 0.0)) k = 0;
        else if (gs >= (double)(n[a] - 2)) k = n[a] - 2;
        else k = (int)gs;
        double gk = g[k], gk1 = g[k + 1];
        while (k > 0 && gk > p[a]) { --k; gk1 = gk; gk = g[k]; }
        while (k < n[a] - 2 && gk1 <= p[a]) { ++k; gk = gk1; gk1 = g[k + 1]; }
        if (!(gk <= p[a] && p[a] < gk1)) return false;
        idx[a] = k;
        inv[a] = 1.0 / (gk1 - gk);
        r[a] = (p[a] - gk) * inv[a];
    }
    cell = ((long)idx[0] * (M.ny - 1) + idx[1]) * (M.nz - 1) + idx[2];
    return true;
}

/* Total field B = sum_k s_k B_k: components, gradient G[c][j] = dB_c/dx_j and, with
   hess, second derivatives H[c][jk] (jk = xx xy xz yy yz zz) */
struct Field2 { double B[3], G[3][3], H[3][6]; bool in; long cell; double r[3], inv[3]; };

PTA void field_total(const MapSet &M, const double *p, bool hess, Field2 &F) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        F.B[c] = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) F.G[c][j] = 0;
#pragma unroll
        for (int j = 0; j < 6; ++j) F.H[c][j] = 0;
    }
    F.in = locate(M, p, F.cell, F.r, F.inv);
    if (!F.in) return;
    const double ix = F.inv[0], iy = F.inv[1], iz = F.inv[2];
    const bool coop = coop_fits(M);
    CoopIdx ci;
    if (coop) coop_setup(F.cell, ci);
#pragma unroll 1
    for (int k = 0; k < M.nmaps; ++k) {
        const double sk = M.s[k];
        D2 dm[3];
        if (coop) tric_d2_map_coop(M.coef[k], ci, F.r[0], F.r[1], F.r[2], hess, dm);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            D2 d;
            if (coop) d = dm[c];
            else tric_d2(M.coef[k] + ((size_t)F.cell * 3 + c) * 64, F.r[0], F.r[1], F.r[2], hess, d);
            F.B[c] += sk * d.f;
            F.G[c][0] += sk * (d.fx * ix);
            F.G[c][1] += sk * (d.fy * iy);
            F.G[c][2] += sk * (d.fz * iz);
            if (hess) {
                F.H[c][0] += sk * (d.fxx * ix * ix);
                F.H[c][1] += sk * (d.fxy * ix * iy);
                F.H[c][2] += sk * (d.fxz * ix * iz);
                F.H[c][3] += sk * (d.fyy * iy * iy);
                F.H[c][4] += sk * (d.fyz * iy * iz);
                F.H[c][5] += sk * (d.fzz * iz * iz);
            }
        }
    }
}

/* grad|B| from the total field (zero where |B| = 0) */
PTA void grad_absB(const Field2 &F, double &absB, double gB[3]) {
    absB = __builtin_sqrt(F.B[0] * F.B[0] + F.B[1] * F.B[1] + F.B[2] * F.B[2]);
#pragma unroll
    for (int j = 0; j < 3; ++j)
        gB[j] = absB > 0 ? (F.B[0] * F.G[0][j] + F.B[1] * F.G[1][j] + F.B[2] * F.G[2][j]) / absB : 0.0;
}

/* right-hand side f(y) = (v, -g z + coupling * grad|B|) */
PTA void rhs(const AdjParams &A, const double *y, double *f) {
    Field2 F;
    field_total(A.maps, y, false, F);
    double absB, gB[3];
    grad_absB(F, absB, gB);
    f[0] = y[3]; f[1] = y[4]; f[2] = y[5];
    f[3] = A.coupling * gB[0];
    f[4] = A.coupling * gB[1];
    f[5] = A.coupling * gB[2] - A.grav;
}

/* one RK4 step; stage slopes returned for the adjoint */
PTA void rk4(const AdjParams &A, const double *y, double *yn, double k1[6], double k2[6], double k3[6]) {
    double k4[6], u[6];
    const double h = A.dt, h2 = 0.5 * A.dt;
    rhs(A, y, k1);
#pragma unroll
    for (int q = 0; q < 6; ++q) u[q] = y[q] + h2 * k1[q];
    rhs(A, u, k2);
#pragma unroll
    for (int q = 0; q < 6; ++q) u[q] = y[q] + h2 * k2[q];
    rhs(A, u, k3);
#pragma unroll
    for (int q = 0; q < 6; ++q) u[q] = y[q] + h * k3[q];
    rhs(A, u, k4);
#pragma unroll
    for (int q = 0; q < 6; ++q) yn[q] = y[q] + (h / 6.0) * (k1[q] + 2.0 * k2[q] + 2.0 * k3[q] + k4[q]);
}

/* specular box walls as a fixed map after each step: bit a (low wall) / a + 3 (high wall) */
PTA int reflect(const AdjParams &A, double *y) {
    int flags = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (y[a] < A.lo[a]) { y[a] = 2.0 * A.lo[a] - y[a]; y[3 + a] = -y[3 + a]; flags |= 1 << a; }
        else if (y[a] > A.hi[a]) { y[a] = 2.0 * A.hi[a] - y[a]; y[3 + a] = -y[3 + a]; flags |= 1 << (a + 3); }
    }
    return flags;
}

/*
 * Vector-Jacobian product of f at u: given w = (wx, wv) (adjoint of the slope),
 * returns ubar = (df/du)^T w and adds (df/ds)^T w to sbar.
 *   df_x/dv = I                       -> ubar_v += wx
 *   df_v/dx = coupling * Hess|B|       -> ubar_x += coupling * Hess|B| wv
 *   df_v/ds_k = coupling * d(grad|B|)/ds_k
 * With b = B/|B| and g = grad|B|: Hess|B|_jk = (sum_c G_cj G_ck + B_c H_c,jk - g_j g_k)/|B|,
 * d g_j/ds_k = (sum_c B^k_c G_cj + B_c G^k_cj - g_j b.B^k)/|B| for map k's field B^k, G^k,
 * so sbar_k += sum_c alpha_c B^k_c + sum_cj beta_cj G^k_cj with
 * alpha_c = (sum_j w'_j G_cj - (w'.g) b_c)/|B|, beta_cj = w'_j b_c, w' = coupling wv.
 */
PTA void vjp(const AdjParams &A, const double *u, const double *w, double *ubar, double *sbar) {
    Field2 F;
    field_total(A.maps, u, true, F);
    double absB, g[3];
    grad_absB(F, absB, g);
    ubar[3] = w[0]; ubar[4] = w[1]; ubar[5] = w[2];
    ubar[0] = 0; ubar[1] = 0; ubar[2] = 0;
    if (!F.in || !(absB > 0)) return;
    const double wp[3] = {A.coupling * w[3], A.coupling * w[4], A.coupling * w[5]};
    const double inv = 1.0 / absB;
    /* Hess|B| wp */
    const int hidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double acc = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double h = -g[j] * g[k];
#pragma unroll
            for (int c = 0; c < 3; ++c) h += F.G[c][j] * F.G[c][k] + F.B[c] * F.H[c][hidx[j][k]];
            acc += h * inv * wp[k];
        }
        ubar[j] = acc;
    }
    /* parameter part: second pass over the maps with value and gradient only */
    double b[3] = {F.B[0] * inv, F.B[1] * inv, F.B[2] * inv};
    double wg = wp[0] * g[0] + wp[1] * g[1] + wp[2] * g[2];
    double alpha[3], beta[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        alpha[c] = (wp[0] * F.G[c][0] + wp[1] * F.G[c][1] + wp[2] * F.G[c][2] - wg * b[c]) * inv;
#pragma unroll
        for (int j = 0; j < 3; ++j) beta[c][j] = wp[j] * b[c];
    }
    const bool coop = coop_fits(A.maps);
    CoopIdx ci;
    if (coop) coop_setup(F.cell, ci);
#pragma unroll 1
    for (int k = 0; k < A.maps.nmaps; ++k) {
        double acc = 0;
        D2 dm[3];
        if (coop) tric_d2_map_coop(A.maps.coef[k], ci, F.r[0], F.r[1], F.r[2], false, dm);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            D2 d;
            if (coop) d = dm[c];
            else tric_d2(A.maps.coef[k] + ((size_t)F.cell * 3 + c) * 64, F.r[0], F.r[1], F.r[2], false, d);
            acc += alpha[c] * d.f + beta[c][0] * (d.fx * F.inv[0]) + beta[c][1] * (d.fy * F.inv[1]) +
                   beta[c][2] * (d.fz * F.inv[2]);
        }
        sbar[k] += acc;
    }
}

/* ---- kernels ---------------------------------------------------------------- */

/* forward: y (SoA [6][n]) advanced nsteps; checkpoint c = state before step c*K ([nc][6][n]) */
__global__ void __launch_bounds__(256) forward_kernel(AdjParams A, const double *y0, double *yT, double *ckpt,
                                                      int n, int nsteps, int K) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double y[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) y[q] = y0[(size_t)q * n + i];
    for (int s = 0; s < nsteps; ++s) {
        if (ckpt && s % K == 0) {
            size_t c = (size_t)(s / K);
#pragma unroll
            for (int q = 0; q < 6; ++q) ckpt[(c * 6 + q) * n + i] = y[q];
        }
        double yn[6], k1[6], k2[6], k3[6];
        rk4(A, y, yn, k1, k2, k3);
        reflect(A, yn);
#pragma unroll
        for (int q = 0; q < 6; ++q) y[q] = yn[q];
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) yT[(size_t)q * n + i] = y[q];
}

/*
 * adjoint: lam (SoA [6][n]) holds dJ/dy_T on entry and dJ/dy_0 on exit; sbar ([nmaps][n])
 * receives each particle's dJ/ds_k.  seg: [K][24][n] stage data, flags [K][n].
 */
__global__ void __launch_bounds__(256) adjoint_kernel(AdjParams A, const double *ckpt, double *lam, double *sbar,
                                                      double *seg, int *flags, int n, int nsteps, int K) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double l[6], sb[PTA_MAX_MAPS];
#pragma unroll
    for (int q = 0; q < 6; ++q) l[q] = lam[(size_t)q * n + i];
    for (int k = 0; k < PTA_MAX_MAPS; ++k) sb[k] = 0;
    const int nc = (nsteps + K - 1) / K;
    const double h = A.dt, h2 = 0.5 * A.dt;
    for (int c = nc - 1; c >= 0; --c) {
        const int s0 = c * K, s1 = (s0 + K < nsteps) ? s0 + K : nsteps;
        /* recompute the segment, keeping y_j and k1..k3 of each step */
        double y[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) y[q] = ckpt[((size_t)c * 6 + q) * n + i];
        for (int s = s0; s < s1; ++s) {
            const size_t j = (size_t)(s - s0);
            double yn[6], k1[6], k2[6], k3[6];
            rk4(A, y, yn, k1, k2, k3);
            int f = reflect(A, yn);
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                seg[(j * 24 + q) * n + i] = y[q];
                seg[(j * 24 + 6 + q) * n + i] = k1[q];
                seg[(j * 24 + 12 + q) * n + i] = k2[q];
                seg[(j * 24 + 18 + q) * n + i] = k3[q];
            }
            flags[j * n + i] = f;
#pragma unroll
            for (int q = 0; q < 6; ++q) y[q] = yn[q];
        }
        /* reverse the steps */
        for (int s = s1 - 1; s >= s0; --s) {
            const size_t j = (size_t)(s - s0);
            /* the wall map after the step: x -> 2w - x, v -> -v on each flagged axis */
            int f = flags[j * n + i];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if (f & ((1 << a) | (1 << (a + 3)))) { l[a] = -l[a]; l[3 + a] = -l[3 + a]; }
            double yj[6], k1[6], k2[6], k3[6], u[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                yj[q] = seg[(j * 24 + q) * n + i];
                k1[q] = seg[(j * 24 + 6 + q) * n + i];
                k2[q] = seg[(j * 24 + 12 + q) * n + i];
                k3[q] = seg[(j * 24 + 18 + q) * n + i];
            }
            /* adjoints of the slopes from y_{j+1} = y + h/6 (k1 + 2k2 + 2k3 + k4) */
            double a1[6], a2[6], a3[6], a4[6], ub[6], ly[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                a4[q] = (h / 6.0) * l[q];
                a3[q] = (h / 6.0) * 2.0 * l[q];
                a2[q] = (h / 6.0) * 2.0 * l[q];
                a1[q] = (h / 6.0) * l[q];
                ly[q] = l[q];
            }
            /* k4 = f(y + h k3) */
#pragma unroll
            for (int q = 0; q < 6; ++q) u[q] = yj[q] + h * k3[q];
            vjp(A, u, a4, ub, sb);
#pragma unroll
            for (int q = 0; q < 6; ++q) { ly[q] += ub[q]; a3[q] += h * ub[q]; }
            /* k3 = f(y + h/2 k2) */
#pragma unroll
            for (int q = 0; q < 6; ++q) u[q] = yj[q] + h2 * k2[q];
            vjp(A, u, a3, ub, sb);
#pragma unroll
            for (int q = 0; q < 6; ++q) { ly[q] += ub[q]; a2[q] += h2 * ub[q]; }
            /* k2 = f(y + h/2 k1) */
#pragma unroll
            for (int q = 0; q < 6; ++q) u[q] = yj[q] + h2 * k1[q];
            vjp(A, u, a2, ub, sb);
#pragma unroll
            for (int q = 0; q < 6; ++q) { ly[q] += ub[q]; a1[q] += h2 * ub[q]; }
            /* k1 = f(y) */
            vjp(A, yj, a1, ub, sb);
#pragma unroll
            for (int q = 0; q < 6; ++q) l[q] = ly[q] + ub[q];
        }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) lam[(size_t)q * n + i] = l[q];
    for (int k = 0; k < A.maps.nmaps; ++k) sbar[(size_t)k * n + i] = sb[k];
}

} // namespace pta

/* ---- C ABI ------------------------------------------------------------------ */

extern "C" int pt_rk4_forward_ckpt(void *stream, const pt_adj_params *params, int n, const double *y0,
                                   double *yT, double *ckpt, int nsteps, int K) {
    if (!params || n < 0 || nsteps < 0 || K < 1 || (n > 0 && (!y0 || !yT))) return PT_EINVAL;
    if (params->maps.nmaps < 0 || params->maps.nmaps > PTA_MAX_MAPS) return PT_EINVAL;
    if (n == 0) return PT_OK;
    hipLaunchKernelGGL(pta::forward_kernel, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       *params, y0, yT, ckpt, n, nsteps, K);
    return hipGetLastError() == hipSuccess ? PT_OK : PT_EDEVICE;
}

extern "C" int pt_rk4_adjoint(void *stream, const pt_adj_params *params, int n, const double *ckpt, double *lam,
                              double *sbar, double *seg, int *flags, int nsteps, int K) {
    if (!params || n < 0 || nsteps < 0 || K < 1 || (n > 0 && (!ckpt || !lam || !sbar || !seg || !flags)))
        return PT_EINVAL;
    if (params->maps.nmaps < 0 || params->maps.nmaps > PTA_MAX_MAPS) return PT_EINVAL;
    if (n == 0) return PT_OK;
    hipLaunchKernelGGL(pta::adjoint_kernel, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       *params, ckpt, lam, sbar, seg, flags, n, nsteps, K);
    return hipGetLastError() == hipSuccess ? PT_OK : PT_EDEVICE;
}

/* ---- the adjoint without allocation ---------------------------------------- */
/*
 * pt_rk4_adjoint_noalloc keeps the workspace pt_rk4_adjoint takes from its caller (the
 * stage data of one checkpoint segment, [rows][24][n] doubles, and the wall flags,
 * [rows][n] ints, rows = min(K, nsteps)) in this file-scope device array instead.
 * Every GPU the library runs on holds its own copy until the process ends.  In host code
 * the name `scratch` is not a device address: the address of a device's copy comes from
 * hipGetSymbolAddress with that device current, once per device, and only that address
 * is handed to kernels.  The adjoint kernel writes every workspace entry of a segment
 * (forward recomputation) before it reads it (reverse sweep), so nothing needs clearing.
 */
static __device__ __align__(256) unsigned char scratch[PT_ADJ_NOALLOC_MAX_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

/* device address of `scratch` on device dev (set-up on the first call for that device) */
int scratch_address(int dev, void **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return pt_fail(PT_EDEVICE, "pt_rk4_adjoint_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return pt_fail(PT_EDEVICE, (std::string("pt_rk4_adjoint_noalloc: hipGetSymbolAddress: ") +
                                        hipGetErrorString(e)).c_str());
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return PT_OK;
}
} // namespace

extern "C" int pt_rk4_adjoint_noalloc(void *stream, const pt_adj_params *params, int n, const double *ckpt,
                                      double *lam, double *sbar, int nsteps, int K) {
    if (!params || n < 0 || nsteps < 0 || K < 1 || (n > 0 && (!ckpt || !lam || !sbar))) return PT_EINVAL;
    if (params->maps.nmaps < 0 || params->maps.nmaps > PTA_MAX_MAPS) return PT_EINVAL;
    if (n == 0) return PT_OK;
    const size_t rows = (size_t)(nsteps < K ? nsteps : K);
    const size_t seg_bytes = rows * 24 * (size_t)n * sizeof(double);
    const size_t flag_off = (seg_bytes + 255) / 256 * 256;
    const size_t need = flag_off + rows * (size_t)n * sizeof(int);
    if (need > (size_t)PT_ADJ_NOALLOC_MAX_BYTES)
        return pt_fail(PT_EINVAL, ("pt_rk4_adjoint_noalloc: workspace of " + std::to_string(need) +
                                   " bytes (min(K, nsteps) x n x 196) exceeds PT_ADJ_NOALLOC_MAX_BYTES = " +
                                   std::to_string((size_t)PT_ADJ_NOALLOC_MAX_BYTES)).c_str());
    int dev = -1;
    hipError_t e = stream ? hipStreamGetDevice(static_cast<hipStream_t>(stream), &dev) : hipGetDevice(&dev);
    if (e != hipSuccess) return pt_fail(PT_EDEVICE, hipGetErrorString(e));
    void *base = nullptr;
    int rc = scratch_address(dev, &base);
    if (rc != PT_OK) return rc;
    double *seg = static_cast<double *>(base);
    int *flags = reinterpret_cast<int *>(static_cast<unsigned char *>(base) + flag_off);
    hipLaunchKernelGGL(pta::adjoint_kernel, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                       *params, ckpt, lam, sbar, seg, flags, n, nsteps, K);
    e = hipGetLastError();
    return e == hipSuccess ? PT_OK : pt_fail(PT_EDEVICE, hipGetErrorString(e));
}

/*
 * pt_track.hip -- the tracking kernels of the MI355X engine.
 *
 * One thread per particle runs TTracker::IntegrateParticle's loop
 * (src/tracking.cpp:17-173) on resumable state kept in HBM: each launch
 * advances every live particle by up to max_steps accepted dopri5 steps.
 * Built with -ffp-contract=off: the FP64 path evaluates every expression in
 * the reference's order, with the reference's long double constants rounded
 * to double.
 */
#include "pt_device.h"
#include "pt_kernels.h"

namespace ptd {

/* ------------------------------------------------------------------------ */
/* fields                                                                    */
/* ------------------------------------------------------------------------ */

/* TabField3::Interpolate cell search (std::upper_bound semantics, src/field_3d.cpp:490-500) */
PTD bool tric_locate(const Field &f, double x, double y, double z, long &cell, double r[3], double dist[3]) {
    double p[3] = {x, y, z};
    int n[3] = {f.nx, f.ny, f.nz};
    int idx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double *g = f.g[a];
        double gs = (p[a] - f.x0[a]) * f.inv_h[a];
        int k;
        if (!(gs >= 0.0)) k = 0;
        else if (gs >= (double)(n[a] - 2)) k = n[a] - 2;
        else k = (int)gs;
        double gk = g[k], gk1 = g[k + 1];
        while (k > 0 && gk > p[a]) { --k; gk1 = gk; gk = g[k]; }
        while (k < n[a] - 2 && gk1 <= p[a]) { ++k; gk = gk1; gk1 = g[k + 1]; }
        if (!(gk <= p[a] && p[a] < gk1)) return false;
        idx[a] = k;
        dist[a] = gk1 - gk;
        r[a] = (p[a] - gk) / dist[a];
    }
    cell = ((long)idx[0] * (f.ny - 1) + idx[1]) * (f.nz - 1) + idx[2];
    return true;
}

/*
 * Value and gradient of one tricubic block: the four Horner evaluations of
 * tricubic_eval_fast (src/field_3d.cpp:36-64, 82-121) fused over one pass of the
 * coefficients.  Multiplications by fact[..][0] = 1 are exact, so sharing the
 * inner z- and y-Horner partial sums between value and derivatives leaves every
 * result bit-identical to the four separate evaluations.
 */
PTD void tric_eval_grad64(const double *__restrict__ a, double x, double y, double z, bool grad,
                          double &F, double &Fx, double &Fy, double &Fz) {
    double ry[4], ryd[4], ryz[4];
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        double rY = 0., rYd = 0., rYz = 0.;
#pragma unroll
        for (int j = 3; j >= 0; --j) {
            const double2 *c2 = reinterpret_cast<const double2 *>(a + i + 4 * j);
            double a0 = a[i + 4 * j], a1 = a[i + 4 * j + 16], a2 = a[i + 4 * j + 32], a3 = a[i + 4 * j + 48];
            (void)c2;
            double rz = 0.;
            rz += a3; rz *= z;
            rz += a2; rz *= z;
            rz += a1; rz *= z;
            rz += a0;
            rY += rz;
            if (grad) {
                double rzd = 0.;
                rzd += 3.0 * a3; rzd *= z;
                rzd += 2.0 * a2; rzd *= z;
                rzd += a1;
                rYz += rzd;
                if (j > 0) {
                    rYd += (double)j * rz;
                    if (j > 1) rYd *= y;
                }
                if (j > 0) rYz *= y;
            }
            if (j > 0) rY *= y;
        }
        ry[i] = rY; ryd[i] = rYd; ryz[i] = rYz;
    }
    double rx = 0.;
#pragma unroll
    for (int i = 3; i >= 0; --i) { rx += ry[i]; if (i > 0) rx *= x; }
    F = rx;
    if (grad) {
        double dx = 0.;
        dx += 3.0 * ry[3]; dx *= x;
        dx += 2.0 * ry[2]; dx *= x;
        dx += ry[1];
        Fx = dx;
        double dy = 0.;
#pragma unroll
        for (int i = 3; i >= 0; --i) { dy += ryd[i]; if (i > 0) dy *= x; }
        Fy = dy;
        double dz = 0.;
#pragma unroll
        for (int i = 3; i >= 0; --i) { dz += ryz[i]; if (i > 0) dz *= x; }
        Fz = dz;
    }
}

/* FP32 variant for the mixed-precision tracker: coefficients and Horner in FP32 */
PTD void tric_eval_grad32(const float *__restrict__ a, float x, float y, float z, bool grad,
                          float &F, float &Fx, float &Fy, float &Fz) {
    float ry[4], ryd[4], ryz[4];
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        float rY = 0.f, rYd = 0.f, rYz = 0.f;
#pragma unroll
        for (int j = 3; j >= 0; --j) {
            float a0 = a[i + 4 * j], a1 = a[i + 4 * j + 16], a2 = a[i + 4 * j + 32], a3 = a[i + 4 * j + 48];
            float rz = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(a3, z, a2), z, a1), z, a0);
            rY = __builtin_fmaf(rY, y, rz);
            if (grad) {
                float rzd = __builtin_fmaf(__builtin_fmaf(3.0f * a3, z, 2.0f * a2), z, a1);
                rYz = __builtin_fmaf(rYz, y, rzd);
                if (j > 0) rYd = __builtin_fmaf(rYd, y, (float)j * rz);
            }
        }
        ry[i] = rY; ryd[i] = rYd; ryz[i] = rYz;
    }
    F = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(ry[3], x, ry[2]), x, ry[1]), x, ry[0]);
    if (grad) {
        Fx = __builtin_fmaf(__builtin_fmaf(3.0f * ry[3], x, 2.0f * ry[2]), x, ry[1]);
        Fy = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(ryd[3], x, ryd[2]), x, ryd[1]), x, ryd[0]);
        Fz = __builtin_fmaf(__builtin_fmaf(__builtin_fmaf(ryz[3], x, ryz[2]), x, ryz[1]), x, ryz[0]);
    }
}

PTD bool in_bounds(const Field &f, double x, double y, double z) {
    const double *b = f.box;
    bool has = b[1] < b[0] && b[3] < b[2] && b[5] < b[4];
    if (!has) return true;
    return x >= b[1] && x < b[0] && y >= b[3] && y < b[2] && z >= b[5] && z < b[4];
}
PTD double smth(double x) { return ((6 * x - 15) * x + 10) * x * x * x; }
PTD double smthd(double x) { return ((30 * x - 60) * x + 30) * x * x; }

/* TFieldBoundaryBox::scaleScalarFieldAtBounds (src/field.cpp:71-112) */
PTD_NOINLINE void scale_at_bounds(const Field &f, double x, double y, double z, double *F, double *dF) {
    const double *b = f.box;
    double xmax = b[0], xmin = b[1], ymax = b[2], ymin = b[3], zmax = b[4], zmin = b[5], bw = f.bw;
    bool has = xmin < xmax && ymin < ymax && zmin < zmax;
    if (!has || (*F == 0 && dF == nullptr) || (*F == 0 && dF[0] == 0 && dF[1] == 0 && dF[2] == 0)) return;
    if (!in_bounds(f, x, y, z)) {
        *F = 0.;
        if (dF) dF[0] = dF[1] = dF[2] = 0.;
    } else if (bw > 0) {
        double Fscale = 1.;
        double dFadd[3] = {0., 0., 0.};
        double lo[3] = {(x - xmin) / bw, (y - ymin) / bw, (z - zmin) / bw};
        double hi[3] = {(xmax - x) / bw, (ymax - y) / bw, (zmax - z) / bw};
        for (int i = 0; i < 3; ++i) {
            if (0 <= lo[i] && lo[i] <= 1) {
                Fscale *= smth(lo[i]);
                if (dF) dFadd[i] += *F * smthd(lo[i]) / bw / smth(lo[i]);
            } else if (0 <= hi[i] && hi[i] <= 1) {
                Fscale *= smth(hi[i]);
                if (dF) dFadd[i] += -*F * smthd(hi[i]) / bw / smth(hi[i]);
            }
        }
        if (Fscale != 1.) {
            *F *= Fscale;
            if (dF) for (int i = 0; i < 3; i++) dF[i] = dF[i] * Fscale + dFadd[i] * Fscale;
        }
    }
}

PTD bool has_box(const Field &f) { return f.box[1] < f.box[0] && f.box[3] < f.box[2] && f.box[5] < f.box[4]; }

/* TFieldContainer::BField (src/field.cpp:139-153) accumulated into the manager sum
   (TFieldManager::BField, src/fields.cpp:139-164) */
template <bool MIXED>
PTD void container_B(const Field &f, double x, double y, double z, double B[3], double dB[3][3], bool grad, long long &nfield) {
    double Bt[3] = {0, 0, 0};
    double dBt[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (in_bounds(f, x, y, z) && f.bscale != 0.) {
        if (f.type == 1 && f.has_b) {
            long cell; double r[3], dist[3];
            if (tric_locate(f, x, y, z, cell, r, dist)) {
                ++nfield;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    double Fv, Fx = 0, Fy = 0, Fz = 0;
                    if (MIXED) {
                        float fv, fx, fy, fz;
                        tric_eval_grad32(f.coef32 + ((size_t)cell * f.ncomp + i) * 64, (float)r[0], (float)r[1], (float)r[2], grad, fv, fx, fy, fz);
                        Fv = fv; Fx = fx; Fy = fy; Fz = fz;
                    } else {
                        tric_eval_grad64(f.coef + ((size_t)cell * f.ncomp + i) * 64, r[0], r[1], r[2], grad, Fv, Fx, Fy, Fz);
                    }
                    Bt[i] = Fv;
                    if (grad) { dBt[i][0] = Fx / dist[0]; dBt[i][1] = Fy / dist[1]; dBt[i][2] = Fz / dist[2]; }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            Bt[i] *= f.bscale;
            if (grad) { dBt[i][0] *= f.bscale; dBt[i][1] *= f.bscale; dBt[i][2] *= f.bscale; }
        }
        if (has_box(f))
            for (int i = 0; i < 3; ++i) scale_at_bounds(f, x, y, z, &Bt[i], grad ? dBt[i] : nullptr);
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        B[i] += Bt[i];
        if (grad) {
#pragma unroll
            for (int j = 0; j < 3; j++) dB[i][j] += dBt[i][j];
        }
    }
}

template <bool MIXED>
PTD void field_B(const Params &P, double x, double y, double z, double B[3], double dB[3][3], bool grad, long long &nfield) {
#pragma unroll
    for (int i = 0; i < 3; i++) { B[i] = 0; if (grad) { dB[i][0] = 0; dB[i][1] = 0; dB[i][2] = 0; } }
    for (int k = 0; k < P.n_fields; ++k) container_B<MIXED>(P.fields[k], x, y, z, B, dB, grad, nfield);
}

/* TFieldContainer::EField / TFieldManager::EField (src/field.cpp:155-169, src/fields.cpp:167-182) */
template <bool MIXED>
PTD_NOINLINE void field_E(const Params &P, double x, double y, double z, double *V, double E[3], long long &nfield) {
    *V = 0; E[0] = E[1] = E[2] = 0;
    for (int k = 0; k < P.n_fields; ++k) {
        const Field &f = P.fields[k];
        double Vt = 0, Et[3] = {0, 0, 0};
        if (in_bounds(f, x, y, z) && f.escale != 0.) {
            if (f.type == 2) {
                Et[0] = f.e[0]; Et[1] = f.e[1]; Et[2] = f.e[2];
                Vt = Et[2] * z;
            } else if (f.type == 1 && f.has_v) {
                long cell; double r[3], dist[3];
                double dV[3] = {0., 0., 0.};
                if (tric_locate(f, x, y, z, cell, r, dist)) {
                    ++nfield;
                    double Fv, Fx, Fy, Fz;
                    int comp = f.has_b ? 3 : 0;
                    if (MIXED) {
                        float fv, fx, fy, fz;
                        tric_eval_grad32(f.coef32 + ((size_t)cell * f.ncomp + comp) * 64, (float)r[0], (float)r[1], (float)r[2], true, fv, fx, fy, fz);
                        Fv = fv; Fx = fx; Fy = fy; Fz = fz;
                    } else {
                        tric_eval_grad64(f.coef + ((size_t)cell * f.ncomp + comp) * 64, r[0], r[1], r[2], true, Fv, Fx, Fy, Fz);
                    }
                    Vt = Fv;
                    dV[0] = Fx / dist[0]; dV[1] = Fy / dist[1]; dV[2] = Fz / dist[2];
                }
                for (int i = 0; i < 3; i++) Et[i] = -dV[i];
            }
            Vt *= f.escale;
            Et[0] *= f.escale; Et[1] *= f.escale; Et[2] *= f.escale;
            if (has_box(f)) scale_at_bounds(f, x, y, z, &Vt, Et);
        }
        *V += Vt;
        for (int i = 0; i < 3; i++) E[i] += Et[i];
    }
}

/* ------------------------------------------------------------------------ */
/* particle                                                                  */
/* ------------------------------------------------------------------------ */

/* TParticle::EquationOfMotion, src/particle.cpp:105-135 */
PTD void eom(const Species &sp, const double *y, const double B[3], const double dB[3][3], const double E[3], double *dydx) {
    dydx[0] = y[3]; dydx[1] = y[4]; dydx[2] = y[5];
    double F[3] = {0, 0, 0};
    F[2] += -GRAV * sp.m * ELE_E;
    if (sp.q != 0) {
        F[0] += sp.q * (E[0] + y[4] * B[2] - y[5] * B[1]);
        F[1] += sp.q * (E[1] + y[5] * B[0] - y[3] * B[2]);
        F[2] += sp.q * (E[2] + y[3] * B[1] - y[4] * B[0]);
    }
    if (sp.mu != 0 && y[7] != 0 && (B[0] != 0 || B[1] != 0 || B[2] != 0)) {
        double Babs = __builtin_sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
        double d0 = (B[0] * dB[0][0] + B[1] * dB[1][0] + B[2] * dB[2][0]) / Babs;
        double d1 = (B[0] * dB[0][1] + B[1] * dB[1][1] + B[2] * dB[2][1]) / Babs;
        double d2 = (B[0] * dB[0][2] + B[1] * dB[1][2] + B[2] * dB[2][2]) / Babs;
        F[0] += y[7] * sp.mu * d0;
        F[1] += y[7] * sp.mu * d1;
        F[2] += y[7] * sp.mu * d2;
    }
    double v2 = y[3] * y[3] + y[4] * y[4] + y[5] * y[5];
    double inversegamma = __builtin_sqrt(1 - v2 / (C_0 * C_0));
    double f = inversegamma / sp.m / ELE_E;
    dydx[3] = f * (F[0] - (y[3] * y[3] * F[0] + y[3] * y[4] * F[1] + y[3] * y[5] * F[2]) / C_0 / C_0);
    dydx[4] = f * (F[1] - (y[4] * y[3] * F[0] + y[4] * y[4] * F[1] + y[4] * y[5] * F[2]) / C_0 / C_0);
    dydx[5] = f * (F[2] - (y[5] * y[3] * F[0] + y[5] * y[4] * F[1] + y[5] * y[5] * F[2]) / C_0 / C_0);
    dydx[6] = inversegamma;
    dydx[7] = 0;
    dydx[8] = __builtin_sqrt(v2);
}

/* evaluation counters kept per thread */
struct Ctr {
    long long rhs, field, tri;
};

/* TParticle::derivs, src/particle.cpp:96-103 */
template <bool MIXED>
PTD void derivs(const Params &P, const Species &sp, const double *y, double *dydx, double t, Ctr &c) {
    double B[3] = {0, 0, 0}, dB[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, E[3] = {0, 0, 0}, V = 0;
    (void)t;
    ++c.rhs;
    if (sp.q != 0 || (sp.mu != 0 && y[7] != 0)) field_B<MIXED>(P, y[0], y[1], y[2], B, dB, true, c.field);
    if (sp.q != 0) field_E<MIXED>(P, y[0], y[1], y[2], &V, E, c.field);
    eom(sp, y, B, dB, E, dydx);
}

/* TParticle::GetKineticEnergy, src/particle.cpp:237-251 */
PTD double kinetic_energy(const Species &sp, const double *v) {
    double beta2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / C_0 / C_0;
    double gm1;
    if (ptm_pow(beta2, 5) < DBL_EPS)
        gm1 = (0.5 + (3.0 / 8.0 + (5.0 / 16.0 + 35.0 / 128.0 * beta2) * beta2) * beta2) * beta2;
    else
        gm1 = 1.0 / __builtin_sqrt(1.0 - beta2) - 1.0;
    return C_0 * C_0 * sp.m * gm1;
}

/* TParticle::GetPotentialEnergy, src/particle.cpp:254-269 */
template <bool MIXED>
PTD_NOINLINE double potential_energy(const Params &P, const Species &sp, double t, const double *y, Ctr &c) {
    double result = 0;
    (void)t;
    if (sp.q != 0 || sp.mu != 0) {
        double B[3], dB[3][3], E[3], V;
        if (sp.mu != 0) {
            field_B<MIXED>(P, y[0], y[1], y[2], B, dB, false, c.field);
            result += -y[7] * sp.mu / ELE_E * __builtin_sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
        }
        if (sp.q != 0) {
            field_E<MIXED>(P, y[0], y[1], y[2], &V, E, c.field);
            result += sp.q / ELE_E * V;
        }
    }
    result += sp.m * GRAV * y[2];
    return result;
}

/* ------------------------------------------------------------------------ */
/* Boost.Odeint dense-output dopri5 (runge_kutta_dopri5 + controlled + dense) */
/* ------------------------------------------------------------------------ */

struct Dopri {
    double xc[9], dxc[9];            /* current state / derivative (FSAL) */
    double xo[9], dxo[9];            /* previous state / derivative */
    double k3[9], k4[9], k5[9], k6[9];
    double t, t_old, dt;
    bool deriv_init;
};

/* returns 1 accepted, 0 rejected (stepper state updated like try_step) */
template <bool MIXED>
PTD int dopri_try(const Params &P, const Species &sp, Dopri &s, double xnew[9], double dxnew[9], Ctr &c) {
    constexpr double a2 = 1.0 / 5, a3 = 3.0 / 10, a4 = 4.0 / 5, a5 = 8.0 / 9;
    constexpr double b21 = 1.0 / 5;
    constexpr double b31 = 3.0 / 40, b32 = 9.0 / 40;
    constexpr double b41 = 44.0 / 45, b42 = -56.0 / 15, b43 = 32.0 / 9;
    constexpr double b51 = 19372.0 / 6561, b52 = -25360.0 / 2187, b53 = 64448.0 / 6561, b54 = -212.0 / 729;
    constexpr double b61 = 9017.0 / 3168, b62 = -355.0 / 33, b63 = 46732.0 / 5247, b64 = 49.0 / 176, b65 = -5103.0 / 18656;
    constexpr double c1 = 35.0 / 384, c3 = 500.0 / 1113, c4 = 125.0 / 192, c5 = -2187.0 / 6784, c6 = 11.0 / 84;
    constexpr double dc1 = c1 - 5179.0 / 57600, dc3 = c3 - 7571.0 / 16695, dc4 = c4 - 393.0 / 640,
                     dc5 = c5 - (-92097.0 / 339200), dc6 = c6 - 187.0 / 2100, dc7 = -1.0 / 40;
    const double dt = s.dt, t = s.t;
    double tmp[9], k2[9];
    const double h21 = dt * b21;
#pragma unroll
    for (int i = 0; i < 9; ++i) tmp[i] = s.xc[i] + h21 * s.dxc[i];
    derivs<MIXED>(P, sp, tmp, k2, t + dt * a2, c);
    const double h31 = dt * b31, h32 = dt * b32;
#pragma unroll
    for (int i = 0; i < 9; ++i) tmp[i] = s.xc[i] + h31 * s.dxc[i] + h32 * k2[i];
    derivs<MIXED>(P, sp, tmp, s.k3, t + dt * a3, c);
    const double h41 = dt * b41, h42 = dt * b42, h43 = dt * b43;
#pragma unroll
    for (int i = 0; i < 9; ++i) tmp[i] = s.xc[i] + h41 * s.dxc[i] + h42 * k2[i] + h43 * s.k3[i];
    derivs<MIXED>(P, sp, tmp, s.k4, t + dt * a4, c);
    const double h51 = dt * b51, h52 = dt * b52, h53 = dt * b53, h54 = dt * b54;
#pragma unroll
    for (int i = 0; i < 9; ++i) tmp[i] = s.xc[i] + h51 * s.dxc[i] + h52 * k2[i] + h53 * s.k3[i] + h54 * s.k4[i];
    derivs<MIXED>(P, sp, tmp, s.k5, t + dt * a5, c);
    const double h61 = dt * b61, h62 = dt * b62, h63 = dt * b63, h64 = dt * b64, h65 = dt * b65;
#pragma unroll
    for (int i = 0; i < 9; ++i) tmp[i] = s.xc[i] + h61 * s.dxc[i] + h62 * k2[i] + h63 * s.k3[i] + h64 * s.k4[i] + h65 * s.k5[i];
    derivs<MIXED>(P, sp, tmp, s.k6, t + dt, c);
    const double g1 = dt * c1, g3 = dt * c3, g4 = dt * c4, g5 = dt * c5, g6 = dt * c6;
#pragma unroll
    for (int i = 0; i < 9; ++i) xnew[i] = s.xc[i] + g1 * s.dxc[i] + g3 * s.k3[i] + g4 * s.k4[i] + g5 * s.k5[i] + g6 * s.k6[i];
    derivs<MIXED>(P, sp, xnew, dxnew, t + dt, c);
    const double e1 = dt * dc1, e3 = dt * dc3, e4 = dt * dc4, e5 = dt * dc5, e6 = dt * dc6, e7 = dt * dc7;
    double err = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double xe = e1 * s.dxc[i] + e3 * s.k3[i] + e4 * s.k4[i] + e5 * s.k5[i] + e6 * s.k6[i] + e7 * dxnew[i];
        double e = dabs(xe) / (1e-9 + 1e-9 * (dabs(s.xc[i]) + dt * dabs(s.dxc[i])));
        err = (err < e) ? e : err;
    }
    if (err > 1.0) {
        double f = 9.0 / 10 * ptm_pow(err, -1.0 / 3);
        double m = 1.0 / 5;
        s.dt = dt * ((f < m) ? m : f);
        return 0;
    }
    s.t_old = t;
    s.t = t + dt;
    if (err < 0.5) {
        double emin = ptm_pow(5.0, -5.0);
        err = (emin < err) ? err : emin;
        s.dt = dt * (9.0 / 10 * ptm_pow(err, -1.0 / 5));
    }
    return 1;
}

/* dense_output_runge_kutta::do_step; returns false when the failed-step checker throws */
template <bool MIXED>
PTD bool dopri_do_step(const Params &P, const Species &sp, Dopri &s, long long &nacc, long long &nrej, Ctr &c) {
    if (!s.deriv_init) {
        derivs<MIXED>(P, sp, s.xc, s.dxc, s.t, c);
        s.deriv_init = true;
    }
    int tries = 0;
    for (;;) {
        double xnew[9], dxnew[9];
        int ok = dopri_try<MIXED>(P, sp, s, xnew, dxnew, c);
        if (ok) {
#pragma unroll
            for (int i = 0; i < 9; ++i) { s.xo[i] = s.xc[i]; s.dxo[i] = s.dxc[i]; s.xc[i] = xnew[i]; s.dxc[i] = dxnew[i]; }
            ++nacc;
        } else {
            ++nrej;
        }
        if (tries++ >= 500) return false;
        if (ok) return true;
    }
}

/* runge_kutta_dopri5::calc_state */
PTD void dopri_calc_state(const Dopri &s, double t, double *x) {
    constexpr double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192, b5 = -2187.0 / 6784, b6 = 11.0 / 84;
    const double dt = (s.t - s.t_old);
    const double theta = (t - s.t_old) / dt;
    const double X1 = 5.0 * (2558722523.0 - 31403016.0 * theta) / 11282082432.0;
    const double X3 = 100.0 * (882725551.0 - 15701508.0 * theta) / 32700410799.0;
    const double X4 = 25.0 * (443332067.0 - 31403016.0 * theta) / 1880347072.0;
    const double X5 = 32805.0 * (23143187.0 - 3489224.0 * theta) / 199316789632.0;
    const double X6 = 55.0 * (29972135.0 - 7076736.0 * theta) / 822651844.0;
    const double X7 = 10.0 * (7414447.0 - 829305.0 * theta) / 29380423.0;
    const double theta_m_1 = theta - 1.0;
    const double theta_sq = theta * theta;
    const double A = theta_sq * (3.0 - 2.0 * theta);
    const double B = theta_sq * theta_m_1;
    const double C = theta_sq * theta_m_1 * theta_m_1;
    const double D = theta * theta_m_1 * theta_m_1;
    const double h1 = dt * (A * b1 - C * X1 + D);
    const double h3 = dt * (A * b3 + C * X3);
    const double h4 = dt * (A * b4 - C * X4);
    const double h5 = dt * (A * b5 + C * X5);
    const double h6 = dt * (A * b6 - C * X6);
    const double h7 = dt * (B + C * X7);
#pragma unroll
    for (int i = 0; i < 9; ++i)
        x[i] = s.xo[i] + h1 * s.dxo[i] + h3 * s.k3[i] + h4 * s.k4[i] + h5 * s.k5[i] + h6 * s.k6[i] + h7 * s.dxc[i];
}

/* ------------------------------------------------------------------------ */
/* tracker                                                                   */
/* ------------------------------------------------------------------------ */

struct Track {
    int i;                      /* particle index */
    Species sp;
    uint64_t key, ctr;
    unsigned long long solids;  /* packed solid stack (TTracker::currentsolids) */
    int nsolid;
    int stop, Nhit, Nstep, Nspinflip, nhitlog;
    long long nacc, nrej;
    double Hmax, noflip;
    Ctr c;
};

PTD double rnd_uniform(Track &T, double a, double b) {
    return ptm_canonical(ptm_draw_u64(T.key, &T.ctr)) * (b - a) + a;
}
PTD double rnd_sincos(Track &T, double amin, double amax) {
    double x = ptm_canonical(ptm_draw_u64(T.key, &T.ctr));
    double cmax = ptm_cos(amax), cmin = ptm_cos(amin);
    return ptm_acos(__builtin_sqrt(x * (cmax * cmax - cmin * cmin) + cmin * cmin));
}

PTD bool state_eq(const double *a, const double *b) {
    bool eq = true;
#pragma unroll
    for (int i = 0; i < 9; ++i) eq = eq && (a[i] == b[i]);
    return eq;
}

/* RotateVector, src/globals.cpp:40-64 */
PTD_NOINLINE void rotate_vector(double v[3], const double n[3], const double *x) {
    double xv[3], zv[3] = {n[0], n[1], n[2]};
    double pa = zv[0], pb = zv[1], pc = zv[2];
    double pd = -zv[0] * 0.0 - zv[1] * 0.0 - zv[2] * 0.0;
    bool use_base = (x == nullptr);
    if (!use_base) {
        double num = pa * x[0] + pb * x[1] + pc * x[2] + pd;
        double den = pa * pa + pb * pb + pc * pc;
        double lambda = num / den;
        xv[0] = x[0] - lambda * pa;
        xv[1] = x[1] - lambda * pb;
        xv[2] = x[2] - lambda * pc;
        if (xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2] < 1e-30) use_base = true;
    }
    if (use_base) {
        if (pa == 0) { xv[0] = 1; xv[1] = 0; xv[2] = 0; }
        else if (pb == 0) { xv[0] = 0; xv[1] = 1; xv[2] = 0; }
        else if (pc == 0) { xv[0] = 0; xv[1] = 0; xv[2] = 1; }
        else {
            double a = dabs(pa), b = dabs(pb), c = dabs(pc);
            if (a <= b && a <= c) { xv[0] = 0; xv[1] = -pc; xv[2] = pb; }
            else if (b <= a && b <= c) { xv[0] = -pc; xv[1] = 0; xv[2] = pa; }
            else { xv[0] = -pb; xv[1] = pa; xv[2] = 0; }
        }
    }
    double lx = __builtin_sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    xv[0] = xv[0] / lx; xv[1] = xv[1] / lx; xv[2] = xv[2] / lx;
    double lz = __builtin_sqrt(zv[0] * zv[0] + zv[1] * zv[1] + zv[2] * zv[2]);
    zv[0] = zv[0] / lz; zv[1] = zv[1] / lz; zv[2] = zv[2] / lz;
    double yv[3] = {zv[1] * xv[2] - zv[2] * xv[1], zv[2] * xv[0] - zv[0] * xv[2], zv[0] * xv[1] - zv[1] * xv[0]};
    double v0 = v[0], v1 = v[1], v2 = v[2];
    v[0] = xv[0] * v0 + yv[0] * v1 + zv[0] * v2;
    v[1] = xv[1] * v0 + yv[1] * v1 + zv[1] * v2;
    v[2] = xv[2] * v0 + yv[2] * v1 + zv[2] * v2;
}

/* neutron wall interaction, TNeutron::OnHit and helpers (src/neutron.cpp:25-264) */
PTD double material_potential(const Species &sp, const double *mat, const double *y) {
    return mat[0] * 1e-9 - mat[6] * sp.mu * y[7] / ELE_E;
}

PTD void neutron_reflect(const double *y1, double &x2, double x1, double *y2, const double *normal) {
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    x2 = x1;
    y2[0] = y1[0]; y2[1] = y1[1]; y2[2] = y1[2];
    y2[3] -= 2 * vnormal * normal[0];
    y2[4] -= 2 * vnormal * normal[1];
    y2[5] -= 2 * vnormal * normal[2];
    y2[6] = y1[6];
    y2[8] = y1[8];
}

PTD_NOINLINE void neutron_reflect_lambert(Track &T, const double *y1, double &x2, double x1, double *y2, const double *normal, const double *mat) {
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    double n[3] = {normal[0], normal[1], normal[2]};
    if (vnormal > 0) { vnormal *= -1; n[0] *= -1; n[1] *= -1; n[2] *= -1; }
    x2 = x1;
    y2[0] = y1[0]; y2[1] = y1[1]; y2[2] = y1[2];
    double vabs = __builtin_sqrt(y1[3] * y1[3] + y1[4] * y1[4] + y1[5] * y1[5]);
    do {
        double phi_r = rnd_uniform(T, 0, 2. * PI);
        double theta_r = rnd_sincos(T, 0, PI / 2.);
        y2[3] = vabs * ptm_cos(phi_r) * ptm_sin(theta_r);
        y2[4] = vabs * ptm_sin(phi_r) * ptm_sin(theta_r);
        y2[5] = vabs * ptm_cos(theta_r);
        if (mat[2] > 0) {
            rotate_vector(y2 + 3, n, y1 + 3);
        } else {
            double sr[3] = {y1[3] - 2 * vnormal * n[0], y1[4] - 2 * vnormal * n[1], y1[5] - 2 * vnormal * n[2]};
            rotate_vector(y2 + 3, sr, y1 + 3);
        }
    } while (y2[3] * n[0] + y2[4] * n[1] + y2[5] * n[2] <= 0);
    y2[6] = y1[6];
    y2[8] = y1[8];
}

PTD void neutron_transmit(const Species &sp, const double *y1, double *y2, const double *normal, double Estep) {
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    double Enormal = 0.5 * sp.m * vnormal * vnormal;
    double k1 = __builtin_sqrt(Enormal);
    double k2 = __builtin_sqrt(Enormal - Estep);
    for (int i = 0; i < 3; i++) y2[i + 3] += (k2 / k1 - 1) * (normal[i] * vnormal);
}

PTD_NOINLINE void neutron_transmit_lambert(Track &T, const double *y1, double *y2, const double *normal, double Estep, const double *mat) {
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    double n[3] = {normal[0], normal[1], normal[2]};
    if (vnormal < 0) { vnormal *= -1; n[0] *= -1; n[1] *= -1; n[2] *= -1; }
    double Enormal = 0.5 * T.sp.m * vnormal * vnormal;
    double k1 = __builtin_sqrt(Enormal);
    double k2 = __builtin_sqrt(Enormal - Estep);
    double st[3] = {y2[3], y2[4], y2[5]};
    for (int i = 0; i < 3; ++i) st[i] += ((k2 / k1 - 1) * n[i]) * vnormal;
    double vabs = __builtin_sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]);
    do {
        double phi_t = rnd_uniform(T, 0, 2. * PI);
        double theta_t = rnd_sincos(T, 0, PI / 2.);
        y2[3] = vabs * ptm_cos(phi_t) * ptm_sin(theta_t);
        y2[4] = vabs * ptm_sin(phi_t) * ptm_sin(theta_t);
        y2[5] = vabs * ptm_cos(theta_t);
        if (mat[2] > 0) rotate_vector(y2 + 3, n, y1 + 3);
        else rotate_vector(y2 + 3, st, y1 + 3);
    } while (y2[3] * n[0] + y2[4] * n[1] + y2[5] * n[2] <= 0);
}

PTD double refl_prob(double E1r, double E1i, double E2r, double E2i) {
    double k1r, k1i, k2r, k2i, qr, qi;
    ptm_csqrt(E1r, E1i, &k1r, &k1i);
    ptm_csqrt(E2r, E2i, &k2r, &k2i);
    ptm_cdiv(k1r - k2r, k1i - k2i, k1r + k2r, k1i + k2i, &qr, &qi);
    double a = ptm_hypot(qr, qi);
    return a * a;
}

PTD_NOINLINE void neutron_onhit(const Params &P, Track &T, double x1, const double *y1, double &x2, double *y2,
                                const double *normal, int leaving, int entering) {
    const Geom &G = P.geom;
    double vnormal = y1[3] * normal[0] + y1[4] * normal[1] + y1[5] * normal[2];
    double Enormal = 0.5 * T.sp.m * vnormal * vnormal;
    const double *lmat = solid_mat(G, leaving), *emat = solid_mat(G, entering);
    const double *mat = vnormal < 0 ? emat : lmat;
    if (rnd_uniform(T, 0, 1) < mat[3]) y2[7] *= -1;
    double Estep = material_potential(T.sp, emat, y2) - material_potential(T.sp, lmat, y2);
    if (mat[4] != 0 && mat[5] != 0) { T.stop = ID_UNSUPPORTED; return; }
    double prob = rnd_uniform(T, 0, 1);
    double reflprob = refl_prob(Enormal, -lmat[1] * 1e-9, Enormal - Estep, -emat[1] * 1e-9);
    if (Enormal > Estep) {
        if (prob < 0 + 0 + reflprob * (1 - 0 - 0)) {
            if (rnd_uniform(T, 0, 1) < mat[2] + mat[7]) neutron_reflect_lambert(T, y1, x2, x1, y2, normal, mat);
            else neutron_reflect(y1, x2, x1, y2, normal);
        } else {
            if (rnd_uniform(T, 0, 1) < mat[2] + mat[7]) neutron_transmit_lambert(T, y1, y2, normal, Estep, mat);
            else neutron_transmit(T.sp, y1, y2, normal, Estep);
        }
    } else {
        double absprob = 1 - reflprob + mat[8];
        if (prob < 0 + 0 + absprob * (1 - 0 - 0)) {
            T.stop = ID_ABSORBED_ON_SURFACE;
        } else {
            if (rnd_uniform(T, 0, 1) < mat[2] + mat[7]) neutron_reflect_lambert(T, y1, x2, x1, y2, normal, mat);
            else neutron_reflect(y1, x2, x1, y2, normal);
        }
    }
}

/* OnStep hooks (src/neutron.cpp:267-282, src/proton.cpp:25-33, src/electron.cpp:25-33) */
PTD_NOINLINE void on_step(const Params &P, Track &T, double x1, const double *y1, double &x2, double *y2, const Dopri &st, int cursolid) {
    if (T.sp.kind == 0) {
        const double *mat = solid_mat(P.geom, cursolid);
        if (mat[1] > 0) {
            double mn = T.sp.m, e = ELE_E, hb = HBAR;
            double Er = 0.5 * mn * (y1[3] * y1[3] + y1[4] * y1[4] + y1[5] * y1[5]);
            double Ei = mat[1] * 1e-9;
            double kr, ki;
            ptm_csqrt((2 * mn) * Er, (2 * mn) * Ei, &kr, &ki);
            ki = ki * e;
            ki = ki / hb;
            double l = __builtin_sqrt((y2[0] - y1[0]) * (y2[0] - y1[0]) + (y2[1] - y1[1]) * (y2[1] - y1[1]) + (y2[2] - y1[2]) * (y2[2] - y1[2]));
            double abspath = -ptm_log(1.0 - ptm_canonical(ptm_draw_u64(T.key, &T.ctr))) / (2 * ki);
            if (abspath < l) {
                x2 = x1 + abspath / l * (x2 - x1);
                dopri_calc_state(st, x2, y2);
                T.stop = ID_ABSORBED_IN_MATERIAL;
            }
        }
    } else {
        if (solid_ID(P.geom, cursolid) > 1) {
            x2 = x1;
            for (int k = 0; k < 9; ++k) y2[k] = y1[k];
            T.stop = ID_ABSORBED_IN_MATERIAL;
        }
    }
}

/* TTracker::DoStep + TParticle::DoStep (src/tracking.cpp:268-278, src/particle.cpp:196-206) */
template <bool MIXED>
PTD_NOINLINE bool do_step_hook(const Params &P, Track &T, double x1, const double *y1, double &x2, double *y2, const Dopri &st, int cursolid) {
    double x2temp = x2, y2temp[9];
    for (int k = 0; k < 9; ++k) y2temp[k] = y2[k];
    on_step(P, T, x1, y1, x2, y2, st, cursolid);
    double H = kinetic_energy(T.sp, y2 + 3) + potential_energy<MIXED>(P, T.sp, x2, y2, T.c);
    T.Hmax = (H < T.Hmax) ? T.Hmax : H;
    if (y2temp[7] != y2[7]) T.Nspinflip++;
    T.Nstep++;
    return !(x2temp == x2 && state_eq(y2temp, y2));
}

/* TTracker::DoHit (src/tracking.cpp:280-365) */
PTD_NOINLINE bool do_hit(const Params &P, Track &T, double x1, const double *y1, double &x2, double *y2) {
    const Geom &G = P.geom;
    HitList L;
    L.n = 0;
    int nc = seg_query(G, y1, y2, &L, &T.c.tri);
    if (nc <= 0) { T.stop = nc == -1 ? ID_CGAL_ERROR : (nc == -2 ? ID_UNSUPPORTED : ID_INVARIANT); return true; }
    double sv[3] = {y2[0] - y1[0], y2[1] - y1[1], y2[2] - y1[2]};
    unsigned long long ns = T.solids;
    int nn = T.nsolid;
    bool trajectoryaltered = false, traversed = true;
    for (int k = 0; k < nc; ++k) {
        unsigned id = L.id[k];
        const double *n = G.nrm + 3 * L.tri[k];
        double distnormal = sv[0] * n[0] + sv[1] * n[1] + sv[2] * n[2];
        int solid = 0;
        for (int m = 0; m < G.n_solid; ++m) if (G.solid_id[m] == id) { solid = m; break; }
        int found = -1;
        for (int i = 0; i < nn; ++i) if (solid_ID(G, sk_sidx(ns, i)) == id) { found = i; break; }
        if (distnormal < 0) {
            if (found >= 0) {
                if (L.s[k] > 0) {
                    if (nn >= MAX_SOLIDS) { T.stop = ID_UNSUPPORTED; return true; }
                    ns = sk_set(ns, nn, solid, sk_ign(ns, found)); nn++;
                }
            } else {
                if (nn >= MAX_SOLIDS) { T.stop = ID_UNSUPPORTED; return true; }
                int ign = solid_ignored(G, solid, x1 + (x2 - x1) * L.s[k]);
                ns = sk_set(ns, nn, solid, ign); nn++;
            }
        } else if (distnormal > 0) {
            if (found < 0) {
                if (L.s[k] > 0) { T.stop = ID_GEOMETRY_ERROR; return true; }
            } else {
                ns = sk_erase(ns, found); nn--;
            }
        } else {
            T.stop = ID_GEOMETRY_ERROR;
            return true;
        }
    }
    int leaving = current_solid(G, T.solids, T.nsolid);
    int entering = -1;
    for (int i = 0; i < nn; ++i)
        if (!sk_ign(ns, i) && solid_ID(G, sk_sidx(ns, i)) > solid_ID(G, entering)) entering = sk_sidx(ns, i);
    unsigned lid = solid_ID(G, leaving), eid = solid_ID(G, entering);
    if (lid != eid) {
        int sel = -1;
        for (int k = 0; k < nc; ++k) {
            if (L.id[k] == lid) { sel = k; break; }
            if (L.id[k] == eid) {
                int solid = 0;
                for (int m = 0; m < G.n_solid; ++m) if (G.solid_id[m] == eid) { solid = m; break; }
                if (!solid_ignored(G, solid, x1 + (x2 - x1) * L.s[k])) { sel = k; break; }
            }
        }
        if (sel < 0) { T.stop = ID_INVARIANT; return true; }
        const double *np = G.nrm + 3 * L.tri[sel];
        double normal[3] = {np[0], np[1], np[2]};
        double x2temp = x2, y2temp[9];
        for (int k = 0; k < 9; ++k) y2temp[k] = y2[k];
        if (T.sp.kind == 0) neutron_onhit(P, T, x1, y1, x2, y2, normal, leaving, entering);
        if (y2temp[7] != y2[7]) T.Nspinflip++;
        T.Nhit++;
        if (x2temp == x2 && state_eq(y2temp, y2)) {
            trajectoryaltered = false; traversed = true;
        } else {
            trajectoryaltered = true;
            if (x2 == x2temp && y2[0] == y2temp[0] && y2[1] == y2temp[1] && y2[2] == y2temp[2]) traversed = true;
            else if (x2 == x1 && y2[0] == y1[0] && y2[1] == y1[1] && y2[2] == y1[2]) traversed = false;
            else { T.stop = ID_INVARIANT; return true; }
        }
        if (T.nhitlog < P.st.hit_cap) {
            Hit &h = P.st.hits[(size_t)T.i * P.st.hit_cap + T.nhitlog];
            h.t = x1; h.tri = L.orig[sel]; h.leaving = (int)lid; h.entering = (int)eid;
            for (int k = 0; k < 3; ++k) { h.normal[k] = normal[k]; h.v1[k] = y1[3 + k]; h.v2[k] = y2[3 + k]; }
            h.pol1 = y1[7]; h.pol2 = y2[7];
        }
        T.nhitlog++;
    }
    if (traversed) { T.solids = ns; T.nsolid = nn; }
    return trajectoryaltered || T.stop != ID_UNKNOWN;
}

/* number of intersections of segment p1->p2 (-1 coplanar, -2 overflow) */
PTD int count_hits(const Params &P, Track &T, const double *p1, const double *p2) {
    return seg_query(P.geom, p1, p2, nullptr, &T.c.tri);
}

/*
 * TTracker::iterate_collision (src/tracking.cpp:228-265) without recursion.
 * The recursion is a depth-first search over dyadic sub-intervals; a frame is
 * identified by the bit string of first/second-half choices leading to it, so
 * the search keeps only that bit string (depth <= 100) and recomputes interval
 * ends by replaying the halvings and states by dense output.  Returns 1 with the
 * final bracket in (x1,y1,x2,y2), 0 if every branch failed, -1 on a geometry error.
 */
PTD_NOINLINE int iterate_collision(const Params &P, Track &T, const Dopri &st, double &X1, double *Y1, double &X2, double *Y2) {
    unsigned long long ph[2] = {0ull, 0ull};   /* phase bit per level: 0 first half, 1 second half */
    int d = 0;
    double lo = X1, hi = X2;
    double plo[3] = {Y1[0], Y1[1], Y1[2]}, phi[3] = {Y2[0], Y2[1], Y2[2]};
    double pmid[3], ymid[9];
    double mid;
    int branch; /* 0: evaluate frame, 1: try second half of frame d */
    branch = 0;
    for (;;) {
        if (branch == 0) {
            double dx0 = phi[0] - plo[0], dx1 = phi[1] - plo[1], dx2 = phi[2] - plo[2];
            if (dx0 * dx0 + dx1 * dx1 + dx2 * dx2 < REFLECT_TOLERANCE * REFLECT_TOLERANCE) break;
            if (hi - lo < 4 * (lo + hi) * DBL_EPS) break;
            if (d >= 100) break;
            mid = lo + (hi - lo) * 0.5;
            dopri_calc_state(st, mid, ymid);
            pmid[0] = ymid[0]; pmid[1] = ymid[1]; pmid[2] = ymid[2];
            int n1 = count_hits(P, T, plo, pmid);
            if (n1 < 0) return -1;
            if (n1 > 0) {
                ph[d >> 6] &= ~(1ull << (d & 63));
                ++d;
                hi = mid; phi[0] = pmid[0]; phi[1] = pmid[1]; phi[2] = pmid[2];
                continue;
            }
        }
        branch = 0;
        int n2 = count_hits(P, T, pmid, phi);
        if (n2 < 0) return -1;
        if (n2 > 0) {
            ph[d >> 6] |= (1ull << (d & 63));
            ++d;
            lo = mid; plo[0] = pmid[0]; plo[1] = pmid[1]; plo[2] = pmid[2];
            continue;
        }
        /* backtrack to the nearest ancestor whose first half was tried */
        for (;;) {
            if (d == 0) return 0;
            --d;
            if (((ph[d >> 6] >> (d & 63)) & 1ull) == 0ull) break;
        }
        /* rebuild frame d: interval by replaying the halvings, states by dense output */
        lo = X1; hi = X2;
        bool loroot = true, hiroot = true;
        for (int k = 0; k < d; ++k) {
            double m = lo + (hi - lo) * 0.5;
            if ((ph[k >> 6] >> (k & 63)) & 1ull) { lo = m; loroot = false; }
            else { hi = m; hiroot = false; }
        }
        (void)loroot;
        mid = lo + (hi - lo) * 0.5;
        dopri_calc_state(st, mid, ymid);
        pmid[0] = ymid[0]; pmid[1] = ymid[1]; pmid[2] = ymid[2];
        if (hiroot) { phi[0] = Y2[0]; phi[1] = Y2[1]; phi[2] = Y2[2]; }
        else { double yh[9]; dopri_calc_state(st, hi, yh); phi[0] = yh[0]; phi[1] = yh[1]; phi[2] = yh[2]; }
        branch = 1;
    }
    /* success: the final bracket; its ends are root states or dense-output states */
    bool loroot = true, hiroot = true;
    for (int k = 0; k < d; ++k) {
        if ((ph[k >> 6] >> (k & 63)) & 1ull) loroot = false;
        else hiroot = false;
    }
    if (!loroot) { X1 = lo; dopri_calc_state(st, lo, Y1); }
    if (!hiroot) { X2 = hi; dopri_calc_state(st, hi, Y2); }
    return 1;
}

/* TTracker::CheckHit (src/tracking.cpp:176-226); the tail recursion is a loop */
template <bool MIXED>
PTD_NOINLINE bool check_hit(const Params &P, Track &T, double x1, const double *y1in, double &x2, double *y2, const Dopri &st) {
    double y1[9];
    for (int k = 0; k < 9; ++k) y1[k] = y1in[k];
    for (;;) {
        if (!check_segment(P.geom, y1, y2)) { T.stop = ID_HIT_BOUNDARIES; return true; }
        if (x2 == x1) return false;
        int cursolid = current_solid(P.geom, T.solids, T.nsolid);
        int n = count_hits(P, T, y1, y2);
        if (n < 0) { T.stop = n == -1 ? ID_CGAL_ERROR : ID_UNSUPPORTED; return true; }
        if (n == 0) return do_step_hook<MIXED>(P, T, x1, y1, x2, y2, st, cursolid);
        double xc1 = x1, xc2 = x2, yc1[9], yc2[9];
        for (int k = 0; k < 9; ++k) { yc1[k] = y1[k]; yc2[k] = y2[k]; }
        int r = iterate_collision(P, T, st, xc1, yc1, xc2, yc2);
        if (r < 0) { T.stop = ID_CGAL_ERROR; return true; }
        if (r == 0) return false;
        if (xc1 > x1 && do_step_hook<MIXED>(P, T, x1, y1, xc1, yc1, st, cursolid)) {
            x2 = xc1;
            for (int k = 0; k < 9; ++k) y2[k] = yc1[k];
            return true;
        }
        if (do_hit(P, T, xc1, yc1, xc2, yc2)) {
            x2 = xc2;
            for (int k = 0; k < 9; ++k) y2[k] = yc2[k];
            return true;
        }
        x1 = xc2;
        for (int k = 0; k < 9; ++k) y1[k] = yc2[k];
    }
}

/* TParticle::DoPolarize (src/particle.cpp:225-234) */
PTD void do_polarize(Track &T, double *y, double polarization, int flipspin) {
    double flipprob = 0.5 * (1 - y[7] * polarization);
    T.noflip *= 1. - flipprob;
    if (flipspin) {
        double prevpol = y[7];
        y[7] = (ptm_canonical(ptm_draw_u64(T.key, &T.ctr)) < 0.5 * (polarization + 1)) ? 1 : -1;
        if (y[7] != prevpol) ++T.Nspinflip;
    }
}

/* TTracker::IntegrateSpin without the BMT branch (src/tracking.cpp:373-394, 473-483) */
template <bool MIXED>
PTD_NOINLINE double integrate_spin(const Params &P, Track &T, double *spin, const Dopri &st, double x2, const double *y2) {
    double x1 = st.t_old;
    const double *y1 = st.xo;
    double B1[3], B2[3], dB[3][3];
    field_B<MIXED>(P, y1[0], y1[1], y1[2], B1, dB, false, T.c.field);
    field_B<MIXED>(P, y2[0], y2[1], y2[2], B2, dB, false, T.c.field);
    double Babs1 = __builtin_sqrt(B1[0] * B1[0] + B1[1] * B1[1] + B1[2] * B1[2]);
    double Babs2 = __builtin_sqrt(B2[0] * B2[0] + B2[1] * B2[1] + B2[2] * B2[2]);
    double Bmax = T.sp.Bmax;
    if (T.sp.gamma == 0 || Babs2 == 0) {
        for (int i = 0; i < 5; ++i) spin[i] = 0;
        return y1[7];
    } else if (Babs1 == 0 && Babs2 > 0) {
        int sp = ptm_uniform_int01(T.key, &T.ctr) * 2 - 1;
        for (int i = 0; i < 5; ++i) spin[i] = 0;
        return sp;
    } else if ((x2 > x1) && (Babs1 < Bmax || Babs2 < Bmax)) {
        if (spin[0] != 0 || spin[1] != 0 || spin[2] != 0) {
            double sp = (spin[0] * B1[0] + spin[1] * B1[1] + spin[2] * B1[2]) / Babs1 / __builtin_sqrt(spin[0] * spin[0] + spin[1] * spin[1] + spin[2] * spin[2]);
            if (sp * sp >= 1) { spin[0] = 0; spin[1] = 0; }
            else {
                double spinaz = rnd_uniform(T, 0, 2. * PI);
                spin[0] = __builtin_sqrt(1 - sp * sp) * ptm_sin(spinaz);
                spin[1] = __builtin_sqrt(1 - sp * sp) * ptm_cos(spinaz);
            }
            spin[2] = sp;
            rotate_vector(spin, B2, nullptr);
        }
    }
    if (Babs2 > Bmax && (spin[0] != 0 || spin[1] != 0 || spin[2] != 0)) {
        double sp = (spin[0] * B2[0] + spin[1] * B2[1] + spin[2] * B2[2]) / Babs2 / __builtin_sqrt(spin[0] * spin[0] + spin[1] * spin[1] + spin[2] * spin[2]);
        for (int i = 0; i < 5; ++i) spin[i] = 0;
        return sp;
    }
    return y1[7];
}

/* flags bits of the resumable state */
constexpr int FL_STARTED = 1, FL_DERIV = 2, FL_RESET = 4;

/*
 * Advance particle i by up to max_steps accepted steps of
 * TTracker::IntegrateParticle's loop (src/tracking.cpp:71-148).
 * Returns true when the particle has stopped.
 */
template <bool MIXED>
PTD bool track_particle(const Params &P, int i, int max_steps, long long *acc_out) {
    const State &S = P.st;
    const int n = S.n;
    Track T;
    T.i = i;
    T.sp = P.species[S.species[i]];
    T.key = S.key[i];
    T.ctr = S.ctr[i];
    T.stop = S.stop[i];
    T.Nhit = S.Nhit[i]; T.Nstep = S.Nstep[i]; T.Nspinflip = S.Nspinflip[i];
    T.nhitlog = S.nhitlog[i];
    T.nacc = S.n_acc[i]; T.nrej = S.n_rej[i];
    T.Hmax = S.Hmax[i]; T.noflip = S.noflip[i];
    T.solids = S.solids[i]; T.nsolid = S.nsolid[i];
    T.c.rhs = 0; T.c.field = 0; T.c.tri = 0;
    int flags = S.flags[i];
    double x = S.t[i];
    double y[9], spin[5];
    Dopri st;
#pragma unroll
    for (int k = 0; k < 9; ++k) { y[k] = S.y[k * n + i]; st.xc[k] = S.xc[k * n + i]; st.dxc[k] = S.dxc[k * n + i]; }
#pragma unroll
    for (int k = 0; k < 5; ++k) spin[k] = S.spin[k * n + i];
    st.t = S.ts[i];
    st.dt = S.dt[i];
    st.t_old = st.t;
    st.deriv_init = (flags & FL_DERIV) != 0;
    double tau = S.tau[i];
    const double tmax = P.tmax_global;

    if (!(flags & FL_STARTED)) {
        /* TParticle ctor Hmax (src/particle.cpp:89) and IntegrateParticle set-up (src/tracking.cpp:19-68) */
        T.Hmax = kinetic_energy(T.sp, y + 3) + potential_energy<MIXED>(P, T.sp, x, y, T.c);
        if (T.sp.tau > 0) tau = -ptm_log(1.0 - ptm_canonical(ptm_draw_u64(T.key, &T.ctr))) / (1. / T.sp.tau);
        else tau = T.sp.tmax;
        if (x > tmax) {
            T.stop = ID_INVARIANT;
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) st.xc[k] = y[k];
            st.t = x;
            st.dt = 10. * MAX_TRACK_DEVIATION / __builtin_sqrt(y[3] * y[3] + y[4] * y[4] + y[5] * y[5]);
            st.deriv_init = false;
            unsigned long long par = ray_parity(P.geom, y);
            unsigned long long sk = sk_set(0ull, 0, -1, 0);
            int ns = 1;
            for (int m = 0; m < P.geom.n_solid; ++m) {
                if ((par >> m) & 1ull) {
                    if (ns >= MAX_SOLIDS) { T.stop = ID_UNSUPPORTED; break; }
                    sk = sk_set(sk, ns, m, solid_ignored(P.geom, m, x));
                    ++ns;
                }
            }
            T.solids = sk; T.nsolid = ns;
        }
        flags = FL_STARTED;
    }

    int steps = 0;
    while (T.stop == ID_UNKNOWN && (max_steps <= 0 || steps < max_steps)) {
        if (flags & FL_RESET) {
#pragma unroll
            for (int k = 0; k < 9; ++k) st.xc[k] = y[k];
            st.t = x;
            st.deriv_init = false;
        }
        double x1 = x, y1[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) y1[k] = y[k];
        if (!dopri_do_step<MIXED>(P, T.sp, st, T.nacc, T.nrej, T.c)) T.stop = ID_ODEINT_ERROR;
        ++steps;
        x = st.t;
#pragma unroll
        for (int k = 0; k < 9; ++k) y[k] = st.xc[k];
        if (tau > 0 && y[6] > tau) {
            x = x1 + (x - x1) * (tau - y1[6]) / (y[6] - y1[6]);
            dopri_calc_state(st, x, y);
        }
        if (x > tmax) {
            x = tmax;
            dopri_calc_state(st, x, y);
        }
        bool reset = (flags & FL_RESET) != 0;
        while (x1 < x) {
            double l2 = (y[8] - y1[8]) * (y[8] - y1[8]);
            double d2 = (y[0] - y1[0]) * (y[0] - y1[0]) + (y[1] - y1[1]) * (y[1] - y1[1]) + (y[2] - y1[2]) * (y[2] - y1[2]);
            double dev2 = 0.25 * (l2 - d2);
            double x2 = x, y2[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) y2[k] = y[k];
            if (dev2 > MAX_TRACK_DEVIATION * MAX_TRACK_DEVIATION) {
                x2 = x1 + (x - x1) / __builtin_ceil(__builtin_sqrt(dev2) / MAX_TRACK_DEVIATION);
                dopri_calc_state(st, x2, y2);
            }
            reset = check_hit<MIXED>(P, T, x1, y1, x2, y2, st);
            if (reset) {
                x = x2;
#pragma unroll
                for (int k = 0; k < 9; ++k) y[k] = y2[k];
            }
            x1 = x2;
#pragma unroll
            for (int k = 0; k < 9; ++k) y1[k] = y2[k];
        }
        flags = reset ? (flags | FL_RESET) : (flags & ~FL_RESET);
        double spinProjection = integrate_spin<MIXED>(P, T, spin, st, x, y);
        bool spinHasFlipped = st.xo[7] != y[7];
        if (spinProjection != st.xo[7]) {
            do_polarize(T, y, spinProjection, T.sp.flipspin);
            if (T.sp.flipspin && spinHasFlipped) y[7] *= -1;
        } else if (spinHasFlipped) {
            for (int k = 0; k < 5; ++k) spin[k] = 0;
        }
        if (T.stop == ID_UNKNOWN && y[6] >= tau) T.stop = ID_DECAYED;
        else if (T.stop == ID_UNKNOWN && (x >= tmax || y[8] >= T.sp.lmax)) T.stop = ID_NOT_FINISH;
    }

    /* write back */
    if (st.deriv_init) flags |= FL_DERIV; else flags &= ~FL_DERIV;
    S.t[i] = x;
#pragma unroll
    for (int k = 0; k < 9; ++k) { S.y[k * n + i] = y[k]; S.xc[k * n + i] = st.xc[k]; S.dxc[k * n + i] = st.dxc[k]; }
#pragma unroll
    for (int k = 0; k < 5; ++k) S.spin[k * n + i] = spin[k];
    S.ts[i] = st.t;
    S.dt[i] = st.dt;
    S.tau[i] = tau;
    S.flags[i] = flags;
    S.ctr[i] = T.ctr;
    S.stop[i] = T.stop;
    S.Nhit[i] = T.Nhit; S.Nstep[i] = T.Nstep; S.Nspinflip[i] = T.Nspinflip;
    S.nhitlog[i] = T.nhitlog;
    S.n_acc[i] = T.nacc; S.n_rej[i] = T.nrej;
    S.Hmax[i] = T.Hmax; S.noflip[i] = T.noflip;
    S.solids[i] = T.solids; S.nsolid[i] = T.nsolid;
    if (acc_out) *acc_out = T.c.rhs;
    return T.stop != ID_UNKNOWN;
}

/* plain kernel: one thread per particle, max_steps accepted steps per launch */
template <bool MIXED>
__global__ void __launch_bounds__(256) track_kernel(Params P) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.st.n) return;
    if (*P.quit) return;
    if (P.st.stop[i] != ID_UNKNOWN && (P.st.flags[i] & FL_STARTED)) return;
    track_particle<MIXED>(P, i, P.max_steps, nullptr);
}

} // namespace ptd

/* ---- launchers (host side, called from pt_host.cpp) ---------------------- */

hipError_t pt_launch_track(const ptd::Params &P, int variant, int block, hipStream_t stream) {
    int grid = (P.st.n + block - 1) / block;
    if (grid == 0) return hipSuccess;
    if (variant == 1) hipLaunchKernelGGL(ptd::track_kernel<true>, dim3(grid), dim3(block), 0, stream, P);
    else hipLaunchKernelGGL(ptd::track_kernel<false>, dim3(grid), dim3(block), 0, stream, P);
    return hipGetLastError();
}

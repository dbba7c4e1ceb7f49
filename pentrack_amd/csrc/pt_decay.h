/*
 * pt_decay.h -- free-neutron decay into a proton and an electron
 * (TNeutron::Decay, src/neutron.cpp:285-362), host and device.
 *
 * The proton's kinetic energy is drawn from the recoil spectrum
 * ProtonBetaSpectrum (src/mc.cpp:84-92) through the reference's
 * std::piecewise_linear_distribution (include/mc.h:281-296, 337): the table
 * (abscissas, normalised densities, cumulative probabilities, slopes) is built
 * on the host by pt_host.cpp exactly as libstdc++ builds its parameter, and
 * sampled here as libstdc++'s operator() does (lower_bound of the uniform
 * variate in the cumulative probabilities, then the root of the quadratic of
 * the linear density).  The rest follows the reference step by step: an
 * isotropic proton, the virtual W state, its two-body decay into an electron
 * and a neutrino of mass 1 eV/c^2, Lorentz boosts (BOOST, src/globals.cpp:67-78)
 * and the species' energies and angles.  All random draws come from the
 * neutron's own stream after its tracking draws, in the reference's order.
 *
 * Arithmetic: FP64 with the double roundings of the constants; sin, cos, acos,
 * atan2 are the deterministic ones of pt_detmath.h.
 */
#ifndef PT_DECAY_H_
#define PT_DECAY_H_

#include <stdint.h>

#include "pt_detmath.h"

#if defined(__HIPCC__)
#define PTK_FN static inline __host__ __device__
#else
#define PTK_FN static inline
#endif

/* the sampled proton spectrum: n_int intervals between n_int + 1 abscissas */
typedef struct {
    const double *x;     /* [n_int + 1] abscissas (eV) */
    const double *den;   /* [n_int + 1] densities divided by the integral */
    const double *cp;    /* [n_int] cumulative probabilities (last one exactly 1) */
    const double *m;     /* [n_int] slopes divided by the integral */
    int n_int;
} pt_beta_table;

/* one decay's products (energies in eV, angles in rad, polarisations +-1) */
typedef struct {
    double Ep, phip, thetap;
    double Ee, phie, thetae;
    int polp, pole;
} pt_decay_out;

/* std::piecewise_linear_distribution<double>::operator() for the variate p in [0, 1) */
PTK_FN double ptk_beta_sample(const pt_beta_table *T, double p) {
    int lo = 0, hi = T->n_int;                 /* lower_bound: first cp[i] >= p */
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (T->cp[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    const int i = lo;
    const double pref = i > 0 ? T->cp[i - 1] : 0.0;
    const double a = 0.5 * T->m[i];
    const double b = T->den[i];
    const double cm = p - pref;
    double x = T->x[i];
    if (a == 0) x += cm / b;
    else {
        const double d = b * b + 4.0 * a * cm;
        x += 0.5 * (__builtin_sqrt(d) - b) / a;
    }
    return x;
}

/* BOOST (src/globals.cpp:67-78): boost the four-vector p by the velocity beta */
PTK_FN void ptk_boost(const double beta[3], double p[4]) {
    double b2 = beta[0] * beta[0] + beta[1] * beta[1] + beta[2] * beta[2];
    double gamma = 1.0 / __builtin_sqrt(1.0 - b2);
    double bp = beta[0] * p[1] + beta[1] * p[2] + beta[2] * p[3];
    double gamma2 = b2 > 0 ? (gamma - 1.0) / b2 : 0.0;
    p[1] = (p[1] + gamma2 * bp * beta[0] + gamma * beta[0] * p[0]);
    p[2] = (p[2] + gamma2 * bp * beta[1] + gamma * beta[1] * p[0]);
    p[3] = (p[3] + gamma2 * bp * beta[2] + gamma * beta[2] * p[0]);
    p[0] = (gamma * (p[0] + bp));
}

/* TNeutron::Decay for a neutron with velocity v[3] (m/s); draws continue the
   neutron's stream (key, *ctr) */
PTK_FN void ptk_neutron_decay(const pt_beta_table *T, uint64_t key, uint64_t *ctr, const double v[3],
                              pt_decay_out *o) {
    const double ele_e = 1.602176487E-19, c_0 = 299792458.;
    const double pi = 3.1415926535897932384626;
    const double m_n = 1.674927211E-27 / ele_e, m_p = 1.672621637E-27 / ele_e, m_e = 9.10938215e-31 / ele_e;
    const double m_nue = 1 / (c_0 * c_0);                       /* 1 eV/c^2 */
    const double twopi = 2 * pi;                                 /* uniform_real_distribution(0, 2 pi) */
    const double half_pi = pi / 2.;                              /* sin_distribution(0, pi/2) */
    double pp[4], pe[4];
    /* step 1: proton energy and momentum */
    double Ekin = ptk_beta_sample(T, ptm_canonical(ptm_draw_u64(key, ctr)));
    pp[0] = Ekin / c_0 + m_p * c_0;
    double pabs = __builtin_sqrt(pp[0] * pp[0] - m_p * m_p * c_0 * c_0);
    /* step 2: isotropic emission in the neutron frame */
    double beta1 = ptm_canonical(ptm_draw_u64(key, ctr)) * (twopi - 0.0) + 0.0;
    double u1 = ptm_canonical(ptm_draw_u64(key, ctr));
    double delta1 = ptm_acos(ptm_cos(0.0) - u1 * (ptm_cos(0.0) - ptm_cos(half_pi)));
    pp[1] = pabs * ptm_sin(delta1) * ptm_cos(beta1);
    pp[2] = pabs * ptm_sin(delta1) * ptm_sin(beta1);
    pp[3] = pabs * ptm_cos(delta1);
    /* step 3: intermediate virtual state */
    double virt[4] = {m_n * c_0 - pp[0], -pp[1], -pp[2], -pp[3]};
    double m2_virt = virt[0] * virt[0] - virt[1] * virt[1] - virt[2] * virt[2] - virt[3] * virt[3];
    /* step 4: electron energy from the two-body decay of the virtual state */
    pe[0] = (m2_virt + m_e * m_e * c_0 * c_0 - m_nue * m_nue * c_0 * c_0) / 2 / __builtin_sqrt(m2_virt);
    pabs = __builtin_sqrt(pe[0] * pe[0] - m_e * m_e * c_0 * c_0);
    /* step 5: isotropic emission in the virtual state's frame */
    double beta2 = ptm_canonical(ptm_draw_u64(key, ctr)) * (twopi - 0.0) + 0.0;
    double u2 = ptm_canonical(ptm_draw_u64(key, ctr));
    double delta2 = ptm_acos(ptm_cos(0.0) - u2 * (ptm_cos(0.0) - ptm_cos(half_pi)));
    pe[1] = pabs * ptm_sin(delta2) * ptm_cos(beta2);
    pe[2] = pabs * ptm_sin(delta2) * ptm_sin(beta2);
    pe[3] = pabs * ptm_cos(delta2);
    /* step 6: electron into the frame of the virtual state */
    double bv[3] = {virt[1] / virt[0], virt[2] / virt[0], virt[3] / virt[0]};
    ptk_boost(bv, pe);
    /* step 7: neutrino by four-momentum conservation */
    double pnue[4] = {virt[0] - pe[0], virt[1] - pe[1], virt[2] - pe[2], virt[3] - pe[3]};
    /* step 8: everything into the moving neutron's frame */
    double bn[3] = {v[0] / c_0, v[1] / c_0, v[2] / c_0};
    ptk_boost(bn, pe);
    ptk_boost(bn, pp);
    ptk_boost(bn, pnue);
    (void)pnue;
    /* energies and angles */
    double beta_p2 = (pp[1] * pp[1] + pp[2] * pp[2] + pp[3] * pp[3]) / pp[0] / pp[0];
    o->Ep = m_p * c_0 * c_0 * beta_p2 * (0.5 + beta_p2 * (3. / 8. + beta_p2 * (5. / 16.)));
    o->Ee = pe[0] * c_0 - m_e * c_0 * c_0;
    o->phip = ptm_atan2(pp[2], pp[1]);
    o->phie = ptm_atan2(pe[2], pe[1]);
    o->thetap = ptm_acos(pp[3] / __builtin_sqrt(pp[1] * pp[1] + pp[2] * pp[2] + pp[3] * pp[3]));
    o->thetae = ptm_acos(pe[3] / __builtin_sqrt(pe[1] * pe[1] + pe[2] * pe[2] + pe[3] * pe[3]));
    o->polp = ptm_canonical(ptm_draw_u64(key, ctr)) < 0.5 ? 1 : -1;
    o->pole = ptm_canonical(ptm_draw_u64(key, ctr)) < 0.5 ? 1 : -1;
}

#endif

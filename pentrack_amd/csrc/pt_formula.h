/*
 * pt_formula.h -- evaluation of compiled field-scaling formulas f(t).
 *
 * PENTrack scales every field container by exprtk formulas of the time t
 * (TFieldScaler, src/field.cpp:7-45, built by TFieldManager from the FIELDS and
 * FORMULAS sections, src/fields.cpp:19-136).  pt_formula.cpp compiles a formula
 * into the postfix program evaluated here on the host and on the GPU; the
 * compiler reproduces the value semantics of the exprtk version vendored with
 * the reference (literal parsing, operator precedence, constant folding,
 * integer powers by fixed multiplication chains, and the constant
 * re-association of its optimiser) for the formula forms it accepts.
 *
 * Plain C++ usable from host and device translation units.
 */
#ifndef PT_FORMULA_H_
#define PT_FORMULA_H_

#include <stdint.h>

#include "pt_detmath.h"

#if defined(__HIPCC__)
#define PTF_FN static inline __host__ __device__
#else
#define PTF_FN static inline
#endif

enum {
    PF_CONST = 0,   /* push constant arg */
    PF_T,           /* push t */
    PF_NEG,
    PF_ADD, PF_SUB, PF_MUL, PF_DIV,
    PF_POW,         /* std::pow semantics (x^y for a non-integer or large exponent) */
    PF_IPOW,        /* x^arg, arg in 1..60, by exprtk's multiplication chains */
    PF_IPOWINV,     /* 1 / x^arg */
    PF_LT, PF_LE, PF_GT, PF_GE, PF_EQ, PF_NE,
    PF_AND, PF_OR, PF_NOT,
    PF_SEL,         /* c a b -> c != 0 ? a : b */
    PF_ABS, PF_SQRT, PF_EXP, PF_LOG, PF_SIN, PF_COS,
    PF_MIN, PF_MAX,
    PF_PBETA,       /* ProtonBetaSpectrum(x) (src/mc.cpp:84-92), host programs only */
    PF_NOPS
};

/* the proton recoil spectrum of free-neutron decay (a host function, pt_host.cpp) */
double ptf_proton_beta_spectrum(double E);

#define PT_FORMULA_MAX_STACK 16
#define PT_FORMULA_MAX_CODE 256

/* x^n for n >= 0: the specialised chains for n <= 10, binary powering above
   (exprtk fast_exp) */
PTF_FN double ptf_fast_exp(double v, int n) {
    switch (n) {
    case 0: return 1.0;
    case 1: return v;
    case 2: return v * v;
    case 3: return v * v * v;
    case 4: { double v2 = v * v; return v2 * v2; }
    case 5: { double v2 = v * v; double v4 = v2 * v2; return v4 * v; }
    case 6: { double v3 = v * v * v; return v3 * v3; }
    case 7: { double v3 = v * v * v; double v6 = v3 * v3; return v6 * v; }
    case 8: { double v2 = v * v; double v4 = v2 * v2; return v4 * v4; }
    case 9: { double v2 = v * v; double v4 = v2 * v2; double v8 = v4 * v4; return v8 * v; }
    case 10: { double v2 = v * v; double v4 = v2 * v2; double v5 = v4 * v; return v5 * v5; }
    default: {
        unsigned k = (unsigned)n;
        double l = 1.0;
        while (k) {
            if (k & 1) { l *= v; --k; }
            v *= v;
            k >>= 1;
        }
        return l;
    }
    }
}

/* x^y for a non-integer or large exponent: exp(y log|x|), with the sign of an odd
   integer power of a negative base and NaN for a negative base otherwise (C pow) */
PTF_FN double ptf_pow(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (x == 1.0) return 1.0;
    if (x < 0.0) {
        double ay = ptm_fabs(y);
        int odd;
        if (ay >= 4503599627370496.0) {                 /* |y| >= 2^52: an integer */
            odd = ay < 9007199254740992.0 && (((uint64_t)ay) & 1u);
        } else {
            double r = ptm_rint(y);
            if (r != y) return (x - x) / (x - x);       /* NaN */
            odd = (int)(((int64_t)r) & 1);
        }
        double v = ptm_pow(-x, y);
        return odd ? -v : v;
    }
    return ptm_pow(x, y);
}

/* program: code[i] = op | arg << 8; constants in cst.  The compiler guarantees the
   stack depth stays within PT_FORMULA_MAX_STACK and the program leaves one value. */
PTF_FN double ptf_run(const uint32_t *code, int n, const double *cst, double t) {
    double s[PT_FORMULA_MAX_STACK];
    int sp = 0;
    for (int i = 0; i < n; ++i) {
        const uint32_t op = code[i] & 0xffu, arg = code[i] >> 8;
        double a, b;
        switch (op) {
        case PF_CONST: s[sp++] = cst[arg]; break;
        case PF_T: s[sp++] = t; break;
        case PF_NEG: s[sp - 1] = -s[sp - 1]; break;
        case PF_NOT: s[sp - 1] = (s[sp - 1] == 0.0) ? 1.0 : 0.0; break;
        case PF_ABS: a = s[sp - 1]; s[sp - 1] = (a < 0.0) ? -a : a; break;
        case PF_SQRT: s[sp - 1] = __builtin_sqrt(s[sp - 1]); break;
        case PF_EXP: s[sp - 1] = ptm_exp(s[sp - 1]); break;
        case PF_LOG: s[sp - 1] = ptm_log(s[sp - 1]); break;
        case PF_SIN: s[sp - 1] = ptm_sin(s[sp - 1]); break;
        case PF_COS: s[sp - 1] = ptm_cos(s[sp - 1]); break;
        case PF_IPOW: s[sp - 1] = ptf_fast_exp(s[sp - 1], (int)arg); break;
#if defined(__HIP_DEVICE_COMPILE__)
        case PF_PBETA: s[sp - 1] = __builtin_nan(""); break;
#else
        case PF_PBETA: s[sp - 1] = ptf_proton_beta_spectrum(s[sp - 1]); break;
#endif
        case PF_IPOWINV: s[sp - 1] = 1.0 / ptf_fast_exp(s[sp - 1], (int)arg); break;
        case PF_SEL: {
            double c = s[sp - 3];
            s[sp - 3] = (c != 0.0) ? s[sp - 2] : s[sp - 1];
            sp -= 2;
            break;
        }
        default:
            b = s[--sp];
            a = s[sp - 1];
            switch (op) {
            case PF_ADD: a = a + b; break;
            case PF_SUB: a = a - b; break;
            case PF_MUL: a = a * b; break;
            case PF_DIV: a = a / b; break;
            case PF_POW: a = ptf_pow(a, b); break;
            case PF_LT: a = (a < b) ? 1.0 : 0.0; break;
            case PF_LE: a = (a <= b) ? 1.0 : 0.0; break;
            case PF_GT: a = (a > b) ? 1.0 : 0.0; break;
            case PF_GE: a = (a >= b) ? 1.0 : 0.0; break;
            case PF_EQ: a = (a == b) ? 1.0 : 0.0; break;
            case PF_NE: a = (a != b) ? 1.0 : 0.0; break;
            case PF_AND: a = (a != 0.0 && b != 0.0) ? 1.0 : 0.0; break;
            case PF_OR: a = (a != 0.0 || b != 0.0) ? 1.0 : 0.0; break;
            case PF_MIN: a = (b < a) ? b : a; break;       /* std::min(a, b) */
            case PF_MAX: a = (a < b) ? b : a; break;       /* std::max(a, b) */
            default: break;
            }
            s[sp - 1] = a;
            break;
        }
    }
    return s[0];
}

#endif

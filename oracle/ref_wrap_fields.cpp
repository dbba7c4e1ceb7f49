/*
 * ref_wrap_fields.cpp -- TEST INFRASTRUCTURE ONLY: a C interface to the
 * reference's own field classes that build without Boost, compiled from their
 * sources where they lie (oracle/ref_build.sh -> oracle/_ref/libref_islands.so):
 *
 *   TFieldScaler, TFieldBoundaryBox, TFieldContainer   src/field.cpp (exprtk formulas)
 *   TExponentialFieldX, TLinearFieldZ, TB0GradZ, TB0GradX2, TB0GradXY, TB0_XY
 *                                                      src/analyticFields.cpp
 *
 * Used to pin the formula compiler, the analytic fields and the container's
 * scaling and boundary taper of the engine and of the oracle against the
 * reference code itself.  Nothing in the product links or loads this.
 */
#include <cstdio>
#include <memory>
#include <stdexcept>

#include "analyticFields.h"
#include "field.h"

namespace {
/* field type codes of include/pentrack_gpu.h */
std::unique_ptr<TField> make_field(int type, const double *p) {
    switch (type) {
    case 4: return std::unique_ptr<TField>(new TExponentialFieldX(p[0], p[1], p[2], p[3], p[4]));
    case 5: return std::unique_ptr<TField>(new TLinearFieldZ(p[0], p[1]));
    case 6: return std::unique_ptr<TField>(new TB0GradZ(p[0], p[1], p[2]));
    case 7: return std::unique_ptr<TField>(new TB0GradX2(p[0], p[1], p[2], p[3]));
    case 8: return std::unique_ptr<TField>(new TB0GradXY(p[0], p[1], p[2]));
    case 9: return std::unique_ptr<TField>(new TB0_XY(p[0], p[1]));
    }
    return nullptr;
}
}  // namespace

extern "C" {

/* TFieldScaler(formula).scalingFactor(t[i]); -1 if exprtk rejects the formula */
int ref_scaler_eval(const char *formula, int n, const double *t, double *out) {
    try {
        TFieldScaler s(formula);
        for (int i = 0; i < n; ++i) out[i] = s.scalingFactor(t[i]);
        return 0;
    } catch (std::exception &e) {
        fprintf(stderr, "ref_scaler_eval: %s\n", e.what());
        return -1;
    }
}

/* TFieldContainer(field, bscale, "0", box..., bw).BField at pts[i] and t[i];
   dB == nullptr requests no derivatives */
int ref_container_B(int type, const double *p, const char *bscale, const double *box, double bw, int n,
                    const double *pts, const double *t, double *B, double *dB) {
    try {
        std::unique_ptr<TField> f = make_field(type, p);
        if (!f) return -2;
        TFieldContainer c(std::move(f), bscale, "0", box[0], box[1], box[2], box[3], box[4], box[5], bw);
        for (int i = 0; i < n; ++i) {
            double b[3], db[3][3];
            c.BField(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], t[i], b, dB ? db : nullptr);
            for (int k = 0; k < 3; ++k) {
                B[3 * i + k] = b[k];
                if (dB)
                    for (int j = 0; j < 3; ++j) dB[9 * i + 3 * k + j] = db[k][j];
            }
        }
        return 0;
    } catch (std::exception &e) {
        fprintf(stderr, "ref_container_B: %s\n", e.what());
        return -1;
    }
}

}  // extern "C"

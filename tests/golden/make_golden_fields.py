"""Writes tests/golden/reference_fields.npz: input/output vectors of the reference's
own field classes, computed with oracle/_ref/libref_islands.so built by
oracle/ref_build.sh from the sources under /root/reference:

  * TFieldScaler::scalingFactor (src/field.cpp:7-10, exprtk) for a fixed list of
    formulas and random formulas of the grammar the engine accepts, at fixed times;
  * TFieldContainer::BField (src/field.cpp:139-153) around TExponentialFieldX,
    TLinearFieldZ, TB0GradZ, TB0GradX2, TB0GradXY, TB0_XY (src/analyticFields.cpp),
    with time-dependent scale formulas and boundary boxes with a taper.

Run here, where the reference is mounted:

    bash oracle/ref_build.sh && python tests/golden/make_golden_fields.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_islands.so")

FIXED_FORMULAS = [
    "1", "0", "0.5", "t", "2*t", "0.01*(t-400)", "(t-400)/100", "1 - t/100", "0.5 + 0.5*t",
    "t < 400 ? 0 : (t < 500 ? 0.01*(t-400) : (t < 700 ? 1 : (t<800?0.01*(800-t) : 0)))",
    "(t + 0.3) + 0.7", "0.3 + (t + 0.7)", "(0.3 - t) - 0.7", "(t - 0.3) + 0.7", "(t * 0.3) * 0.7",
    "(t / 0.3) * 0.7", "0.3 / (t / 0.7)", "0.3 / (0.7 * t)", "(0.3 * t) / 0.7", "t^2", "t^3", "t^6", "t^11",
    "t^(-2)", "(t+1)^5", "-t^2", "2^3^2", "t^2^3", "1e-3*t", "2.5e+2*t", "49057e23*t", "44866e-25*t",
    "0.000123456789012345678*t", ".5*t", "5.*t", "t*1", "1*t", "t/1", "t+0", "0+t", "t*0", "0*t", "0/t",
    "t/0", "T*PI", "pi*t*2", "epsilon*t", "t == 1", "t != 400", "t > 1 and t < 3 or t > 500",
    "not(t > 1)", "if(t < 1, 5, 6)", "min(t, 2)", "max(t, 2)", "abs(t - 500)/100", "sqrt(t)",
    "t <= 500 & t >= 400", "t < 500 | t > 700", "t <> 400", "t = 400", "(t+1)*(t-1)", "t*t*t",
    "(t >= 400) * (t < 500)", "1 + 2 * t ^ 2",
]
# formulas with exp/log/sin/cos or a non-integer power: last bits follow the C library
TRANSCENDENTAL_FORMULAS = [
    "sin(2*pi*t)", "exp(-t/3)", "log(t + 11)", "cos(t)^3", "t^0.5", "(t+1)^2.5", "t^61", "(t^2.5)^1.5",
    "1 - exp(-t/100)", "0.5 + 0.5*cos(pi*t/1000)",
]


def load():
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    L.ref_scaler_eval.argtypes = [C.c_char_p, C.c_int, dp, dp]
    L.ref_container_B.argtypes = [C.c_int, dp, C.c_char_p, dp, C.c_double, C.c_int, dp, dp, dp, dp]
    return L


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def scaler(L, f, t):
    out = np.empty_like(t)
    if L.ref_scaler_eval(f.encode(), len(t), dp(t), dp(out)) != 0:
        raise ValueError(f"exprtk rejects {f!r}")
    return out


def main():
    from oracle import formula as OF
    from pentrack_amd import _native
    L = load()
    rng = np.random.default_rng(20261204)
    t = np.concatenate([[0.0, 1.0, 2.5, 400.0, 450.0, 500.0, 650.0, 700.0, 750.0, 800.0, 1e-9],
                        rng.uniform(0, 1000, 60), rng.uniform(0, 1, 20)])
    # random formulas of the accepted grammar (the engine's compiler decides acceptance)
    P = _native.lib()
    fuzz = []
    frng = np.random.default_rng(11)
    while len(fuzz) < 300:
        f = OF.random_formula(frng, depth=int(frng.integers(1, 5)))
        out = np.empty(1)
        one = np.zeros(1)
        if P.pt_formula_eval(f.encode(), 1, dp(one), dp(out)) == 0:
            fuzz.append(f)
    formulas = FIXED_FORMULAS + fuzz + TRANSCENDENTAL_FORMULAS
    vals = np.stack([scaler(L, f, t) for f in formulas])
    exact = np.array([f not in TRANSCENDENTAL_FORMULAS for f in formulas])

    # containers around the analytic fields
    specs = {
        4: lambda: [rng.uniform(1e-5, 1e-4), rng.uniform(0.5, 2.0), rng.uniform(-5, -3), rng.uniform(-1e-6, 1e-6),
                    rng.uniform(-1e-6, 1e-6)],
        5: lambda: [rng.uniform(-1e-5, 1e-5), rng.uniform(-1e-5, 1e-5)],
        6: lambda: [rng.uniform(-1e-5, 1e-5), rng.uniform(-1e-6, 1e-6), rng.uniform(1e-6, 1e-5)],
        7: lambda: [rng.uniform(-1e-7, 1e-7), rng.uniform(-1e-7, 1e-7), rng.uniform(-1e-7, 1e-7), rng.uniform(1e-6, 1e-5)],
        8: lambda: [rng.uniform(-1e-7, 1e-7), rng.uniform(-1e-6, 1e-6), rng.uniform(1e-6, 1e-5)],
        9: lambda: [rng.uniform(-1e-6, 1e-6), rng.uniform(1e-6, 1e-5)],
    }
    scales = ["1", "2.5", "t < 0.5 ? 0 : 0.01*(t-0.5)", "0.5 + 0.5*t"]
    boxes = [((0, 0, 0, 0, 0, 0), 0.0), ((1, -1, 1, -1, 1, -1), 0.0), ((1, -1, 0.8, -0.6, 1.2, -0.5), 0.2)]
    c_type, c_par, c_scale, c_box, c_bw, c_pts, c_t, c_B, c_dB, c_B0 = [], [], [], [], [], [], [], [], [], []
    n = 24
    for ty, gen in specs.items():
        for si, sc in enumerate(scales):
            for bi, (box, bw) in enumerate(boxes):
                par = np.zeros(8)
                p = gen()
                par[:len(p)] = p
                pts = rng.uniform(-1.3, 1.3, (n, 3))
                tt = rng.uniform(0, 2, n)
                B = np.zeros((n, 3))
                dB = np.zeros((n, 3, 3))
                B0 = np.zeros((n, 3))
                bx = np.array(box, dtype=np.float64)
                assert L.ref_container_B(ty, dp(par), sc.encode(), dp(bx), bw, n, dp(pts), dp(tt), dp(B), dp(dB)) == 0
                assert L.ref_container_B(ty, dp(par), sc.encode(), dp(bx), bw, n, dp(pts), dp(tt), dp(B0),
                                         C.POINTER(C.c_double)()) == 0
                c_type.append(ty); c_par.append(par); c_scale.append(sc); c_box.append(bx); c_bw.append(bw)
                c_pts.append(pts); c_t.append(tt); c_B.append(B); c_dB.append(dB); c_B0.append(B0)
    np.savez_compressed(
        os.path.join(HERE, "reference_fields.npz"),
        formulas=np.array(formulas), t=t, values=vals, exact=exact,
        c_type=np.array(c_type), c_par=np.array(c_par), c_scale=np.array(c_scale), c_box=np.array(c_box),
        c_bw=np.array(c_bw), c_pts=np.array(c_pts), c_t=np.array(c_t), c_B=np.array(c_B), c_dB=np.array(c_dB),
        c_B0=np.array(c_B0))
    print("wrote", os.path.join(HERE, "reference_fields.npz"), len(formulas), "formulas,", len(c_type), "containers")


if __name__ == "__main__":
    main()

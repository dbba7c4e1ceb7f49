"""Log filters and formula log variables (EvalFormula, src/config.cpp:195-211; TLogger::Log,
src/logger.cpp:286-308): known values of exprtk expressions and the filtered columns."""
import numpy as np
import pytest

from pentrack_amd import logformula as F


@pytest.mark.parametrize("expr,want", [
    ("1 + 2*3", 7.0), ("(1 + 2)*3", 9.0), ("2^3^2", 512.0), ("2^-1", 0.5), ("-2 + 3", 1.0), ("10 % 3", 1.0),
    ("7/2", 3.5), ("1 < 2 and 3 > 4", 0.0), ("1 < 2 or 3 > 4", 1.0), ("not (1 = 1)", 0.0), ("2 = 2", 1.0),
    ("1 != 1", 0.0), ("1 xor 1", 0.0), ("0 ? 5 : 6", 6.0), ("if(1, 5, 6)", 5.0), ("min(3, -1) + max(2, 4)", 3.0),
    ("sqrt(3^2 + 4^2)", 5.0), ("hypot(5, 12)", 13.0), ("abs(-2.5)", 2.5), ("floor(-1.5)", -2.0),
    ("round(2.5)", 3.0), ("round(-2.5)", -3.0), ("sgn(-4)", -1.0), ("cos(pi)", -1.0), ("atan2(1, 1)*4", np.pi),
    ("log10(1000)", 3.0), ("exp(0)", 1.0), ("1e3 + .5", 1000.5), ("2 > 1 & 3 > 2", 1.0), ("0 | 0", 0.0),
])
def test_known_values(expr, want):
    assert F.evaluate(expr, {}, 1)[0] == pytest.approx(want, rel=1e-15)


def test_columns_filter_and_formula_titles():
    vals = {"t": np.array([0.1, 0.6, 0.9]), "vx": np.array([3.0, 0.0, 6.0]), "vy": np.array([4.0, 1.0, 8.0]),
            "stopID": np.array([-1.0, 1.0, -1.0])}
    formulas = {"speed": "sqrt(vx^2 + vy^2)", "late": "t > 0.5 and stopID = -1"}
    cols = F.apply(formulas, "late", ["t", "speed"], vals)
    assert list(cols) == ["t", "speed"]
    assert np.array_equal(cols["t"], [0.9]) and np.array_equal(cols["speed"], [10.0])
    cols = F.apply(formulas, "", ["speed", "stopID"], vals)
    assert np.array_equal(cols["speed"], [5.0, 1.0, 10.0]) and np.array_equal(cols["stopID"], vals["stopID"])
    with pytest.raises(F.FormulaError, match="not found"):
        F.apply(formulas, "", ["t", "energy"], vals)
    with pytest.raises(F.FormulaError, match="unknown variable"):
        F.apply({"bad": "q + 1"}, "bad", ["t"], vals)
    with pytest.raises(F.FormulaError):
        F.evaluate("1 +", {}, 1)

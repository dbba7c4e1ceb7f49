"""Log filters and formula log variables: `<log>logfilter` and names in `<log>logvars`
that are not log variables refer to [FORMULAS] entries, exprtk expressions over the log's
variables (EvalFormula, src/config.cpp:195-211, used by TLogger::Log,
src/logger.cpp:286-308).  Evaluated here over whole columns with numpy.

The accepted expression language is the part of exprtk that such formulas use: numbers,
variables, + - * / % ^ (right-associative power), unary + -, comparisons < <= > >= == !=
and = (equality), logical and / or / xor / not / nand / nor and & |, the ternary
`c ? a : b`, if(c, a, b), and the functions abs sqrt exp log log10 log2 sin cos tan asin
acos atan sinh cosh tanh atan2 pow min max floor ceil round trunc sgn hypot frac, and the
constants pi, epsilon, inf.  Anything else raises FormulaError.
"""
import re

import numpy as np


class FormulaError(ValueError):
    pass


_TOKEN = re.compile(r"\s*(?:(\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?)|([A-Za-z_][A-Za-z0-9_.]*)"
                    r"|(<=|>=|==|!=|<>|&&|\|\||[-+*/%^()<>=,?:&|!]))")
_F1 = {"abs": np.abs, "sqrt": np.sqrt, "exp": np.exp, "log": np.log, "log10": np.log10, "log2": np.log2,
       "sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos, "atan": np.arctan,
       "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh, "floor": np.floor, "ceil": np.ceil,
       "round": lambda a: np.where(a < 0, -np.floor(-a + 0.5), np.floor(a + 0.5)), "trunc": np.trunc,
       "sgn": np.sign, "frac": lambda a: a - np.trunc(a)}
_F2 = {"atan2": np.arctan2, "pow": np.power, "min": np.minimum, "max": np.maximum, "hypot": np.hypot}
_CONST = {"pi": np.pi, "epsilon": np.finfo(float).eps, "inf": np.inf}
_B = lambda x: np.asarray(x) != 0
_BIN = {  # operator: (precedence, right-assoc, function)
    "or": (1, False, lambda a, b: (_B(a) | _B(b)) * 1.0), "||": (1, False, lambda a, b: (_B(a) | _B(b)) * 1.0),
    "|": (1, False, lambda a, b: (_B(a) | _B(b)) * 1.0), "nor": (1, False, lambda a, b: (~(_B(a) | _B(b))) * 1.0),
    "xor": (1, False, lambda a, b: (_B(a) ^ _B(b)) * 1.0),
    "and": (2, False, lambda a, b: (_B(a) & _B(b)) * 1.0), "&&": (2, False, lambda a, b: (_B(a) & _B(b)) * 1.0),
    "&": (2, False, lambda a, b: (_B(a) & _B(b)) * 1.0), "nand": (2, False, lambda a, b: (~(_B(a) & _B(b))) * 1.0),
    "<": (3, False, lambda a, b: (a < b) * 1.0), "<=": (3, False, lambda a, b: (a <= b) * 1.0),
    ">": (3, False, lambda a, b: (a > b) * 1.0), ">=": (3, False, lambda a, b: (a >= b) * 1.0),
    "==": (3, False, lambda a, b: (a == b) * 1.0), "=": (3, False, lambda a, b: (a == b) * 1.0),
    "!=": (3, False, lambda a, b: (a != b) * 1.0), "<>": (3, False, lambda a, b: (a != b) * 1.0),
    "+": (4, False, np.add), "-": (4, False, np.subtract),
    "*": (5, False, np.multiply), "/": (5, False, np.divide), "%": (5, False, np.fmod),
    "^": (7, True, np.power),
}


def _tokens(text):
    out, pos = [], 0
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise FormulaError(f"cannot read formula at {text[pos:]!r}")
        num, name, op = m.groups()
        out.append(("num", float(num)) if num else ("name", name) if name else ("op", op))
        pos = m.end()
    return out


class _Parser:
    def __init__(self, text, variables):
        self.t = _tokens(text)
        self.i = 0
        self.v = variables

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, op=None):
        tok = self.peek()
        if op is not None and tok[1] != op:
            raise FormulaError(f"expected {op!r}, found {tok[1]!r}")
        self.i += 1
        return tok

    def expr(self):
        c = self.binary(1)
        if self.peek()[1] == "?":
            self.take("?")
            a = self.expr()
            self.take(":")
            b = self.expr()
            c = np.where(_B(c), a, b)
        return c

    def binary(self, minp):
        lhs = self.unary()
        while True:
            kind, op = self.peek()
            if kind not in ("op", "name") or op not in _BIN or _BIN[op][0] < minp:
                return lhs
            p, right, f = _BIN[op]
            self.take()
            rhs = self.binary(p if right else p + 1) if p != 7 else self.power_rhs()
            with np.errstate(all="ignore"):
                lhs = f(lhs, rhs)

    def power_rhs(self):
        """the right operand of ^: a unary expression, itself raised by a further ^"""
        base = self.unary()
        if self.peek()[1] == "^":
            self.take()
            with np.errstate(all="ignore"):
                return np.power(base, self.power_rhs())
        return base

    def unary(self):
        kind, tok = self.peek()
        if tok in ("-", "+"):
            self.take()
            a = self.binary(6)
            return -a if tok == "-" else a
        if tok in ("not", "!"):
            self.take()
            return (~_B(self.binary(3))) * 1.0
        return self.atom()

    def atom(self):
        kind, tok = self.take()
        if kind == "num":
            return tok
        if tok == "(":
            a = self.expr()
            self.take(")")
            return a
        if kind != "name":
            raise FormulaError(f"unexpected {tok!r}")
        if self.peek()[1] == "(":
            self.take("(")
            args = [self.expr()]
            while self.peek()[1] == ",":
                self.take(",")
                args.append(self.expr())
            self.take(")")
            with np.errstate(all="ignore"):
                if tok in _F1 and len(args) == 1:
                    return _F1[tok](np.asarray(args[0], dtype=float))
                if tok in _F2 and len(args) == 2:
                    return _F2[tok](args[0], args[1])
                if tok == "if" and len(args) == 3:
                    return np.where(_B(args[0]), args[1], args[2])
            raise FormulaError(f"function {tok} with {len(args)} arguments is not supported")
        if tok in self.v:
            return self.v[tok]
        if tok in _CONST:
            return _CONST[tok]
        raise FormulaError(f"unknown variable {tok!r}")


def evaluate(text, variables, n):
    """The expression over the variables (name -> array of n, or number): an array of n."""
    p = _Parser(text, variables)
    r = p.expr()
    if p.i != len(p.t):
        raise FormulaError(f"unexpected {p.t[p.i][1]!r} in {text!r}")
    return np.broadcast_to(np.asarray(r, dtype=float), (n,)).copy()


def apply(formulas, filt, titles, vals):
    """TLogger::Log: rows for which the filter formula is nonzero (all when filt is empty),
    and every title that is not a variable evaluated as the formula of that name.
    formulas: {name: expression} of [FORMULAS].  Returns the columns, titles in order."""
    n = len(next(iter(vals.values()))) if vals else 0

    def formula(name):
        if name not in formulas:
            raise FormulaError(f"Formula {name} not found in config file")
        return evaluate(formulas[name], vals, n)

    cols = {t: (np.asarray(vals[t], dtype=float) if t in vals else formula(t)) for t in titles}
    if filt:
        keep = formula(filt) != 0
        cols = {t: c[keep] for t, c in cols.items()}
    return cols

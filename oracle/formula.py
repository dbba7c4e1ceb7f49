"""ORACLE (test infrastructure only): field-scaling formulas f(t) as exprtk reads them.

The reference scales each field container by an exprtk formula of the time t
(TFieldScaler, src/field.cpp:7-45; exprtk/exprtk.hpp vendored with the reference).
This module restates, for the formula forms the engine accepts, the value exprtk
gives such a formula:

  * literals: exprtk's string_to_real (exprtk.hpp:1735-1890) -- integer digits
    accumulated as d*10+digit, fraction digits likewise and divided by 10^k
    (compute_pow10, exprtk.hpp:1400-1455), then one multiplication or division
    by 10^|e| for an exponent;
  * precedence (exprtk.hpp:19927-20005): or < and < comparisons < + - < * / < ^,
    ^ right associative; unary minus applies to the power; two adjacent signs are
    joined as exprtk's lexer does ('+-', '-+' -> '-', '--' -> '+', exprtk.hpp:3195-3235);
  * folding of constant sub-trees with the C library (Python's math calls it);
  * x^c, c integer, |c| <= 60 (exprtk.hpp:27105-27150): fast_exp chains
    (exprtk.hpp:1463-1496), 1/x^|c| for c < 0, 1 for c = 0; (t^c0)^c1 with a
    non-integer c1 becomes t^(c0*c1) (exprtk.hpp:29312);
  * c o t and t o c with c = 0 or 1 (exprtk.hpp:28117-28195);
  * the re-association of (c0 o t) o c1, c0 o (t o c1), c0 o (c1 o t),
    (t o c0) o c1 for + - and * / (exprtk.hpp:28882-29320).

parse(formula) returns a tree of tuples; nodes() flattens it into the node array
the C oracle evaluates (ptref_fnode in oracle/ptref.h); evaluate() evaluates a
tree in Python (with the C library's transcendental functions).
"""
import math

# node codes shared with oracle/ptref.c (PTREF_F_*)
CONST, T, NEG, ADD, SUB, MUL, DIV, POW, IPOW, IPOWINV = range(10)
LT, LE, GT, GE, EQ, NE, AND, OR, NOT, SEL = range(10, 20)
ABS, SQRT, EXP, LOG, SIN, COS, MIN, MAX = range(20, 28)

_UNARY = {"abs": ABS, "sqrt": SQRT, "exp": EXP, "log": LOG, "sin": SIN, "cos": COS, "not": NOT}


def fast_exp(v, n):
    """exprtk fast_exp<T,N>::result"""
    if n == 0:
        return 1.0
    if n == 1:
        return v
    if n == 2:
        return v * v
    if n == 3:
        return v * v * v
    if n == 4:
        v2 = v * v
        return v2 * v2
    if n == 5:
        return fast_exp(v, 4) * v
    if n == 6:
        v3 = fast_exp(v, 3)
        return v3 * v3
    if n == 7:
        return fast_exp(v, 6) * v
    if n == 8:
        v4 = fast_exp(v, 4)
        return v4 * v4
    if n == 9:
        return fast_exp(v, 8) * v
    if n == 10:
        v5 = fast_exp(v, 5)
        return v5 * v5
    k, r = n, 1.0
    while k:
        if k & 1:
            r *= v
            k -= 1
        v *= v
        k >>= 1
    return r


def _pow10(e):
    return float("1e%d" % e)


def _compute_pow10(d, exponent):
    e = abs(exponent)
    if exponent >= -307:
        if e <= 308:
            return d * _pow10(e) if exponent > 0 else d / _pow10(e)
        return d * math.pow(10.0, 10.0 * exponent)
    d /= _pow10(307)
    return d / _pow10(-exponent - 307)


def _cpow(a, b):
    """C pow (Python's math.pow raises where C returns an infinity)."""
    odd = b == math.floor(b) and abs(b) < 2 ** 53 and int(b) % 2 == 1
    try:
        return math.pow(a, b)
    except OverflowError:
        return -math.inf if (a < 0 and odd) else math.inf
    except ValueError:
        if a == 0.0 and b < 0:
            return math.copysign(math.inf, a) if odd else math.inf
        return math.nan


def _fold(op, a, b=None):
    if op == NEG:
        return -a
    if op == NOT:
        return 1.0 if a == 0.0 else 0.0
    if op == ABS:
        return -a if a < 0.0 else a
    if op in (SQRT, EXP, LOG, SIN, COS):
        f = {SQRT: math.sqrt, EXP: math.exp, LOG: math.log, SIN: math.sin, COS: math.cos}[op]
        try:
            return f(a)
        except (ValueError, OverflowError):
            return {SQRT: math.nan, EXP: math.inf, LOG: (-math.inf if a == 0 else math.nan)}.get(op, math.nan)
    if op == ADD:
        return a + b
    if op == SUB:
        return a - b
    if op == MUL:
        return a * b
    if op == DIV:
        if b == 0.0:
            if a == 0.0 or a != a:
                return math.nan
            return math.copysign(math.inf, a) * math.copysign(1.0, b)
        return a / b
    if op == POW:
        return _cpow(a, b)
    if op in (LT, LE, GT, GE, EQ, NE):
        return 1.0 if {LT: a < b, LE: a <= b, GT: a > b, GE: a >= b, EQ: a == b, NE: a != b}[op] else 0.0
    if op == AND:
        return 1.0 if (a != 0.0 and b != 0.0) else 0.0
    if op == OR:
        return 1.0 if (a != 0.0 or b != 0.0) else 0.0
    if op == MIN:
        return b if b < a else a
    if op == MAX:
        return b if a < b else a
    raise ValueError(op)


class _Parser:
    def __init__(self, s):
        self.s, self.i = s, 0

    def ws(self):
        while self.i < len(self.s) and self.s[self.i].isspace():
            self.i += 1

    def peek(self, tok):
        self.ws()
        return self.s.startswith(tok, self.i)

    def take(self, tok):
        if self.peek(tok):
            self.i += len(tok)
            return True
        return False

    def word(self, w):
        self.ws()
        n = len(w)
        if self.s[self.i:self.i + n].lower() != w:
            return False
        nxt = self.s[self.i + n:self.i + n + 1]
        if nxt and (nxt.isalnum() or nxt == "_"):
            return False
        self.i += n
        return True

    # --- node synthesis
    @staticmethod
    def is_c(x):
        return x[0] == CONST

    @staticmethod
    def is_t(x):
        return x[0] == T

    def pair_tc(self, x):
        return x[0] in (ADD, SUB, MUL, DIV, POW) and len(x) == 3 and (
            (self.is_t(x[1]) and self.is_c(x[2])) or (self.is_c(x[1]) and self.is_t(x[2])))

    def bin(self, op, a, b):
        if self.is_c(a) and self.is_c(b):
            return (CONST, _fold(op, a[1], b[1]))
        if op == POW and self.is_c(b):
            c = b[1]
            if abs(c) <= 60 and c == math.floor(c):
                p = int(abs(c))
                if p == 0:
                    return (CONST, 1.0)
                if c == 1:
                    return a
                return (IPOW if c > 0 else IPOWINV, a, p)
            if a[0] == POW and self.is_t(a[1]) and self.is_c(a[2]):
                return (POW, (T,), (CONST, a[2][1] * c))
            return (POW, a, b)
        addsub, muldiv = (ADD, SUB), (MUL, DIV)
        same = lambda o1, o2: (o1 in addsub and o2 in addsub) or (o1 in muldiv and o2 in muldiv)
        if op in addsub + muldiv:
            if self.is_c(a) and self.pair_tc(b) and same(op, b[0]):
                c0, o1 = a[1], b[0]
                if self.is_t(b[1]):                       # c0 op (t o1 c1)
                    c1 = b[2][1]
                    r = {(ADD, ADD): (ADD, c0 + c1), (ADD, SUB): (ADD, c0 - c1), (SUB, ADD): (SUB, c0 - c1),
                         (SUB, SUB): (SUB, c0 + c1), (MUL, MUL): (MUL, c0 * c1), (MUL, DIV): (MUL, _fold(DIV, c0, c1)),
                         (DIV, MUL): (DIV, _fold(DIV, c0, c1)), (DIV, DIV): (DIV, c0 * c1)}[(op, o1)]
                else:                                     # c0 op (c1 o1 t)
                    c1 = b[1][1]
                    r = {(ADD, ADD): (ADD, c0 + c1), (ADD, SUB): (SUB, c0 + c1), (SUB, ADD): (SUB, c0 - c1),
                         (SUB, SUB): (ADD, c0 - c1), (MUL, MUL): (MUL, c0 * c1), (MUL, DIV): (DIV, c0 * c1),
                         (DIV, MUL): (DIV, _fold(DIV, c0, c1)), (DIV, DIV): (MUL, _fold(DIV, c0, c1))}[(op, o1)]
                return (r[0], (CONST, r[1]), (T,))
            if self.is_c(b) and self.pair_tc(a) and same(op, a[0]):
                c1, o0 = b[1], a[0]
                if self.is_c(a[1]):                       # (c0 o0 t) op c1
                    c0 = a[1][1]
                    r = {(ADD, ADD): (ADD, c0 + c1), (ADD, SUB): (ADD, c0 - c1), (SUB, ADD): (SUB, c0 + c1),
                         (SUB, SUB): (SUB, c0 - c1), (MUL, MUL): (MUL, c0 * c1), (MUL, DIV): (MUL, _fold(DIV, c0, c1)),
                         (DIV, MUL): (DIV, c0 * c1), (DIV, DIV): (DIV, _fold(DIV, c0, c1))}[(o0, op)]
                    return (r[0], (CONST, r[1]), (T,))
                c0 = a[2][1]                              # (t o0 c0) op c1
                r = {(ADD, ADD): (ADD, c0 + c1), (ADD, SUB): (ADD, c0 - c1), (SUB, ADD): (ADD, c1 - c0),
                     (SUB, SUB): (SUB, c0 + c1), (MUL, MUL): (MUL, c0 * c1), (MUL, DIV): (MUL, _fold(DIV, c0, c1)),
                     (DIV, MUL): (MUL, _fold(DIV, c1, c0)), (DIV, DIV): (DIV, c0 * c1)}[(o0, op)]
                return (r[0], (T,), (CONST, r[1]))
            if self.is_c(a) and self.is_t(b):
                c = a[1]
                if c == 0.0 and op in (MUL, DIV):
                    return (CONST, 0.0)
                if (c == 0.0 and op == ADD) or (c == 1.0 and op == MUL):
                    return b
            if self.is_c(b) and self.is_t(a):
                c = b[1]
                if c == 0.0 and op == MUL:
                    return (CONST, 0.0)
                if c == 0.0 and op == DIV:
                    return (CONST, math.nan)
                if (c == 0.0 and op == ADD) or (c == 1.0 and op in (MUL, DIV)):
                    return a
        return (op, a, b)

    def un(self, op, a):
        if self.is_c(a):
            return (CONST, _fold(op, a[1]))
        return (op, a)

    def sel(self, c, a, b):
        if self.is_c(c):
            return a if c[1] != 0.0 else b
        return (SEL, c, a, b)

    # --- grammar
    def cond(self):
        c = self.or_()
        if self.take("?"):
            a = self.cond()
            if not self.take(":"):
                raise ValueError("expected ':'")
            return self.sel(c, a, self.cond())
        return c

    def or_(self):
        a = self.and_()
        while self.word("or") or self.take("|"):
            a = self.bin(OR, a, self.and_())
        return a

    def and_(self):
        a = self.cmp()
        while self.word("and") or self.take("&"):
            a = self.bin(AND, a, self.cmp())
        return a

    def cmp(self):
        a = self.add()
        while True:
            for tok, op in (("<=", LE), (">=", GE), ("==", EQ), ("!=", NE), ("<>", NE), ("<", LT), (">", GT),
                            ("=", EQ)):
                if self.take(tok):
                    a = self.bin(op, a, self.add())
                    break
            else:
                return a

    def sign(self):
        """A + or - operator with exprtk's joining of two adjacent ones ('+-', '-+' -> '-',
        '--' -> '+'): +1, -1 or 0."""
        self.ws()
        if self.i >= len(self.s) or self.s[self.i] not in "+-":
            return 0
        c = self.s[self.i]
        self.i += 1
        self.ws()
        if self.i < len(self.s) and self.s[self.i] in "+-":
            d = self.s[self.i]
            self.i += 1
            return 1 if c == d else -1
        return 1 if c == "+" else -1

    def add(self):
        a = self.mul()
        while True:
            sg = self.sign()
            if sg == 0:
                return a
            a = self.bin(ADD if sg > 0 else SUB, a, self.mul())

    def mul(self):
        a = self.unary()
        while True:
            if self.take("*"):
                a = self.bin(MUL, a, self.unary())
            elif self.take("/"):
                a = self.bin(DIV, a, self.unary())
            else:
                return a

    def unary(self):
        sg = self.sign()
        if sg < 0:
            return self.un(NEG, self.unary())
        if sg > 0:
            return self.unary()
        a = self.primary()
        if self.take("^"):
            return self.bin(POW, a, self.unary())
        return a

    def number(self):
        s, i = self.s, self.i
        d, instate = 0.0, False
        if s[i] != ".":
            st = i
            while i < len(s) and s[i] == "0":
                i += 1
            while i < len(s) and s[i].isdigit():
                d = d * 10.0 + float(ord(s[i]) - 48)
                i += 1
            instate = i != st
        if i < len(s) and s[i] == ".":
            i += 1
            st, f = i, 0.0
            while i < len(s) and s[i].isdigit():
                f = f * 10.0 + float(ord(s[i]) - 48)
                i += 1
            if i != st:
                instate = True
                d += _compute_pow10(f, -(i - st))
        exponent = 0
        if i < len(s) and s[i] in "eE":
            i += 1
            sign = 1
            if s[i] in "+-":
                sign = -1 if s[i] == "-" else 1
                i += 1
            st = i
            while i < len(s) and s[i].isdigit():
                i += 1
            exponent = sign * int(s[st:i])
        if not instate:
            raise ValueError("bad number")
        self.i = i
        if exponent:
            d = _compute_pow10(d, exponent)
        return (CONST, d)

    def args(self, n):
        if not self.take("("):
            raise ValueError("expected (")
        out = []
        for k in range(n):
            out.append(self.cond())
            if k + 1 < n and not self.take(","):
                raise ValueError("expected ,")
        if not self.take(")"):
            raise ValueError("expected )")
        return out

    def primary(self):
        self.ws()
        c = self.s[self.i]
        if c.isdigit() or c == ".":
            return self.number()
        if self.take("("):
            a = self.cond()
            if not self.take(")"):
                raise ValueError("expected )")
            return a
        st = self.i
        while self.i < len(self.s) and (self.s[self.i].isalnum() or self.s[self.i] == "_"):
            self.i += 1
        w = self.s[st:self.i].lower()
        if w == "t":
            return (T,)
        if w == "pi":
            return (CONST, math.pi)
        if w == "epsilon":
            return (CONST, 1e-10)
        if w == "inf":
            return (CONST, math.inf)
        if w in _UNARY:
            return self.un(_UNARY[w], self.args(1)[0])
        if w in ("min", "max"):
            a, b = self.args(2)
            return self.bin(MIN if w == "min" else MAX, a, b)
        if w == "if":
            return self.sel(*self.args(3))
        raise ValueError(f"unknown name {w!r}")


def parse(formula):
    p = _Parser(formula)
    tree = p.cond()
    p.ws()
    if p.i != len(formula):
        raise ValueError("trailing input")
    return tree


def evaluate(tree, t):
    """Python evaluation of a tree (C-library transcendental functions)."""
    k = tree[0]
    if k == CONST:
        return tree[1]
    if k == T:
        return t
    if k in (IPOW, IPOWINV):
        v = fast_exp(evaluate(tree[1], t), tree[2])
        return v if k == IPOW else _fold(DIV, 1.0, v)
    if k == SEL:
        return evaluate(tree[2], t) if evaluate(tree[1], t) != 0.0 else evaluate(tree[3], t)
    if len(tree) == 2:
        return _fold(k, evaluate(tree[1], t))
    return _fold(k, evaluate(tree[1], t), evaluate(tree[2], t))


def nodes(tree):
    """Flatten into [(op, a, b, c, v)] with children before parents; the root is last."""
    out = []

    def walk(x):
        k = x[0]
        if k == CONST:
            out.append((CONST, -1, -1, -1, x[1]))
        elif k == T:
            out.append((T, -1, -1, -1, 0.0))
        elif k in (IPOW, IPOWINV):
            a = walk(x[1])
            out.append((k, a, -1, -1, float(x[2])))
        elif k == SEL:
            a, b, c = walk(x[1]), walk(x[2]), walk(x[3])
            out.append((SEL, a, b, c, 0.0))
        elif len(x) == 2:
            a = walk(x[1])
            out.append((k, a, -1, -1, 0.0))
        else:
            a, b = walk(x[1]), walk(x[2])
            out.append((k, a, b, -1, 0.0))
        return len(out) - 1

    walk(tree)
    return out


_POOL = ["0", "1", "2", "0.5", "0.3", "0.01", "400", "3.7", "1e-3", "2.5e2", "pi", "12500", "0.7"]


def random_formula(rng, depth=3, transcendental=False):
    """A random formula of the accepted grammar (for comparisons with exprtk).
    With transcendental=False only exactly rounded operations appear."""
    def leaf():
        return "t" if rng.random() < 0.5 else _POOL[rng.integers(len(_POOL))]

    def gen(d):
        if d == 0:
            return leaf()
        r = rng.random()
        if r < 0.45:
            op = ["+", "-", "*", "/"][rng.integers(4)]
            return f"({gen(d - 1)}{op}{gen(d - 1)})"
        if r < 0.55:
            op = ["<", "<=", ">", ">=", "==", "!="][rng.integers(6)]
            return f"({gen(d - 1)} {op} {gen(d - 1)})"
        if r < 0.62:
            return f"({gen(d - 1)} ? {gen(d - 1)} : {gen(d - 1)})"
        if r < 0.7:
            fs = ["abs", "sqrt"] + (["exp", "log", "sin", "cos"] if transcendental else [])
            return f"{fs[rng.integers(len(fs))]}({gen(d - 1)})"
        if r < 0.76:
            return f"{['min', 'max'][rng.integers(2)]}({gen(d - 1)}, {gen(d - 1)})"
        if r < 0.82:
            e = [2, 3, 4, 5, 6, 7, 9, 11, -1, -2, 0][rng.integers(11)]
            return f"({gen(d - 1)})^{e}" if e >= 0 else f"({gen(d - 1)})^({e})"
        if r < 0.88:
            return f"-{gen(d - 1)}"
        if r < 0.94:
            return f"({gen(d - 1)} {['and', 'or'][rng.integers(2)]} {gen(d - 1)})"
        return gen(d - 1)

    return gen(depth)

/*
 * pt_formula.cpp -- compiler of field-scaling formulas f(t) into the postfix
 * programs of pt_formula.h.
 *
 * The reference evaluates these formulas with exprtk (TFieldScaler,
 * src/field.cpp:7-45; exprtk/exprtk.hpp vendored with the reference).  The
 * value a formula takes depends on details of that library, which this
 * compiler reproduces:
 *   - numeric literals are read as exprtk's string_to_real does (integer and
 *     fraction digits accumulated in double, the fraction divided by 10^k, the
 *     exponent applied by one multiplication or division by 10^|e|);
 *   - precedence: or < and < comparisons < + - < * / < ^ (right associative);
 *     unary minus binds looser than ^;
 *   - constant sub-expressions are folded at compile time with the C library's
 *     functions, as exprtk folds them when it builds the expression tree;
 *   - x^c for an integer constant |c| <= 60 is a fixed multiplication chain
 *     (x^-c = 1 / x^c, x^0 = 1);
 *   - a constant 0 or 1 operand simplifies c*x, x*c, c/x, x/c, c+x, x+c as
 *     exprtk's node synthesis does (0*x = 0, x/0 = NaN, 0/x = 0, 0+x = x, ...);
 *   - constants on both sides of a variable in one +/- or * / chain are
 *     combined first: (t + 0.3) + 0.7 is evaluated as t + (0.3 + 0.7), and so
 *     on for the 32 forms exprtk re-associates;
 *   - == and != compare exactly; and/or/not/?: treat any non-zero value (NaN
 *     included) as true.
 * exprtk re-associates further chains of + - or * / (an inner operation with a
 * constant operand, or a division); a formula containing such a chain is rejected
 * with an error naming it instead of being evaluated differently.
 * Accepted functions: abs sqrt exp log sin cos min max if not; constants pi,
 * epsilon, inf; the variable t (field scales) or x (distributions, which may also
 * call ProtonBetaSpectrum(x) on the host).  Names are case-insensitive.
 */
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/pentrack_gpu.h"
#include "pt_formula.h"

namespace {

enum Kind { N_CONST, N_VAR, N_UN, N_BIN, N_IPOW, N_SEL };

struct Node {
    int kind;
    int op;          /* PF_* */
    double v;        /* N_CONST value; N_IPOW: exponent count in p */
    int a, b, c;
    int p;           /* N_IPOW exponent (> 0), inverse if inv */
    bool inv;
};

bool addsub(int op) { return op == PF_ADD || op == PF_SUB; }
bool muldiv(int op) { return op == PF_MUL || op == PF_DIV; }

/* exprtk compute_pow10: one multiplication or division by 10^|e| */
double pow10_table(int e) {
    char buf[16];
    snprintf(buf, sizeof buf, "1e%d", e);
    return strtod(buf, nullptr);
}

double compute_pow10(double d, int exponent) {
    int e = exponent < 0 ? -exponent : exponent;
    if (exponent >= -307) {
        if (e <= 308) return exponent > 0 ? d * pow10_table(e) : d / pow10_table(e);
        return d * std::pow(10.0, 10.0 * exponent);
    }
    d /= pow10_table(307);
    return d / pow10_table(-exponent - 307);
}

class Compiler {
public:
    Compiler(const char *s, const char *var, bool host_functions) : s_(s), var_(var), host_(host_functions) {}

    bool compile(std::vector<uint32_t> &code, std::vector<double> &cst, std::string &err) {
        int root = parse_cond();
        if (ok_ && !at_end()) fail_at("unexpected input");
        if (!ok_) { err = err_; return false; }
        if (!check(root)) { err = err_; return false; }
        int depth = 0, maxd = 0;
        emit(root, code, cst, depth, maxd);
        if (maxd > PT_FORMULA_MAX_STACK) { err = "formula too deeply nested"; return false; }
        if ((int)code.size() > PT_FORMULA_MAX_CODE) { err = "formula too long"; return false; }
        return true;
    }

private:
    const char *s_;
    std::string var_;        /* the variable's name: t (field scales) or x (distributions) */
    bool host_;              /* allow functions evaluated on the host only */
    size_t i_ = 0;
    bool ok_ = true;
    std::string err_;
    std::vector<Node> n_;

    /* ---- lexing --------------------------------------------------------- */
    void skip() { while (s_[i_] && isspace((unsigned char)s_[i_])) ++i_; }
    bool at_end() { skip(); return s_[i_] == 0; }
    bool peek(const char *tok) { skip(); return strncmp(s_ + i_, tok, strlen(tok)) == 0; }
    bool accept(const char *tok) {
        if (!peek(tok)) return false;
        i_ += strlen(tok);
        return true;
    }
    /* keyword operator (and, or): a whole word */
    bool accept_word(const char *w) {
        skip();
        size_t n = strlen(w);
        if (strncasecmp(s_ + i_, w, n) != 0) return false;
        char c = s_[i_ + n];
        if (isalnum((unsigned char)c) || c == '_') return false;
        i_ += n;
        return true;
    }
    void fail_at(const std::string &m) {
        if (!ok_) return;
        ok_ = false;
        err_ = m + " at position " + std::to_string(i_) + " of '" + s_ + "'";
    }

    int add(const Node &x) { n_.push_back(x); return (int)n_.size() - 1; }
    int mk_const(double v) { return add(Node{N_CONST, 0, v, -1, -1, -1, 0, false}); }
    int mk_var() { return add(Node{N_VAR, PF_T, 0, -1, -1, -1, 0, false}); }
    bool is_const(int k) const { return n_[k].kind == N_CONST; }
    bool is_var(int k) const { return n_[k].kind == N_VAR; }
    double cv(int k) const { return n_[k].v; }

    /* ---- node synthesis (folding and exprtk's rewrites) ------------------- */
    static double fold_un(int op, double a) {
        switch (op) {
        case PF_NEG: return -a;
        case PF_NOT: return a == 0.0 ? 1.0 : 0.0;
        case PF_ABS: return a < 0.0 ? -a : a;
        case PF_SQRT: return std::sqrt(a);
        case PF_EXP: return std::exp(a);
        case PF_LOG: return std::log(a);
        case PF_SIN: return std::sin(a);
        case PF_COS: return std::cos(a);
        case PF_PBETA: return ptf_proton_beta_spectrum(a);
        }
        return a;
    }
    static double fold_bin(int op, double a, double b) {
        switch (op) {
        case PF_ADD: return a + b;
        case PF_SUB: return a - b;
        case PF_MUL: return a * b;
        case PF_DIV: return a / b;
        case PF_POW: return std::pow(a, b);
        case PF_LT: return a < b ? 1.0 : 0.0;
        case PF_LE: return a <= b ? 1.0 : 0.0;
        case PF_GT: return a > b ? 1.0 : 0.0;
        case PF_GE: return a >= b ? 1.0 : 0.0;
        case PF_EQ: return a == b ? 1.0 : 0.0;
        case PF_NE: return a != b ? 1.0 : 0.0;
        case PF_AND: return (a != 0.0 && b != 0.0) ? 1.0 : 0.0;
        case PF_OR: return (a != 0.0 || b != 0.0) ? 1.0 : 0.0;
        case PF_MIN: return b < a ? b : a;
        case PF_MAX: return a < b ? b : a;
        }
        return 0.0;
    }

    int mk_un(int op, int a) {
        if (is_const(a)) return mk_const(fold_un(op, cv(a)));
        return add(Node{N_UN, op, 0, a, -1, -1, 0, false});
    }

    int raw_bin(int op, int a, int b) { return add(Node{N_BIN, op, 0, a, b, -1, 0, false}); }

    /* a binary node of two leaves, one of them t (exprtk's cov / voc nodes) */
    bool is_pair_vc(int k) const {
        const Node &x = n_[k];
        return x.kind == N_BIN && ((is_var(x.a) && is_const(x.b)) || (is_const(x.a) && is_var(x.b)));
    }

    int mk_bin(int op, int a, int b) {
        if (is_const(a) && is_const(b)) return mk_const(fold_bin(op, cv(a), cv(b)));
        if (op == PF_POW && is_const(b)) {
            double c = cv(b);
            if (std::fabs(c) <= 60.0 && c == std::floor(c)) {       /* cardinal power */
                int p = (int)std::fabs(c);
                if (p == 0) return mk_const(1.0);
                if (p == 1 && c > 0) return a;
                return add(Node{N_IPOW, c > 0 ? PF_IPOW : PF_IPOWINV, 0, a, -1, -1, p, c < 0});
            }
            const Node &L = n_[a];
            if (L.kind == N_BIN && L.op == PF_POW && is_var(L.a) && is_const(L.b))   /* (t ^ c0) ^ c1 */
                return voc(PF_POW, cv(L.b) * c);
            return raw_bin(op, a, b);
        }
        if (!(addsub(op) || muldiv(op))) return raw_bin(op, a, b);
        /* three-node forms: a constant and a (t o c) / (c o t) node of the same class */
        if (is_const(a) && is_pair_vc(b) && same_class(op, n_[b].op)) {
            const Node &R = n_[b];
            return is_var(R.a) ? rule_c_vc(op, cv(a), R.op, cv(R.b)) : rule_c_cv(op, cv(a), R.op, cv(R.a));
        }
        if (is_const(b) && is_pair_vc(a) && same_class(op, n_[a].op)) {
            const Node &L = n_[a];
            return is_const(L.a) ? rule_cv_c(L.op, cv(L.a), op, cv(b)) : rule_vc_c(L.op, cv(L.b), op, cv(b));
        }
        if (is_const(a) && is_var(b)) {                          /* c o t */
            double c = cv(a);
            if (c == 0.0 && op == PF_MUL) return mk_const(0.0);
            if (c == 0.0 && op == PF_DIV) return mk_const(0.0);
            if (c == 0.0 && op == PF_ADD) return b;
            if (c == 1.0 && op == PF_MUL) return b;
            return raw_bin(op, a, b);
        }
        if (is_const(b) && is_var(a)) {                          /* t o c */
            double c = cv(b);
            if (c == 0.0 && op == PF_MUL) return mk_const(0.0);
            if (c == 0.0 && op == PF_DIV) return mk_const(std::numeric_limits<double>::quiet_NaN());
            if (c == 0.0 && op == PF_ADD) return a;
            if (c == 1.0 && op == PF_MUL) return a;
            if (c == 1.0 && op == PF_DIV) return a;
            return raw_bin(op, a, b);
        }
        if ((is_const(a) && cv(a) == 0.0) || (is_const(b) && cv(b) == 0.0)) {
            if (op != PF_SUB) {
                fail_at("a literal 0 operand of + * / on a sub-expression is simplified by exprtk; remove it");
                return mk_const(0);
            }
        }
        return raw_bin(op, a, b);
    }

    static bool same_class(int o1, int o2) { return (addsub(o1) && addsub(o2)) || (muldiv(o1) && muldiv(o2)); }

    int cov(int op, double c) { int k = mk_const(c); return raw_bin(op, k, mk_var()); }
    int voc(int op, double c) { int v = mk_var(); return raw_bin(op, v, mk_const(c)); }

    /* (c0 o0 t) o1 c1 */
    int rule_cv_c(int o0, double c0, int o1, double c1) {
        if (o0 == PF_ADD && o1 == PF_ADD) return cov(PF_ADD, c0 + c1);
        if (o0 == PF_ADD && o1 == PF_SUB) return cov(PF_ADD, c0 - c1);
        if (o0 == PF_SUB && o1 == PF_ADD) return cov(PF_SUB, c0 + c1);
        if (o0 == PF_SUB && o1 == PF_SUB) return cov(PF_SUB, c0 - c1);
        if (o0 == PF_MUL && o1 == PF_MUL) return cov(PF_MUL, c0 * c1);
        if (o0 == PF_MUL && o1 == PF_DIV) return cov(PF_MUL, c0 / c1);
        if (o0 == PF_DIV && o1 == PF_MUL) return cov(PF_DIV, c0 * c1);
        return cov(PF_DIV, c0 / c1);                                     /* (c0 / t) / c1 */
    }
    /* c0 o0 (t o1 c1) */
    int rule_c_vc(int o0, double c0, int o1, double c1) {
        if (o0 == PF_ADD && o1 == PF_ADD) return cov(PF_ADD, c0 + c1);
        if (o0 == PF_ADD && o1 == PF_SUB) return cov(PF_ADD, c0 - c1);
        if (o0 == PF_SUB && o1 == PF_ADD) return cov(PF_SUB, c0 - c1);
        if (o0 == PF_SUB && o1 == PF_SUB) return cov(PF_SUB, c0 + c1);
        if (o0 == PF_MUL && o1 == PF_MUL) return cov(PF_MUL, c0 * c1);
        if (o0 == PF_MUL && o1 == PF_DIV) return cov(PF_MUL, c0 / c1);
        if (o0 == PF_DIV && o1 == PF_MUL) return cov(PF_DIV, c0 / c1);
        return cov(PF_DIV, c0 * c1);                                     /* c0 / (t / c1) */
    }
    /* c0 o0 (c1 o1 t) */
    int rule_c_cv(int o0, double c0, int o1, double c1) {
        if (o0 == PF_ADD && o1 == PF_ADD) return cov(PF_ADD, c0 + c1);
        if (o0 == PF_ADD && o1 == PF_SUB) return cov(PF_SUB, c0 + c1);
        if (o0 == PF_SUB && o1 == PF_ADD) return cov(PF_SUB, c0 - c1);
        if (o0 == PF_SUB && o1 == PF_SUB) return cov(PF_ADD, c0 - c1);
        if (o0 == PF_MUL && o1 == PF_MUL) return cov(PF_MUL, c0 * c1);
        if (o0 == PF_MUL && o1 == PF_DIV) return cov(PF_DIV, c0 * c1);
        if (o0 == PF_DIV && o1 == PF_MUL) return cov(PF_DIV, c0 / c1);
        return cov(PF_MUL, c0 / c1);                                     /* c0 / (c1 / t) */
    }
    /* (t o0 c0) o1 c1 */
    int rule_vc_c(int o0, double c0, int o1, double c1) {
        if (o0 == PF_ADD && o1 == PF_ADD) return voc(PF_ADD, c0 + c1);
        if (o0 == PF_ADD && o1 == PF_SUB) return voc(PF_ADD, c0 - c1);
        if (o0 == PF_SUB && o1 == PF_ADD) return voc(PF_ADD, c1 - c0);
        if (o0 == PF_SUB && o1 == PF_SUB) return voc(PF_SUB, c0 + c1);
        if (o0 == PF_MUL && o1 == PF_MUL) return voc(PF_MUL, c0 * c1);
        if (o0 == PF_MUL && o1 == PF_DIV) return voc(PF_MUL, c0 / c1);
        if (o0 == PF_DIV && o1 == PF_MUL) return voc(PF_MUL, c1 / c0);
        return voc(PF_DIV, c0 * c1);                                     /* (t / c0) / c1 */
    }

    /* chains exprtk re-associates beyond the rules above: rejected */
    bool check(int k) {
        const Node &x = n_[k];
        switch (x.kind) {
        case N_CONST: case N_VAR: return true;
        case N_UN: case N_IPOW: return check(x.a);
        case N_SEL: return check(x.a) && check(x.b) && check(x.c);
        default: break;
        }
        if (!check(x.a) || !check(x.b)) return false;
        if (!(addsub(x.op) || muldiv(x.op))) return true;
        for (int side = 0; side < 2; ++side) {
            int ch = side ? x.b : x.a, other = side ? x.a : x.b;
            const Node &c = n_[ch];
            if (c.kind != N_BIN || !same_class(x.op, c.op)) continue;
            (void)other;
            bool has_div = x.op == PF_DIV || c.op == PF_DIV;
            bool child_const = is_const(c.a) || is_const(c.b);
            if (has_div || child_const) {
                ok_ = false;
                err_ = std::string("formula '") + s_ +
                       "': a chain of " + (addsub(x.op) ? "+/-" : "*//") +
                       " operations with a constant or a division is re-associated by exprtk; "
                       "combine the constants into one or regroup with a function";
                return false;
            }
        }
        return true;
    }

    /* ---- grammar --------------------------------------------------------- */
    int parse_cond() {
        int c = parse_or();
        if (ok_ && accept("?")) {
            int a = parse_cond();
            if (!accept(":")) { fail_at("expected ':'"); return c; }
            int b = parse_cond();
            return mk_sel(c, a, b);
        }
        return c;
    }
    int mk_sel(int c, int a, int b) {
        if (is_const(c)) return cv(c) != 0.0 ? a : b;
        return add(Node{N_SEL, PF_SEL, 0, c, a, b, 0, false});
    }
    int parse_or() {
        int a = parse_and();
        while (ok_) {
            if (peek("||")) { fail_at("'||' is not supported; use 'or'"); break; }
            if (accept_word("or") || accept("|")) a = mk_bin(PF_OR, a, parse_and());
            else break;
        }
        return a;
    }
    int parse_and() {
        int a = parse_cmp();
        while (ok_) {
            if (peek("&&")) { fail_at("'&&' is not supported; use 'and'"); break; }
            if (accept_word("and") || accept("&")) a = mk_bin(PF_AND, a, parse_cmp());
            else break;
        }
        return a;
    }
    int parse_cmp() {
        int a = parse_add();
        while (ok_) {
            int op;
            if (accept("<=")) op = PF_LE;
            else if (accept(">=")) op = PF_GE;
            else if (accept("==")) op = PF_EQ;
            else if (accept("!=") || accept("<>")) op = PF_NE;
            else if (accept("<")) op = PF_LT;
            else if (accept(">")) op = PF_GT;
            else if (accept("=")) op = PF_EQ;
            else break;
            a = mk_bin(op, a, parse_add());
        }
        return a;
    }
    /* a + or - operator; exprtk's lexer joins two adjacent ones ('+ -' and '- +' to
       '-', '- -' to '+'): returns +1, -1, or 0 when no sign follows */
    int take_sign() {
        skip();
        char c = s_[i_];
        if (c != '+' && c != '-') return 0;
        ++i_;
        skip();
        char d = s_[i_];
        if (d == '+' || d == '-') {
            ++i_;
            skip();
            if (s_[i_] == '+' || s_[i_] == '-') { fail_at("three adjacent signs"); return 0; }
            if (c == '+' && d == '+') return 2;          /* '+' then a unary '+' */
            return (c == d) ? 1 : -1;
        }
        return c == '+' ? 1 : -1;
    }
    int parse_add() {
        int a = parse_mul();
        while (ok_) {
            int sg = take_sign();
            if (sg == 0) break;
            a = mk_bin(sg > 0 ? PF_ADD : PF_SUB, a, parse_mul());
        }
        return a;
    }
    int parse_mul() {
        int a = parse_unary();
        while (ok_) {
            if (accept("*")) a = mk_bin(PF_MUL, a, parse_unary());
            else if (accept("/")) a = mk_bin(PF_DIV, a, parse_unary());
            else if (peek("%")) { fail_at("'%' is not supported"); break; }
            else break;
        }
        return a;
    }
    int parse_unary() {
        int sg = take_sign();
        if (sg < 0) return mk_un(PF_NEG, parse_unary());
        if (sg > 0) return parse_unary();
        return parse_pow();
    }
    int parse_pow() {
        int a = parse_primary();
        if (ok_ && accept("^")) return mk_bin(PF_POW, a, parse_unary());
        return a;
    }
    int parse_number() {
        /* exprtk string_to_real on the literal token */
        double d = 0.0;
        bool instate = false;
        if (s_[i_] != '.') {
            size_t st = i_;
            while (s_[i_] == '0') ++i_;
            while (isdigit((unsigned char)s_[i_])) { d = d * 10.0 + (double)(unsigned)(s_[i_] - '0'); ++i_; }
            if (i_ != st) instate = true;
        }
        if (s_[i_] == '.') {
            size_t st = ++i_;
            double f = 0.0;
            while (isdigit((unsigned char)s_[i_])) { f = f * 10.0 + (double)(unsigned)(s_[i_] - '0'); ++i_; }
            if (i_ != st) {
                instate = true;
                d += compute_pow10(f, -(int)(i_ - st));
            }
        }
        int exponent = 0;
        if (s_[i_] == 'e' || s_[i_] == 'E') {
            ++i_;
            bool neg = false;
            if (s_[i_] == '+' || s_[i_] == '-') { neg = s_[i_] == '-'; ++i_; }
            if (!isdigit((unsigned char)s_[i_])) { fail_at("bad number"); return mk_const(0); }
            long e = 0;
            while (isdigit((unsigned char)s_[i_])) { e = e * 10 + (s_[i_] - '0'); if (e > 100000) e = 100000; ++i_; }
            if (e > 308 + 20) { fail_at("number exponent out of range"); return mk_const(0); }
            exponent = neg ? -(int)e : (int)e;
        }
        if (!instate) { fail_at("bad number"); return mk_const(0); }
        if (isalpha((unsigned char)s_[i_]) || s_[i_] == '_') { fail_at("implicit multiplication is not supported"); return mk_const(0); }
        if (exponent) d = compute_pow10(d, exponent);
        return mk_const(d);
    }
    int parse_args(int n, int *out) {
        if (!accept("(")) { fail_at("expected '('"); return 0; }
        for (int k = 0; k < n; ++k) {
            out[k] = parse_cond();
            if (!ok_) return 0;
            if (k + 1 < n && !accept(",")) { fail_at("expected ','"); return 0; }
        }
        if (!accept(")")) { fail_at("expected ')'"); return 0; }
        return 1;
    }
    int parse_primary() {
        skip();
        char c = s_[i_];
        if (isdigit((unsigned char)c) || c == '.') return parse_number();
        if (accept("(")) {
            int a = parse_cond();
            if (!accept(")")) fail_at("expected ')'");
            return a;
        }
        if (isalpha((unsigned char)c) || c == '_') {
            size_t st = i_;
            while (isalnum((unsigned char)s_[i_]) || s_[i_] == '_') ++i_;
            std::string w(s_ + st, i_ - st);
            for (auto &ch : w) ch = (char)tolower((unsigned char)ch);
            if (w == var_) return mk_var();
            if (host_ && w == "protonbetaspectrum") {
                int a[1];
                if (!parse_args(1, a)) return mk_const(0);
                return mk_un(PF_PBETA, a[0]);
            }
            if (w == "pi") return mk_const(3.141592653589793238462643383279502);
            if (w == "epsilon") return mk_const(0.0000000001);
            if (w == "inf") return mk_const(std::numeric_limits<double>::infinity());
            static const struct { const char *name; int op; } un[] = {
                {"abs", PF_ABS}, {"sqrt", PF_SQRT}, {"exp", PF_EXP}, {"log", PF_LOG},
                {"sin", PF_SIN}, {"cos", PF_COS}, {"not", PF_NOT}};
            for (auto &u : un)
                if (w == u.name) {
                    int a[1];
                    if (!parse_args(1, a)) return mk_const(0);
                    return mk_un(u.op, a[0]);
                }
            if (w == "min" || w == "max") {
                int a[2];
                if (!parse_args(2, a)) return mk_const(0);
                return mk_bin(w == "min" ? PF_MIN : PF_MAX, a[0], a[1]);
            }
            if (w == "if") {
                int a[3];
                if (!parse_args(3, a)) return mk_const(0);
                return mk_sel(a[0], a[1], a[2]);
            }
            i_ = st;
            fail_at("unknown name '" + w + "'");
            return mk_const(0);
        }
        fail_at("unexpected character");
        return mk_const(0);
    }

    /* ---- code generation ------------------------------------------------- */
    void emit(int k, std::vector<uint32_t> &code, std::vector<double> &cst, int &depth, int &maxd) {
        const Node &x = n_[k];
        auto push = [&](int d) { depth += d; if (depth > maxd) maxd = depth; };
        switch (x.kind) {
        case N_CONST:
            code.push_back(PF_CONST | ((uint32_t)cst.size() << 8));
            cst.push_back(x.v);
            push(1);
            break;
        case N_VAR: code.push_back(PF_T); push(1); break;
        case N_UN: emit(x.a, code, cst, depth, maxd); code.push_back((uint32_t)x.op); break;
        case N_IPOW: emit(x.a, code, cst, depth, maxd); code.push_back((uint32_t)x.op | ((uint32_t)x.p << 8)); break;
        case N_BIN:
            emit(x.a, code, cst, depth, maxd);
            emit(x.b, code, cst, depth, maxd);
            code.push_back((uint32_t)x.op);
            push(-1);
            break;
        case N_SEL:
            emit(x.a, code, cst, depth, maxd);
            emit(x.b, code, cst, depth, maxd);
            emit(x.c, code, cst, depth, maxd);
            code.push_back(PF_SEL);
            push(-2);
            break;
        }
    }
};

}  // namespace

namespace ptf {
/* used by pt_fields_create (variable t); returns false and sets err on a rejected formula */
bool compile(const char *formula, std::vector<uint32_t> &code, std::vector<double> &cst, std::string &err) {
    code.clear();
    cst.clear();
    if (!formula) { err = "null formula"; return false; }
    Compiler c(formula, "t", false);
    return c.compile(code, cst, err);
}
/* distributions of the source and the spectra (variable x, host functions allowed) */
bool compile_x(const char *formula, std::vector<uint32_t> &code, std::vector<double> &cst, std::string &err) {
    code.clear();
    cst.clear();
    if (!formula) { err = "null formula"; return false; }
    Compiler c(formula, "x", true);
    return c.compile(code, cst, err);
}
}  // namespace ptf

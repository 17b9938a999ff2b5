"""Parse the CH4-CO2 induced-dipole surface out of the reference's generated C source text.

The reference ships the dipole (PES-IDS/ai_ids_ch4_co2.cpp) as computer-algebra output: the
molecular-frame dipole is a sum of 320 products

    dip = sum_b Re(idsbf_b(theta1, phi1, theta2, phi2)) * radcoeff_b(R)        (:672-1066)

of an angular basis function (a 3-vector) and a radial function.  Nothing here compiles or runs
that code.  Every function body is read as arithmetic and evaluated in exact rational arithmetic:

* an angular basis function becomes, per Cartesian component, a polynomial in the symbols
  LegP1[i] (associated Legendre P_l^m(cos theta1), GSL_SF_LEGENDRE_NONE with the Condon-Shortley
  phase, index i = l(l+1)/2 + m, l <= 16), LegP2[j] (the same at cos theta2, l <= 12), cosT1,
  cosP1_k = cos(k phi1), sinP1_k, cosP2_k, sinP2_k (:672-760);
* a radial function, a chain of ternaries `R < b ? piece : ...`, becomes a list of breakpoints
  and pieces, each a Laurent polynomial in R (cubic pieces below 30 bohr, c R^-6 (a + b R +
  c R^2) beyond).

`parse(path)` returns both; `gen_ch4co2_tables.py` turns them into the engine's tables and the
oracle's, `eval_reference(...)` evaluates the parsed expressions directly (an independent path
from the tables, used to write the test fixtures).
"""
import re
from fractions import Fraction

# ------------------------------------------------------------------ polynomials over symbols
# A polynomial is a dict {monomial: Fraction}; a monomial is a sorted tuple of symbol names.


def _padd(a, b, sign=1):
    out = dict(a)
    for k, v in b.items():
        out[k] = out.get(k, Fraction(0)) + sign * v
        if out[k] == 0:
            del out[k]
    return out


def _pmul(a, b):
    out = {}
    for ka, va in a.items():
        for kb, vb in b.items():
            k = tuple(sorted(ka + kb))
            out[k] = out.get(k, Fraction(0)) + va * vb
    return {k: v for k, v in out.items() if v != 0}


def _const(c):
    return {(): Fraction(c)} if c != 0 else {}


_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\d+[eE][-+]?\d+|\d+)|([A-Za-z_][A-Za-z_0-9]*)|(.))")


def _tokens(s):
    out = []
    for num, ident, op in _TOKEN.findall(s):
        if num:
            out.append(("num", num))
        elif ident:
            out.append(("id", ident))
        elif op.strip():
            out.append(("op", op))
    return out


class _Parser:
    """Recursive descent over + - * ( ) [ ] , numbers and identifiers; `lookup(name, index)`
    maps an identifier (with its [index] if any) to a value; `call(name, args)` a function."""

    def __init__(self, toks, lookup, call=None, add=_padd, mul=_pmul, const=_const):
        self.t, self.i = toks, 0
        self.lookup, self.call = lookup, call
        self.add, self.mul, self.const = add, mul, const

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self, kind=None, val=None):
        tok = self.peek()
        if (kind and tok[0] != kind) or (val and tok[1] != val):
            raise SyntaxError(f"expected {kind} {val}, got {tok}")
        self.i += 1
        return tok

    def expr(self):
        v = self.term()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            w = self.term()
            v = self.add(v, w, 1 if op == "+" else -1)
        return v

    def term(self):
        v = self.factor()
        while self.peek() == ("op", "*"):
            self.take()
            v = self.mul(v, self.factor())
        return v

    def factor(self):
        tok = self.peek()
        if tok == ("op", "-"):
            self.take()
            return self.mul(self.const(-1), self.factor())
        if tok == ("op", "+"):
            self.take()
            return self.factor()
        if tok == ("op", "("):
            self.take()
            v = self.expr()
            self.take("op", ")")
            return v
        if tok[0] == "num":
            self.take()
            return self.const(Fraction(tok[1]))
        if tok[0] == "id":
            name = self.take()[1]
            if self.peek() == ("op", "["):
                self.take()
                idx = int(self.take("num")[1])
                self.take("op", "]")
                return self.lookup(name, idx)
            if self.peek() == ("op", "("):
                self.take()
                args = [self.expr()]
                while self.peek() == ("op", ","):
                    self.take()
                    args.append(self.expr())
                self.take("op", ")")
                return self.call(name, args)
            return self.lookup(name, None)
        raise SyntaxError(f"unexpected {tok}")


def _function_bodies(src, prefix):
    """{name: body text} of every definition `static <type> prefix...(...) {` ... `}`."""
    out = {}
    for m in re.finditer(r"^static \w+ (" + prefix + r"[0-9_]+)\([^)]*\) \{\n(.*?)^\}", src, re.M | re.S):
        out[m.group(1)] = m.group(2)
    return out


# ------------------------------------------------------------------ angular basis functions
_SYM = re.compile(r"^(cosP1_\d+|sinP1_\d+|cosP2_\d+|sinP2_\d+|cosT1|cosT2)$")


def parse_basis(body):
    """[poly_x, poly_y, poly_z] of one idsbf function body."""
    env = {}

    def lookup(name, idx):
        if idx is not None:
            if name in ("LegP1", "LegP2"):
                return {(f"{name}[{idx}]",): Fraction(1)}
            raise KeyError(name)
        if name in env:
            return env[name]
        if _SYM.match(name):
            return {(name,): Fraction(1)}
        raise KeyError(name)

    bf = [None, None, None]
    for stmt in body.split(";"):
        stmt = stmt.strip()
        if not stmt or stmt.startswith("double "):
            continue
        lhs, rhs = stmt.split("=", 1)
        lhs = lhs.strip()
        val = _Parser(_tokens(rhs), lookup).expr()
        m = re.match(r"bf\[(\d)\]$", lhs)
        if m:
            bf[int(m.group(1))] = val
        else:
            env[lhs] = val
    if any(b is None for b in bf):
        raise ValueError("basis function without all three components")
    return bf


# ------------------------------------------------------------------ radial functions
# A Laurent polynomial in R is a dict {power: Fraction}.


def _lmul(a, b):
    out = {}
    for pa, va in a.items():
        for pb, vb in b.items():
            out[pa + pb] = out.get(pa + pb, Fraction(0)) + va * vb
    return {k: v for k, v in out.items() if v != 0}


def _lpow(a, n):
    if n >= 0:
        out = {0: Fraction(1)}
        for _ in range(n):
            out = _lmul(out, a)
        return out
    if len(a) != 1:
        raise ValueError("negative power of a non-monomial")
    (p, v), = a.items()
    return {p * n: v ** n}


def _laurent(expr):
    def lookup(name, idx):
        if name == "R" and idx is None:
            return {1: Fraction(1)}
        raise KeyError(name)

    def call(name, args):
        if name != "pow" or len(args) != 2:
            raise KeyError(name)
        e = args[1]
        if set(e) - {0}:
            raise ValueError("non-constant exponent")
        n = e.get(0, Fraction(0))
        if n.denominator != 1:
            raise ValueError("fractional exponent")
        return _lpow(args[0], int(n))

    return _Parser(_tokens(expr), lookup, call, mul=_lmul, const=lambda c: {0: Fraction(c)} if c != 0 else {}).expr()


def _split_ternary(s):
    """`cond ? a : b` at the top level of s (outer parentheses removed) -> (cond, a, b), or None."""
    s = s.strip()
    while s.startswith("(") and _matching(s, 0) == len(s) - 1:
        s = s[1:-1].strip()
    depth, q = 0, -1
    for i, ch in enumerate(s):
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        elif ch == "?" and depth == 0:
            q = i
            break
    if q < 0:
        return None
    depth, nested = 0, 0
    for i in range(q + 1, len(s)):
        ch = s[i]
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        elif ch == "?" and depth == 0:
            nested += 1
        elif ch == ":" and depth == 0:
            if nested == 0:
                return s[:q].strip(), s[q + 1:i].strip(), s[i + 1:].strip()
            nested -= 1
    raise SyntaxError("unbalanced ternary")


def _matching(s, i):
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j
    return -1


def parse_radial(body):
    """[(upper_breakpoint or None, laurent poly)] in increasing R: piece k holds for
    R < upper_k (and R >= upper_{k-1}); the last piece has no upper bound."""
    m = re.search(r"return\s+(.*);", body, re.S)
    expr = m.group(1)
    pieces = []

    def walk(e):
        t = _split_ternary(e)
        if t is None:
            pieces.append([None, _laurent(e)])
            return
        cond, a, b = t
        cm = re.match(r"R\s*<\s*([0-9.eE+-]+)$", cond)
        if not cm:
            raise SyntaxError("condition " + cond)
        bound = Fraction(cm.group(1))
        walk(a)
        # every piece of `a` lies below `bound`: the last of them is bounded by it
        if pieces[-1][0] is None:
            pieces[-1][0] = bound
        walk(b)

    walk(expr)
    return [(p[0], p[1]) for p in pieces]


# ------------------------------------------------------------------ the whole surface


def parse(path):
    src = open(path).read()
    body = re.search(r"^void dipole_vector\(double \*q, double dip\[3\]\)\n\{(.*?)^\}", src, re.M | re.S).group(1)
    # the multiple-angle convention: cosP1_k = cos(k*phi1), ... (:695-748)
    for side in ("1", "2"):
        for fn in ("cos", "sin"):
            for k, kk in re.findall(rf"{fn}P{side}_(\d+) = std::{fn}\((\d*)\*?phi{side}\);", body):
                if int(kk or 1) != int(k):
                    raise ValueError(f"{fn}P{side}_{k} is not {fn}({k} phi{side})")
    pairs = re.findall(r"(idsbf_[0-9_]+)\(tmp\);\s*res \+= tmp\.real\(\) \* (radcoeff_[0-9_]+)\(R\);", body)
    if len(pairs) != 320:
        raise ValueError(f"expected 320 terms, found {len(pairs)}")
    basis_src = _function_bodies(src, "idsbf_")
    rad_src = _function_bodies(src, "radcoeff_")
    lmax1 = int(re.search(r"static int lmax_bond\s*=\s*(\d+);", src).group(1))
    lmax2 = int(re.search(r"static int lmax_n2\s*=\s*(\d+);", src).group(1))
    terms = []
    for b, r in pairs:
        terms.append((b, parse_basis(basis_src[b]), r, parse_radial(rad_src[r])))
    return {"terms": terms, "lmax1": lmax1, "lmax2": lmax2}


def laurent_value(poly, R):
    return float(sum(v * Fraction(R) ** p for p, v in poly.items())) if isinstance(R, Fraction) else \
        sum(float(v) * R ** p for p, v in poly.items())


def radial_value(pieces, R):
    for upper, poly in pieces:
        if upper is None or R < upper:
            return laurent_value(poly, R)
    raise AssertionError


def eval_reference(parsed, R, phi1, theta1, phi2, theta2, legendre):
    """dipole_vector(q) from the parsed expressions, term by term in the reference's order:
    legendre(l, m, x) must return P_l^m(x) with the Condon-Shortley phase."""
    import math
    sym = {"cosT1": math.cos(theta1), "cosT2": math.cos(theta2)}
    x1, x2 = math.cos(theta1), math.cos(theta2)
    for l in range(parsed["lmax1"] + 1):
        for m in range(l + 1):
            sym[f"LegP1[{l * (l + 1) // 2 + m}]"] = legendre(l, m, x1)
    for l in range(parsed["lmax2"] + 1):
        for m in range(l + 1):
            sym[f"LegP2[{l * (l + 1) // 2 + m}]"] = legendre(l, m, x2)
    for k in range(1, 14):
        sym[f"cosP1_{k}"], sym[f"sinP1_{k}"] = math.cos(k * phi1), math.sin(k * phi1)
        sym[f"cosP2_{k}"], sym[f"sinP2_{k}"] = math.cos(k * phi2), math.sin(k * phi2)
    res = [0.0, 0.0, 0.0]
    for _, bf, _, rad in parsed["terms"]:
        rv = radial_value(rad, R)
        for c in range(3):
            s = 0.0
            for mono, coef in bf[c].items():
                v = float(coef)
                for name in mono:
                    v *= sym[name]
                s += v
            res[c] += s * rv
    return res

"""Parse the CO2-Ar PES / induced-dipole surfaces out of the reference C source text.

The reference ships the CO2-Ar surfaces as generated C code: one ``if (R < b)``
branch per radial interval, each an arithmetic expression in ``R`` and in the
Legendre values ``legP[l]`` (PES) or ``LegP[index(l, m)]`` (dipole).  Nothing
here compiles or runs that code.  Each branch's statements are read as
arithmetic, evaluated symbolically in exact rational arithmetic, and expanded
into a radial table:

    coefficient(l[, m])(R) = sum_k a_k R^k            (k in -7 .. 3)
                           + A exp(alpha R)           (short-range branch only)

Sources followed:
  PES-IDS/ai_pes_co2ar.c:27-459   pes_co2ar      (value, cm^-1, /HTOCM on return)
  PES-IDS/ai_pes_co2ar.c:461-852  dpes_dTheta    (only used as a check)
  PES-IDS/ai_pes_co2ar.c:854-1028 dpes_dR        (only used as a check)
  PES-IDS/ai_ids_co2ar.cpp:21-255   dipx  (P_l^1 terms, Condon-Shortley phase)
  PES-IDS/ai_ids_co2ar.cpp:257-492  dipz  (P_l^0 terms)
"""
import re
import sympy as sp

R = sp.Symbol("R", positive=True)
COST = sp.Symbol("cosT")
SINT = sp.Symbol("sinT")
PSYM = {}


def psym(l, m):
    key = (l, m)
    if key not in PSYM:
        PSYM[key] = sp.Symbol(f"P{l}_{m}")
    return PSYM[key]


_num = re.compile(r"(?<![A-Za-z_0-9.])(\d+\.\d*(?:[eE][-+]?\d+)?|\d+[eE][-+]?\d+)")


def _c_to_py(expr):
    expr = re.sub(r"LegP\[index\((\d+),\s*(\d+)\)\]", lambda m: f"P{m.group(1)}_{m.group(2)}", expr)
    expr = re.sub(r"legP\[(\d+)\]", lambda m: f"P{m.group(1)}_0", expr)
    # exact rationals for every decimal literal
    expr = _num.sub(lambda m: f"Rational('{m.group(1)}')", expr)
    return expr


def _eval(expr, env):
    ns = dict(env)
    ns.update({"R": R, "cosT": COST, "sinT": SINT, "exp": sp.exp, "Rational": _exact})
    for (l, m), s in PSYM.items():
        ns[str(s)] = s
    # create any P symbols referenced
    for l, m in re.findall(r"P(\d+)_(\d+)", expr):
        ns[f"P{l}_{m}"] = psym(int(l), int(m))
    return eval(expr, {"__builtins__": {}}, ns)


def _exact(s):
    """Exact rational value of a C decimal literal."""
    s = s.lower()
    if "e" in s:
        mant, ex = s.split("e")
        ex = int(ex)
    else:
        mant, ex = s, 0
    if "." in mant:
        ip, fp = mant.split(".")
    else:
        ip, fp = mant, ""
    digits = int((ip or "0") + (fp or "")) if (ip or fp) else 0
    scale = len(fp)
    return sp.Rational(digits) * sp.Rational(10) ** (ex - scale)


def _function_body(src, signature_regex):
    m = re.search(signature_regex, src)
    if not m:
        raise ValueError("function not found: " + signature_regex)
    i = src.index("{", m.end() - 1)
    depth = 0
    for j in range(i, len(src)):
        if src[j] == "{":
            depth += 1
        elif src[j] == "}":
            depth -= 1
            if depth == 0:
                return src[i + 1:j]
    raise ValueError("unbalanced braces")


def _strip_comments(s):
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    s = re.sub(r"//[^\n]*", "", s)
    return s


def split_branches(body):
    """Return [(upper_bound or None, statements_text)] for the if/else-if/else chain on R."""
    body = _strip_comments(body)
    out = []
    heads = []
    for m in re.finditer(r"\b(?:else\s+if|if)\s*\(\s*R\s*<\s*([0-9.]+)\s*\)\s*\{|\belse\s*\{", body):
        bound = m.group(1)
        heads.append((m.start(), m.end(), float(bound) if bound else None))
    for k, (s, e, bound) in enumerate(heads):
        # find matching close brace
        depth = 1
        j = e
        while depth:
            if body[j] == "{":
                depth += 1
            elif body[j] == "}":
                depth -= 1
            j += 1
        out.append((bound, body[e:j - 1]))
    return out, body


def eval_statements(text, env=None, tail_return=None):
    env = dict(env or {})
    result = None
    for stmt in text.split(";"):
        stmt = " ".join(stmt.replace("{", " ").replace("}", " ").split())
        if not stmt:
            continue
        if stmt.startswith("return"):
            result = _eval(_c_to_py(stmt[len("return"):].strip()), env)
            break
        m = re.match(r"([A-Za-z_]\w*)\s*=\s*(.*)$", stmt, flags=re.S)
        if not m:
            raise ValueError("cannot parse statement: " + stmt[:80])
        env[m.group(1)] = _eval(_c_to_py(m.group(2)), env)
    if result is None and tail_return is not None:
        result = env[tail_return]
    return result


def radial_table(expr, angular_symbols):
    """Expand expr = sum_s f_s(R) * s over the angular symbols s into
    {s: {"poly": {k: Rational}, "exp": [(A, alpha)]}}."""
    expr = sp.expand(expr)
    table = {}
    for term in sp.Add.make_args(expr):
        ang = [f for f in sp.Mul.make_args(term) if f in angular_symbols]
        if len(ang) > 1:
            raise ValueError("product of angular functions in term: " + str(term))
        key = ang[0] if ang else None
        radial = term / key if key is not None else term
        entry = table.setdefault(key, {"poly": {}, "exp": []})
        coeff, rest = radial.as_coeff_Mul()
        if rest == 1:
            entry["poly"][0] = entry["poly"].get(0, 0) + coeff
            continue
        if rest.func == sp.exp or (rest.is_Mul and any(f.func == sp.exp for f in sp.Mul.make_args(rest))):
            fac = [f for f in sp.Mul.make_args(rest) if f.func == sp.exp]
            others = [f for f in sp.Mul.make_args(rest) if f.func != sp.exp]
            if len(fac) != 1 or others:
                raise ValueError("unsupported exponential term: " + str(term))
            arg = fac[0].args[0]
            alpha = sp.expand(arg / R)
            if alpha.free_symbols:
                raise ValueError("non-linear exponent: " + str(arg))
            entry["exp"].append((coeff, alpha))
            continue
        k = sp.degree(rest, R) if rest.is_polynomial(R) else None
        if rest == R:
            k = 1
        elif rest.is_Pow and rest.base == R:
            k = int(rest.exp)
        elif k is None:
            raise ValueError("unsupported radial factor: " + str(rest))
        entry["poly"][k] = entry["poly"].get(k, 0) + coeff
    return table


def parse_pes(src):
    """PES value branches: list of (upper, {l: table})."""
    body = _function_body(src, r"double\s+pes_co2ar\s*\(\s*double\s+R\s*,\s*double\s+Theta\s*\)\s*\{")
    branches, _ = split_branches(body)
    syms = {psym(l, 0) for l in range(0, 11)}
    out = []
    for upper, text in branches:
        expr = eval_statements(text, tail_return="t48")
        tab = radial_table(expr, syms)
        out.append((upper, {(None if k is None else int(str(k)[1:].split("_")[0])): v for k, v in tab.items()}))
    return out


def parse_pes_dR(src):
    body = _function_body(src, r"double\s+dpes_dR\s*\(\s*double\s+R\s*,\s*double\s+Theta\s*\)\s*\{")
    branches, _ = split_branches(body)
    return [(upper, eval_statements(text, tail_return="t29")) for upper, text in branches]


def parse_pes_dTheta(src):
    body = _function_body(src, r"double\s+dpes_dTheta\s*\(\s*double\s+R\s*,\s*double\s+Theta\s*\)\s*\{")
    branches, _ = split_branches(body)
    return [(upper, eval_statements(text, tail_return="t86")) for upper, text in branches]


def parse_dipole(src, func):
    body = _function_body(src, r"double\s+AI_IDS_co2_ar::" + func + r"\s*\([^)]*\)\s*\{")
    branches, _ = split_branches(body)
    syms = {psym(l, m) for l in range(0, 11) for m in (0, 1)}
    out = []
    for upper, text in branches:
        expr = eval_statements(text)
        tab = radial_table(expr, syms)
        conv = {}
        for k, v in tab.items():
            if k is None:
                if func != "dipz":
                    raise ValueError("constant term without an angular factor in " + func)
                conv[0] = v  # P_0 = 1 constant term
            else:
                conv[int(str(k)[1:].split("_")[0])] = v
        out.append((upper, conv))
    return out

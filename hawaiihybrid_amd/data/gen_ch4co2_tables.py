"""Generate the CH4-CO2 induced-dipole tables from the reference's generated source text.

    python hawaiihybrid_amd/data/gen_ch4co2_tables.py [/path/to/PES-IDS/ai_ids_ch4_co2.cpp]

Reads PES-IDS/ai_ids_ch4_co2.cpp (default: under /root/reference) through ch4co2_parse.py and
writes

  hawaiihybrid_amd/csrc/ch4co2_dipole_tables.h   the engine's tables (FP64; rounded to FP32 on upload)
  oracle/ch4co2_dipole_table.h                   the oracle's tables (the same expansion, FP64)
  tests/golden/ch4co2_dipole_ref.npz             dipole_vector at 200 molecular-frame points,
                                                  evaluated from the parsed expressions with
                                                  scipy's associated Legendre functions (an
                                                  evaluation path independent of the tables)

The dipole in the CH4 frame is  mu_c = sum_b r_b(R) sum_(x, y) C[b, c, x, y] X_x Y_y  with
  X_x = P_l^m(cos theta1) {1 | cos m phi1 | sin m phi1}, l <= 13:  x = l^2 (m = 0),
        l^2 + 2m - 1 (cos), l^2 + 2m (sin)                                    (196 entries)
  Y_y = the same at (theta2, phi2) for even l <= 12: y = l(l-1)/2 + (0 | 2m-1 | 2m)  (91 entries)
  r_b = a cubic in (R - centre_k) on interval k of the common radial grid (breakpoints 5, 5.25,
        ..., 30 bohr; interval 0 is R < 5), and c4 R^-4 + c5 R^-5 + c6 R^-6 for R >= 30.
Each cubic is re-expanded about its interval's centre in exact rational arithmetic (the
reference's pieces are written about the interval's left end, with a linear term in R itself).
"""
import os
import sys
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import ch4co2_parse as P   # noqa: E402

L1MAX, L2MAX = 13, 12


def x_index(l, m, kind):
    return l * l if m == 0 else l * l + 2 * m - 1 + kind


def y_index(l, m, kind):
    assert l % 2 == 0
    return l * (l - 1) // 2 if m == 0 else l * (l - 1) // 2 + 2 * m - 1 + kind


def lm_of(i):
    l = 0
    while (l + 1) * (l + 2) // 2 <= i:
        l += 1
    return l, i - l * (l + 1) // 2


def factor_index(mono):
    """(x, y) of a monomial LegP1[i] {cosP1_m|sinP1_m} LegP2[j] {cosP2_m|sinP2_m} (cosT1 = P_1^0)."""
    leg1 = [n for n in mono if n.startswith("LegP1")]
    leg2 = [n for n in mono if n.startswith("LegP2")]
    t1 = [n for n in mono if "P1_" in n]
    t2 = [n for n in mono if "P2_" in n]
    if "cosT1" in mono:
        leg1.append("LegP1[1]")
    if "cosT2" in mono:
        leg2.append("LegP2[1]")
    assert len(leg1) <= 1 and len(leg2) <= 1 and len(t1) <= 1 and len(t2) <= 1, mono
    l1, m1 = lm_of(int(leg1[0][6:-1])) if leg1 else (0, 0)
    l2, m2 = lm_of(int(leg2[0][6:-1])) if leg2 else (0, 0)
    k1 = int(t1[0].split("_")[1]) if t1 else 0
    k2 = int(t2[0].split("_")[1]) if t2 else 0
    assert k1 == m1 and k2 == m2, mono
    kind1 = 0 if not t1 or t1[0].startswith("cos") else 1
    kind2 = 0 if not t2 or t2[0].startswith("cos") else 1
    assert l1 <= L1MAX and l2 <= L2MAX
    return x_index(l1, m1, kind1), y_index(l2, m2, kind2)


def recentre(poly, c):
    """Laurent polynomial with powers 0..3 -> coefficients of (R - c)^0..3, exact."""
    assert set(poly) <= {0, 1, 2, 3}, poly
    out = [Fraction(0)] * 4
    # R^p = (t + c)^p
    from math import comb
    for p, v in poly.items():
        for k in range(p + 1):
            out[k] += v * comb(p, k) * c ** (p - k)
    return out


def build(path):
    parsed = P.parse(path)
    terms = parsed["terms"]
    bounds = [u for u, _ in terms[0][3][:-1]]
    for _, _, _, rad in terms:
        assert [u for u, _ in rad[:-1]] == bounds and rad[-1][0] is None
    nint = len(bounds)   # intervals below the last bound; the last piece is the tail
    # interval k: [bounds[k-1], bounds[k]) (k = 0 extends below bounds[0]); centre: midpoint, for
    # k = 0 the midpoint of [bounds[0] - 0.25, bounds[0]]
    centres = []
    for k in range(nint):
        lo = bounds[k - 1] if k > 0 else bounds[0] - Fraction(1, 4)
        centres.append((lo + bounds[k]) / 2)
    radial = np.zeros((nint, len(terms), 4))
    tail = np.zeros((len(terms), 3))
    for b, (_, _, _, rad) in enumerate(terms):
        for k in range(nint):
            radial[k, b] = [float(v) for v in recentre(rad[k][1], centres[k])]
        tp = rad[-1][1]
        assert set(tp) <= {-4, -5, -6}, tp
        tail[b] = [float(tp.get(-4, 0)), float(tp.get(-5, 0)), float(tp.get(-6, 0))]
    # terms sorted by basis function, then component, then (x, y)
    rows = []
    for b, (_, bf, _, _) in enumerate(terms):
        for c in range(3):
            for mono, coef in sorted(bf[c].items()):
                x, y = factor_index(mono)
                rows.append((b, c, x, y, float(coef)))
    return parsed, np.array([float(v) for v in bounds]), np.array([float(v) for v in centres]), radial, tail, rows


def write_header(fname, guard, bounds, centres, radial, tail, rows, ctype, note):
    nb = radial.shape[1]
    # per (basis function, component): [begin, end) in the term list
    starts = np.zeros((nb, 4), dtype=np.int64)
    idx = 0
    for b in range(nb):
        for c in range(3):
            starts[b, c] = idx
            while idx < len(rows) and rows[idx][0] == b and rows[idx][1] == c:
                idx += 1
        starts[b, 3] = idx
    with open(fname, "w") as f:
        f.write(f"// {note}\n")
        f.write("// Generated by hawaiihybrid_amd/data/gen_ch4co2_tables.py from the reference's\n")
        f.write("// PES-IDS/ai_ids_ch4_co2.cpp (read as text; every piece expanded in exact arithmetic).\n")
        f.write(f"#ifndef {guard}\n#define {guard}\n\n")
        f.write(f"#define HH_CH4CO2_NB {nb}            /* basis functions (radial functions) */\n")
        f.write(f"#define HH_CH4CO2_NINT {len(bounds)}          /* radial intervals below the tail */\n")
        f.write(f"#define HH_CH4CO2_NTERMS {len(rows)}      /* angular monomials */\n")
        f.write("#define HH_CH4CO2_NX 196          /* X_x: l <= 13 at (theta1, phi1) */\n")
        f.write("#define HH_CH4CO2_NY 91           /* Y_y: even l <= 12 at (theta2, phi2) */\n\n")
        f.write("/* interval k covers [HH_CH4CO2_BOUNDS[k-1], HH_CH4CO2_BOUNDS[k]) (k = 0: R < BOUNDS[0]); R >= the\n"
                " * last bound uses the tail */\n")
        f.write("static const double HH_CH4CO2_BOUNDS[HH_CH4CO2_NINT] = {" +
                ", ".join(repr(float(v)) for v in bounds) + "};\n")
        f.write("static const double HH_CH4CO2_CENTRE[HH_CH4CO2_NINT] = {" +
                ", ".join(repr(float(v)) for v in centres) + "};\n\n")
        f.write("/* r_b(R) = sum_j HH_CH4CO2_RADIAL[k][b][j] (R - centre_k)^j on interval k */\n")
        f.write(f"static const {ctype} HH_CH4CO2_RADIAL[HH_CH4CO2_NINT][HH_CH4CO2_NB][4] = {{\n")
        for k in range(radial.shape[0]):
            f.write(" {" + ",\n  ".join("{" + ", ".join(repr(float(v)) for v in radial[k, b]) + "}" for b in range(nb)) + "},\n")
        f.write("};\n\n")
        f.write("/* R >= the last bound: r_b = c4 R^-4 + c5 R^-5 + c6 R^-6 */\n")
        f.write(f"static const {ctype} HH_CH4CO2_TAIL[HH_CH4CO2_NB][3] = {{\n")
        f.write(",\n".join(" {" + ", ".join(repr(float(v)) for v in tail[b]) + "}" for b in range(nb)) + "\n};\n\n")
        f.write("/* terms of basis function b, component c: HH_CH4CO2_START[b][c] .. HH_CH4CO2_START[b][c + 1] */\n")
        f.write("static const int HH_CH4CO2_START[HH_CH4CO2_NB][4] = {\n")
        f.write(",\n".join(" {" + ", ".join(str(int(v)) for v in starts[b]) + "}" for b in range(nb)) + "\n};\n\n")
        f.write("/* term: X index, Y index (packed x | y << 8), coefficient */\n")
        f.write("static const unsigned short HH_CH4CO2_XY[HH_CH4CO2_NTERMS] = {\n")
        for i in range(0, len(rows), 16):
            f.write(" " + ", ".join(str(r[2] | (r[3] << 8)) for r in rows[i:i + 16]) + ",\n")
        f.write("};\n")
        f.write(f"static const {ctype} HH_CH4CO2_COEF[HH_CH4CO2_NTERMS] = {{\n")
        for i in range(0, len(rows), 6):
            f.write(" " + ", ".join(repr(r[4]) for r in rows[i:i + 6]) + ",\n")
        f.write("};\n\n#endif\n")


def fixtures(parsed, out):
    from scipy.special import lpmv
    rng = np.random.default_rng(20240521)
    n = 200
    R = np.concatenate([rng.uniform(4.6, 30.0, n - 20), rng.uniform(30.0, 45.0, 20)])
    phi1, phi2 = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    th1, th2 = np.arccos(rng.uniform(-1, 1, n)), np.arccos(rng.uniform(-1, 1, n))
    # scipy's lpmv carries the Condon-Shortley phase, as GSL's csphase = -1 arrays do
    leg = lambda l, m, x: float(lpmv(m, l, x))
    dip = np.array([P.eval_reference(parsed, R[i], phi1[i], th1[i], phi2[i], th2[i], leg) for i in range(n)])
    np.savez_compressed(out, R=R, phi1=phi1, theta1=th1, phi2=phi2, theta2=th2, dipole=dip)


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/PES-IDS/ai_ids_ch4_co2.cpp"
    parsed, bounds, centres, radial, tail, rows = build(src)
    write_header(os.path.join(REPO, "hawaiihybrid_amd", "csrc", "ch4co2_dipole_tables.h"), "HH_CH4CO2_DIPOLE_TABLES_H",
                 bounds, centres, radial, tail, rows, "double",
                 "CH4-CO2 induced dipole (molecular frame of CH4): radial tables and angular terms.")
    write_header(os.path.join(REPO, "oracle", "ch4co2_dipole_table.h"), "ORC_CH4CO2_DIPOLE_TABLE_H",
                 bounds, centres, radial, tail, rows, "double",
                 "CH4-CO2 induced dipole for the oracle (test infrastructure): the same expansion in FP64.")
    fixtures(parsed, os.path.join(REPO, "tests", "golden", "ch4co2_dipole_ref.npz"))
    print(f"{len(rows)} terms, {radial.shape[1]} radial functions, {radial.shape[0]} intervals")


if __name__ == "__main__":
    main()

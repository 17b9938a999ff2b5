"""Permutation-invariant polynomial (PIP) bases of the N2-Ar PIP-NN surfaces.

    python -m hawaiihybrid_amd.data.gen_pip_n2ar     # writes csrc/pip_n2ar.h and oracle/pip_n2ar_table.h

The surfaces of the reference feed monomial-symmetrised polynomials of Morse-like
variables y_ij into their networks:
  * PES: 3 atoms (N1 N2 Ar), atoms {N1, N2} alike, degree <= 4, 17 polynomials
    (PES-IDS/c_basis_2_1_4_purify.cc:3-24, used by PES-IDS/ai_pes_n2_ar_pip_nn.cpp:96-99);
  * dipole: 5 sites (N1 N2 N1' N2' Ar), three bases of degree <= 3 concatenated in this order
    (PES-IDS/ai_ids_n2_ar_pip_nn.cpp:41-48):
      {N1'<->N2'} alike        119 polynomials (PES-IDS/c_basis_1_1_2_1_3_purify.cc)
      {N1<->N2} alike          119 polynomials (PES-IDS/c_basis_2_1_1_1_3_purify.cc)
      both pairs alike          70 polynomials (PES-IDS/c_basis_2_2_1_3_purify.cc)
Every polynomial is the sum of the distinct monomials in the orbit of one monomial under the
basis' permutation group (the purified constant term is identically 0).  The networks were
trained on the polynomials in a fixed order, so the order is part of the model: the tables
below give, per polynomial, one monomial of its orbit as the list of its distance indices
(index k of pair (i, j), i < j, counts pairs in the order (0,1), (0,2), ..., (n-2, n-1)), and
this script expands each orbit.

Outputs:
  csrc/pip_n2ar.h        straight-line evaluation in blocks of 16 polynomials, for the device
                         kernels (and host code), generic in the arithmetic type
  oracle/pip_n2ar_table.h  the representative monomials and the permutation generators as C
                         tables; the oracle expands the orbits itself at run time
"""
import itertools
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))

# one representative monomial per polynomial ("-": the purified constant, identically 0)
PES_REPS = "1 12 01 11 012 112 011 001 111 0112 0012 1122 1112 0111 0011 0001 1111"
DIP_REPS = [
    # N1' <-> N2' alike
    "- 8 6 3 89 78 68 67 49 48 46 38 37 36 34 19 18 16 13 08 06 03 88 66 33 789 689 678 489 479 478 469 "
    "468 467 458 456 389 378 368 367 349 348 347 346 345 189 179 178 169 168 167 159 158 156 149 148 146 "
    "139 138 137 136 135 134 128 126 123 089 078 068 067 049 048 046 038 037 036 034 019 018 016 013 889 "
    "788 778 688 677 668 667 499 488 466 449 448 446 388 377 366 344 338 337 336 334 199 188 166 133 119 "
    "118 116 113 088 066 033 008 006 003 888 666 333",
    # N1 <-> N2 alike
    "- 9 8 3 89 79 78 39 38 37 36 29 28 26 23 19 18 16 13 09 08 03 99 88 33 789 389 379 378 369 368 367 "
    "289 279 278 269 268 267 259 258 239 238 237 236 235 189 179 178 169 168 167 159 158 156 149 148 139 "
    "138 137 136 135 134 129 128 126 123 089 079 078 039 038 037 036 029 028 026 023 019 018 016 013 899 "
    "889 799 788 779 778 399 388 377 339 338 337 336 299 288 266 233 229 228 226 223 199 188 166 133 119 "
    "118 116 113 099 088 033 009 008 003 999 888 333",
    # both pairs alike
    "- 8 3 89 78 38 37 36 19 18 16 13 08 03 88 33 789 389 378 368 367 189 179 178 169 168 167 158 149 "
    "148 139 138 137 136 135 134 128 126 123 089 078 038 037 036 019 018 016 013 889 788 778 388 377 338 "
    "337 336 199 188 166 133 119 118 116 113 088 033 008 003 888 333",
]
# atom permutations generating each basis' group
PES_GENS = [(1, 0, 2)]
DIP_GENS = [[(0, 1, 3, 2, 4)], [(1, 0, 2, 3, 4)], [(1, 0, 2, 3, 4), (0, 1, 3, 2, 4)]]


def parse_reps(s):
    return [None if t == "-" else tuple(int(c) for c in t) for t in s.split()]


def pair_index(n):
    return {p: k for k, p in enumerate(itertools.combinations(range(n), 2))}


def closure(n, gens):
    G = {tuple(range(n))}
    todo = list(G)
    while todo:
        g = todo.pop()
        for s in gens:
            h = tuple(s[g[i]] for i in range(n))
            if h not in G:
                G.add(h)
                todo.append(h)
    return sorted(G)


def distance_perms(n, gens):
    idx = pair_index(n)
    out = []
    for g in closure(n, gens):
        out.append([idx[tuple(sorted((g[i], g[j])))] for (i, j) in itertools.combinations(range(n), 2)])
    return out


def orbit(mono, dperms):
    return sorted({tuple(sorted(dp[x] for x in mono)) for dp in dperms})


def expand(reps, n, gens):
    dps = distance_perms(n, gens)
    return [[] if r is None else orbit(r, dps) for r in reps]


def term(mono):
    return "*".join(f"y[{i}]" for i in mono)


def emit_blocks(name, polys, lines):
    nb = (len(polys) + 15) // 16
    lines.append(f"// {len(polys)} polynomials in {nb} blocks of 16 (zero-padded)")
    lines.append("template <typename T>")
    lines.append(f"HH_PIP_FN void {name}(int kb, const T* __restrict__ y, T* __restrict__ p)")
    lines.append("{")
    lines.append("    switch (kb) {")
    for b in range(nb):
        lines.append(f"    case {b}:")
        for j in range(16):
            k = 16 * b + j
            if k < len(polys) and polys[k]:
                expr = " + ".join(term(m) for m in polys[k])
            else:
                expr = "T(0)"
            lines.append(f"        p[{j}] = {expr};")
        lines.append("        break;")
    lines.append("    default:")
    lines.append("        break;")
    lines.append("    }")
    lines.append("}")
    lines.append("")


def write_device_header(path):
    pes = expand(parse_reps(PES_REPS), 3, PES_GENS)
    dip = []
    for reps, gens in zip(DIP_REPS, DIP_GENS):
        dip += expand(parse_reps(reps), 5, gens)
    lines = [
        "// Generated by hawaiihybrid_amd/data/gen_pip_n2ar.py -- do not edit.",
        "// Permutation-invariant polynomial bases of the N2-Ar PIP-NN surfaces, evaluated in",
        "// blocks of 16 (the k-blocks of the first network layer).  Each polynomial is the orbit",
        "// sum of one monomial of the distance variables y under the basis' permutation group;",
        "// the order is the reference's (PES-IDS/c_basis_2_1_4_purify.cc:3-24; dipole bases",
        "// PES-IDS/c_basis_1_1_2_1_3_purify.cc, c_basis_2_1_1_1_3_purify.cc,",
        "// c_basis_2_2_1_3_purify.cc concatenated as PES-IDS/ai_ids_n2_ar_pip_nn.cpp:41-48).",
        "#pragma once",
        "#include <hip/hip_runtime.h>",
        "",
        "#define HH_PIP_FN __host__ __device__ __forceinline__",
        "",
        "namespace hh {",
        "",
        f"constexpr int kPipPes = {len(pes)};",
        f"constexpr int kPipDip = {len(dip)};",
        "",
    ]
    emit_blocks("pip_n2ar_pes_block", pes, lines)
    emit_blocks("pip_n2ar_dip_block", dip, lines)
    lines.append("}  // namespace hh")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_oracle_table(path):
    lines = [
        "/* Generated by hawaiihybrid_amd/data/gen_pip_n2ar.py -- do not edit.",
        " * Representative monomials (distance indices, -1 = unused) of the N2-Ar PIP bases and",
        " * the atom permutations generating each basis' group; the oracle expands the orbits. */",
        "#ifndef PIP_N2AR_TABLE_H",
        "#define PIP_N2AR_TABLE_H",
        "",
    ]

    def table(name, reps):
        lines.append(f"static const signed char {name}[{len(reps)}][4] = {{")
        for r in reps:
            v = list(r) if r is not None else []
            v = v + [-1] * (4 - len(v))
            lines.append("    {" + ", ".join(str(x) for x in v) + "},")
        lines.append("};")

    table("ORC_PIP_PES_REPS", parse_reps(PES_REPS))
    lines.append("static const int ORC_PIP_PES_GENS[1][3] = {{1, 0, 2}};")
    for b, (reps, gens) in enumerate(zip(DIP_REPS, DIP_GENS)):
        table(f"ORC_PIP_DIP{b}_REPS", parse_reps(reps))
        g = ", ".join("{" + ", ".join(str(x) for x in p) + "}" for p in gens)
        lines.append(f"static const int ORC_PIP_DIP{b}_NGENS = {len(gens)};")
        lines.append(f"static const int ORC_PIP_DIP{b}_GENS[{len(gens)}][5] = {{{g}}};")
    lines.append("")
    lines.append("#endif")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    write_device_header(os.path.join(REPO, "hawaiihybrid_amd", "csrc", "pip_n2ar.h"))
    write_oracle_table(os.path.join(REPO, "oracle", "pip_n2ar_table.h"))


if __name__ == "__main__":
    main()

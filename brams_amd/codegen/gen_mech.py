#!/usr/bin/env python3
"""Generate a mechanism's device functions for the chemistry kernels.

Reads brams_amd/mechanism/<mech>.json (relacs_tuv by default; relacs_mx,
cb07_tuv, racm_tuv with --mech) and writes
brams_amd/csrc/generated/<mech>_gen.h with straight-line HIP device code:

  mech_kinetic  rk[128] from T, p, rv, J[17]       (chem_spack_kinetic.f90)
  mech_fun      f[47] = S w(rk, y)                  (rates.f90 + fexchem.f90)
  mech_mask     value-based sparsity mask of one cell's matrix
                1/(gamma h) I - J                   (rodas3:449-482)
  mech_factor   masked matrix assembled from rk, y0 and factored in place
                (Crout, reciprocal pivots, as spFactor.c:336-401)
  mech_solve    forward / backward substitution     (spSolve.c:183-218)

Sums keep the reference's term order (reactions ascending).  The LU uses one
pivot order fixed here: Markowitz on the structural pattern with diagonal
pivots, ties to the lowest species index (the oracle uses the same rule).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
MECH = "relacs_tuv"
GEN = os.path.join(PKG, "csrc", "generated")


def json_path():
    return os.path.join(PKG, "mechanism", MECH + ".json")


# J relations that tuvDriver applies after the TUV mapping, per
# chemical_mechanism (ModTuvDriver2.7.f90:1333-1368), in the reference's
# order: (target, expression over v[] with 0-based indices).  Constants are
# REAL literals.
J_RELATIONS = {
    "RELACS": [(13, "(double)0.962055f * v[12] + (double)0.0106247f * v[8]"),
               (15, "12.0 * v[12] + 208.0 * v[14]")],
    "RACM": [(14, "(double)0.28f * v[8]"), (19, "(double)0.2f * v[0]"), (13, "v[12]"),
             (15, "(v[15] + (double)8.7e-07f) / 2.0"), (22, "v[15]")],
    "CB07": [(14, "v[13]")],
}


def lit(x):
    x = float(x)
    s = repr(x)
    if "e" not in s and "." not in s and "inf" not in s and "nan" not in s:
        s += ".0"
    return s


def load():
    m = json.load(open(json_path()))
    ns, nr = m["nspecies"], m["nr"]
    react = [[x - 1 for x in r["reactants"]] for r in m["reactions"]]
    prod = {sp - 1: [(c, r - 1) for c, r in terms] for sp, terms in m["production"]}
    return m, ns, nr, react, prod


def jac_terms(ns, react, prod):
    present = {}
    for sp, terms in prod.items():
        for c, r in terms:
            present[(sp, r)] = c
    jt = {}
    for i in range(ns):
        for j in range(ns):
            terms = []
            for r, rs in enumerate(react):
                if (i, r) not in present:
                    continue
                for slot, s in enumerate(rs):
                    if s == j:
                        other = rs[1 - slot] if len(rs) == 2 else -1
                        terms.append((present[(i, r)], r, other))
            if terms:
                jt[(i, j)] = terms
    return jt


def markowitz(ns, struct):
    """Diagonal Markowitz order; returns perm (internal -> species) and the
    filled pattern in species indices."""
    P = [row[:] for row in struct]
    active = [True] * ns
    perm = []
    for _ in range(ns):
        best, bestprod = -1, None
        for c in range(ns):
            if not active[c]:
                continue
            rc = sum(1 for j in range(ns) if active[j] and P[c][j]) - 1
            cc = sum(1 for i in range(ns) if active[i] and P[i][c]) - 1
            prod = rc * cc
            if best < 0 or prod < bestprod:
                best, bestprod = c, prod
        perm.append(best)
        active[best] = False
        for i in range(ns):
            if active[i] and P[i][best]:
                for j in range(ns):
                    if active[j] and P[best][j]:
                        P[i][j] = True
    return perm, P


# rate-law families -> device helpers in mech_families.h
def rate_expr(fam, p, r):
    if fam == "unset":
        return "0.0"
    if fam == "photolysis":
        return "J[%d]" % (int(p[0]) - 1)
    if fam == "constant":
        return lit(p[0])
    if fam == "arrhenius":
        return "rl_arr(T, %s, %s)" % (lit(p[0]), lit(p[1]))
    if fam == "arrhenius_tn":
        return "rl_arr_tn(T, lnT, %s, %s, %s)" % (lit(p[0]), lit(p[1]), lit(p[2]))
    if fam == "exp_sum":
        c0 = p[0]
        terms = [p[1 + 3 * t: 4 + 3 * t] for t in range((len(p) - 1) // 3)]
        parts = []
        for c, n, b in terms:
            if n != 0.0:
                parts.append("%s * exp(%s * lnT - %s / T)" % (lit(c), lit(n), lit(b)))
            else:
                parts.append("%s * exp(-(%s / T))" % (lit(c), lit(b)))
        return "(" + " + ".join(parts) + ") + %s" % lit(c0)
    if fam == "m_power":
        return "M * %s * pow(T / 3.0e2, %s)" % (lit(p[0]), lit(p[1]))
    if fam == "exp_plus_m_exp":
        return "%s * exp(%s / T) + %s * M * exp(%s / T)" % tuple(lit(x) for x in p)
    if fam == "troe":
        return "rl_troe(T, M, %s, %s, %s, %s, %s)" % tuple(lit(x) for x in p)
    if fam == "hno3_oh":
        return "rl_hno3_oh(T, M, %s, %s, %s, %s, %s, %s)" % tuple(lit(x) for x in p)
    if fam == "linear_m":
        return "%s * (1.0 + %s * M)" % (lit(p[0]), lit(p[1]))
    if fam == "troe_ca":
        return "rl_troe_ca(T, M, %s, %s, %s, %s, %s)" % tuple(lit(x) for x in p)
    if fam == "h2o_sq":
        return "%s * H2O * H2O" % lit(p[0])
    if fam == "m_exp_h2o":
        return "%s * M * exp(%s / T) * (1.0 + %s * H2O * exp(%s / T))" % tuple(lit(x) for x in p)
    raise ValueError(fam)


def nflop_sum_pl(coefs):
    """FP64 operations of a sum of terms c * x (c = 1 not multiplied)."""
    if not coefs:
        return 0
    return (len(coefs) - 1) + sum(1 for c in coefs if abs(c) != 1.0)


def main():
    m, ns, nr, react, prod = load()
    jt = jac_terms(ns, react, prod)
    struct = [[False] * ns for _ in range(ns)]
    for (i, j) in jt:
        struct[i][j] = True
    for i in range(ns):
        struct[i][i] = True
    # structural entries in row-major order: these carry the mask bits
    sentries = [(i, j) for i in range(ns) for j in range(ns) if struct[i][j]]
    sidx = {e: k for k, e in enumerate(sentries)}
    perm, filled = markowitz(ns, struct)
    pos = {s: k for k, s in enumerate(perm)}
    F = [[filled[perm[I]][perm[J]] for J in range(ns)] for I in range(ns)]
    # LU storage slot of each (I,J) in the filled pattern (internal indices)
    slot = {}
    for J in range(ns):
        for I in range(ns):
            if F[I][J]:
                slot[(I, J)] = len(slot)
    nlu = len(slot)
    nwords = (len(sentries) + 31) // 32

    L = []
    w = L.append
    w("// Generated by brams_amd/codegen/gen_mech.py from brams_amd/mechanism/%s.json." % MECH)
    w("// %s: %d species, %d reactions, %d Jacobian entries, LU nnz %d." % (m["name"], ns, nr, len(jt), nlu))
    w("#pragma once")
    w("#include \"mech_families.h\"")
    w("#define MECH_NAME \"%s\"" % m["name"])
    w("#define MECH_NS %d" % ns)
    w("#define MECH_NR %d" % nr)
    w("#define MECH_NPHOT %d" % m["nr_photo"])
    w("#define MECH_NSTRUCT %d" % len(sentries))
    w("#define MECH_MASK_WORDS %d" % nwords)
    w("#define MECH_NLU %d" % nlu)
    w("// pivot order (internal -> species, 0-based): %s" % perm)
    w("static const int mech_perm_host[MECH_NS] = {%s};" % ", ".join(map(str, perm)))
    w("static const int mech_struct_i_host[MECH_NSTRUCT] = {%s};" % ", ".join(str(i) for i, _ in sentries))
    w("static const int mech_struct_j_host[MECH_NSTRUCT] = {%s};" % ", ".join(str(j) for _, j in sentries))
    w("")
    # ---------------- kinetic
    w("// chem_spack_kinetic.f90:7-435")
    w("__device__ __forceinline__ void mech_kinetic(double T, double P, double vapp, const double* J, double* rk) {")
    w("  const double M = P * 7.243e16 / T;")
    w("  const double H2O = 29.0 * M * vapp / (18.0 + 11.0 * vapp);")
    w("  const double lnT = log(T);")
    w("  (void)H2O; (void)lnT;")
    for r, rx in enumerate(m["reactions"]):
        law = rx["rate"]
        e = rate_expr(law["family"], law["params"], r)
        post = law["post"]
        if post:
            kind = post[0][0]
            if kind == "mul_M_const":
                w("  rk[%d] = (%s) * M * %s;" % (r, e, lit(post[0][1])))
            elif kind == "mul_M":
                w("  rk[%d] = (%s) * M;" % (r, e))
            elif kind == "mul_H2O":
                w("  rk[%d] = (%s) * H2O;" % (r, e))
            elif kind == "mul_troe":
                w("  rk[%d] = rl_troe(T, M, %s) * (%s);" % (r, ", ".join(lit(x) for x in post[0][1:]), e))
            else:
                raise ValueError(kind)
        else:
            w("  rk[%d] = %s;" % (r, e))
    w("}")
    w("")
    # ---------------- rates + fun
    def wexpr(r):
        rs = react[r]
        s = "rk[%d]" % r
        for x in rs:
            s = "%s * y[%d]" % (s, x)
        return s

    def sum_terms(terms):
        out = ""
        for k, (c, ref) in enumerate(terms):
            if c == 1.0:
                t = ref if k == 0 else " + " + ref
            elif c == -1.0:
                t = "-" + ref if k == 0 else " - " + ref
            elif c < 0:
                t = ("%s * %s" % (lit(c), ref)) if k == 0 else " - %s * %s" % (lit(-c), ref)
            else:
                t = ("%s * %s" % (lit(c), ref)) if k == 0 else " + %s * %s" % (lit(c), ref)
            out += t
        return out

    w("// chem_spack_rates.f90:63-192 and chem_spack_fexchem.f90:73-676")
    w("__device__ __forceinline__ void mech_fun(const double* rk, const double* y, double* f) {")
    used = sorted({r for terms in prod.values() for _, r in terms})
    for r in used:
        w("  const double w%d = %s;" % (r, wexpr(r)))
    for s in range(ns):
        terms = prod.get(s, [])
        if not terms:
            w("  f[%d] = 0.0;" % s)
        else:
            w("  f[%d] = %s;" % (s, sum_terms([(c, "w%d" % r) for c, r in terms])))
    w("}")
    w("")

    # ---------------- production / loss of the QSSA split
    def dwexpr(r, s):
        rs = react[r]
        if s not in rs:          # negative product (CB07_TUV): dratedc never assigns dw(r,s)
            return "0.0"
        if len(rs) == 1:
            return "rk[%d]" % r
        o = rs[1] if rs[0] == s else rs[0]
        return "(rk[%d] * y[%d])" % (r, o)

    w("// QSSA's P - L c split: chem_spack_rates.f90:63-192, chem_spack_dratedc.f90:62-296,")
    w("// chem_spack_fexprod.f90:64-448, chem_spack_fexloss.f90:64-284.  P: the positive")
    w("// stoichiometric terms c w_r, L: -c dw_rs over the negative ones (dw_rs = rk_r y_other,")
    w("// y_s itself in a self-reaction), in reaction order; rows with no term are 0.")
    w("__device__ __forceinline__ void mech_prod_loss(const double* rk, const double* y, double* P, double* L) {")
    upos = sorted({r for terms in prod.values() for c, r in terms if c > 0})
    for r in upos:
        w("  const double w%d = %s;" % (r, wexpr(r)))
    fl_pl = sum(len(react[r]) for r in upos)
    for s in range(ns):
        pt = [(c, "w%d" % r) for c, r in prod.get(s, []) if c > 0]
        lt = [(-c, dwexpr(r, s)) for c, r in prod.get(s, []) if c < 0]
        w("  P[%d] = %s;" % (s, sum_terms(pt) if pt else "0.0"))
        w("  L[%d] = %s;" % (s, sum_terms(lt) if lt else "0.0"))
        fl_pl += nflop_sum_pl([c for c, _ in pt]) + nflop_sum_pl([c for c, _ in lt])
        fl_pl += sum(1 for _, e in lt if "*" in e)
    w("}")
    w("")

    def jexpr(i, j):
        terms = []
        for c, r, o in jt[(i, j)]:
            ref = "(rk[%d] * y[%d])" % (r, o) if o >= 0 else "rk[%d]" % r
            terms.append((c, ref))
        return sum_terms(terms)

    def mval(i, j):
        """entry of 1/(gamma h) I - J for species (i,j)"""
        if (i, j) in jt:
            if i == j:
                return "igdt - (%s)" % jexpr(i, j)
            return "-(%s)" % jexpr(i, j)
        return "igdt" if i == j else None

    def mbit(e):
        return "((mw[%d] >> %d) & 1u)" % (e // 32, e % 32)

    # ---------------- mask of one cell
    w("// rodas3:470-482: keep the entries of DLmat that are nonzero in this cell")
    w("__device__ __forceinline__ void mech_mask(const double* rk, const double* y, double igdt, unsigned* mw) {")
    w("  unsigned acc[MECH_MASK_WORDS] = {0};")
    for e, (i, j) in enumerate(sentries):
        w("  acc[%d] |= ((%s) != 0.0 ? 1u : 0u) << %d;" % (e // 32, mval(i, j), e % 32))
    w("  for (int k = 0; k < MECH_MASK_WORDS; ++k) mw[k] = acc[k];")
    w("}")
    w("")
    # ---------------- factor
    w("// spFactor.c:336-401 on the fixed pivot order; lu[] keeps L (reciprocal")
    w("// diagonal) and unit-diagonal U.  Returns nonzero on a zero pivot.  With")
    w("// COUNT, *viol is increased by the pivots that fail Sparse1.3's threshold")
    w("// rule, which accepts a pivot only if RelThreshold (1e-3, spAllocate.c:191)")
    w("// times the largest other entry of its reduced column is below |pivot|")
    w("// (spFactor.c:202-203, FindBiggestInColExclude spFactor.c:2019-2050).")
    w("template <bool COUNT>")
    w("__device__ __forceinline__ int mech_factor(const double* rk, const double* y, double igdt,")
    w("                                           const unsigned* mw, double* lu, int* viol) {")
    w("  int bad = 0, nv = 0;")
    for J in range(ns):
        rows = [I for I in range(ns) if F[I][J]]
        w("  { // column %d (species %d)" % (J, perm[J]))
        for I in rows:
            i, j = perm[I], perm[J]
            v = mval(i, j)
            if v is None:
                w("    double d%d = 0.0;" % I)
            else:
                w("    double d%d = %s ? (%s) : 0.0;" % (I, mbit(sidx[(i, j)]), v))
        for K in [K for K in rows if K < J]:
            w("    { const double u = d%d * lu[%d]; lu[%d] = u;" % (K, slot[(K, K)], slot[(K, J)]))
            for I in range(K + 1, ns):
                if F[I][K]:
                    w("      d%d -= u * lu[%d];" % (I, slot[(I, K)]))
            w("    }")
        below = [I for I in rows if I > J]
        for I in below:
            w("    lu[%d] = d%d;" % (slot[(I, J)], I))
        if below:
            w("    if constexpr (COUNT) { double cm = fabs(d%d);" % below[0])
            for I in below[1:]:
                w("      cm = fmax(cm, fabs(d%d));" % I)
            w("      nv += (cm * 1.0e-3 >= fabs(d%d)) ? 1 : 0; }" % J)
        else:
            w("    if constexpr (COUNT) nv += (d%d == 0.0) ? 1 : 0;" % J)
        w("    bad |= (d%d == 0.0); lu[%d] = 1.0 / d%d;" % (J, slot[(J, J)], J))
        w("  }")
    w("  if constexpr (COUNT) *viol += nv;")
    w("  return bad;")
    w("}")
    w("")
    # ---------------- solve
    w("// spSolve.c:183-218: x (species order) overwritten by the solution")
    w("__device__ __forceinline__ void mech_solve(const double* lu, double* x) {")
    for I in range(ns):
        w("  double c%d = x[%d];" % (I, perm[I]))
    for I in range(ns):
        w("  c%d *= lu[%d];" % (I, slot[(I, I)]))
        for R in range(I + 1, ns):
            if F[R][I]:
                w("  c%d -= c%d * lu[%d];" % (R, I, slot[(R, I)]))
    for I in range(ns - 1, -1, -1):
        for Jc in range(I + 1, ns):
            if F[I][Jc]:
                w("  c%d -= lu[%d] * c%d;" % (I, slot[(I, Jc)], Jc))
    for I in range(ns):
        w("  x[%d] = c%d;" % (perm[I], I))
    w("}")
    # ---------------- algorithmic FP64 operation counts of the generated code
    def nflop_sum(coefs):
        return (len(coefs) - 1) + sum(1 for c in coefs if abs(c) != 1.0)

    fl_fun = sum(len(react[r]) for r in used) + sum(
        nflop_sum([c for c, _ in prod[s]]) for s in prod)
    fl_jac = 0
    for (i, j), terms in jt.items():
        fl_jac += sum(1 for _, _, o in terms if o >= 0) + nflop_sum([c for c, _, _ in terms])
        fl_jac += 1 if i == j else 0
    fl_elim = 0
    for J in range(ns):
        for K in range(J):
            if F[K][J]:
                fl_elim += 1 + 2 * sum(1 for I in range(K + 1, ns) if F[I][K])
        fl_elim += 1                                    # reciprocal pivot
    nl = sum(1 for (I, J) in slot if I > J)
    nu = sum(1 for (I, J) in slot if I < J)
    fl_solve = ns + 2 * nl + 2 * nu
    tr = {"exp": 0, "log": 1, "log10": 0, "pow": 0}
    ar_kin = 7
    for rx in m["reactions"]:
        fam, p, post = rx["rate"]["family"], rx["rate"]["params"], rx["rate"]["post"]
        if fam == "arrhenius":
            tr["exp"] += 1; ar_kin += 2
        elif fam == "arrhenius_tn":
            tr["exp"] += 1; ar_kin += 4
        elif fam == "exp_sum":
            nt = (len(p) - 1) // 3
            tr["exp"] += nt; ar_kin += 4 * nt + 1
        elif fam == "m_power":
            tr["pow"] += 1; ar_kin += 3
        elif fam == "exp_plus_m_exp":
            tr["exp"] += 2; ar_kin += 6
        elif fam == "troe":
            tr["pow"] += 3; tr["log10"] += 1; ar_kin += 14
        elif fam == "hno3_oh":
            tr["exp"] += 3; ar_kin += 10
        elif fam == "linear_m":
            ar_kin += 3
        for q in post:
            if q[0] == "mul_M_const":
                ar_kin += 2
            elif q[0] == "mul_H2O":
                ar_kin += 1
            elif q[0] == "mul_troe":
                tr["pow"] += 3; tr["log10"] += 1; ar_kin += 15
    flops = {"fun": fl_fun, "prod_loss": fl_pl, "jacobian_assembly": fl_jac, "elimination": fl_elim,
             "factor": fl_jac + fl_elim, "solve": fl_solve, "nlu": nlu, "nl": nl, "nu": nu,
             "kinetic_arith": ar_kin, "kinetic_transcendentals": tr,
             "note": "FP64 add/sub/mul/div counted once each (an FMA is 2), "
                     "transcendentals listed separately"}
    for k in ("fun", "factor", "solve"):
        w("#define MECH_FLOPS_%s %d" % (k.upper(), flops[k]))

    w("")
    # ---------------- J relations after the TUV mapping
    rel = J_RELATIONS.get(m.get("chemical_mechanism", ""), [])
    w("// J relations of tuvDriver for %s (ModTuvDriver2.7.f90:1333-1368), each" % m.get("chemical_mechanism"))
    w("// product and sum rounded as the reference's x86-64 build does (no FMA)")
    w("#define MECH_J_RELATIONS %d" % len(rel))
    w("__device__ __forceinline__ void mech_j_relations(double* v) {")
    w("#pragma clang fp contract(off)")
    if not rel:
        w("  (void)v;")
    for t, e in rel:
        w("  v[%d] = %s;" % (t, e))
    w("}")

    os.makedirs(GEN, exist_ok=True)
    with open(os.path.join(GEN, MECH + "_flops.json"), "w") as fh:
        json.dump(flops, fh, indent=1)
    text = "\n".join(L) + "\n"
    emit_slab(m, ns, react, prod, jt, perm, F, slot, sentries, sidx, wexpr, sum_terms, jexpr, rate_expr)
    out = os.path.join(GEN, MECH + "_gen.h")
    old = open(out).read() if os.path.exists(out) else None
    if old != text:
        with open(out, "w") as fh:
            fh.write(text)
    print("generated %s: nstruct=%d nlu=%d" % (out, len(sentries), nlu))


def slab_out():
    return os.path.join(GEN, MECH + "_slab.h")
RL_REGS = 120        # factorisation entries kept in registers at most


def emit_slab(m, ns, react, prod, jt, perm, F, slot, sentries, sidx, wexpr, sum_terms, jexpr, rate_expr):
    """Device functions of the slab kernel (rodas3_slab.h), templated over
    memory accessors: an accessor is anything with operator[](int) returning
    the lane's element, so the same generated code runs on registers, LDS and
    global scratch.

      mech_kinetic_t  rk[128] written through an accessor
      mech_fun_t      f = S w(rk, y), each rate constant loaded once
      mech_jac_t      the Jacobian entries J[e] (e in mech_jac order), column
                      by column with the column's rate constants and
                      concentrations loaded once
      mech_factor_rl  factorisation of the masked 1/(gamma h) I - J read from
                      J[e], right-looking with lazy assembly: every entry is
                      a register from its first update to its final value,
                      which is written once to the LU accessor; at most 169
                      entries are live at a time.  Each entry receives its
                      updates in ascending pivot order and U(K,J) =
                      A(K,J) * (1/pivot_K), the operations of mech_factor in
                      the same order, so the factors are bit-identical.
      mech_solve_t    mech_solve reading the LU through an accessor
    """
    L = []
    w = L.append
    jorder = sorted(jt.keys(), key=lambda e: (e[1], e[0]))          # column-major
    jidx = {e: k for k, e in enumerate(jorder)}
    nlu = len(slot)
    w("// Generated by brams_amd/codegen/gen_mech.py from brams_amd/mechanism/%s.json." % MECH)
    w("// Accessor-templated device functions of the slab kernel (rodas3_slab.h).")
    w("#pragma once")
    w("#include \"%s_gen.h\"" % MECH)
    w("#define MECH_NJAC %d" % len(jorder))
    w("")
    # ---- kinetic
    w("// chem_spack_kinetic.f90:7-435")
    w("template <class RK>")
    w("__device__ __forceinline__ void mech_kinetic_t(double T, double P, double vapp, const double* J, RK& rk) {")
    w("  const double M = P * 7.243e16 / T;")
    w("  const double H2O = 29.0 * M * vapp / (18.0 + 11.0 * vapp);")
    w("  const double lnT = log(T);")
    w("  (void)H2O; (void)lnT;")
    for r, rx in enumerate(m["reactions"]):
        law = rx["rate"]
        e = rate_expr(law["family"], law["params"], r)
        post = law["post"]
        if post:
            kind = post[0][0]
            if kind == "mul_M_const":
                w("  rk[%d] = (%s) * M * %s;" % (r, e, lit(post[0][1])))
            elif kind == "mul_M":
                w("  rk[%d] = (%s) * M;" % (r, e))
            elif kind == "mul_H2O":
                w("  rk[%d] = (%s) * H2O;" % (r, e))
            elif kind == "mul_troe":
                w("  rk[%d] = rl_troe(T, M, %s) * (%s);" % (r, ", ".join(lit(x) for x in post[0][1:]), e))
            else:
                raise ValueError(kind)
        else:
            w("  rk[%d] = %s;" % (r, e))
    w("}")
    w("")
    # ---- fun, reaction-major: each rate w_r is added to the sums of its
    # species as soon as it is formed, so only the 47 sums are live.  Every
    # sum still receives its terms in ascending reaction order with the same
    # coefficients, the evaluation order of mech_fun.
    used = sorted({r for terms in prod.values() for _, r in terms})
    by_r = {}
    for sp in range(ns):
        for k, (c, r) in enumerate(prod.get(sp, [])):
            by_r.setdefault(r, []).append((sp, c, k == 0))
    w("// chem_spack_rates.f90:63-192 and chem_spack_fexchem.f90:73-676")
    w("template <class RK, class Y, class FO>")
    w("__device__ __forceinline__ void mech_fun_t(const RK& rk, const Y& y, FO& f) {")
    for sp in range(ns):
        if prod.get(sp):
            w("  double f%d;" % sp)
    for r in used:
        w("  { const double w = %s;" % wexpr(r))
        for sp, c, first in by_r[r]:
            if first:
                if c == 1.0:
                    w("    f%d = w;" % sp)
                elif c == -1.0:
                    w("    f%d = -w;" % sp)
                else:
                    w("    f%d = %s * w;" % (sp, lit(c)))
            else:
                if c == 1.0:
                    w("    f%d = f%d + w;" % (sp, sp))
                elif c == -1.0:
                    w("    f%d = f%d - w;" % (sp, sp))
                elif c < 0:
                    w("    f%d = f%d - %s * w;" % (sp, sp, lit(-c)))
                else:
                    w("    f%d = f%d + %s * w;" % (sp, sp, lit(c)))
        w("  }")
    for sp in range(ns):
        w("  f[%d] = %s;" % (sp, ("f%d" % sp) if prod.get(sp) else "0.0"))
    w("}")
    w("")
    # ---- jacobian entries, column by column
    w("// chem_spack_dratedc.f90 + chem_spack_jacdchemdc.f90: entry e of the")
    w("// column-major list mech_jac_i/j (0-based species) is dF_i/dy_j")
    w("static const int mech_jac_i_host[MECH_NJAC] = {%s};" % ", ".join(str(i) for i, _ in jorder))
    w("static const int mech_jac_j_host[MECH_NJAC] = {%s};" % ", ".join(str(j) for _, j in jorder))
    w("// The call's value-based mask (rodas3:470-482) is applied here, once per")
    w("// sub-step: an off-diagonal entry outside the mask is stored as -0.0, so")
    w("// that its matrix entry -J is the +0.0 of the plain kernel; diagonal")
    w("// entries are stored unmasked (their mask bit is applied per attempt,")
    w("// with 1/(gamma h)).")
    w("template <class RK, class Y, class JO>")
    w("__device__ __forceinline__ void mech_jac_t(const RK& rk, const Y& y, const unsigned* mw, JO& jo) {")
    col = None
    for (i, j) in jorder:
        if j != col:
            if col is not None:
                w("  }")
            col = j
            rs = sorted({r for (ii, jj) in jorder if jj == j for _, r, _ in jt[(ii, jj)]})
            ys = sorted({o for (ii, jj) in jorder if jj == j for _, _, o in jt[(ii, jj)] if o >= 0})
            w("  { // column %d" % j)
            for r in rs:
                w("    const double k%d = rk[%d];" % (r, r))
            for o in ys:
                w("    const double c%d = y[%d];" % (o, o))
        terms = []
        for c, r, o in jt[(i, j)]:
            ref = "(k%d * c%d)" % (r, o) if o >= 0 else "k%d" % r
            terms.append((c, ref))
        if i == j:
            w("    jo[%d] = %s;" % (jidx[(i, j)], sum_terms(terms)))
        else:
            e = sidx[(i, j)]
            w("    jo[%d] = ((mw[%d] >> %d) & 1u) ? (%s) : -0.0;" % (jidx[(i, j)], e // 32, e % 32,
                                                             sum_terms(terms)))
    w("  }")
    w("}")
    w("")
    # ---- value-based mask of one cell (rodas3:470-482) from its Jacobian
    # entries: an off-diagonal entry is kept where -J != 0, a diagonal one
    # where 1/(gamma h) - J_ii != 0 (the values mech_mask tests)
    w("// rodas3:470-482 from stored Jacobian entries: the off-diagonal bits")
    w("// (constant over a sub-step) and the diagonal bits (one per attempt)")
    w("template <class JR>")
    w("__device__ __forceinline__ void mech_mask_offdiag(const JR& jac, unsigned* mw) {")
    w("  unsigned acc[MECH_MASK_WORDS] = {0};")
    for e, (i, j) in enumerate(sentries):
        if i != j:
            w("  acc[%d] |= (jac[%d] != 0.0 ? 1u : 0u) << %d;" % (e // 32, jidx[(i, j)], e % 32))
    w("  for (int k = 0; k < MECH_MASK_WORDS; ++k) mw[k] = acc[k];")
    w("}")
    w("template <class JR>")
    w("__device__ __forceinline__ void mech_mask_diag(const JR& jac, double igdt, unsigned* mw) {")
    w("  unsigned acc[MECH_MASK_WORDS] = {0};")
    for e, (i, j) in enumerate(sentries):
        if i == j:
            v = "igdt - jac[%d]" % jidx[(i, i)] if (i, i) in jt else "igdt"
            w("  acc[%d] |= ((%s) != 0.0 ? 1u : 0u) << %d;" % (e // 32, v, e % 32))
    w("  for (int k = 0; k < MECH_MASK_WORDS; ++k) mw[k] = acc[k];")
    w("}")
    w("")
    # ---- right-looking lazy factorisation
    Lcol = [[I for I in range(K + 1, ns) if F[I][K]] for K in range(ns)]
    Urow = [[J for J in range(K + 1, ns) if F[K][J]] for K in range(ns)]
    birth = {}
    for (I, J) in slot:
        k0 = min(I, J)
        for K in range(min(I, J)):
            if F[I][K] and F[K][J]:
                k0 = K
                break
        birth.setdefault(k0, []).append((I, J))

    def mbit(e):
        return "((mw[%d] >> %d) & 1u)" % (e // 32, e % 32)

    # Statements in emission order; each names the entries it reads or
    # writes, so that registers can be budgeted: when more than RL_REGS
    # entries are live, the one whose next use is farthest (Belady) is parked
    # in its own LU slot (the factorisation never reads LU slots, and the
    # slot's final value is written later) and reloaded before its next use.
    stmts = []          # (kind, text, entries used, entries born, entries dying)
    for K in range(ns):
        stmts.append(("code", "  // step %d (species %d)" % (K, perm[K]), [], [], []))
        for (I, J) in sorted(birth.get(K, [])):
            i, j = perm[I], perm[J]
            # diagonal mask bits as a 0/1 factor: with the bit set the value
            # is exact; a cleared bit gives a zero pivot either way, which
            # ends the block at its first attempt as in mech_factor
            if (i, j) in jt and i == j:
                txt = "  double a%d_%d = (igdt - jac[%d]) * (double)%s;" % (I, J, jidx[(i, j)], mbit(sidx[(i, j)]))
            elif (i, j) in jt:
                txt = "  double a%d_%d = -jac[%d];" % (I, J, jidx[(i, j)])
            elif i == j:
                txt = "  double a%d_%d = igdt * (double)%s;" % (I, J, mbit(sidx[(i, j)]))
            else:
                txt = "  double a%d_%d = 0.0;" % (I, J)
            stmts.append(("code", txt, [], [(I, J)], []))
        below = Lcol[K]
        for I in below:
            stmts.append(("code", "  lu[%d] = a%d_%d;" % (slot[(I, K)], I, K), [(I, K)], [], []))
        if below:
            stmts.append(("code", "  { double cm = fabs(a%d_%d);" % (below[0], K), [(below[0], K)], [], []))
            for I in below[1:]:
                stmts.append(("code", "    cm = fmax(cm, fabs(a%d_%d));" % (I, K), [(I, K)], [], []))
            stmts.append(("code", "    nv += (cm * 1.0e-3 >= fabs(a%d_%d)) ? 1 : 0; }" % (K, K), [(K, K)], [], []))
        else:
            stmts.append(("code", "  nv += (a%d_%d == 0.0) ? 1 : 0;" % (K, K), [(K, K)], [], []))
        stmts.append(("code", "  bad |= (a%d_%d == 0.0);" % (K, K), [(K, K)], [], []))
        stmts.append(("code", "  const double r%d = 1.0 / a%d_%d; lu[%d] = r%d;" % (K, K, K, slot[(K, K)], K),
                      [(K, K)], [], [(K, K)]))
        for J in Urow[K]:
            stmts.append(("code", "  { const double u = a%d_%d * r%d; lu[%d] = u;" % (K, J, K, slot[(K, J)]),
                          [(K, J)], [], [(K, J)]))
            for I in below:
                if F[I][J]:
                    stmts.append(("code", "    a%d_%d -= u * a%d_%d;" % (I, J, I, K), [(I, J), (I, K)], [], []))
            stmts.append(("code", "  }", [], [], []))
        stmts.append(("die", "", [], [], [(I, K) for I in below]))
    # next-use tables
    uses = {}
    for n, (_, _, used_e, born_e, _) in enumerate(stmts):
        for e in used_e + born_e:
            uses.setdefault(e, []).append(n)
    ptr = {e: 0 for e in uses}

    def next_use(e, n):
        lst, k = uses[e], ptr[e]
        while k < len(lst) and lst[k] <= n:
            k += 1
        ptr[e] = k
        return lst[k] if k < len(lst) else 1 << 30

    live, parked, nload, nstore = set(), set(), 0, 0
    body = []
    for n, (kind, txt, used_e, born_e, dying) in enumerate(stmts):
        if kind == "die":
            for e in dying:
                live.discard(e)
            continue
        for e in used_e:
            if e in parked:
                parked.discard(e)
                body.append("  a%d_%d = lu[%d];" % (e[0], e[1], slot[e]))
                nload += 1
            live.add(e)
        for e in born_e:
            live.add(e)
        while len(live) > RL_REGS:
            cands = [x for x in live if x not in used_e and x not in born_e]
            victim = max(cands, key=lambda x: next_use(x, n))
            live.discard(victim)
            parked.add(victim)
            body.append("  lu[%d] = a%d_%d;" % (slot[victim], victim[0], victim[1]))
            nstore += 1
        body.append(txt)
        for e in dying:
            live.discard(e)

    w("// spFactor.c:336-401 arithmetic, right-looking with lazy assembly: A(I,J) =")
    w("// masked 1/(gamma h) I - J in pivot order; lu[] gets L (reciprocal pivots)")
    w("// and unit-diagonal U.  At most %d entries are kept in registers; the" % RL_REGS)
    w("// others wait in their own LU slot (%d parkings, %d reloads per call)." % (nstore, nload))
    w("// *viol counts pivots failing Sparse1.3's threshold rule")
    w("// (spFactor.c:202-203).  Returns nonzero on a zero pivot.")
    w("#define MECH_RL_PARK_STORES %d" % nstore)
    w("#define MECH_RL_PARK_LOADS %d" % nload)
    w("template <class JR, class LU>")
    w("__device__ __forceinline__ int mech_factor_rl(const JR& jac, double igdt, const unsigned* mw, LU& lu, int* viol) {")
    w("  int bad = 0, nv = 0;")
    L.extend(body)
    w("  *viol += nv;")
    w("  return bad;")
    w("}")
    w("")
    # ---- solve
    w("// spSolve.c:183-218: x (species order) overwritten by the solution")
    w("template <class LU, class X>")
    w("__device__ __forceinline__ void mech_solve_t(const LU& lu, X& x) {")
    for I in range(ns):
        w("  double c%d = x[%d];" % (I, perm[I]))
    for I in range(ns):
        w("  c%d *= lu[%d];" % (I, slot[(I, I)]))
        for R in range(I + 1, ns):
            if F[R][I]:
                w("  c%d -= c%d * lu[%d];" % (R, I, slot[(R, I)]))
    for I in range(ns - 1, -1, -1):
        for Jc in range(I + 1, ns):
            if F[I][Jc]:
                w("  c%d -= lu[%d] * c%d;" % (I, slot[(I, Jc)], Jc))
    for I in range(ns):
        w("  x[%d] = c%d;" % (perm[I], I))
    w("}")
    text = "\n".join(L) + "\n"
    old = open(slab_out()).read() if os.path.exists(slab_out()) else None
    if old != text:
        with open(slab_out(), "w") as fh:
            fh.write(text)


if __name__ == "__main__":
    if len(sys.argv) > 2 and sys.argv[1] == "--mech":
        MECH = sys.argv[2].lower()
    main()

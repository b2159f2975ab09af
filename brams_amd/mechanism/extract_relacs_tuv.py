#!/usr/bin/env python3
"""Extract the RELACS_TUV mechanism tables from the reference's SPACK output.

This is a one-off data-extraction step, run in a container that holds the
reference source tree.  It reads the generated Fortran kernels as *text* and
writes ``relacs_tuv.json`` beside this file: species (order, molar weight,
transport flag), reactions (ordered reactants, rate-law family + parameters)
and the stoichiometric table in the order the reference sums it.  Nothing in
the library or its tests runs this script; they read the JSON.

Sources (paths relative to the reference root):
  src/brams/ccatt/RELACS_TUV/chem1_list.f90          species, weights, flags
  src/brams/ccatt/RELACS_TUV/chem_spack_rates.f90    reactants  (w = rk*y*y)
  src/brams/ccatt/RELACS_TUV/chem_spack_fexchem.f90  stoichiometry (f = S w)
  src/brams/ccatt/RELACS_TUV/chem_spack_dratedc.f90  partials (checked only)
  src/brams/ccatt/RELACS_TUV/chem_spack_jacdchemdc.f90 Jacobian (checked only)
  src/brams/ccatt/RELACS_TUV/chem_spack_kinetic.f90  rate laws

The Jacobian is not stored: the generator derives it from the reactants and
the stoichiometry.  This script checks that the derivation reproduces the
reference's Jacobian term for term (same entries, same term order, same
coefficients), so the derived code sums in the reference's order.

Usage:  python extract_relacs_tuv.py [/root/reference] [MECHANISM_DIR]
        (MECHANISM_DIR under src/brams/ccatt: RELACS_TUV by default, or
        RACM_TUV, CB07_TUV, RELACS_MX; the output is <mechanism_dir>.json)
"""
import json
import os
import re
import sys
from collections import OrderedDict

MECH_DIR = "src/brams/ccatt/RELACS_TUV"


def read_statements(path):
    """Return Fortran statements with comments stripped and '&' lines joined."""
    out, cur = [], ""
    with open(path, encoding="latin-1") as fh:
        for raw in fh:
            line = raw.split("!")[0].rstrip()
            if not line.strip():
                continue
            s = line.strip()
            if cur:
                if s.startswith("&"):
                    s = s[1:]
                cur += " " + s
            else:
                cur = s
            if cur.endswith("&"):
                cur = cur[:-1]
            else:
                out.append(cur)
                cur = ""
    return out


def fnum(tok):
    return float(tok.lower().replace("d", "e"))


# ---------------------------------------------------------------- species
def parse_species(root):
    stmts = read_statements(os.path.join(root, MECH_DIR, "chem1_list.f90"))
    names = []
    for s in stmts:
        m = re.match(r"INTEGER,PARAMETER :: (\w+)\s*=\s*0*(\d+)$", s)
        if m and m.group(1) not in ("maxnspecies", "nspecies", "src", "ddp", "wdp",
                                    "fdda", "offline", "transport", "on", "off",
                                    "nr", "nrt", "nrh2o", "maxJcomb"):
            idx = int(m.group(2))
            if idx == len(names) + 1:
                names.append(m.group(1))
    weights, alloc, nr, nphot, ns = None, None, None, None, None
    for s in stmts:
        if s.startswith("REAL,PARAMETER,DIMENSION(nspecies) :: weight="):
            body = s.split("(/", 1)[1].rsplit("/)", 1)[0]
            weights = [t.strip() for t in body.split(",") if t.strip()]
        if "spc_alloc=RESHAPE" in s:
            body = s.split("(/", 1)[1].split("/)", 1)[0]
            alloc = [int(t) for t in body.split(",") if t.strip()]
        m = re.match(r"INTEGER,PARAMETER :: nr\s*=\s*(\d+)$", s)
        if m:
            nr = int(m.group(1))
        m = re.match(r"INTEGER,PARAMETER :: nspecies\s*=\s*0*(\d+)$", s)
        if m:
            ns = int(m.group(1))
        m = re.search(r"nr_photo=\s*(\d+)", s)
        if m:
            nphot = int(m.group(1))
    names = names[:ns]
    assert len(names) == ns, (len(names), ns)
    assert len(weights) == ns and len(alloc) == 6 * ns, (len(weights), len(alloc), ns)
    species = []
    for i, n in enumerate(names):
        species.append(OrderedDict(name=n, weight=weights[i].rstrip("."),
                                   transported=bool(alloc[6 * i + 5])))
    return species, nr, nphot


# ---------------------------------------------------------------- rates
def parse_reactants(root, nr):
    stmts = read_statements(os.path.join(root, MECH_DIR, "chem_spack_rates.f90"))
    reac = {}
    for s in stmts:
        m = re.match(r"w\(ijk,\s*(\d+)\)\s*=\s*rk\(ijk,\s*(\d+)\)((?:\s*\*\s*Y\(ijk,\s*\d+\))*)$", s)
        if m:
            r = int(m.group(1))
            assert int(m.group(2)) == r
            ys = [int(x) for x in re.findall(r"Y\(ijk,\s*(\d+)\)", m.group(3))]
            reac[r] = ys
    assert sorted(reac) == list(range(1, nr + 1))
    return [reac[r] for r in range(1, nr + 1)]


def parse_terms(rhs, ref_pat):
    """Split '- w(..,a) + 0.5D0 * w(..,b) ...' into [(coef, idx...)] in order."""
    terms = []
    pat = re.compile(r"([+-])\s*(?:([0-9.]+[dDeE][+-]?\d+)\s*\*\s*)?" + ref_pat)
    pos = 0
    rhs = rhs.strip()
    for m in pat.finditer(rhs):
        assert rhs[pos:m.start()].strip() == "", (rhs[pos:m.start()], rhs)
        sign = -1.0 if m.group(1) == "-" else 1.0
        mag = fnum(m.group(2)) if m.group(2) else 1.0
        terms.append((sign * mag, m.group(2) is not None) + tuple(int(g) for g in m.groups()[2:]))
        pos = m.end()
    assert rhs[pos:].strip() == "", rhs[pos:]
    return terms


def parse_production(root):
    stmts = read_statements(os.path.join(root, MECH_DIR, "chem_spack_fexchem.f90"))
    prod = OrderedDict()
    for s in stmts:
        m = re.match(r"chem\(ijk,\s*(\d+)\)\s*=(.*)$", s)
        if m:
            sp = int(m.group(1))
            assert sp not in prod
            prod[sp] = [(c, r) for c, explicit, r in parse_terms(m.group(2), r"w\(ijk,\s*(\d+)\)")]
    return prod


def parse_jacobian(root):
    stmts = read_statements(os.path.join(root, MECH_DIR, "chem_spack_jacdchemdc.f90"))
    jac = OrderedDict()
    for s in stmts:
        m = re.match(r"JacC\((?:ijkbeg:ijkend|ijk),\s*(\d+),\s*(\d+)\)\s*=(.*)$", s)
        if m:
            key = (int(m.group(1)), int(m.group(2)))
            assert key not in jac, key          # every entry assigned once
            jac[key] = [(c, r, sp) for c, _, r, sp in
                        parse_terms(m.group(3), r"dw\((?:ijkbeg:ijkend|ijk),\s*(\d+),\s*(\d+)\)")]
    return jac


def parse_dratedc(root):
    stmts = read_statements(os.path.join(root, MECH_DIR, "chem_spack_dratedc.f90"))
    dw = {}
    for s in stmts:
        m = re.match(r"dw\(ijk,\s*(\d+),\s*(\d+)\)\s*=\s*rk\(ijk,\s*(\d+)\)(?:\s*\*\s*Y\(ijk,\s*(\d+)\))?$", s)
        if m:
            r, sp = int(m.group(1)), int(m.group(2))
            assert int(m.group(3)) == r
            other = int(m.group(4)) if m.group(4) else 0
            if (r, sp) in dw:
                assert dw[(r, sp)] == other   # self-reactions assign the same value twice
            dw[(r, sp)] = other
    return dw


# ---------------------------------------------------------------- kinetic
NUM = r"(?:[0-9]+\.?[0-9]*|\.[0-9]+)(?:[dDeE][+-]?\d+)?"


def shape_of(expr):
    """Replace variables by short names and literals by '#'; return shape, numbers.

    A '-' or '+' directly before a literal is folded into the literal when it is
    a unary sign (start of expression or after '(' / operator)."""
    e = expr
    e = re.sub(r"\(\s*(?:ijkbeg\s*:\s*ijkend|ijk)\s*\)", "", e, flags=re.I)
    e = re.sub(r"Jphoto\(\s*(?:ijkbeg\s*:\s*ijkend|ijk)\s*,\s*(\d+)\)", r"J[\1]", e, flags=re.I)
    e = re.sub(r"rk\(\s*(?:ijkbeg\s*:\s*ijkend|ijk)\s*,\s*(\d+)\)", r"RK[\1]", e, flags=re.I)
    e = re.sub(r"\bxlw\b", "vapp", e, flags=re.I)
    e = e.replace(" ", "")
    e = re.sub(r"\btemp\b", "T", e, flags=re.I)
    e = re.sub(r"\bSumM\b", "M", e, flags=re.I)
    e = re.sub(r"\bYlH2O\b", "W", e, flags=re.I)
    e = re.sub(r"\bdexp\b", "EXP", e, flags=re.I)
    e = re.sub(r"\bLOG10\b", "LGT", e)
    e = re.sub(r"\bLOG\b", "LN", e)
    nums = []
    out = ""
    i = 0
    tokre = re.compile(NUM)
    while i < len(e):
        c = e[i]
        prev = out[-1] if out else "="
        if c in "+-" and prev in "=(*/+-" and tokre.match(e, i + 1):
            m = tokre.match(e, i + 1)
            nums.append((-1.0 if c == "-" else 1.0) * fnum(m.group(0)))
            out += "#"
            i = m.end()
            continue
        if (c.isdigit() or c == ".") and not (out and (out[-1].isalpha() or out[-1] == "[" or out[-1].isdigit())):
            m = tokre.match(e, i)
            if m:
                nums.append(fnum(m.group(0)))
                out += "#"
                i = m.end()
                continue
        out += c
        i += 1
    return out, nums


def parse_kinetic(root, nr, nphot):
    stmts = read_statements(os.path.join(root, MECH_DIR, "chem_spack_kinetic.f90"))
    troe_shape = "(Effko*M/(#+Effko*M/Rapk))*#**(#/(#+(LGT(Effko*M/Rapk))**#))"
    troe_shape2 = "(Effko*M/(#+Effko*M/Rapk))*#**(#/(#+(LGT(Effko*M/Rapk))**#))"
    laws = {}
    tmp = {}
    for s in stmts:
        m = re.match(r"(Effko|Rapk|facteur)\((?:ijkbeg:ijkend|ijk)\)\s*=(.*)$", s, flags=re.I)
        if m:
            tmp[m.group(1).lower()] = shape_of(m.group(2))
            continue
        m = re.match(r"rk\((?:ijkbeg:ijkend|ijk),\s*(\d+)\)\s*=(.*)$", s, flags=re.I)
        if not m:
            continue
        r = int(m.group(1))
        sh, nums = shape_of(m.group(2))
        if r in laws:
            law = laws[r]
            if sh == "RK[%d]*M*#" % r:
                law["post"].append(["mul_M_const", nums[0]])
            elif sh == "RK[%d]*M" % r:
                law["post"].append(["mul_M"])
            elif sh == "RK[%d]*W" % r:
                law["post"].append(["mul_H2O"])
            elif sh == "facteur*RK[%d]" % r:
                # Troe fall-off factor times the Arrhenius already in rk
                law["post"].append(["mul_troe"] + troe_params(tmp))
            else:
                raise ValueError("post %d %s" % (r, sh))
            continue
        law = OrderedDict(family=None, params=[], post=[])
        if sh == "J[%d]" % r and r <= nphot:
            law["family"], law["params"] = "photolysis", [r]
        elif sh == "#":
            law["family"], law["params"] = "constant", nums
        elif sh == "EXP(#-(#)/T)":
            law["family"], law["params"] = "arrhenius", nums           # exp(a - b/T)
        elif sh in ("EXP(#+(#*LN(T))-#/T)", "EXP(#+(#*LN(T))+#/T)"):
            a, n, b = nums
            if sh.endswith("+#/T)"):
                b = -b                                                  # x + b/T == x - (-b)/T
            law["family"], law["params"] = "arrhenius_tn", [a, n, b]  # exp(a + n ln T - b/T)
        elif sh == "M*#*(T/#)**(#)":
            c, t0, n = nums
            assert t0 == 300.0
            law["family"], law["params"] = "m_power", [c, n]           # (M c)(T/300)^n
        elif sh == "#*EXP(#/T)+#*M*EXP(#/T)":
            law["family"], law["params"] = "exp_plus_m_exp", nums      # c1 e^(b1/T) + (c2 M) e^(b2/T)
        elif sh == "(Effko*M/(#+Effko*M/Rapk))*#**(#/(#+(LGT(Effko*M/Rapk))**#))":
            one1, fc, one2, one3, two = nums
            assert (one1, one2, one3, two) == (1.0, 1.0, 1.0, 2.0)
            law["family"], law["params"] = "troe", troe_params(tmp, fc)
        elif sh == "Effko+facteur/(#+facteur/Rapk)":
            e_sh, e_n = tmp["effko"]
            r_sh, r_n = tmp["rapk"]
            f_sh, f_n = tmp["facteur"]
            assert e_sh == "#*EXP(#/T)" and r_sh == "#*EXP(#/T)" and f_sh == "#*EXP(#/T)*M", (e_sh, r_sh, f_sh)
            law["family"], law["params"] = "hno3_oh", e_n + r_n + f_n
        elif sh.startswith("#*EXP("):
            law["family"], law["params"] = "exp_sum", parse_exp_sum(sh, nums)
        elif sh == "#*M*EXP(#/T)*(#+#*W*EXP(#/T))":
            c1, b1, one, c2, b2 = nums
            assert one == 1.0
            law["family"], law["params"] = "m_exp_h2o", [c1, b1, c2, b2]  # c1 M e^(b1/T) (1 + c2 H2O e^(b2/T))
        elif sh == "(Effko/(#+Effko/(Rapk/M)))*#**(#+(DLOG10(Effko/(Rapk/M))**#)**(#))":
            one1, fc, one2, two, mone = nums
            assert (one1, one2, two, mone) == (1.0, 1.0, 2.0, -1.0)
            (e_sh, e_n), (r_sh, r_n) = tmp["effko"], tmp["rapk"]
            assert e_sh == "#*(T/#)**(#)" and r_sh == "#*(T/#)**(#)", (e_sh, r_sh)
            assert e_n[1] == 300.0 and r_n[1] == 300.0
            # Effko = k0 (T/300)^n, Rapk = kinf (T/300)^m,
            # k = Effko/(1 + Effko/(Rapk/M)) Fc^(1 + (log10(Effko/(Rapk/M))^2)^-1)
            # (the reference's parentheses put the -1 power on the square only)
            law["family"], law["params"] = "troe_ca", [e_n[0], e_n[2], r_n[0], r_n[2], fc]
        elif sh == "#*W*W":
            law["family"], law["params"] = "h2o_sq", nums                # c H2O H2O
        elif sh == "#*(#+#*M)":
            c, one, d = nums
            assert one == 1.0
            law["family"], law["params"] = "linear_m", [c, d]          # c (1 + d M)
        else:
            raise ValueError("unknown shape r=%d %s %s" % (r, sh, nums))
        laws[r] = law
    out = []
    for r in range(1, nr + 1):
        out.append(laws.get(r, OrderedDict(family="unset", params=[], post=[])))
    return out


def troe_params(tmp, fc=None):
    e_sh, e_n = tmp["effko"]
    r_sh, r_n = tmp["rapk"]
    assert e_sh == "#*(T/#)**(-(#))" and r_sh == "#*(T/#)**(-(#))", (e_sh, r_sh)
    assert e_n[1] == 300.0 and r_n[1] == 300.0
    if fc is None:
        f_sh, f_n = tmp["facteur"]
        assert f_sh == "(Effko*M/(#+Effko*M/Rapk))*#**(#/(#+(LGT(Effko*M/Rapk))**#))", f_sh
        fc = f_n[1]
    # k0 (T/300)^-(n), kinf (T/300)^-(m), broadening Fc
    return [e_n[0], e_n[2], r_n[0], r_n[2], fc]


def parse_exp_sum(sh, nums):
    terms = []
    body = sh
    i = 0
    # split at top-level '+' between terms
    parts, depth, cur = [], 0, ""
    for ch in body:
        if ch == "(":
            depth += 1
        if ch == ")":
            depth -= 1
        if ch == "+" and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    parts.append(cur)
    const = 0.0
    for p in parts:
        k = p.count("#")
        vals = nums[i:i + k]
        i += k
        if p == "#*EXP(-(#)/T)":
            terms.append([vals[0], 0.0, vals[1]])          # c exp(-b/T)
        elif p == "#*EXP(#*LN(T)-#/T)":
            terms.append([vals[0], vals[1], vals[2]])      # c exp(n ln T - b/T)
        elif p == "#":
            const = vals[0]
        else:
            raise ValueError("exp_sum part %s" % p)
    assert i == len(nums)
    return [const] + [x for t in terms for x in t]


# ---------------------------------------------------------------- checks
def derive_jacobian(reactants, prod, nspec):
    """J(i,j) terms: for r ascending, S(i,r) * d w_r / d y_j per reactant slot."""
    S = {}
    for sp, terms in prod.items():
        for c, r in terms:
            S[(sp, r)] = c
    jac = {}
    for (sp, r), c in sorted(S.items(), key=lambda kv: (kv[0][1], kv[0][0])):
        pass
    for i in range(1, nspec + 1):
        for j in range(1, nspec + 1):
            terms = []
            for r, rs in enumerate(reactants, start=1):
                c = S.get((i, r))
                if c is None:
                    continue
                for slot in rs:
                    if slot == j:
                        terms.append((c, r, j))
            if terms:
                jac[(i, j)] = terms
    return jac


def parse_prod_loss(root):
    """fexprod (prod = + c w_r) and fexloss (loss = + c dw_rs), term lists."""
    out = []
    for fname, lhs, ref in (("chem_spack_fexprod.f90", "prod", r"w\(ijk,\s*(\d+)\)"),
                            ("chem_spack_fexloss.f90", "loss", r"dw\(ijk,\s*(\d+),\s*(\d+)\)")):
        rows = {}
        for s in read_statements(os.path.join(root, MECH_DIR, fname)):
            m = re.match(lhs + r"\(ijk,\s*(\d+)\)\s*=(.*)$", s)
            if m:
                sp = int(m.group(1))
                assert sp not in rows
                rows[sp] = parse_terms(m.group(2), ref)
        out.append(rows)
    return out


def main():
    global MECH_DIR
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    mech = sys.argv[2] if len(sys.argv) > 2 else "RELACS_TUV"
    MECH_DIR = "src/brams/ccatt/" + mech
    species, nr, nphot = parse_species(root)
    reactants = parse_reactants(root, nr)
    prod = parse_production(root)
    jac_ref = parse_jacobian(root)
    dw = parse_dratedc(root)
    laws = parse_kinetic(root, nr, nphot)

    # production terms are listed in increasing reaction order
    for sp, terms in prod.items():
        rs = [r for _, r in terms]
        assert rs == sorted(rs) and len(set(rs)) == len(rs), sp
    # partials: dw(r, s) = rk(r) * y(other reactant)
    for r, rs in enumerate(reactants, start=1):
        for k, s in enumerate(rs):
            others = rs[:k] + rs[k + 1:]
            assert dw[(r, s)] == (others[0] if others else 0), (r, s)
    assert len(dw) == sum(len(set(rs)) for rs in reactants)
    # derived Jacobian == reference Jacobian, term for term
    jac = derive_jacobian(reactants, prod, len(species))
    assert set(jac) == set(jac_ref), set(jac) ^ set(jac_ref)
    for key, terms in jac_ref.items():
        assert [(c, r, s) for c, r, s in terms] == jac[key], (key, terms, jac[key])
    # QSSA's split: fexprod = the positive net terms c w_r, fexloss = -c dw_rs
    # over the negative ones, both in reaction order
    p_ref, l_ref = parse_prod_loss(root)
    for sp in range(1, len(species) + 1):
        terms = prod.get(sp, [])
        assert [(c, r) for c, r in terms if c > 0] == [(c, r) for c, _, r in p_ref.get(sp, [])], sp
        assert [(-c, r) for c, r in terms if c < 0] == [(c, r) for c, _, r, s2 in l_ref.get(sp, [])], sp
        assert all(s2 == sp for _, _, _, s2 in l_ref.get(sp, []))

    data = OrderedDict()
    data["name"] = mech
    cm = [re.search(r"chemical_mechanism\s*=\s*'(\w+)'", x) for x in
          read_statements(os.path.join(root, MECH_DIR, "chem1_list.f90"))]
    data["chemical_mechanism"] = [c.group(1) for c in cm if c][0]
    data["note"] = ("Indices are 1-based as in the reference. The rate-law families are the device "
                    "functions rl_<family> of brams_amd/csrc/mech_families.h; extract_relacs_tuv.py "
                    "documents the kinetic shapes it recognises.")
    data["nspecies"] = len(species)
    data["nr"] = nr
    data["nr_photo"] = nphot
    data["species"] = species
    data["reactions"] = [OrderedDict(reactants=reactants[r], rate=laws[r]) for r in range(nr)]
    data["production"] = [[sp, [[c, r] for c, r in prod[sp]]] for sp in sorted(prod)]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), mech.lower() + ".json")
    with open(out, "w") as fh:
        json.dump(data, fh, indent=1)
    nj = len(jac)
    print("species=%d reactions=%d jacobian entries=%d production rows=%d -> %s"
          % (len(species), nr, nj, len(prod), out))


if __name__ == "__main__":
    main()

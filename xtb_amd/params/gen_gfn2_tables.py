#!/usr/bin/env python3
"""Extract the GFN2-xTB, DFT-D4 and EEQ parameter tables from the xtb sources.

This is a data-extraction recipe, not code translation: it reads the numeric
literals of the parameter arrays out of the Fortran sources of xtb 6.7.1 and
writes them, unchanged, as C arrays into ``xtb_amd/csrc/gfn2_params.h``.  All
arithmetic that xtb applies to these tables (unit conversion, shell mapping,
scaling) is done by the consumers of the header, in the same operation order
as xtb, so that the resulting doubles are bit-identical.

Sources read (paths relative to the reference root):
  src/xtb/gfn2.f90                      GFN2 element parameters (gfn2.f90:73-730)
  src/xtb/data.f90                      valenceCN, multiRad (data.f90:331-365)
  src/param/paulingen.f90               Pauling electronegativities
  src/param/atomicrad.f90               atomic radii (Angstrom, scaled by aatoau)
  src/param/covalentradd3.f90           D3 covalent radii (Angstrom)
  src/param/sqrtzr4r2.f90               <r4>/<r2> expectation values
  src/mctc/param/chemical_hardnesses.f90  chemical hardness (D4 zeta function)
  src/disp/dftd4_parameters.f90         D4 effective nuclear charges
  src/disp/ncoord.f90                   Pyykko radii / EN used by the EEQ CN
  src/charge_model.f90                  EEQ 2019 parameters
  include/param_ref.fh                  D4 reference polarisabilities

Usage:  python gen_gfn2_tables.py [/root/reference] [output.h]
"""
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = sys.argv[2] if len(sys.argv) > 2 else os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "..", "csrc", "gfn2_params.h")

MULT = {}
NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][-+]?\d+)?(?:_wp)?")


def _clean(text):
    lines = []
    for ln in text.splitlines():
        s = ln.split("!", 1)[0]
        lines.append(s)
    return "\n".join(lines)


def _numbers(body):
    out = []
    for tok in NUM.findall(body):
        t = tok.replace("_wp", "").replace("d", "e").replace("D", "e")
        out.append(float(t) if ("." in t or "e" in t) else int(t))
    return out


def array(path, name, after=None):
    """Numeric literals of the array constructor assigned to ``name``
    (first declaration following the text ``after`` if given)."""
    src = _clean(open(os.path.join(REF, path)).read())
    if after is not None:
        src = src[src.index(after):]
    m = re.search(r"::\s*" + re.escape(name) + r"\s*(\([^)]*\))?\s*=", src)
    if m is None:
        raise KeyError(f"{name} not found in {path}")
    i = m.end()
    # find the constructor start: '[' or '(/'
    j1 = src.find("[", i)
    j2 = src.find("(/", i)
    cands = [(j, k) for j, k in ((j1, "["), (j2, "(/")) if j >= 0]
    j, kind = min(cands)
    end = src.find("]" if kind == "[" else "/)", j)
    body = src[j + len(kind):end]
    body = body.replace("&", " ")
    # a scalar factor applied to the whole constructor, e.g. "[...] * 0.1_wp"
    tail = re.match(r"\s*\*\s*([0-9.eEdD+-]+)_wp", src[end + (1 if kind == "[" else 2):])
    MULT[name] = float(tail.group(1)) if tail else None
    # shape(...) trailing reshape arguments are outside the constructor
    return _numbers(body)


def data_statements(path):
    """Parse ``data name(idx) / values /`` statements of a Fortran include."""
    src = _clean(open(os.path.join(REF, path)).read()).replace("&", " ")
    src = src.replace(";", "\n")
    stmt = re.compile(r"data\s+(\w+)\s*\(([^)]*)\)\s*/([^/]*)/", re.S)
    out = {}
    for m in stmt.finditer(src):
        name, idx, vals = m.group(1), m.group(2), m.group(3)
        idx = [s.strip() for s in idx.split(",")]
        out.setdefault(name, []).append((idx, _numbers(vals)))
    return out


def cdbl(x):
    return float(x).hex()


def emit_1d(f, ctype, name, vals, per_line=4):
    f.write(f"static const {ctype} {name}[{len(vals)}] = {{\n")
    for k in range(0, len(vals), per_line):
        chunk = vals[k:k + per_line]
        if ctype == "double":
            f.write("  " + ", ".join(cdbl(v) for v in chunk) + ",\n")
        else:
            f.write("  " + ", ".join(str(int(v)) for v in chunk) + ",\n")
    f.write("};\n\n")


def main():
    g2 = "src/xtb/gfn2.f90"
    tabs = {}
    tabs["rep_alpha"] = array(g2, "repAlpha")
    tabs["rep_zeff"] = array(g2, "repZeff")
    tabs["chemical_hardness_gfn2"] = array(g2, "chemicalHardness")
    tabs["third_order_atom"] = array(g2, "thirdOrderAtom")        # * K_THIRD_ORDER_ATOM_MULT
    tabs["shell_hardness_l"] = array(g2, "shellHardness")      # (3, 86): l = 0..2
    tabs["dip_kernel_raw"] = array(g2, "dipKernel")             # * K_..._KERNEL_MULT (0.01)
    tabs["quad_kernel_raw"] = array(g2, "quadKernel")           # * K_..._KERNEL_MULT (0.01)
    tabs["n_shell"] = array(g2, "nShell")
    tabs["ang_shell"] = array(g2, "angShell")                   # (3, 86)
    tabs["pqn_shell"] = array(g2, "principalQuantumNumber")     # (3, 86)
    tabs["reference_occ_l"] = array(g2, "referenceOcc")         # (0:2, 86)
    tabs["shell_poly_l"] = array(g2, "shellPoly")               # (4, 86): l = 0..3
    tabs["kcn_l"] = array(g2, "kCN")                            # (4, 86): l = 0..3
    tabs["self_energy"] = array(g2, "selfEnergy")               # (3, 86): per shell
    tabs["slater_exponent"] = array(g2, "slaterExponent")       # (3, 86): per shell
    tabs["gfn2_kinds"] = array(g2, "gfn2Kinds")                 # (118)
    tabs["valence_cn"] = array("src/xtb/data.f90", "valenceCN")
    tabs["multi_rad"] = array("src/xtb/data.f90", "multiRad")
    tabs["pauling_en"] = array("src/param/paulingen.f90", "paulingEN")
    tabs["atomic_rad_aa"] = array("src/param/atomicrad.f90", "atomicRad")
    tabs["covalent_rad_d3_aa"] = array("src/param/covalentradd3.f90", "covalentRadD3")
    tabs["r4_over_r2"] = array("src/param/sqrtzr4r2.f90", "r4Overr2")
    tabs["chemical_hardness_d4"] = array("src/mctc/param/chemical_hardnesses.f90",
                                         "chemical_hardness")
    tabs["d4_zeff"] = array("src/disp/dftd4_parameters.f90", "zeff")
    tabs["ncoord_rad_aa"] = array("src/disp/ncoord.f90", "rad")
    tabs["ncoord_en"] = array("src/disp/ncoord.f90", "en")
    cm = "src/charge_model.f90"
    m19 = "subroutine new_charge_model_2019"     # src/charge_model.f90:147-245
    tabs["eeq_en"] = array(cm, "en", m19)
    tabs["eeq_gam"] = array(cm, "gam", m19)
    tabs["eeq_kappa"] = array(cm, "kappa", m19)
    tabs["eeq_alpha"] = array(cm, "alpha", m19)

    # Stewart STO-nG expansion tables (J. Chem. Phys. 52, 431 (1970)),
    # layout (ng, 15) column-major: t[ityp*ng + k], ityp = 0..14 (1s..5g)
    sl = "src/slater.f90"
    tabs["sto_alpha1"] = array(sl, "pAlpha1")
    for ng in range(2, 7):
        tabs[f"sto_alpha{ng}"] = array(sl, f"pAlpha{ng}")
        tabs[f"sto_coeff{ng}"] = array(sl, f"pCoeff{ng}")
    tabs["sto_alpha6s"] = array(sl, "pAlpha6s")
    tabs["sto_coeff6s"] = array(sl, "pCoeff6s")
    tabs["sto_alpha6p"] = array(sl, "pAlpha6p")
    tabs["sto_coeff6p"] = array(sl, "pCoeff6p")
    for ng in range(1, 7):
        n = len(tabs[f"sto_alpha{ng}"])
        if n != 15 * ng:
            raise SystemExit(f"sto_alpha{ng}: parsed {n} values")

    expect = {
        "rep_alpha": 86, "rep_zeff": 86, "chemical_hardness_gfn2": 86,
        "third_order_atom": 86, "shell_hardness_l": 258, "dip_kernel_raw": 86,
        "quad_kernel_raw": 86, "n_shell": 86, "ang_shell": 258, "pqn_shell": 258,
        "reference_occ_l": 258, "shell_poly_l": 344, "kcn_l": 344,
        "self_energy": 258, "slater_exponent": 258, "gfn2_kinds": 118,
        "valence_cn": 118, "multi_rad": 118, "pauling_en": 118,
        "atomic_rad_aa": 118, "covalent_rad_d3_aa": 118, "r4_over_r2": 118,
        "chemical_hardness_d4": 118, "d4_zeff": 118, "ncoord_rad_aa": 118,
        "ncoord_en": 118, "eeq_en": 86, "eeq_gam": 86, "eeq_kappa": 86,
        "eeq_alpha": 86,
    }
    for k, n in expect.items():
        if len(tabs[k]) != n:
            raise SystemExit(f"table {k}: expected {n} values, parsed {len(tabs[k])}")

    # ---- D4 reference systems (param_ref.fh) -------------------------------
    d = data_statements("include/param_ref.fh")
    MAXREF, NFREQ, NSEC = 7, 23, 17
    refn = [0] * 118
    refq = [[0.0] * MAXREF for _ in range(118)]
    refh = [[0.0] * MAXREF for _ in range(118)]
    hcount = [[0.0] * MAXREF for _ in range(118)]
    ascale = [[0.0] * MAXREF for _ in range(118)]
    refcovcn = [[0.0] * MAXREF for _ in range(118)]
    refcn = [[0.0] * MAXREF for _ in range(118)]
    refsys = [[0] * MAXREF for _ in range(118)]
    alphaiw = [[[0.0] * NFREQ for _ in range(MAXREF)] for _ in range(118)]
    sscale = [0.0] * NSEC
    secaiw = [[0.0] * NFREQ for _ in range(NSEC)]
    for idx, v in d["refn"]:
        refn[int(idx[0]) - 1] = v[0]
    for name, tab in (("refq", refq), ("refh", refh), ("hcount", hcount),
                      ("ascale", ascale), ("refcovcn", refcovcn), ("refcn", refcn),
                      ("refsys", refsys)):
        for idx, v in d[name]:
            tab[int(idx[1]) - 1][int(idx[0]) - 1] = v[0]
    for idx, v in d["alphaiw"]:
        assert idx[0] == ":" and len(v) == NFREQ
        alphaiw[int(idx[2]) - 1][int(idx[1]) - 1] = v
    for idx, v in d["sscale"]:
        sscale[int(idx[0]) - 1] = v[0]
    for idx, v in d["secaiw"]:
        assert idx[0] == ":" and len(v) == NFREQ
        secaiw[int(idx[1]) - 1] = v

    with open(OUT, "w") as f:
        f.write("/* Generated by xtb_amd/params/gen_gfn2_tables.py from the xtb 6.7.1\n"
                " * parameter sources (LGPL-3.0 data).  Do not edit by hand.\n"
                " * Element-indexed tables are 0-based (index Z-1).  Tables named *_l\n"
                " * are indexed by angular momentum, the others by shell position.\n"
                " * 2-D tables keep the Fortran column-major order: t[col*ld + row]. */\n")
        f.write("#pragma once\n\n")
        f.write("#define GFN2_MAX_ELEM 86\n#define D4_MAX_ELEM 118\n")
        f.write(f"#define D4_MAX_REF {MAXREF}\n#define D4_NFREQ {NFREQ}\n#define D4_NSEC {NSEC}\n\n")
        for src_name, c_name in (("thirdOrderAtom", "THIRD_ORDER_ATOM"), ("dipKernel", "DIP_KERNEL"),
                                 ("quadKernel", "QUAD_KERNEL")):
            f.write(f"/* {src_name} is multiplied by this factor in xtb (src/xtb/gfn2.f90) */\n")
            f.write(f"#define K_{c_name}_MULT {cdbl(MULT[src_name])}\n")
        f.write("\n")
        ints = {"n_shell", "ang_shell", "pqn_shell", "gfn2_kinds"}
        for k, v in tabs.items():
            emit_1d(f, "int" if k in ints else "double", "k_" + k, v)
        emit_1d(f, "int", "k_d4_refn", refn)
        for name, tab in (("refq", refq), ("refh", refh), ("hcount", hcount),
                          ("ascale", ascale), ("refcovcn", refcovcn), ("refcn", refcn)):
            emit_1d(f, "double", "k_d4_" + name, [x for row in tab for x in row])
        emit_1d(f, "int", "k_d4_refsys", [x for row in refsys for x in row])
        emit_1d(f, "double", "k_d4_alphaiw",
                [x for el in alphaiw for r in el for x in r])
        emit_1d(f, "double", "k_d4_sscale", sscale)
        emit_1d(f, "double", "k_d4_secaiw", [x for r in secaiw for x in r])
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Generate the D3 physics tables used by the library and by the oracle.

The element data of DFT-D3 (reference C6 coefficients, reference coordination
numbers, <r4>/<r2> expectation values, pairwise van-der-Waals radii and the
damping-parameter database) is physics data, not code.  This script extracts
the numbers from the Fortran sources of simple-dftd3 v1.6.0 and writes them
into two data files that are committed next to it:

  d3_tables.npz   numeric element tables (numpy arrays, no pickles)
  d3_params.json  method-name -> damping parameters of every family

It only runs where the reference checkout exists (``--reference`` option,
default ``/root/reference``); everything else loads the committed files.

Sources read (reference file:line):
  src/dftd3/reference.f90:30-45     number_of_references(103)
  src/dftd3/reference.f90:47-151    reference_cn(7, 103)
  src/dftd3/reference.f90:207-66221 copy_c6: reference_c6(7, 7, 5356)
  src/dftd3/data/r4r2.f90:42-66     r4_over_r2(118)
  src/dftd3/data/vdwrad.f90:38-1111 vdwrad(5356) in Angstrom
  src/dftd3/param.f90:96-300        get_method_id (aliases)
  src/dftd3/param.f90:302-800       get_rational_damping
  src/dftd3/param.f90:805-1090      get_zero_damping
  src/dftd3/param.f90:1091-1326     get_mrational_damping, get_mzero_damping,
                                    get_optimizedpower_damping, get_cso_damping,
                                    get_z_damping (no entries)

The covalent radii come from mctc-lib (grimme-lab/mctc-lib, rev 06db4d8,
``mctc_data_covrad``), which is not part of the reference checkout.  They are
the Pyykkoe-Atsumi single-bond radii (Chem. Eur. J. 15, 2009, 186-197) with
metals reduced by 10 %, in Angstrom, as published; the D3 model scales them
by 4/3 and converts them to bohr at run time.
"""
import argparse
import hashlib
import json
import os
import re

import numpy as np

MAX_ELEM = 103
MAX_REF = 7
NPAIR = MAX_ELEM * (MAX_ELEM + 1) // 2

# Pyykkoe & Atsumi (2009) covalent radii in Angstrom, metals reduced by 10 %,
# elements 1-118 (the table mctc-lib names covalent_rad_2009).
COVALENT_RAD_2009_AA = [
    0.32, 0.46,  # H, He
    1.20, 0.94, 0.77, 0.75, 0.71, 0.63, 0.64, 0.67,  # Li-Ne
    1.40, 1.25, 1.13, 1.04, 1.10, 1.02, 0.99, 0.96,  # Na-Ar
    1.76, 1.54,  # K, Ca
    1.33, 1.22, 1.21, 1.10, 1.07, 1.04, 1.00, 0.99, 1.01, 1.09,  # Sc-Zn
    1.12, 1.09, 1.15, 1.10, 1.14, 1.17,  # Ga-Kr
    1.89, 1.67,  # Rb, Sr
    1.47, 1.39, 1.32, 1.24, 1.15, 1.13, 1.13, 1.08, 1.15, 1.23,  # Y-Cd
    1.28, 1.26, 1.26, 1.23, 1.32, 1.31,  # In-Xe
    2.09, 1.76,  # Cs, Ba
    1.62, 1.47, 1.58, 1.57, 1.56, 1.55, 1.51,  # La-Eu
    1.52, 1.51, 1.50, 1.49, 1.49, 1.48, 1.53,  # Gd-Yb
    1.46, 1.37, 1.31, 1.23, 1.18, 1.16, 1.11, 1.12, 1.13, 1.32,  # Lu-Hg
    1.30, 1.30, 1.36, 1.31, 1.38, 1.42,  # Tl-Rn
    2.01, 1.81,  # Fr, Ra
    1.67, 1.58, 1.52, 1.53, 1.54, 1.55, 1.49,  # Ac-Am
    1.49, 1.51, 1.51, 1.48, 1.50, 1.56, 1.58,  # Cm-No
    1.45, 1.41, 1.34, 1.29, 1.27, 1.21, 1.16, 1.15, 1.09, 1.22,  # Lr-Cn
    1.36, 1.43, 1.46, 1.58, 1.48, 1.57,  # Nh-Og
]
assert len(COVALENT_RAD_2009_AA) == 118


def _fortran_join(text):
    """Join Fortran free-form continuation lines ('&' at end / start)."""
    text = re.sub(r"!.*", "", text)  # comments (no '!' inside our literals)
    text = re.sub(r"&[ \t]*\n[ \t]*&?", " ", text)
    return text


_NUM = r"[+-]?\d+\.?\d*(?:[eEdD][+-]?\d+)?(?:_wp)?"


def _floats(chunk):
    vals = re.findall(_NUM, chunk)
    return [float(v.replace("_wp", "").replace("d", "e").replace("D", "e")) for v in vals]


def parse_reference(src):
    text = open(os.path.join(src, "reference.f90")).read()
    joined = _fortran_join(text)
    m = re.search(r"number_of_references\(max_elem\)\s*=\s*\[(.*?)\]", joined, re.S)
    nref = [int(v) for v in re.findall(r"\d+", m.group(1))]
    assert len(nref) == MAX_ELEM, len(nref)
    m = re.search(r"reference_cn\(max_ref, max_elem\)\s*=\s*reshape\(\[(.*?)\]\s*,", joined, re.S)
    refcn = _floats(m.group(1))
    assert len(refcn) == MAX_REF * MAX_ELEM, len(refcn)
    refcn = np.array(refcn).reshape(MAX_ELEM, MAX_REF)  # [Z-1][ref]
    c6 = np.zeros(MAX_REF * MAX_REF * NPAIR)
    filled = np.zeros(c6.size, dtype=bool)
    for m in re.finditer(r"c6ab_view\((\d+):(\d+)\)\s*=\s*\[(.*?)\]", joined, re.S):
        lo, hi = int(m.group(1)), int(m.group(2))
        vals = _floats(m.group(3))
        assert len(vals) == hi - lo + 1, (lo, hi, len(vals))
        c6[lo - 1:hi] = vals
        filled[lo - 1:hi] = True
    assert filled.all()
    # Fortran reference_c6(iref, jref, ic) is column-major: iref fastest.
    # Stored here as [ic][jref][iref] so that c6[ic, jref, iref] is the entry.
    c6 = c6.reshape(NPAIR, MAX_REF, MAX_REF)
    return np.array(nref, dtype=np.int32), refcn, c6


def parse_r4r2(src):
    text = _fortran_join(open(os.path.join(src, "data", "r4r2.f90")).read())
    m = re.search(r"r4_over_r2\(max_elem\)\s*=\s*\[(.*?)\]", text, re.S)
    vals = _floats(m.group(1))
    assert len(vals) == 118, len(vals)
    return np.array(vals)


def parse_vdwrad(src):
    text = _fortran_join(open(os.path.join(src, "data", "vdwrad.f90")).read())
    m = re.search(r"vdwrad\(max_elem\*\(1\+max_elem\)/2\)\s*=\s*aatoau\s*\*\s*\[(.*?)\]", text, re.S)
    vals = _floats(m.group(1))
    assert len(vals) == NPAIR, len(vals)
    return np.array(vals)


def parse_params(src):
    text = _fortran_join(open(os.path.join(src, "param.f90")).read())
    lines = text.split("\n")

    def block(name):
        start = None
        for k, line in enumerate(lines):
            if re.match(r"\s*(function|subroutine)\s+" + name + r"\b", line):
                start = k
            if start is not None and k > start and re.match(r"\s*end\s+(function|subroutine)\s+" + name, line):
                return lines[start:k]
        raise KeyError(name)

    aliases = {}
    for line in block("get_method_id"):
        m = re.match(r"\s*case\((.*?)\);\s*id\s*=\s*(p_\w+)", line)
        if m:
            for name in re.findall(r'"([^"]*)"', m.group(1)):
                aliases[name] = m.group(2)

    def damping(name):
        table = {}
        current = None
        for line in block(name):
            m = re.match(r"\s*case\((p_[\w, ]+)\)", line)
            if m:
                current = [c.strip() for c in m.group(1).split(",")]
                continue
            m = re.match(r"\s*param\s*=\s*d3_param\((.*)\)", line)
            if m and current:
                kw = {}
                for key, val in re.findall(r"(\w+)\s*=\s*(" + _NUM + ")", m.group(1)):
                    kw[key] = float(val.replace("_wp", ""))
                for c in current:
                    table[c] = kw
        return table

    families = {"rational": "get_rational_damping", "zero": "get_zero_damping",
                "mrational": "get_mrational_damping", "mzero": "get_mzero_damping",
                "optimizedpower": "get_optimizedpower_damping", "cso": "get_cso_damping",
                "z": "get_z_damping"}
    return aliases, {fam: damping(fn) for fam, fn in families.items()}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    args = ap.parse_args()
    src = os.path.join(args.reference, "src", "dftd3")
    nref, refcn, c6 = parse_reference(src)
    r4r2 = parse_r4r2(src)
    vdwrad = parse_vdwrad(src)
    rcov = np.array(COVALENT_RAD_2009_AA)
    out = os.path.join(args.out, "d3_tables.npz")
    np.savez_compressed(out, nref=nref, refcn=refcn, c6ref=c6, r4_over_r2=r4r2,
                        vdwrad_aa=vdwrad, covrad_2009_aa=rcov)
    aliases, tables = parse_params(src)
    pout = os.path.join(args.out, "d3_params.json")
    with open(pout, "w") as fh:
        json.dump(dict({"source": "simple-dftd3 v1.6.0 src/dftd3/param.f90",
                        "defaults": {"s6": 1.0, "s8": 1.0, "s9": 0.0, "rs6": 1.0, "rs8": 1.0,
                                     "a1": 0.4, "a2": 5.0, "alp": 14.0, "bet": 0.0},
                        "aliases": aliases}, **tables),
                  fh, indent=1, sort_keys=True)
    for path in (out, pout):
        digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
        print(f"{os.path.basename(path)}: sha256 {digest}")
    print(f"aliases {len(aliases)}, " + ", ".join(f"{k} {len(v)}" for k, v in tables.items()))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Extract the reference's golden vectors into committed JSON fixtures.

Runs only where the reference checkout exists (default /root/reference);
the JSON files it writes (tests/golden/*.json) are the fixtures the tests
load.  Every fixture holds data only: atomic numbers, positions in bohr,
lattice and periodicity, damping parameters, cutoffs, the expected values,
the tolerance the reference test uses, and the file:line it came from.

Sources:
  test/unit/test_regression.f90:44-196     Ta90 3D cell (dftbplus#871), BJ s9=0
  test/unit/test_periodic_2d.f90:78-211    BP86-D3(0) 2D and gh185 2D cells
  test/validation/01,02,06,07 (*.resp + geometry + *.json)  CLI goldens
  python/dftd3/test_interface.py:78-100,236-356  16-atom model energies
  test/validation/11-13 (*.resp + geometry + *.json)  pair energy matrices
  python/dftd3/test_interface.py:569-762   S8 pair energy matrices
  test/validation/03-05,08-10,14-20         modified rational / zero, optimized
                                            power damping and parameter files
"""
import argparse
import json
import os
import re

# Angstrom -> bohr of mctc-lib: Bohr radius hbar/(m_e c alpha), CODATA 2018
# constants; xyz and gen inputs are in Angstrom.
_HBAR = 6.62607015e-34 / (2.0 * 3.141592653589793)
AATOAU = 1.0 / (_HBAR / (9.1093837015e-31 * 299792458.0 * 7.2973525693e-3) * 1.0e10)

SYMBOLS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn "
    "Ga Ge As Se Br Kr Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce "
    "Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn "
    "Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr"
).split()


def to_number(sym):
    sym = sym.strip()
    sym = sym[0].upper() + sym[1:].lower()
    return SYMBOLS.index(sym) + 1


def _join(text):
    text = re.sub(r"!.*", "", text)
    return re.sub(r"&[ \t]*\n[ \t]*&?", " ", text)


def _floats(chunk):
    return [float(v.replace("_wp", "")) for v in re.findall(r"[+-]?\d+\.\d*(?:[eE][+-]?\d+)?(?:_wp)?", chunk)]


def _subroutine(text, name):
    m = re.search(r"subroutine\s+" + name + r"\b(.*?)end\s+subroutine\s+" + name, text, re.S)
    return m.group(1)


def _array(body, name):
    m = re.search(name + r"\([^)]*\)\s*=\s*reshape\(\[(.*?)\]\s*,", body, re.S)
    return _floats(m.group(1))


def _triples(vals):
    return [vals[i:i + 3] for i in range(0, len(vals), 3)]


def unit_fixtures(ref):
    out = []
    text = _join(open(os.path.join(ref, "test/unit/test_regression.f90")).read())
    body = _subroutine(text, "structure_dftbplus_871")
    syms = re.findall(r'"(\w+)"', re.search(r"sym\(nat\)\s*=\s*\[(.*?)\]", body, re.S).group(1))
    out.append(dict(
        name="ta90_3d_bj", source="test/unit/test_regression.f90:44-73,77-196",
        numbers=[to_number(s) for s in syms], positions=_triples(_array(body, "xyz")),
        lattice=_triples(_array(body, "lattice")), periodic=[True, True, True], wrap=False,
        kind="rational", param=dict(s6=1.0, s8=2.34, a1=6.30, a2=5.00, s9=0.0, alp=14.0),
        cutoff=[40.0, 60.0, 40.0, 0.0, 0.0], grad=True,
        expected=dict(energy=-8.9964956516666404e-5, gradient_norm=8.9844177273972976e-9),
        atol=100 * 2.220446049250313e-16))

    text = _join(open(os.path.join(ref, "test/unit/test_periodic_2d.f90")).read())
    cut2d = [30.0, 60.0, 15.0, 0.0, 0.0]  # test_periodic_2d.f90:34-35
    body = _subroutine(text, "test_bp86d3zero_2d")
    xyz = _triples(_array(body, "xyz"))
    lat = _triples(_array(body, "lattice"))
    latv = _triples(_array(body, "lattice_vac"))
    par = dict(s6=1.0, s8=1.683, rs6=1.139, rs8=1.0, s9=0.0, alp=14.0)
    for tag, lattice, pbc in (("2d", lat, [True, True, False]), ("vac3d", latv, [True, True, True])):
        out.append(dict(
            name=f"bp86_zero_c4_{tag}", source="test/unit/test_periodic_2d.f90:78-121",
            numbers=[6] * 4, positions=xyz, lattice=lattice, periodic=pbc, wrap=False,
            kind="zero", param=par, cutoff=cut2d, grad=False,
            expected=dict(energy=-1.5983949078995030e-2), atol=100 * 2.220446049250313e-16))
    body = _subroutine(text, "test_gh185")
    syms = re.findall(r'"(\w+)"', re.search(r"sym\(nat\)\s*=\s*\[(.*?)\]", body, re.S).group(1))
    xyz = _triples(_array(body, "xyz"))
    lat = _triples(_array(body, "lattice"))
    latv = _triples(_array(body, "lattice_vac"))
    par = dict(s6=1.0, s8=0.7875, a1=0.4289, a2=4.4407, s9=0.0, alp=14.0)
    for tag, lattice, pbc in (("2d", lat, [True, True, False]), ("vac3d", latv, [True, True, True])):
        out.append(dict(
            name=f"gh185_bj_{tag}", source="test/unit/test_periodic_2d.f90:124-211",
            numbers=[to_number(s) for s in syms], positions=xyz, lattice=lattice, periodic=pbc,
            wrap=False, kind="rational", param=par, cutoff=cut2d, grad=False,
            expected=dict(energy=-0.34811410398591064), atol=100 * 2.220446049250313e-16))
    return out


def read_xyz(path):
    lines = open(path).read().splitlines()
    n = int(lines[0].split()[0])
    nums, pos = [], []
    for line in lines[2:2 + n]:
        f = line.split()
        nums.append(to_number(f[0]))
        pos.append([float(x) * AATOAU for x in f[1:4]])
    return nums, pos, None, None


def read_gen(path):
    lines = [l for l in open(path).read().splitlines() if l.strip() and not l.strip().startswith("#")]
    n, kind = lines[0].split()[:2]
    n = int(n)
    species = lines[1].split()
    nums, pos = [], []
    for line in lines[2:2 + n]:
        f = line.split()
        nums.append(to_number(species[int(f[1]) - 1]))
        pos.append([float(x) for x in f[2:5]])
    lattice = None
    if kind.upper() in ("S", "F"):
        lattice = [[float(x) * AATOAU for x in lines[3 + n + k].split()[:3]] for k in range(3)]
    if kind.upper() == "F":
        pos = [[sum(p[k] * lattice[k][c] for k in range(3)) for c in range(3)] for p in pos]
    else:
        pos = [[x * AATOAU for x in p] for p in pos]
    return nums, pos, lattice, ([True] * 3 if lattice else None)


def read_tmol(path):
    lines = open(path).read().splitlines()
    nums, pos = [], []
    k = lines.index("$coord") + 1
    while not lines[k].startswith("$"):
        f = lines[k].split()
        pos.append([float(x) for x in f[:3]])
        nums.append(to_number(f[3]))
        k += 1
    return nums, pos, None, None


def read_molfile(path):
    """MDL molfile / SD file (V2000): counts line 4, atom block x y z symbol."""
    lines = open(path).read().splitlines()
    n = int(lines[3][0:3])
    nums, pos = [], []
    for line in lines[4:4 + n]:
        pos.append([float(line[0:10]) * AATOAU, float(line[10:20]) * AATOAU, float(line[20:30]) * AATOAU])
        nums.append(to_number(line[31:34]))
    return nums, pos, None, None


def read_pdb(path):
    """PDB ATOM/HETATM records: coordinates in columns 31-54, element 77-78."""
    nums, pos = [], []
    for line in open(path).read().splitlines():
        if line.startswith(("ATOM", "HETATM")):
            pos.append([float(line[30:38]) * AATOAU, float(line[38:46]) * AATOAU, float(line[46:54]) * AATOAU])
            nums.append(to_number(line[76:78]))
    return nums, pos, None, None


def family_fixtures(ref, params):
    """test/validation cases of the modified rational (bjm), modified zero
    (zerom) and optimized power (op) damping: energies (03-05), energies,
    gradients and virials (08-10), pair energy matrices (14, 15) and the
    parameter-database cases (16-20).  Parameters: loaded by method name as
    the CLI does (param.f90:1091-1244, s9 = 0 without --atm), given on the
    command line (app/cli.f90: --zerom-param s6 s8 rs6 bet, --op-param s6 s8 a1
    a2 bet, --atm-scale s9), or, for 16-20, the damping parameters the JSON
    records (assets/parameters.toml; 17 lists none: pkzb zero damping,
    parameters.toml:418-419 with the file's defaults s9 = 1, rs8 = 1)."""
    vdir = os.path.join(ref, "test/validation")

    def loaded(family, method):
        p = dict(params["defaults"])
        p.update(params[family][params["aliases"][method]])
        return p

    cases = [
        ("03-energy-d3-bjm", "03-met-xag.coord", read_tmol, "rational", loaded("mrational", "b97d")),
        ("04-energy-d3-zerom", "04-val-xad.mol", read_molfile, "mzero", loaded("mzero", "bp")),
        ("05-energy-d3-op", "05-cys-xag.sdf", read_molfile, "optimizedpower", loaded("optimizedpower", "tpss")),
        ("08-gradient-d3-bjm", "08-teh2-hbr.mol", read_molfile, "rational", loaded("mrational", "pbe0")),
        ("09-gradient-d3-zerom", "09-0a.pdb", read_pdb, "mzero", loaded("mzero", "lcwpbe")),
        ("10-gradient-d3-op", "10-but14diol.log", read_xyz, "optimizedpower", loaded("optimizedpower", "revpbe0")),
        ("14-pairwise-d3-zerom", "14-anthracene2.xyz", read_xyz, "mzero",
         dict(params["defaults"], s6=1.0, s8=1.841686, rs6=1.279637, bet=0.014370, s9=0.5)),
        ("15-pairwise-d3-op", "15-pxylene.mol", read_molfile, "optimizedpower",
         dict(params["defaults"], s6=1.0, s8=0.27632, a1=0.700, a2=2.500, bet=8.0)),
        ("16-db-d3-bj", "16-i10e.gen", read_gen, "rational", None),
        ("17-db-d3-zero", "17-li2-ch4.sdf", read_molfile, "zero",
         dict(params["defaults"], rs6=0.6327, s8=0.0, s9=1.0, rs8=1.0)),
        ("18-db-d3-bjm", "18-c1.pdb", read_pdb, "rational", None),
        ("19-db-d3-zerom", "19-g1.mol", read_molfile, "mzero", None),
        ("20-db-d3-op", "20-i21p.tmol", read_tmol, "optimizedpower", None),
    ]
    keys = dict(rational=("s6", "s8", "s9", "a1", "a2", "alp"), zero=("s6", "s8", "s9", "rs6", "rs8", "alp"),
                mzero=("s6", "s8", "s9", "rs6", "rs8", "alp", "bet"),
                optimizedpower=("s6", "s8", "s9", "a1", "a2", "alp", "bet"))
    out = []
    for name, geo, reader, kind, par in cases:
        nums, pos, lat, pbc = reader(os.path.join(vdir, geo))
        ref_json = json.load(open(os.path.join(vdir, name + ".json")))
        if par is None:
            par = dict(params["defaults"])
            par.update({k: v for k, v in ref_json["damping parameters"].items() if k != "damping"})
        par = {k: par[k] for k in keys[kind]}
        exp = dict(energy=ref_json["energy"])
        if "gradient" in ref_json:
            exp["gradient"] = _triples(ref_json["gradient"])
            exp["virial"] = _triples(ref_json["virial"])
        n = len(nums)
        if "additive pairwise energy" in ref_json:
            e2, e3 = ref_json["additive pairwise energy"], ref_json["non-additive pairwise energy"]
            exp["pair2"] = [[e2[c * n + r] for c in range(n)] for r in range(n)]
            exp["pair3"] = [[e3[c * n + r] for c in range(n)] for r in range(n)]
        out.append(dict(name="validation_" + name, source=f"test/validation/{name}.json,{name}.resp,{geo}",
                        numbers=nums, positions=pos, lattice=lat, periodic=pbc, wrap=True, kind=kind, param=par,
                        cutoff=[40.0, 60.0, 40.0, 0.0, 0.0], grad="gradient" in ref_json, expected=exp,
                        atol=1.0e-6, version=ref_json["version"]))
    return out


def validation_fixtures(ref, params):
    vdir = os.path.join(ref, "test/validation")
    cases = [
        ("01-energy-d3-bj", "01-tyr-xab.xyz", read_xyz, "rational", "pbe", True),
        ("02-energy-d3-zero", "02-ser-xad.gen", read_gen, "zero", "b3lyp", False),
        ("06-gradient-d3-bj", "06-229.xyz", read_xyz, "rational", "olyp", False),
        ("07-gradient-d3-zero", "07-bih3-nh3.tmol", read_tmol, "zero", "m06l", True),
    ]
    out = []
    for name, geo, reader, kind, method, atm in cases:
        nums, pos, lat, pbc = reader(os.path.join(vdir, geo))
        ref_json = json.load(open(os.path.join(vdir, name + ".json")))
        p = dict(params["defaults"])
        p.update(params[kind][params["aliases"][method]])
        p["s9"] = 1.0 if atm else 0.0
        if kind == "rational":
            p = {k: p[k] for k in ("s6", "s8", "s9", "a1", "a2", "alp")}
        else:
            p = {k: p[k] for k in ("s6", "s8", "s9", "rs6", "rs8", "alp")}
        exp = dict(energy=ref_json["energy"])
        if "gradient" in ref_json:
            exp["gradient"] = _triples(ref_json["gradient"])
            exp["virial"] = _triples(ref_json["virial"])
        out.append(dict(
            name="validation_" + name, source=f"test/validation/{name}.json,{name}.resp,{geo}",
            numbers=nums, positions=pos, lattice=lat, periodic=pbc, wrap=True, kind=kind,
            method=method, atm=atm, param=p, cutoff=[40.0, 60.0, 40.0, 0.0, 0.0],
            grad="gradient" in ref_json, expected=exp, atol=1.0e-6))
    return out


def pairwise_fixtures(ref):
    """test/validation/11-13 (s-dftd3 --pair-resolved): the pair energy
    matrices and the total energy.  Parameters as the CLI builds them from the
    .resp arguments (app/cli.f90:400-427,478-492,530-547): --bj-param /
    --bjm-param / --zero-param give s6 s8 a1|rs6 [a2]; s9 is 0 (param.f90:39)
    unless --atm (1) or --atm-scale sets it; rs8 = 1, alp = 14.  --bjm-param
    builds the rational damping (app/driver.f90:144-165).  Cases 14 and 15 use
    the modified zero and optimized power damping and are not extracted."""
    vdir = os.path.join(ref, "test/validation")
    cases = [
        ("11-pairwise-d3-bj", "11-cf3br-benzene.xyz", read_xyz, "rational",
         dict(s6=1.0, s8=2.9491, a1=0.4466, a2=6.1742, s9=0.0, alp=14.0)),
        ("12-pairwise-d3-zero", "12-butane.gen", read_gen, "zero",
         dict(s6=1.0, s8=1.764, rs6=0.806, rs8=1.0, s9=1.0, alp=14.0)),
        ("13-pairwise-d3-bjm", "13-c60.coord", read_tmol, "rational",
         dict(s6=1.0, s8=1.206988, a1=0.240184, a2=3.864426, s9=2.0, alp=14.0)),
    ]
    out = []
    for name, geo, reader, kind, par in cases:
        nums, pos, lat, pbc = reader(os.path.join(vdir, geo))
        ref_json = json.load(open(os.path.join(vdir, name + ".json")))
        n = len(nums)
        # Fortran (n, n) matrices written column by column; both are symmetric
        e2 = ref_json["additive pairwise energy"]
        e3 = ref_json["non-additive pairwise energy"]
        out.append(dict(
            name="validation_" + name, source=f"test/validation/{name}.json,{name}.resp,{geo}",
            numbers=nums, positions=pos, lattice=lat, periodic=pbc, kind=kind, param=par,
            cutoff=[40.0, 60.0, 40.0, 0.0, 0.0],
            expected=dict(energy=ref_json["energy"],
                          pair2=[[e2[c * n + r] for c in range(n)] for r in range(n)],
                          pair3=[[e3[c * n + r] for c in range(n)] for r in range(n)]),
            atol=1.0e-6))
    return out


def python_pairwise_fixture(ref):
    """python/dftd3/test_interface.py:569-762 test_pair_resolved (S8 ring,
    bohr, RationalDampingParam(a1=0.5545, s8=2.2609, a2=3.2297) with the
    Python defaults s6 = 1, s9 = 1, alp = 14; abs tolerance 1e-8)."""
    text = open(os.path.join(ref, "python/dftd3/test_interface.py")).read()
    body = re.search(r"def test_pair_resolved\(\).*?(?=\ndef |\Z)", text, re.S).group(0)
    nums = json.loads(re.search(r"numbers = np.array\((\[.*?\])\)", body, re.S).group(1))

    def block(name):
        m = re.search(name + r" = np.array\(\s*\[(.*?)\]\s*\)\n", body, re.S)
        rows = re.findall(r"\[([^\[\]]*)\]", m.group(1))
        return [[float(v) for v in re.findall(r"[+-]?\d+\.\d+(?:[eE][+-]?\d+)?", r)] for r in rows]

    par = re.search(r"RationalDampingParam\((.*?)\)", body, re.S).group(1)
    par = {k: float(v) for k, v in re.findall(r"(\w+)=([-\d.]+)", par)}
    par.update(s6=1.0, s9=1.0, alp=14.0)
    thr = float(re.search(r"thr = ([-\d.eE]+)", body).group(1))
    return dict(name="python_pair_resolved", source="python/dftd3/test_interface.py:569-762",
                numbers=nums, positions=block("positions"), lattice=None, periodic=None,
                kind="rational", param=par, cutoff=[40.0, 60.0, 40.0, 0.0, 0.0],
                expected=dict(pair2=block("pair_disp2"), pair3=block("pair_disp3")), atol=thr)


def python_fixtures(ref):
    text = open(os.path.join(ref, "python/dftd3/test_interface.py")).read()
    m = re.search(r"def model\(\).*?numbers = np.array\((\[.*?\])\).*?positions = np.array\(\s*\[(.*?)\]\s*\)", text, re.S)
    nums = json.loads(m.group(1))
    vals = [float(v) for v in re.findall(r"[+-]?\d+\.\d+", m.group(2))]
    pos = _triples(vals)
    refs = {}
    # (test, damping kind, method name; "" for explicit parameters)
    cases = (("test_pbe0_d3_bj", "rational", "pbe0"), ("test_b3lyp_d3_zero", "zero", "b3lyp"),
             ("test_pbe_d3_bjm", "mrational", "pbe"), ("test_bp_d3_zerom", "mzero", "bp"),
             ("test_b97d_d3_op", "optimizedpower", "b97d"), ("test_pbe_d3_z", "z", ""))
    explicit = {}
    for test, kind, method in cases:
        mm = re.search(r"def " + test + r"\(.*?ref = ([-\d.e]+) if atm else ([-\d.e]+)(.*?)\n\n", text, re.S)
        refs[(kind, method)] = (float(mm.group(1)), float(mm.group(2)))
        if not method:  # ZDampingParam(a1=..., s9=1.0 if atm else 0.0), other keywords at their defaults
            explicit[kind] = {k: float(v) for k, v in re.findall(r"(\w+)=([-\d.]+)(?:,|\))", mm.group(3))}
    return dict(name="python_interface_model",
                source="python/dftd3/test_interface.py:78-100,236-240,319-356",
                numbers=nums, positions=pos, lattice=None, periodic=None,
                energies={f"{k}:{m}": {"atm": v[0], "noatm": v[1]} for (k, m), v in refs.items()},
                explicit=explicit, rel=1.0e-6)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    here = os.path.dirname(os.path.abspath(__file__))
    ap.add_argument("--out", default=here)
    args = ap.parse_args()
    params = json.load(open(os.path.join(here, "..", "..", "simple-dftd3_amd", "data", "d3_params.json")))
    fix = unit_fixtures(args.reference) + validation_fixtures(args.reference, params)
    with open(os.path.join(args.out, "reference_goldens.json"), "w") as fh:
        json.dump(fix, fh, indent=1)
    with open(os.path.join(args.out, "python_interface_model.json"), "w") as fh:
        json.dump(python_fixtures(args.reference), fh, indent=1)
    pw = pairwise_fixtures(args.reference) + [python_pairwise_fixture(args.reference)]
    with open(os.path.join(args.out, "pairwise_goldens.json"), "w") as fh:
        json.dump(pw, fh)
    with open(os.path.join(args.out, "family_goldens.json"), "w") as fh:
        json.dump(family_fixtures(args.reference, params), fh)
    print("wrote", len(fix), "golden cases")


if __name__ == "__main__":
    main()

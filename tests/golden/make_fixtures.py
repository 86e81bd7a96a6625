#!/usr/bin/env python3
"""Extract the reference's own golden data into committed fixtures.

Reads (as text) the LSQuantumED test sources and reference_data under
``/root/reference/tests/unit_tests`` and writes

* ``tests/golden/reference_data/*.ref`` -- the reference's dump files, copied
  verbatim (they are data: PETSc ASCII VecView/MatView dumps and DSF tables);
* ``tests/golden/ref_literals.json`` -- every literal the reference's gtest
  files assert on the hot path, with its source file:line, grouped by the
  routine it pins (inputs + expected outputs only).

Run once in a container that has the reference mounted; the outputs are
committed and nothing at test time reads the reference.
"""
from __future__ import annotations

import json
import os
import re
import shutil
import sys

REF = os.environ.get("LSQED_REFERENCE", "/root/reference")
UT = os.path.join(REF, "tests", "unit_tests")
OUT = os.path.dirname(os.path.abspath(__file__))

NUM = r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?"


def rel(path):
    return os.path.relpath(path, REF)


def meson_np(dirpath: str) -> dict[str, int]:
    """test-exe basename -> number of MPI processes from the suite's meson.build."""
    out = {}
    mb = os.path.join(dirpath, "meson.build")
    if not os.path.exists(mb):
        return out
    txt = open(mb).read()
    # pattern: get_variable('exec_' + 'name' + 'k') ... args: ['-np', 'P'
    for m in re.finditer(r"get_variable\('exec_'\s*\+\s*'(\w+)'\s*\+\s*'(\w+)'\)(.*?)'-np',\s*'(\d+)'", txt, re.S):
        out[m.group(1) + m.group(2)] = int(m.group(4))
    return out


def split(size, P, rank):
    rest = size % P
    loc = size // P + (rest > rank)
    return loc, rank * loc + (rest <= rank) * rest


def comb(n, k):
    from math import comb as c
    return c(n, k) if 0 <= k <= n else 0


class Ctx:
    """Parsing state for one test source: environment size, basis, lattices."""

    def __init__(self, path):
        self.path = path
        self.N = None
        self.mag = None
        self.lat = {}     # var -> (type, lx, ly)
        self.rank = 0
        self.ranks_seen = []

    def lattice_decl(self, line):
        m = re.search(r"(chain1D|square2D|honeycomb2D)\s+(\w+)\s*[{(]\s*env\s*(?:,\s*(\d+)\s*,\s*(\d+))?\s*[})]", line)
        if m:
            t = {"chain1D": "chain", "square2D": "square", "honeycomb2D": "honeycomb"}[m.group(1)]
            self.lat[m.group(2)] = (t, int(m.group(3) or 0), int(m.group(4) or 0))
            return True
        return False

    def env_decl(self, line):
        m = re.search(r"Environment\s+env\s*\{[^,]*,[^,]*,\s*(\d+)\s*,", line)
        if m:
            self.N = int(m.group(1))
        m = re.search(r"env\.nspins\s*=\s*(\d+)", line)
        if m:
            self.N = int(m.group(1))
        m = re.search(r"Basis\s+\w+\s*\{\s*env\s*,\s*(-?\d+)\s*\}", line)
        if m:
            self.mag = int(m.group(1))


def parse_coup_elems(path, np_):
    """get_coup_elems(neigh_type::K, b, LOCAL, lat, ncol, coup_elems) blocks."""
    cases = []
    ctx = Ctx(path)
    lines = open(path).read().splitlines()
    test = None
    cur = None
    branch_ranks = []
    for ln_no, line in enumerate(lines, 1):
        s = line.strip()
        if s.startswith("//"):
            continue
        ctx.env_decl(line)
        ctx.lattice_decl(line)
        m = re.match(r"TEST\((\w+),\s*(\w+)\)", s)
        if m:
            test = m.group(2)
            branch_ranks = []
            ctx.rank = 0
        m = re.search(r"if \(b\.mpi_rank == (\d+)\)", s)
        if m:
            ctx.rank = int(m.group(1))
            branch_ranks.append(ctx.rank)
        elif re.match(r"\}?\s*else\s*\{?$", s):
            r = 0
            while r in branch_ranks:
                r += 1
            ctx.rank = r
            branch_ranks.append(r)
        m = re.search(r"get_coup_elems\(neigh_type::(nnn|nn)\s*,\s*b\s*,\s*(\d+)\s*,\s*(\w+)", s)
        if m:
            nup = ctx.mag + ctx.N // 2
            D = comb(ctx.N, nup)
            _, start = split(D, np_, ctx.rank)
            t, lx, ly = ctx.lat[m.group(3)]
            cur = dict(test=test, src=f"{rel(path)}:{ln_no}", N=ctx.N, mag=ctx.mag, lattice=t, lx=lx, ly=ly,
                       kind=m.group(1), mpi_size=np_, mpi_rank=ctx.rank, local_index=int(m.group(2)),
                       global_index=start + int(m.group(2)), ncol=None, cols={})
            cases.append(cur)
            continue
        if cur is not None:
            m = re.search(r"ASSERT_EQ\((\d+),\s*ncol\)", s)
            if m:
                cur["ncol"] = int(m.group(1))
            m = re.search(r"EXPECT_EQ\((-?\d+),\s*coup_elems\[(\d+)\]\)", s)
            if m:
                cur["cols"][int(m.group(2))] = int(m.group(1))
    for c in cases:
        c["cols"] = [c["cols"][k] for k in sorted(c["cols"])]
    return cases


def parse_elem_diag(path):
    out = []
    ctx = Ctx(path)
    test = None
    for ln_no, line in enumerate(open(path).read().splitlines(), 1):
        s = line.strip()
        if s.startswith("//"):
            continue
        ctx.env_decl(line)
        ctx.lattice_decl(line)
        m = re.match(r"TEST\((\w+),\s*(\w+)\)", s)
        if m:
            test = m.group(2)
        m = re.search(r"EXPECT_EQ\((-?\d+),\s*HamiltHelper::get_elem_diag\(neigh_type::(nnn|nn)\s*,\s*(\d+)\s*,\s*(\d+)\s*,\s*(\w+)\)\)", s)
        if m:
            t, lx, ly = ctx.lat[m.group(5)]
            out.append(dict(test=test, src=f"{rel(path)}:{ln_no}", expect=int(m.group(1)), kind=m.group(2),
                            state=int(m.group(3)), N=int(m.group(4)), lattice=t, lx=lx, ly=ly))
    return out


def safe_arith(expr):
    if not re.fullmatch(r"[-+*/(). 0-9eE]+", expr):
        raise ValueError(expr)
    return float(eval(expr, {"__builtins__": {}}, {}))  # arithmetic literal only


def parse_disorder_elem(path):
    out = []
    test = None
    for ln_no, line in enumerate(open(path).read().splitlines(), 1):
        s = line.strip()
        if s.startswith("//"):
            continue
        m = re.match(r"TEST\((\w+),\s*(\w+)\)", s)
        if m:
            test = m.group(2)
        m = re.search(r"EXPECT_NEAR\(([^,]+),\s*\(double\)HamiltHelper::get_disorder_elem\((\d+)\s*,\s*(\d+)\s*,\s*hi\)\s*,\s*(" + NUM + r")\)", s)
        if m:
            out.append(dict(test=test, src=f"{rel(path)}:{ln_no}", expect=safe_arith(m.group(1)),
                            state=int(m.group(2)), N=int(m.group(3)), tol=float(m.group(4)),
                            hi="i/N"))
    return out


def parse_solver(path):
    """One case per TEST: couplings, sector, lattice, builds, EPS dims, literals."""
    txt = open(path).read().splitlines()
    cases = []
    ctx = Ctx(path)
    cur = None
    refeigs = {}
    for ln_no, line in enumerate(txt, 1):
        s = line.strip()
        if s.startswith("//"):
            continue
        ctx.env_decl(line)
        m = re.match(r"TEST\((\w+),\s*(\w+)\)", s)
        if m:
            cur = dict(test=f"{m.group(1)}.{m.group(2)}", src=f"{rel(path)}:{ln_no}", J1=0.0, D1=0.0, J2=0.0,
                       D2=0.0, builds=[], nev=None, ncv=None, mpd=None, eigs=[], tol=None, assert_gt=None,
                       check_gt=None)
            cases.append(cur)
            refeigs = {}
            continue
        if cur is None:
            continue
        m = re.search(r"PetscReal (J1|Delta1|J2|Delta2)\s*=\s*(" + NUM + ")", s)
        if m:
            cur[{"J1": "J1", "Delta1": "D1", "J2": "J2", "Delta2": "D2"}[m.group(1)]] = float(m.group(2))
        if ctx.lattice_decl(line):
            t, lx, ly = list(ctx.lat.values())[-1]
            cur.update(lattice=t, lx=lx, ly=ly)
        m = re.search(r"Basis\s+b\s*\{\s*env\s*,\s*(-?\d+)\s*\}", s)
        if m:
            cur.update(N=ctx.N, mag=int(m.group(1)))
        for b in ("build_diag", "build_off_diag", "build_disorder"):
            if re.search(r"h\." + b + r"\(", s):
                cur["builds"].append(b)
        m = re.search(r"SolverInit\(solver,\s*h\.hamilt,\s*(\d+)\)", s)
        if m:
            cur["nev"] = int(m.group(1))
        m = re.search(r"EPSSetDimensions\(solver,\s*(\w+),\s*(\w+),\s*(\w+)\)", s)
        if m:
            conv = lambda v: None if v == "PETSC_DECIDE" else int(v)
            cur["nev"], cur["ncv"], cur["mpd"] = conv(m.group(1)), conv(m.group(2)), conv(m.group(3))
        m = re.search(r"EXPECT_NEAR\((" + NUM + r"),\s*PetscRealPart\(real_eig\),\s*(" + NUM + r")\)", s)
        if m:
            cur["eigs"].append(float(m.group(1)))
            cur["tol"] = float(m.group(2))
        m = re.search(r"ASSERT_GT\(nconv,\s*(\d+)\)", s)
        if m:
            cur["assert_gt"] = int(m.group(1))
        m = re.search(r"if \(nconv > (\d+)\)", s)
        if m:
            cur["check_gt"] = int(m.group(1))
        if re.search(r"for \(int i = 0; i < nconv; \+\+i\)", s):
            cur["check_gt"] = 0
        m = re.search(r"ref_eigs\[(\d+)\]\s*=\s*(" + NUM + ")", s)
        if m:
            refeigs[int(m.group(1))] = float(m.group(2))
            cur["eigs"] = [refeigs[k] for k in sorted(refeigs)]
        m = re.search(r"EXPECT_NEAR\(ref_eigs\[i\],\s*PetscRealPart\(real_eig\),\s*(" + NUM + r")\)", s)
        if m:
            cur["tol"] = float(m.group(1))
    return [c for c in cases if c["eigs"]]


def parse_correlation(path, opname):
    txt = open(path).read().splitlines()
    ctx = Ctx(path)
    case = dict(src=rel(path), op=opname, J1=0.0, D1=0.0, J2=0.0, D2=0.0, values=[], E0=None)
    for ln_no, line in enumerate(txt, 1):
        s = line.strip()
        if s.startswith("//"):
            continue
        ctx.env_decl(line)
        if ctx.lattice_decl(line):
            t, lx, ly = list(ctx.lat.values())[-1]
            case.update(lattice=t, lx=lx, ly=ly)
        m = re.search(r"PetscReal (J1|Delta1|J2|Delta2)\s*=\s*(" + NUM + ")", s)
        if m:
            case[{"J1": "J1", "Delta1": "D1", "J2": "J2", "Delta2": "D2"}[m.group(1)]] = float(m.group(2))
        m = re.search(r"Basis\s+b\s*\{\s*env\s*,\s*(-?\d+)\s*\}", s)
        if m:
            case.update(N=ctx.N, mag=int(m.group(1)))
        m = re.search(r"ExpectVal\(env,\s*(\d+)\s*,\s*i\s*(?:\+\s*(\d+))?\s*,\s*xr\)", s)
        if m:
            case["i"] = int(m.group(1))
            case["j_offset"] = int(m.group(2) or 0)
        m = re.search(r"EXPECT_NEAR\((" + NUM + r"),\s*PetscRealPart\(real_eig\)", s)
        if m:
            case["E0"] = float(m.group(1))
        m = re.search(r"EXPECT_NEAR\((" + NUM + r"),\s*PetscRealPart\(exp_val\[(\d+)\]\),\s*(" + NUM + r")\)", s)
        if m:
            case["values"].append(float(m.group(1)))
            case["tol"] = float(m.group(3))
            case.setdefault("lines", []).append(ln_no)
    return case


def parse_random(path):
    txt = open(path).read()
    seeds = [int(v) for v in re.findall(r"EXPECT_EQ\((\d+),\s*rd\.srand_seq\[\d+\]\)", txt)]
    m = re.search(r"RandomDisorder rd\{(\d+),\s*PETSC_TRUE,\s*(" + NUM + r"),\s*(" + NUM + r"),\s*(\d+)\}", txt)
    hi_idx = int(re.search(r"rd\.hi_init\((\d+)\)", txt).group(1))
    his = [float(v) for v in re.findall(r"EXPECT_NEAR\((" + NUM + r"),\s*rd\.hi\[\d+\]", txt)]
    return dict(src=rel(path), N=int(m.group(1)), min=float(m.group(2)), max=float(m.group(3)),
                rep=int(m.group(4)), seeds=seeds, hi_index=hi_idx, hi=his, tol=1e-13)


def parse_lattice_tests(paths):
    out = []
    for path in paths:
        ctx = Ctx(path)
        for ln_no, line in enumerate(open(path).read().splitlines(), 1):
            s = line.strip()
            if s.startswith("//"):
                continue
            ctx.env_decl(line)
            ctx.lattice_decl(line)
            m = re.search(r"EXPECT_EQ\((-?\d+),\s*(\w+)\.(nn|nnn)\[(\d+)\]\)", s)
            if m and m.group(2) in ctx.lat:
                t, lx, ly = ctx.lat[m.group(2)]
                out.append(dict(src=f"{rel(path)}:{ln_no}", N=ctx.N, lattice=t, lx=lx, ly=ly, list=m.group(3),
                                index=int(m.group(4)), expect=int(m.group(1))))
            m = re.search(r"EXPECT_EQ\((-?\d+),\s*(\w+)\.(num_nnXsite|num_nnnXsite)\((\d+)\)\)", s)
            if m and m.group(2) in ctx.lat:
                t, lx, ly = ctx.lat[m.group(2)]
                out.append(dict(src=f"{rel(path)}:{ln_no}", N=ctx.N, lattice=t, lx=lx, ly=ly, list=m.group(3),
                                index=int(m.group(4)), expect=int(m.group(1))))
    return out


def parse_sublattice(paths):
    out = []
    for path in paths:
        ctx = Ctx(path)
        subl = {}
        for ln_no, line in enumerate(open(path).read().splitlines(), 1):
            s = line.strip()
            if s.startswith("//"):
                continue
            ctx.env_decl(line)
            ctx.lattice_decl(line)
            m = re.search(r"(AF|Striped)<(\w+)>\s+(\w+)\s*\{\s*(\w+)", s)
            if m:
                subl[m.group(3)] = (m.group(1), m.group(4))
            m = re.search(r"EXPECT_EQ\((-?\d+),\s*(\w+)\.get_sl\((\d+)\)\[(\d+)\]\)", s)
            if m and m.group(2) in subl:
                kind, latv = subl[m.group(2)]
                t, lx, ly = ctx.lat[latv]
                out.append(dict(src=f"{rel(path)}:{ln_no}", N=ctx.N, lattice=t, lx=lx, ly=ly, kind=kind,
                                sl=int(m.group(3)), index=int(m.group(4)), expect=int(m.group(1))))
    return out


def parse_compare_cases():
    """Parameters of every build_* compare test (test source -> .ref file)."""
    out = []
    for root, _, files in os.walk(UT):
        for f in sorted(files):
            if not (f.startswith("test_build_") and f.endswith(".cpp")):
                continue
            path = os.path.join(root, f)
            txt = open(path).read()
            d = {"src": rel(path)}
            for k, key in (("J1", "J1"), ("Delta1", "D1"), ("J2", "J2"), ("Delta2", "D2")):
                d[key] = float(re.search(r"PetscReal " + k + r"\s*=\s*(" + NUM + ")", txt).group(1))
            d["N"] = int(re.search(r"Environment env\{[^,]*,[^,]*,\s*(\d+)", txt).group(1))
            d["mag"] = int(re.search(r"Basis b\{env,\s*(-?\d+)\}", txt).group(1))
            m = re.search(r"(chain1D|square2D|honeycomb2D) lat\{env\s*(?:,\s*(\d+)\s*,\s*(\d+))?\}", txt)
            d["lattice"] = {"chain1D": "chain", "square2D": "square", "honeycomb2D": "honeycomb"}[m.group(1)]
            d["lx"], d["ly"] = int(m.group(2) or 0), int(m.group(3) or 0)
            body = txt.split("int main", 1)[1]
            d["builds"] = [b for b in ("build_diag", "build_off_diag", "build_disorder")
                           if re.search(r"h\." + b + r"\(", body)]
            d["hi"] = "i/N" if "hi[i] = PetscReal(i)/env.nspins" in txt else None
            d["view"] = "vec" if "VecView" in txt else "mat"
            d["ref"] = re.search(r'"(\w+)\.dat"', txt).group(1) + ".ref"
            out.append(d)
    return out


def main():
    if not os.path.isdir(UT):
        sys.exit(f"reference tests not found under {UT}")
    dst = os.path.join(OUT, "reference_data")
    os.makedirs(dst, exist_ok=True)
    for f in sorted(os.listdir(os.path.join(UT, "reference_data"))):
        shutil.copyfile(os.path.join(UT, "reference_data", f), os.path.join(dst, f))

    lit = {"_about": "Literals asserted by LSQuantumED's unit tests (file:line per entry); generated by "
                     "tests/golden/make_fixtures.py from the reference test sources."}
    ce = []
    npmap = meson_np(os.path.join(UT, "get_coup_elems"))
    for k in range(1, 6):
        p = os.path.join(UT, "get_coup_elems", f"test_get_coup_elems{k}.cpp")
        ce += parse_coup_elems(p, npmap.get(f"get_coup_elems{k}", 1))
    lit["get_coup_elems"] = ce
    ed = []
    for k in range(1, 6):
        ed += parse_elem_diag(os.path.join(UT, "get_elem_diag", f"test_get_elem_diag{k}.cpp"))
    lit["get_elem_diag"] = ed
    de = []
    for k in (1, 3):
        de += parse_disorder_elem(os.path.join(UT, "get_disorder_elem", f"test_get_disorder_elem{k}.cpp"))
    lit["get_disorder_elem"] = de
    sv = []
    for f in sorted(os.listdir(os.path.join(UT, "solver"))):
        if f.startswith("test_solver") and f[11].isdigit():
            sv += parse_solver(os.path.join(UT, "solver", f))
    lit["solver"] = sv
    lit["correlation"] = [
        parse_correlation(os.path.join(UT, "correlation", "SzSz", "test_szsz1.cpp"), "SzSz"),
        parse_correlation(os.path.join(UT, "correlation", "SzSz", "test_szsz2.cpp"), "SzSz"),
        parse_correlation(os.path.join(UT, "correlation", "SpSm", "test_spsm1.cpp"), "SpSm"),
    ]
    lit["random"] = parse_random(os.path.join(UT, "tools_random", "test_tools_random1.cpp"))
    lit["lattice"] = parse_lattice_tests(sorted(
        os.path.join(UT, "lattice", f) for f in os.listdir(os.path.join(UT, "lattice")) if f.endswith(".cpp")))
    lit["sublattice"] = parse_sublattice(sorted(
        os.path.join(UT, "sublattice", f) for f in os.listdir(os.path.join(UT, "sublattice")) if f.endswith(".cpp")))
    lit["compare"] = parse_compare_cases()
    # DSF drivers (tests/unit_tests/phys/test_dsf{1,2}.cpp): chain N=6, J1=D1=1,
    # GS in sector mag 0, DSF with nev=10 in sector mag 0 (Sqz) or mag 1 (Sqp).
    lit["dsf"] = [
        dict(src="tests/unit_tests/phys/test_dsf1.cpp:59-88", ref="dsf_d0_1.ref", op="Sqz", N=6, mag0=0, mag1=0,
             lattice="chain", J1=1.0, D1=1.0, J2=0.0, D2=0.0, nev=10, E0=-2.8027756377319952),
        dict(src="tests/unit_tests/phys/test_dsf2.cpp:69-88", ref="dsf_d0_2.ref", op="Sqp", N=6, mag0=0, mag1=1,
             lattice="chain", J1=1.0, D1=1.0, J2=0.0, D2=0.0, nev=10),
    ]
    with open(os.path.join(OUT, "ref_literals.json"), "w") as fh:
        json.dump(lit, fh, indent=1)
    print({k: (len(v) if isinstance(v, list) else 1) for k, v in lit.items()})


if __name__ == "__main__":
    main()

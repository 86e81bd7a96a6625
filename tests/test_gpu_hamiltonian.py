"""GPU parity of the matrix-free Hamiltonian against the oracle.

Bit-exact: basis states (unrank), ranks, column lists in emission order,
the stored diagonal.  FP64 H.v: within 1e-13 relative (summation order
differs from a CSR row sum).  Every case prints its name first so a failure
is attributable."""
import math
import sys

import numpy as np
import pytest

from closed_forms import CLOSED, magnon_levels, polarised_energy

pytestmark = pytest.mark.gpu

LT = {"chain": 0, "square": 1, "honeycomb": 2}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as T
    assert T.cuda.is_available(), "GPU tests need a GPU"
    return T


def make(lat, N, mag, J1=1.0, D1=1.0, J2=0.0, D2=0.0, lx=0, ly=0, builds=("build_diag", "build_off_diag"), hi=None):
    import lsquantumed_amd as Q
    env = Q.Environment(N)
    b = Q.Basis(env, mag)
    L = {"chain": lambda: Q.chain1D(env), "square": lambda: Q.square2D(env, lx, ly),
         "honeycomb": lambda: Q.honeycomb2D(env, lx, ly)}[lat]()
    h = Q.Hamiltonian(env, b, L, J1, D1, J2, D2)
    if "build_diag" in builds:
        h.build_diag(env)
    if "build_off_diag" in builds:
        h.build_off_diag(env)
    if "build_disorder" in builds:
        h.build_disorder(env, hi)
    return h


def oracle_model(O, lat, N, mag, J1=1.0, D1=1.0, J2=0.0, D2=0.0, lx=0, ly=0):
    return O.Model(LT[lat], N, mag + N // 2, J1, D1, J2, D2, lx=lx, ly=ly)


def flags_of(O, builds):
    f = 0
    f |= O.Model.BUILD_DIAG if "build_diag" in builds else 0
    f |= O.Model.BUILD_OFFDIAG if "build_off_diag" in builds else 0
    f |= O.Model.BUILD_DISORDER if "build_disorder" in builds else 0
    return f


@pytest.mark.parametrize("N,mag", [(6, 0), (6, -2), (8, 1), (12, 0), (16, 0), (20, 3), (24, 0), (34, -13), (36, -15)])
def test_unrank_rank_bitexact(O, torch, N, mag):
    say("unrank/rank", N, mag)
    h = make("chain", N, mag)
    D = h.hamilt_size()
    nup = mag + N // 2
    assert D == math.comb(N, nup)
    n = min(D, 200000)
    st = h.unrank(0, n).cpu().numpy().astype(np.uint64)
    assert np.array_equal(st, O.int_basis(N, nup, 0, n))
    r = h.rank(torch.from_numpy(st.astype(np.int64)).cuda()).cpu().numpy()
    assert np.array_equal(r, np.arange(n))
    if D > n:  # the tail of the sector
        st2 = h.unrank(D - 1000, 1000).cpu().numpy().astype(np.uint64)
        assert np.array_equal(st2, O.int_basis(N, nup, D - 1000, 1000))


@pytest.mark.parametrize("N,lx,ly", [(32, 4, 8), (36, 6, 6)])
def test_unrank_rank_large_sector_sample(O, torch, N, lx, ly):
    say("unrank/rank sample", N)
    h = make("square", N, 0, lx=lx, ly=ly)
    D = h.hamilt_size()
    rng = np.random.default_rng(N)
    for row0 in [0, D // 7, D // 2, D - 4096] + [int(v) for v in rng.integers(0, D - 4096, 4)]:
        st = h.unrank(row0, 4096).cpu().numpy()
        ref = O.int_basis(N, N // 2, row0, 4096).astype(np.int64)
        assert np.array_equal(st, ref), row0
        assert np.array_equal(h.rank(torch.from_numpy(st).cuda()).cpu().numpy(), np.arange(row0, row0 + 4096))


@pytest.mark.parametrize("N,mag", [(12, 0), (16, -3), (36, 0)])
def test_rank_rejects_states_outside_sector(O, torch, N, mag):
    """lsqed_rank writes -1 for a state with a bit at or above N or with a
    number of up spins other than the sector's, and the exact rank for the
    sector states mixed in with them."""
    say("rank outside the sector", N, mag)
    h = make("chain", N, mag)
    nup = N // 2 + mag
    good = [O.unrank_ref(r, N, nup) for r in (0, 1, h.hamilt_size() // 2, h.hamilt_size() - 1)]
    top = (1 << N) - 1
    bad = [good[0] | (1 << N), good[1] ^ 1 if good[1] & 1 else good[1] | 1, 0, top, 1 << 63, (1 << 64) - 1,
           (good[2] << 1) & top if bin((good[2] << 1) & top).count("1") != nup else 1 << (N + 2)]
    states = good + bad
    u64 = np.array(states, dtype=np.uint64)
    got = h.rank(torch.from_numpy(u64.view(np.int64)).cuda()).cpu().numpy()
    want = [O.rank_ref(s, N) for s in good] + [-1] * len(bad)
    assert list(got) == want, (got, want)


def test_export_cols_reference_literals(O, torch, literals):
    """Every get_coup_elems literal of the reference's tests (emission order,
    duplicates included), through the device column export."""
    say("export literals", len(literals["get_coup_elems"]))
    cache = {}
    for c in literals["get_coup_elems"]:
        key = (c["lattice"], c["N"], c["mag"], c["lx"], c["ly"])
        if key not in cache:
            cache[key] = make(c["lattice"], c["N"], c["mag"], lx=c["lx"], ly=c["ly"])
        h = cache[key]
        kind = 0 if c["kind"] == "nn" else 1
        cols, ncol = h.export_cols(kind, c["global_index"], 1)
        n = int(ncol[0])
        assert n == c["ncol"], c["src"]
        assert cols[0, :n].cpu().tolist() == c["cols"], c["src"]


CASES = [
    ("chain", 6, 0, 1, 1, 0, 0, 0, 0), ("chain", 6, -1, 1, 1, 0, 0, 0, 0), ("chain", 8, 0, 1, 0.3456, 0.7, 0.2, 0, 0),
    ("chain", 16, 0, 1, 1, 0, 0, 0, 0), ("chain", 18, 0, 1, 1, 0.2, 0.2, 0, 0),
    ("square", 8, 2, 1, 1, 1, 1, 2, 4), ("square", 12, 5, 1, 1, 1, 1, 3, 4), ("square", 16, 0, 1, -0.43, -1, 1, 4, 4),
    ("square", 20, 0, 1, 1, 0.5, 0.5, 4, 5), ("honeycomb", 12, 5, 1, 1, 1, 1, 6, 2),
    ("honeycomb", 8, 2, 1, 1, 0, 0, 2, 4), ("square", 6, 0, 1, 1, 1, 1, 6, 1),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}{c[1]}m{c[2]}" for c in CASES])
def test_export_cols_bitexact(O, torch, case):
    lat, N, mag, J1, D1, J2, D2, lx, ly = case
    say("export", case)
    h = make(lat, N, mag, J1, D1, J2, D2, lx, ly)
    D = h.hamilt_size()
    nup = mag + N // 2
    states = O.int_basis(N, nup)
    for kind in (0, 1):
        bi, bj = O.bonds(LT[lat], N, lx, ly, kind)
        cols, ncol = h.export_cols(kind)
        cols, ncol = cols.cpu().numpy(), ncol.cpu().numpy()
        for r in range(D):
            ref = O.coup_elems(int(states[r]), N, bi, bj)
            assert ncol[r] == len(ref) and list(cols[r, :ncol[r]]) == ref, (kind, r)


def test_diag_fixtures_bitexact(O, torch, literals, refdata):
    """build_diag / build_disorder dumps of the reference, through lsqed_diag."""
    for c in literals["compare"]:
        if c["view"] != "vec" and "build_disorder" not in c["builds"]:
            continue
        say("diag fixture", c["ref"])
        hi = np.arange(c["N"]) / c["N"] if c["hi"] else None
        h = make(c["lattice"], c["N"], c["mag"], c["J1"], c["D1"], c["J2"], c["D2"], c["lx"], c["ly"],
                 builds=c["builds"], hi=hi)
        d = h.diagonal().cpu().numpy()
        m = oracle_model(O, c["lattice"], c["N"], c["mag"], c["J1"], c["D1"], c["J2"], c["D2"], c["lx"], c["ly"])
        dref, _ = m.diag(hi, flags_of(O, c["builds"]))
        assert np.array_equal(d, dref), c["ref"]
        if c["view"] == "vec":
            assert [O.petsc_g(v) for v in d] == O.parse_petsc_vec(refdata(c["ref"]))


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}{c[1]}m{c[2]}" for c in CASES])
@pytest.mark.parametrize("fields", [False, True])
@pytest.mark.parametrize("kernel", [1, 2, 3], ids=["generic", "split", "split_jit"])
def test_hv_matches_oracle_csr(O, torch, case, fields, kernel):
    lat, N, mag, J1, D1, J2, D2, lx, ly = case
    say("hv", case, fields, kernel)
    hi = np.random.default_rng(N).uniform(-2, 2, N) if fields else None
    builds = ("build_diag", "build_off_diag", "build_disorder") if fields else ("build_diag", "build_off_diag")
    h = make(lat, N, mag, J1, D1, J2, D2, lx, ly, builds=builds, hi=hi)
    h.set_hv_kernel(kernel)
    m = oracle_model(O, lat, N, mag, J1, D1, J2, D2, lx, ly)
    A = m.csr(hi, flags_of(O, builds))
    D = h.hamilt_size()
    x = np.random.default_rng(D).uniform(-1, 1, D)
    y = h.matvec(torch.from_numpy(x).cuda()).cpu().numpy()
    yref = A @ x
    scale = np.abs(A).max() * np.abs(x).max()
    assert np.max(np.abs(y - yref)) <= 1e-13 * scale
    # diagonal entries bit-exact as well
    d = h.diagonal().cpu().numpy()
    assert np.array_equal(d, A.diagonal())


def test_build_flags(O, torch):
    """The operator holds only what the build calls added (H starts empty)."""
    say("build flags")
    import lsquantumed_amd as Q
    env = Q.Environment(8)
    h = Q.Hamiltonian(env, Q.Basis(env, 0), Q.chain1D(env), 1.0, 1.0, 0.0, 0.0)
    D = h.hamilt_size()
    x = torch.rand(D, dtype=torch.float64, device="cuda")
    assert torch.count_nonzero(h.matvec(x)) == 0
    h.build_off_diag(env)
    m = O.Model(0, 8, 4, 1.0, 1.0, 0.0, 0.0)
    assert np.allclose(h.matvec(x).cpu().numpy(), m.csr(None, 2) @ x.cpu().numpy(), atol=1e-14)


@pytest.mark.parametrize("kernel", [1, 2, 3], ids=["generic", "split", "split_jit"])
def test_hv_graph_capture_replays(O, torch, kernel):
    """lsqed_hv captured into a HIP graph (torch.cuda.CUDAGraph, on the
    capturing stream) and replayed three times, with eager work and
    synchronisations between the replays: every replay equals the eager
    product bit for bit, also after new contents are copied into x (the
    graph holds the pointers).  4x5 J1-J2 with fields (D = 184,756)."""
    say("hv graph capture", kernel)
    N = 20
    hi = np.random.default_rng(7).uniform(-1, 1, N)
    builds = ("build_diag", "build_off_diag", "build_disorder")
    h = make("square", N, 0, 1.0, 1.0, 0.5, 0.5, 4, 5, builds=builds, hi=hi)
    h.set_hv_kernel(kernel)
    D = h.hamilt_size()
    gen = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(D, dtype=torch.float64, device="cuda", generator=gen)
    x2 = torch.rand(D, dtype=torch.float64, device="cuda", generator=gen)
    ref = h.matvec(x)  # eager; also builds the tables and compiles the kernel
    ref2 = h.matvec(x2)
    y = torch.empty_like(x)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        h.matvec(x, y)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        h.matvec(x, y)
    for rep in range(3):
        y.zero_()
        junk = torch.rand(D, dtype=torch.float64, device="cuda", generator=gen) * 2.0  # eager work in between
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, ref), (kernel, rep)
        del junk
    x.copy_(x2)
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(y, ref2), kernel


@pytest.mark.parametrize("lat,N,mag,lx,ly,builds", [
    ("chain", 16, 0, 0, 0, ("build_diag", "build_off_diag")),
    ("chain", 22, 1, 0, 0, ("build_diag", "build_off_diag", "build_disorder")),
    ("square", 16, 0, 4, 4, ("build_diag", "build_off_diag")),
    ("square", 24, 0, 4, 6, ("build_diag", "build_off_diag")),
    ("honeycomb", 24, 0, 4, 6, ("build_diag", "build_off_diag", "build_disorder")),
    ("chain", 40, -18, 0, 0, ("build_diag", "build_off_diag", "build_disorder")),
    ("square", 36, 17, 6, 6, ("build_diag", "build_off_diag"))])
def test_hv_auto_kernel_is_specialised(torch, lat, N, mag, lx, ly, builds):
    """The automatic choice resolves to the hipRTC-specialised row split on
    every lattice kind, with and without fields, 32- and 64-bit ranks and
    row splits of L = 10..16: a kernel that fails to compile would otherwise
    fall back to the unspecialised kernels unnoticed."""
    hi = np.linspace(-0.5, 0.5, N) if "build_disorder" in builds else None
    h = make(lat, N, mag, 1.0, 1.0, 0.5, 0.5, lx, ly, builds=builds, hi=hi)
    say("auto kernel", lat, N, mag, h.hv_kernel())
    assert h.hv_kernel() == 3


@pytest.mark.parametrize("lat,N,mag,lx,ly,builds", [
    ("square", 24, 0, 4, 6, ("build_diag", "build_off_diag")),
    ("chain", 22, 0, 0, 0, ("build_diag", "build_off_diag", "build_disorder")),
    ("square", 32, 2, 4, 8, ("build_diag", "build_off_diag", "build_disorder"))])
def test_hv_one_row_form_equals_looped_form(O, torch, monkeypatch, lat, N, mag, lx, ly, builds):
    """With one row per block (the default grid) the specialised kernel is
    generated with a row loop that runs once; LSQED_JIT_ROWS_PER_BLOCK=2 / 3
    gives the looped form over a smaller grid.  All three give the same
    product to the last bit (each entry's sum runs in bond order whatever
    block takes the row), and the default one matches the oracle's CSR to
    1e-13 on the smaller sectors."""
    hi = np.random.default_rng(N).uniform(-1, 1, N) if "build_disorder" in builds else None
    ys = []
    x = None
    for k in (None, "2", "3"):
        if k is None:
            monkeypatch.delenv("LSQED_JIT_ROWS_PER_BLOCK", raising=False)
        else:
            monkeypatch.setenv("LSQED_JIT_ROWS_PER_BLOCK", k)
        h = make(lat, N, mag, 1.0, 0.9, 0.5, 0.4, lx, ly, builds=builds, hi=hi)
        assert h.hv_kernel() == 3
        if x is None:
            D = h.hamilt_size()
            x = torch.from_numpy(np.random.default_rng(D).uniform(-1, 1, D)).cuda()
        ys.append(h.matvec(x))
        say("one-row vs looped", lat, N, mag, "rows per block", k or "default")
    for y in ys[1:]:
        assert torch.equal(ys[0], y)
    if h.hamilt_size() <= 3_000_000:
        m = oracle_model(O, lat, N, mag, 1.0, 0.9, 0.5, 0.4, lx, ly)
        A = m.csr(hi, flags_of(O, builds))
        xr = x.cpu().numpy()
        yref = A @ xr
        err = np.abs(ys[0].cpu().numpy() - yref).max() / max(1.0, np.abs(yref).max())
        assert err <= 1e-13, err


@pytest.mark.parametrize("lat,N,mag,lx,ly", [("chain", 38, -17, 0, 0), ("chain", 40, -18, 0, 0),
                                              ("chain", 40, 18, 0, 0), ("square", 40, -17, 4, 10)])
@pytest.mark.parametrize("kernel", [1, 3], ids=["generic", "split_jit"])
def test_hv_largest_n_few_particle_sectors(O, torch, lat, N, mag, lx, ly, kernel):
    """N = 38 and 40 in sectors of 2 or 3 up spins (or 2 down): the row split
    at its widest -- L = N - 24 low sites (14, 16) under 24 high ones, 64-bit
    ranks, 1024-thread row blocks at L = 16, rows that are mostly empty --
    with J2 and fields, against the oracle's CSR to 1e-13."""
    say("hv large N", lat, N, mag, kernel)
    hi = np.random.default_rng(N + mag).uniform(-1, 1, N)
    builds = ("build_diag", "build_off_diag", "build_disorder")
    h = make(lat, N, mag, 1.0, 0.8, 0.4, 0.3, lx, ly, builds=builds, hi=hi)
    h.set_hv_kernel(kernel)
    m = oracle_model(O, lat, N, mag, 1.0, 0.8, 0.4, 0.3, lx, ly)
    A = m.csr(hi, flags_of(O, builds))
    D = h.hamilt_size()
    assert D == A.shape[0]
    x = np.random.default_rng(D).uniform(-1, 1, D)
    y = h.matvec(torch.from_numpy(x).cuda()).cpu().numpy()
    yref = A @ x
    assert np.max(np.abs(y - yref)) <= 1e-13 * np.abs(A).max() * np.abs(x).max()


@pytest.mark.parametrize("lat,N,lx,ly", CLOSED, ids=[f"{c[0]}{c[1]}" for c in CLOSED])
def test_hv_closed_form_sectors(torch, lat, N, lx, ly):
    """Sectors whose spectrum is known in closed form, through every H.v
    kernel that takes N (the unspecialised row split stops at 32): the
    fully polarised sectors with fields (D = 1, polarised_energy) and the
    single-magnon sectors n_up = 1 and N - 1 (D = N; the dense H assembled
    column by column from H e_i has the spectrum of magnon_levels to
    1e-12)."""
    J1, D1, J2, D2 = 1.0, 0.8, 0.4, 0.3
    hi = np.random.default_rng(N).uniform(-1, 1, N)
    fb = ("build_diag", "build_off_diag", "build_disorder")
    for kernel in ([1, 2, 3] if N <= 32 else [1, 3]):
        for mag in (N // 2, -(N // 2)):
            say("closed form: polarised", lat, N, mag, kernel)
            h = make(lat, N, mag, J1, D1, J2, D2, lx, ly, builds=fb, hi=hi)
            h.set_hv_kernel(kernel)
            assert h.hamilt_size() == 1
            e = polarised_energy(lat, N, J1, D1, J2, D2, hi, up=mag > 0)
            y = h.matvec(torch.ones(1, dtype=torch.float64, device="cuda")).item()
            assert abs(y - e) <= 1e-14 * max(1.0, abs(e)), (y, e)
            assert abs(h.diagonal().item() - e) <= 1e-14 * max(1.0, abs(e))
        ref = magnon_levels(lat, N, lx, ly, J1, D1, J2, D2)
        for mag in (1 - N // 2, N // 2 - 1):
            say("closed form: magnon", lat, N, mag, kernel)
            h = make(lat, N, mag, J1, D1, J2, D2, lx, ly)
            h.set_hv_kernel(kernel)
            assert h.hamilt_size() == N
            eye = torch.eye(N, dtype=torch.float64, device="cuda")
            H = torch.stack([h.matvec(eye[i].contiguous()) for i in range(N)], dim=1).cpu().numpy()
            assert np.array_equal(H, H.T), (lat, N, mag, kernel)
            np.testing.assert_allclose(np.linalg.eigvalsh(H), ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("lat,N,lx,ly,J2", [("square", 24, 4, 6, 0.5), ("chain", 22, 0, 0, 0.2),
                                             ("honeycomb", 24, 4, 6, 0.4)])
@pytest.mark.parametrize("kernel", [1, 2, 3], ids=["generic", "split", "split_jit"])
def test_hv_config_sizes_matfree(O, torch, lat, N, lx, ly, J2, kernel):
    """Full-size H.v of the 4x6 J1-J2 lattice, the N=22 chain and the 4x6
    honeycomb (24 sites, D = 2,704,156) against the oracle's matrix-free
    product (same row semantics, exact binomials)."""
    say("hv config", lat, N, kernel)
    h = make(lat, N, 0, 1.0, 1.0, J2, J2, lx, ly)
    h.set_hv_kernel(kernel)
    m = oracle_model(O, lat, N, 0, 1.0, 1.0, J2, J2, lx, ly)
    D = h.hamilt_size()
    x = np.random.default_rng(5).uniform(-1, 1, D)
    y = h.matvec(torch.from_numpy(x).cuda()).cpu().numpy()
    yref = m.hv_matfree(x)
    assert np.max(np.abs(y - yref)) <= 1e-13 * np.abs(yref).max()


@pytest.mark.parametrize("kernel", [1, 2, 3], ids=["generic", "split", "split_jit"])
def test_hv_4x8_properties_and_rows(O, torch, kernel):
    """4x8 J1-J2 (D = 601,080,390): rows sampled across the sector against
    the oracle's matrix-free rows, symmetry <x, H y> = <H x, y>, linearity."""
    say("hv 4x8", kernel)
    h = make("square", 32, 0, 1.0, 1.0, 0.5, 0.5, 4, 8)
    h.set_hv_kernel(kernel)
    m = oracle_model(O, "square", 32, 0, 1.0, 1.0, 0.5, 0.5, 4, 8)
    D = h.hamilt_size()
    assert D == 601080390
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5
    y = h.matvec(x)
    xh = x.cpu().numpy()
    yh = y.cpu().numpy()
    for row0 in (0, 123456789, D // 2, D - 5000):
        yref = m.hv_matfree(xh, row0=row0, nrows=5000)
        assert np.max(np.abs(yh[row0:row0 + 5000] - yref)) <= 1e-12 * np.abs(yref).max(), row0
    del xh, yh
    z = torch.rand(D, dtype=torch.float64, device="cuda", generator=g) - 0.5
    hz = h.matvec(z)
    a = torch.dot(z, y).item()
    b = torch.dot(hz, x).item()
    assert abs(a - b) <= 1e-10 * (abs(a) + abs(b))
    w = h.matvec(2.0 * x - 3.0 * z)
    assert torch.max(torch.abs(w - (2.0 * y - 3.0 * hz))).item() <= 1e-11 * torch.max(torch.abs(w)).item()


def formula_vector(O, torch, D, chunk=1 << 27):
    """x(c) of the oracle's formula vector, built on the GPU in chunks (exact:
    integer arithmetic, then a power-of-two scale)."""
    (p1, k1), (p2, k2) = O.FORMULA_P
    x = torch.empty(D, dtype=torch.float64, device="cuda")
    for s in range(0, D, chunk):
        c = torch.arange(s, min(D, s + chunk), dtype=torch.int64, device="cuda")
        n = (c % p1) * k1 % p1 + (c % p2) * k2 % p2
        x[s:s + c.numel()] = n.double() * (1.0 / 1048576.0) - 1.0
        del c, n
    return x


@pytest.mark.parametrize("kernel", [1, 3], ids=["generic", "split_jit"])
def test_hv_honeycomb32_fields_rows_formula_vector(O, torch, kernel):
    """4x8 honeycomb (N = 32) with J2 and random fields in Sz = 0 and Sz = 1
    (D = 601,080,390 and 565,722,720): rows of H x sampled across each
    sector against the oracle, for the oracle's formula vector."""
    builds = ("build_diag", "build_off_diag", "build_disorder")
    hi = np.random.default_rng(32).uniform(-2, 2, 32)
    for mag in (0, 1):
        h = make("honeycomb", 32, mag, 1.0, 0.7, 0.4, 0.3, 4, 8, builds=builds, hi=hi)
        h.set_hv_kernel(kernel)
        D = h.hamilt_size()
        assert D == math.comb(32, 16 + mag)
        say("hv honeycomb 4x8 Sz", mag, "kernel", h.hv_kernel())
        x = formula_vector(O, torch, D)
        y = h.matvec(x)
        m = oracle_model(O, "honeycomb", 32, mag, 1.0, 0.7, 0.4, 0.3, 4, 8)
        for row0 in (0, 98765432, D // 2, D - 3000):
            yref = m.hv_rows_formula(row0, 3000, hi=hi)
            got = y[row0:row0 + 3000].cpu().numpy()
            assert np.max(np.abs(got - yref)) <= 1e-12 * np.abs(yref).max(), (mag, row0)
        del x, y
        torch.cuda.empty_cache()


@pytest.mark.parametrize("kernel", [1, 3], ids=["generic", "split_jit"])
def test_hv_6x6_rows_formula_vector(O, torch, kernel):
    """6x6 J1-J2 (N = 36, D = 9,075,135,300; 72.6 GB per vector): rows of
    H x sampled across the sector (64-bit ranks, the 2^32 boundary included)
    against the oracle, for x given by a formula of the rank that the oracle
    evaluates itself; the generic kernel and the specialised row split with
    64-bit row bases."""
    import time
    h = make("square", 36, 0, 1.0, 1.0, 0.5, 0.5, 6, 6)
    h.set_hv_kernel(kernel)
    D = h.hamilt_size()
    assert D == 9075135300
    say("hv 6x6: kernel", h.hv_kernel())
    x = formula_vector(O, torch, D)
    xs = x[[0, 1, 2**32 + 7, D - 1]].cpu().numpy()
    assert np.array_equal(xs, O.formula_x(np.array([0, 1, 2**32 + 7, D - 1])))
    y = torch.empty_like(x)
    h.matvec(x, y)
    torch.cuda.synchronize()
    t = time.perf_counter()
    h.matvec(x, y)
    torch.cuda.synchronize()
    say("hv 6x6 ms", round((time.perf_counter() - t) * 1e3, 1))
    m = oracle_model(O, "square", 36, 0, 1.0, 1.0, 0.5, 0.5, 6, 6)
    for row0 in (0, 2**32 - 1000, D // 2, D - 2000):
        yref = m.hv_rows_formula(row0, 2000)
        got = y[row0:row0 + 2000].cpu().numpy()
        assert np.max(np.abs(got - yref)) <= 1e-12 * np.abs(yref).max(), row0
    del x, y
    torch.cuda.empty_cache()

"""CPU checks of the DSF post-processing of the Python host mirror
(lsquantumed_amd.phys: read_dsf, site_coords, q_grid, dsf_weights,
multiplet_weights), which restates the q transform of
postprocessing/dynamicStep0_preprocess.ipynb cells 3-4, 9, 12-13 on the
files Phys::DynStructFactor writes.  Pinned by the reference's DSF fixture
(tests/unit_tests/reference_data/dsf_d0_1.ref), by closed forms of the
lattice Fourier sum, and by the oracle's
dsf_weight / multiplets on the same amplitudes."""
import os

import numpy as np

from test_oracle import _parse_dsf

from conftest import GOLDEN

REF = os.path.join(GOLDEN, "reference_data")


def test_read_dsf_matches_the_fixture():
    from lsquantumed_amd import phys
    for name in ("dsf_d0_1.ref", "dsf_d0_2.ref"):
        path = os.path.join(REF, name)
        de, amp, gs = phys.read_dsf(path)
        gs_ref, blocks = _parse_dsf(open(path).read())
        assert gs == gs_ref
        assert amp.shape == (10, 6) and de.shape == (10,)
        assert np.array_equal(de, [e for e, _ in blocks[0]])
        assert np.array_equal(amp, np.array([[blocks[i][n][1] for i in range(6)] for n in range(10)]))


def test_weights_closed_forms():
    """One site: W = 1/N at every q.  Uniform: N at q = 0, 0 elsewhere.
    Staggered (-1)^(ix+iy) on a 2x4 lattice: N at q = (pi, pi) only, which
    is index 6 = 1*ly + 2 (a transposed lx/ly would move it)."""
    from lsquantumed_amd import phys
    lx, ly = 2, 4
    N = lx * ly
    one = np.zeros((1, N))
    one[0, 5] = 1.0
    np.testing.assert_allclose(phys.dsf_weights(one, lx, ly)[:, 0], np.full(N, 1.0 / N), rtol=0, atol=1e-15)
    uni = phys.dsf_weights(np.ones((1, N)), lx, ly)[:, 0]
    assert abs(uni[0] - N) < 1e-12 and np.max(np.abs(uni[1:])) < 1e-12
    r = phys.site_coords(lx, ly)
    stag = phys.dsf_weights(((-1.0) ** (r[:, 0] + r[:, 1]))[None, :], lx, ly)[:, 0]
    assert abs(stag[6] - N) < 1e-12 and np.max(np.abs(np.delete(stag, 6))) < 1e-12
    q = phys.q_grid(lx, ly)
    assert np.allclose(q[6], (np.pi, np.pi))


def test_weights_and_multiplets_vs_oracle(O):
    """Same amplitudes (the fixture's), every q of the 6-site chain: the
    mirror's W equals the oracle's dsf_weight, and the multiplet sums use the
    oracle's grouping of the energies."""
    from lsquantumed_amd import phys
    de, amp, _ = phys.read_dsf(os.path.join(REF, "dsf_d0_1.ref"))
    W = phys.dsf_weights(amp, 6, 1)
    coords = np.array([[i, 0] for i in range(6)], float)
    for k, q in enumerate(phys.q_grid(6, 1)):
        for n in range(10):
            assert abs(W[k, n] - O.dsf_weight(amp[n], q, coords)) < 1e-14
    levels, Wl = phys.multiplet_weights(de, W)
    groups = O.multiplets(de)
    assert len(levels) == len(groups)
    for j, g in enumerate(groups):
        assert abs(levels[j] - de[g[0]]) == 0.0
        np.testing.assert_allclose(Wl[:, j], W[:, g].sum(axis=1), rtol=0, atol=1e-15)

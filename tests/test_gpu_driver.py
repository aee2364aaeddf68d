"""GPU tests of the host-driver features around the step, through the C ABI:
strand re-boxing (fix_diffusion), a broken bond reported as an error
instead of a crash, and the reference's default of re-boxing during runs.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def test_fix_diffusion_reboxes_strands(hip):
    """Strands whose centre of mass lies outside the box come back by whole
    box lengths (src/Backends/SimBackend.cpp:788-882); the energy does not
    change (nearest images are invariant)."""
    from oxdna_amd import systems
    s = systems.duplex_gas(1, n_duplex=8, bp=12, density=1e-3)
    L = s.box
    t = s.copy()
    moves = {0: np.array([2, -1, 0]), 3: np.array([0, 0, 5]), 9: np.array([-3, 0, 1])}
    for strand, m in moves.items():
        t.pos[t.strand == strand] += m * L
    with hip.HIPBatch(t, seed=3) as b:
        E0 = b.energies()[0, 0]
        n = b.fix_diffusion()
        st = b.state()
        b.compute_forces()
        E1 = b.energies()[0, 0]
    assert n == len(moves)
    shift = (st["pos"] - t.pos) / L
    assert np.allclose(shift, np.rint(shift), atol=1e-9)
    for strand in np.unique(t.strand):
        sel = t.strand == strand
        com = st["pos"][sel].mean(axis=0)
        assert np.all(com >= 0) and np.all(com < L), (strand, com)
        # whole strands move together
        assert np.all(np.abs(shift[sel] - shift[sel][0]) < 1e-9)
    assert abs(E1 - E0) <= 1e-6 * abs(E0)


def test_broken_bond_is_reported(hip):
    """A bond stretched beyond the FENE range is an error of that system
    (the kernel's own check), not a crash."""
    from oxdna_amd import _lib, systems
    s = systems.duplex_20bp(0)
    s.pos[5] += np.array([0.0, 0.0, 1.5]) * s.a3[5] + 1.0 * s.a1[5]
    with hip.HIPBatch(s, thermostat="langevin", diff_coeff=2.5, seed=1) as b:
        with pytest.raises(_lib.OxhError, match="FENE"):
            b.run(10)
        nr, err = b.info(0)
        assert err & 1

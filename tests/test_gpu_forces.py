"""GPU parity of the oxDNA2 force kernel (through the C ABI) against the
reference's FP64 CPU backend (fixtures written by the reference binary)
and the FP64 oracle.

Tolerance (FP32 forces from FP64 positions): max |dF_i| over particles
<= 1e-5 of the largest force in the system, for forces and torques, and
the per-particle error (each particle's |dF_i| over its own |F_i|, with a
floor at 0.1 x RMS) <= 1e-5 as well.  Energies: each term within 1e-5 x
sum|E_terms| + 1e-6 N.  The base-pair count (H-bond energy below -0.1)
equals the reference's hb_list count.
"""
import numpy as np
import pytest

from _common import check_pair_set, golden_names, load_golden, oracle_params, vec_rel_err

pytestmark = pytest.mark.gpu

F_TOL = 1e-5
F_TOL_PER_PARTICLE = 1e-5


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


@pytest.mark.parametrize("name", golden_names())
def test_forces_vs_reference(hip, name):
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, **opts) as b:
        F, T, Ep = b.forces()
        e8 = b.energy_split()
        U = b.energies()[0, 0]
        hb = int(b.hb_count()[0])
    assert hb == ref["hb_count"], (hb, ref["hb_count"])
    ef_big, ef_pp = vec_rel_err(F, ref["F"])
    et_big, et_pp = vec_rel_err(T, ref["T"])
    print(f"{name}: N={s.N} F {ef_big:.2e}/{ef_pp:.2e} T {et_big:.2e}/{et_pp:.2e}")
    assert ef_big <= F_TOL and et_big <= F_TOL
    assert ef_pp <= F_TOL_PER_PARTICLE and et_pp <= F_TOL_PER_PARTICLE
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    assert np.all(np.abs(e8 - ref["E8"]) <= 1e-5 * scale), (e8, ref["E8"])
    assert abs(U - ref["E8"].sum()) <= 1e-5 * scale
    assert abs(Ep.sum() - U) <= 1e-5 * scale


@pytest.mark.parametrize("name", golden_names())
def test_stress_vs_reference(hip, name):
    """Pair virial against the reference's pressure observable (stress
    tensor x volume; zero velocities in the fixtures).  FP32 pair forces:
    each pair term carries ~1e-6 of the largest force, so the bound is
    1e-6 x N x max|F| (|r| ~ 1 for interacting pairs)."""
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, **opts) as b:
        W, _ = b.stress()
    want = ref["stress7"][1:] * np.prod(s.box)
    scale = s.N * np.abs(ref["F"]).max()
    err = np.abs(W - want).max() / scale
    print(f"{name}: virial err {err:.2e} (|W| {np.abs(want).max():.3g}, scale {scale:.3g})")
    assert err <= 1e-6


@pytest.mark.parametrize("name", ["duplex20_md_20C", "bundle_md_30C", "gas8_25C_0.5M", "tangle2_45C_0.1M"])
def test_pair_set_exact(hip, name):
    """The GPU Verlet pair set equals the FP64 predicate's, bit for bit."""
    from oracle import pyoracle as po
    s, opts, _ = load_golden(name)
    P = oracle_params(s, opts)
    with hip.HIPBatch(s, verlet_skin=0.05, **opts) as b:
        got = b.pairs()
        far_back = b.stats()["far_back"] > 0
    check_pair_set(got, s, P, 0.05, far_back, po)


@pytest.mark.parametrize("kind", ["bundle", "gas"])
def test_pair_set_exact_cells(hip, kind):
    """Cell lists (systems above 1,024 nt), with per-particle culling of
    cells out of range: half-size cells and a 125-cell stencil in a dense
    bundle (0.9 nt per unit volume), full-size cells and 27 in a dilute gas
    (half-size cells there would hold under 2 nt on average); the pair set
    equals the brute-force FP64 predicate's."""
    from oracle import pyoracle as po
    from oxdna_amd import systems
    if kind == "bundle":
        s = systems.duplex_bundle(2, nx=6, ny=6, bp=32, spacing=2.6, lz=14.0, tc=30.0, salt=0.2)
    else:
        s = systems.duplex_gas(4, n_duplex=64, bp=12, density=2e-3)
    assert s.N > 1024
    P = po.Params(T=s.T, salt=s.salt)
    with hip.HIPBatch(s, verlet_skin=0.05, thermostat="no", salt_concentration=s.salt) as b:
        got = b.pairs()
        far_back = b.stats()["far_back"] > 0
    n_want, n_got = check_pair_set(got, s, P, 0.05, far_back, po)
    print(f"{kind}: N={s.N}, {n_want} pairs, {n_got} kept (far pairs by backbone sites: {far_back})")


@pytest.mark.parametrize("kind", ["bundle", "gas"])
def test_pair_set_exact_cells_full(hip, monkeypatch, kind):
    """OXH_FAR_BACK=0: every far pair within r_verlet, the reference's
    list, equal to the brute-force FP64 predicate's."""
    from oracle import pyoracle as po
    from oxdna_amd import systems
    monkeypatch.setenv("OXH_FAR_BACK", "0")
    if kind == "bundle":
        s = systems.duplex_bundle(2, nx=6, ny=6, bp=32, spacing=2.6, lz=14.0, tc=30.0, salt=0.2)
    else:
        s = systems.duplex_gas(4, n_duplex=64, bp=12, density=2e-3)
    P = po.Params(T=s.T, salt=s.salt)
    with hip.HIPBatch(s, verlet_skin=0.05, thermostat="no", salt_concentration=s.salt) as b:
        got = b.pairs()
        assert b.stats()["far_back"] == 0
    check_pair_set(got, s, P, 0.05, False, po)


@pytest.mark.parametrize("salt", [0.2, 1.0])
def test_far_pairs_by_backbone_sites_over_a_run(hip, monkeypatch, salt):
    """Far pairs kept by backbone sites through a Langevin run with many
    rebuilds on a dense bundle (half-size cells), at a salt where the
    Debye-Hueckel reach exceeds that of close pairs (0.2 M) and at one where
    close pairs reach farther (1 M: their sites' distance must not drop
    them): after 3,000 steps the
    pair set follows the filter's rule for the current state (a rebuild
    is forced first), the forces of that state equal the FP64 oracle's at
    the bars of evolved states (1e-5 of the largest, 2e-5 per particle, as
    in the full-size runs), and equal those of a batch with every far pair
    (OXH_FAR_BACK=0) on the same state to FP32 rounding."""
    from oracle import pyoracle as po
    from oxdna_amd import systems
    s = systems.duplex_bundle(3, nx=6, ny=6, bp=32, spacing=2.6, lz=14.0, tc=30.0, salt=salt)
    P = po.Params(T=s.T, salt=s.salt)
    with hip.HIPBatch(s, verlet_skin=0.05, thermostat="langevin", diff_coeff=2.5, dt=0.003, seed=5,
                      salt_concentration=s.salt) as b:
        assert b.stats()["far_back"] == 1
        b.run(3000)
        n_reb = b.info()[0]
        st = b.state()
        t = s.copy()
        t.pos, t.a1, t.a3 = st["pos"], st["a1"], st["a3"]
        F, T, _ = b.forces()
        b.update_lists(True)
        n_want, n_got = check_pair_set(b.pairs(), t, P, 0.05, True, po)
    ref = po.forces(P, t)
    fb, fp = vec_rel_err(F, ref["F"])
    tb, tp = vec_rel_err(T, ref["T"])
    print(f"{n_reb} rebuilds in 3000 steps; {n_got} of {n_want} pairs kept; F {fb:.2e}/{fp:.2e} T {tb:.2e}/{tp:.2e}")
    assert n_reb >= 20 and n_got < (0.8 if salt < 0.5 else 1.0) * n_want
    assert max(fb, tb) <= 1e-5 and max(fp, tp) <= 2e-5
    monkeypatch.setenv("OXH_FAR_BACK", "0")
    with hip.HIPBatch(t, verlet_skin=0.05, thermostat="no", salt_concentration=s.salt) as c:
        assert c.stats()["far_back"] == 0
        F0, T0, _ = c.forces()
    assert vec_rel_err(F, F0)[0] <= 2e-6 and vec_rel_err(T, T0)[0] <= 2e-6


def test_deterministic(hip):
    s, opts, _ = load_golden("bundle_md_30C")
    outs = []
    for _ in range(2):
        with hip.HIPBatch(s, thermostat="langevin", diff_coeff=2.5, seed=7, **opts) as b:
            b.run(200)
            st = b.state()
            F, T, _ = b.forces()
        outs.append((st["pos"], st["vel"], F, T))
    for a, c in zip(*outs):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("lanes,far_lanes,edge", [(1, 0, 0), (8, 0, 0), (4, 4, 0), (8, 16, 0), (32, 1, 0), (2, 0, 1), (8, 4, 1)])
@pytest.mark.parametrize("name", ["bundle_md_30C", "tangle3_nohalf_dh", "gas8_25C_0.5M"])
def test_launch_shapes(hip, monkeypatch, name, lanes, far_lanes, edge):
    """Every launch shape the autotuner may pick gives the same physics:
    far pairs inline or in their own kernel, any lanes per nucleotide, the
    close pairs in full-list or half-list form (grid engine)."""
    monkeypatch.setenv("OXH_PERSIST", "0")
    monkeypatch.setenv("OXH_EDGE", str(edge))
    monkeypatch.setenv("OXH_FORCE_LANES", str(lanes))
    monkeypatch.setenv("OXH_FAR_LANES", str(far_lanes))
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, **opts) as b:
        b.compute_forces()
        F, T, _ = b.forces()
        e8 = b.energy_split()
        st = b.stats()
    assert st["lanes"] == lanes and st["far_lanes"] == far_lanes
    assert st["engine"] == ("grid-edge" if edge else "grid")
    assert vec_rel_err(F, ref["F"])[0] <= F_TOL
    assert vec_rel_err(T, ref["T"])[0] <= F_TOL
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    assert np.all(np.abs(e8 - ref["E8"]) <= 1e-5 * scale)


@pytest.mark.parametrize("lanes,far_lanes", [(2, 0), (4, 16), (8, 0), (32, 4)])
def test_packed_form_bitwise(hip, monkeypatch, lanes, far_lanes):
    """The packed form of the close kernel (HB, CRST and CXST of a
    workgroup's close pairs evaluated from LDS work lists) gives the same
    bits as the in-place form: forces, torques, per-particle energies, the
    energy split, the stress tensor and the HB count, on a batch of
    fixtures of different temperatures, salts and Debye-Hueckel options
    (items carry their system), and the same bits for each system as in a
    batch of its own.  (With one lane per nucleotide the packed form is not
    used: OXH_PACK=1 leaves the in-place form.)"""
    monkeypatch.setenv("OXH_PERSIST", "0")
    monkeypatch.setenv("OXH_EDGE", "0")
    monkeypatch.setenv("OXH_FORCE_LANES", str(lanes))
    monkeypatch.setenv("OXH_FAR_LANES", str(far_lanes))
    names = ["bundle_md_30C", "tangle3_nohalf_dh", "gas8_25C_0.5M", "duplex20_md_20C"]
    loaded = [load_golden(n) for n in names]
    systems = [x[0] for x in loaded]
    options = [x[1] for x in loaded]

    def run(pack, idx):
        monkeypatch.setenv("OXH_PACK", str(pack))
        with hip.HIPBatch([systems[k] for k in idx], options=[options[k] for k in idx]) as b:
            assert b.stats()["pack"] == bool(pack) and b.stats()["lanes"] == lanes
            b.compute_forces()
            out = []
            hb = b.hb_count()
            for m in range(len(idx)):
                F, T, Ep = b.forces(m)
                w, _ = b.stress(m)
                out.append((F, T, Ep, b.energy_split(m), w, hb[m]))
        return out

    allidx = list(range(len(names)))
    ref = run(0, allidx)
    got = run(1, allidx)
    for m, n in enumerate(names):
        for a, c in zip(ref[m], got[m]):
            assert np.array_equal(np.asarray(a), np.asarray(c)), n
    alone = run(1, [2])[0]
    for a, c in zip(alone, got[2]):
        assert np.array_equal(np.asarray(a), np.asarray(c))


@pytest.mark.parametrize("lanes,group,far_lanes", [(1, 4, 0), (4, 16, 0), (16, 64, 0), (2, 8, 4)])
@pytest.mark.parametrize("name", golden_names())
def test_edge_forces_vs_reference(hip, monkeypatch, name, lanes, group, far_lanes):
    """Half-list form of the step's forces (each close pair and bond
    evaluated once, one per lane, `group` particles per wave; the partner's
    torque from t_q = -t_p - r x F_q; partner shares gathered in a fixed
    order with `lanes` lanes per particle; far pairs in the gather or in
    their own kernel): the same bar as the full list, and the same
    per-particle energies."""
    monkeypatch.setenv("OXH_PERSIST", "0")
    monkeypatch.setenv("OXH_EDGE", "1")
    monkeypatch.setenv("OXH_EDGE_GROUP", str(group))
    monkeypatch.setenv("OXH_FORCE_LANES", str(lanes))
    monkeypatch.setenv("OXH_FAR_LANES", str(far_lanes))
    s, opts, ref = load_golden(name)
    with hip.HIPBatch(s, **opts) as b:
        assert b.stats()["engine"] == "grid-edge"
        b.compute_forces()
        F, T, Ep = b.forces()
    monkeypatch.setenv("OXH_EDGE", "0")
    with hip.HIPBatch(s, **opts) as b:
        assert b.stats()["engine"] == "grid"
        _, _, Ep_full = b.forces()
    ef_big, ef_pp = vec_rel_err(F, ref["F"])
    et_big, et_pp = vec_rel_err(T, ref["T"])
    print(f"{name} K={lanes} P={group} far={far_lanes}: edge F {ef_big:.2e}/{ef_pp:.2e} T {et_big:.2e}/{et_pp:.2e}")
    assert ef_big <= F_TOL and et_big <= F_TOL
    assert ef_pp <= F_TOL_PER_PARTICLE and et_pp <= F_TOL_PER_PARTICLE
    scale = np.abs(ref["E8"]).sum() + 1e-6 * s.N
    assert abs(Ep.sum() - ref["E8"].sum()) <= 1e-5 * scale
    assert np.abs(Ep - Ep_full).max() <= 1e-5 * (np.abs(Ep_full).max() + 1e-6)


@pytest.mark.parametrize("kind", ["bundle", "bundle_batch", "gas"])
def test_cell_centric_build_same_rows(hip, monkeypatch, kind):
    """The cell-centric list build (one workgroup per home cell, the
    stencil's candidates staged in LDS once for all the cell's particles)
    writes the same rows as the per-particle build, entry for entry and in
    the same order: the same pair set, and forces bitwise equal (their sums
    follow row order), also after 200 Langevin steps with rebuilds."""
    from oxdna_amd import systems
    monkeypatch.setenv("OXH_PERSIST", "0")
    monkeypatch.setenv("OXH_EDGE", "0")
    monkeypatch.setenv("OXH_FORCE_LANES", "4")
    monkeypatch.setenv("OXH_FAR_LANES", "16")
    b0 = systems.duplex_bundle(4, nx=8, ny=10, bp=36, lz=36 * systems.BASE_BASE + 1.0)
    if kind == "bundle":
        ss = [b0]
    elif kind == "bundle_batch":
        ss = [b0, systems.duplex_bundle(7, nx=8, ny=10, bp=36, lz=36 * systems.BASE_BASE + 1.0, tc=45.0)]
    else:
        ss = [systems.duplex_gas(5, n_duplex=64, bp=12, density=2e-3)]
    out = {}
    for cb in ("0", "1"):
        monkeypatch.setenv("OXH_CELL_BUILD", cb)
        with hip.HIPBatch(ss, thermostat="langevin", diff_coeff=2.5, seed=3) as b:
            b.compute_forces()
            first = [b.forces(k) for k in range(len(ss))]
            pairs = [b.pairs(k) for k in range(len(ss))]
            b.run(200)
            nreb, err = b.info()
            assert err == 0 and nreb >= 2
            last = [b.forces(k) for k in range(len(ss))]
        out[cb] = (first, pairs, last)
    for k in range(len(ss)):
        assert np.array_equal(out["0"][1][k], out["1"][1][k])
        for a, c in zip(out["0"][0][k] + out["0"][2][k], out["1"][0][k] + out["1"][2][k]):
            assert np.array_equal(a, c)


def test_edge_run_with_rebuilds(hip, monkeypatch):
    """A 1,536-nt duplex gas (cell lists) run 400 Langevin steps in the
    half-list form, with list rebuilds and Hilbert re-sorts in between (the
    edge rows are rebuilt with the lists): forces of the final state match
    the oracle, and the run matches the full-list run's energy to the
    thermal noise."""
    from oracle import pyoracle as po
    from oxdna_amd import systems
    monkeypatch.setenv("OXH_PERSIST", "0")
    s = systems.duplex_gas(3, n_duplex=64, bp=12, density=2e-3)
    out = {}
    for edge in (1, 0):
        monkeypatch.setenv("OXH_EDGE", str(edge))
        with hip.HIPBatch(s, thermostat="langevin", diff_coeff=2.5, dt=0.003, verlet_skin=0.05, seed=9) as b:
            assert b.stats()["engine"] == ("grid-edge" if edge else "grid")
            b.set_sort_interval(100)
            b.run(400)
            st = b.state()
            F, T, Ep = b.forces()
            nreb, err = b.info()
        out[edge] = (st, F, T, Ep, nreb, err)
    st, F, T, Ep, nreb, err = out[1]
    assert err == 0 and nreb >= 5
    t = s.copy()
    t.pos, t.a1, t.a3 = st["pos"], st["a1"], st["a3"]
    ref = po.forces(po.Params(T=s.T, salt=s.salt), t)
    ef, et = vec_rel_err(F, ref["F"])[0], vec_rel_err(T, ref["T"])[0]
    print(f"edge run: {nreb} rebuilds, F {ef:.2e} T {et:.2e}")
    assert ef <= F_TOL and et <= F_TOL
    assert abs(Ep.sum() - ref["U"]) <= 1e-5 * np.abs(ref["Epart"]).sum()
    # both forms sample the same ensemble: U/N within a few thermal widths
    u_edge, u_full = Ep.sum() / s.N, out[0][3].sum() / s.N
    assert abs(u_edge - u_full) <= 0.05, (u_edge, u_full)

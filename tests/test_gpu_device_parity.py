"""Device-side parity of the pieces of the path that are not forces, through
the C ABI:

  - the thermostats' counter generator (Philox4x32-10 on the device)
    against the Random123 known-answer vectors and the oracle on random
    inputs;
  - the Hilbert key the sort uses, on the device, against the oracle's key
    over whole lattices (and against the defining properties of a Hilbert
    curve directly) and on random cells at the sort's depth;
  - the particle order after oxh_batch_hilbert_sort: the stable order of
    the particles' Hilbert keys (system index in the high bits), computed
    here on the host from the FP64 positions with the same cell mapping
    (reference: src/CUDA/CUDA_sort.cu:106-193 and thrust::sort_by_key,
    src/CUDA/Backends/CUDABaseBackend.cu:277).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

# Random123 known-answer vectors for philox4x32-10 (counter, key, output)
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.fixture(scope="module")
def hip():
    from oxdna_amd import _lib, backend
    assert _lib.device_count() >= 1, "no GPU visible"
    return backend


def test_philox_known_answers_on_device(hip):
    ctr = np.array([c for c, _, _ in PHILOX_KAT], np.uint32)
    key = np.array([k for _, k, _ in PHILOX_KAT], np.uint32)
    want = np.array([w for _, _, w in PHILOX_KAT], np.uint32)
    got = hip.philox(ctr, key)
    assert np.array_equal(got, want), [[hex(x) for x in r] for r in got]


def test_philox_random_inputs_match_oracle(hip):
    rng = np.random.default_rng(5)
    ctr = rng.integers(0, 2 ** 32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, size=(4096, 2), dtype=np.uint64).astype(np.uint32)
    got = hip.philox(ctr, key)
    want = np.array([po.philox(c, k) for c, k in zip(ctr, key)], np.uint32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bits", [1, 2, 3, 4])
def test_hilbert_keys_whole_lattice(hip, bits):
    n = 1 << bits
    xyz = np.array([(x, y, z) for x in range(n) for y in range(n) for z in range(n)], np.uint32)
    got = hip.hilbert_keys(xyz, bits)
    want = np.array([po.hilbert_key(x, y, z, bits) for x, y, z in xyz], np.uint32)
    assert np.array_equal(got, want)
    # and the curve itself: a bijection onto [0, n^3) whose consecutive
    # cells share a face
    assert np.array_equal(np.sort(got), np.arange(n ** 3))
    path = xyz[np.argsort(got)].astype(np.int64)
    assert np.all(np.abs(np.diff(path, axis=0)).sum(1) == 1)


@pytest.mark.parametrize("bits", [8, 10])
def test_hilbert_keys_random_cells(hip, bits):
    rng = np.random.default_rng(bits)
    xyz = rng.integers(0, 1 << bits, size=(3000, 3)).astype(np.uint32)
    got = hip.hilbert_keys(xyz, bits)
    want = np.array([po.hilbert_key(x, y, z, bits) for x, y, z in xyz], np.uint32)
    assert np.array_equal(got, want)


def _sort_keys(positions, boxes, bits, sys_bits):
    """Host restatement of the sort key: each FP64 coordinate's fraction of
    its box, times 2^bits, truncated (clamped to the last cell); the key is
    the system index above the 3 x bits Hilbert index."""
    keys = []
    for s, (pos, box) in enumerate(zip(positions, boxes)):
        n = 1 << bits
        inv = 1.0 / np.asarray(box, dtype=np.float64)
        sc = pos * inv
        f = sc - np.floor(sc)
        c = np.minimum((f * n).astype(np.int64), n - 1)
        h = np.array([po.hilbert_key(x, y, z, bits) for x, y, z in c], np.uint64)
        keys.append((np.uint64(s) << np.uint64(3 * bits)) | h)
    return keys


def test_hilbert_sort_order(hip):
    """One sort from the upload order, then a second one after 300 steps
    (from the first sort's order): every system's device order is the
    stable argsort of its keys, and the systems keep their own ranges."""
    from oxdna_amd import systems
    ss = [systems.duplex_gas(3, n_duplex=27, bp=24, density=0.01), systems.hairpin_set(4, n=3)[2],
          systems.duplex_bundle(2, nx=3, ny=4, bp=16, spacing=2.6, lz=16 * systems.BASE_BASE + 2.0)]
    sys_bits = 2                                    # three systems
    bits = min(10, (32 - sys_bits) // 3)
    with hip.HIPBatch(ss, thermostat="langevin", diff_coeff=2.5, seed=4) as b:
        assert all(np.array_equal(b.order(k), np.arange(s.N)) for k, s in enumerate(ss))
        keys = _sort_keys([s.pos for s in ss], [s.box for s in ss], bits, sys_bits)
        b.hilbert_sort()
        orders = [b.order(k) for k in range(len(ss))]
        for k, s in enumerate(ss):
            assert np.array_equal(orders[k], np.argsort(keys[k], kind="stable")), k
        b.run(300)
        pos = [b.state(k)["pos"] for k in range(len(ss))]
        # the second sort starts from the first one's device order
        keys2 = _sort_keys([p[o] for p, o in zip(pos, orders)], [s.box for s in ss], bits, sys_bits)
        b.hilbert_sort()
        for k, s in enumerate(ss):
            want = orders[k][np.argsort(keys2[k], kind="stable")]
            assert np.array_equal(b.order(k), want), k

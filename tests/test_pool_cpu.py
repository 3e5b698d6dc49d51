"""Host-side pieces of the job pool: the seeded generator and the exported C ABI."""
import ctypes as C

import numpy as np

from euler_kokkos_amd import native, pool


def test_mt19937_64_known_answer():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed
    # std::mt19937_64 (seed 5489) is 9981545732273789042
    g = pool.MT19937_64(5489)
    x = None
    for _ in range(10000):
        x = g.next_u64()
    assert x == 9981545732273789042


def test_uniform_mapping_is_top_53_bits():
    g1, g2 = pool.MT19937_64(7), pool.MT19937_64(7)
    for _ in range(100):
        u = g1.uniform()
        assert u == (g2.next_u64() >> 11) / 2.0 ** 53
        assert 0.0 <= u < 1.0


def test_random_blast_batch_ranges_and_determinism():
    a = pool.random_blast_batch(40, seed=11)
    b = pool.random_blast_batch(40, seed=11)
    for (pa, Ua, da), (pb, Ub, db) in zip(a, b):
        assert da == db and np.array_equal(Ua, Ub)
        assert da["n"] in range(64, 513, 32) and pa.nx == pa.ny == da["n"]
        assert 0.2 <= da["center"][0] <= 0.8 and 0.05 <= da["radius"] <= 0.15
        assert 10 * 0.99999 <= da["pressure_in"] / 0.1 <= 1000 * 1.00001
        assert pa.bc[0] == 1 and pa.riemann == native.RIEMANN["hllc"]
        # blast: rho = 1 everywhere, E = p/(gamma-1)
        assert np.all(Ua[0] == 1.0)
    assert len({d["n"] for _, _, d in a}) > 5


def test_pool_symbols_exported():
    L = native.lib()
    for s in pool.POOL_EXPORTS:
        assert hasattr(L, s), s


def test_pool_rejects_kinds_it_does_not_run():
    # checked on the host before any device allocation
    import pytest
    from euler_kokkos_amd.native import EkError
    from euler_kokkos_amd.params import make_params
    for kind in (make_params("MHD_Muscl_2D", 32, 32, gamma0="1.6666667", riemann="hlld"),
                 make_params("Hydro_Muscl_3D", 16, 16, 16, riemann="llf"),
                 make_params("MHD_Muscl_3D", 16, 16, 16, gamma0="1.6666667", riemann="hlld", slope_type="0")):
        with pytest.raises(EkError, match="job pool runs|slope_type"):
            pool.Pool(kind, 1 << 20, 4)

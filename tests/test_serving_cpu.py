"""Host side of serving: the seeded Poisson stream and the 3D Kelvin-Helmholtz state."""
import numpy as np

from euler_kokkos_amd import serving


def test_poisson_stream_is_seeded_and_in_range():
    a = serving.poisson_stream(200, rate=20.0, seed=3, edge=(64, 192), build_state=False)
    b = serving.poisson_stream(200, rate=20.0, seed=3, edge=(64, 192), build_state=False)
    assert [(j.arrival, j.n, j.seed) for j in a] == [(j.arrival, j.n, j.seed) for j in b]
    t = np.array([j.arrival for j in a])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert 0.035 < gaps.mean() < 0.065  # mean inter-arrival 1/20 s
    assert min(j.n for j in a) >= 64 and max(j.n for j in a) <= 192
    assert len({j.n for j in a}) > 50


def test_kh3d_state():
    j = serving.poisson_stream(1, rate=20.0, seed=1, edge=(16, 16), build_state=False)[0]
    U = serving.kh3d(j.params, 7)
    p = j.params
    z = p.zmin + p.dz / 2 + np.arange(p.nz) * p.dz
    inner = np.abs(z - 0.5) < 0.25
    assert np.all(U[0][inner] == 2.0) and np.all(U[0][~inner] == 1.0)
    vx = U[2] / U[0]
    assert np.allclose(vx[inner], 0.5) and np.allclose(vx[~inner], -0.5)
    vy = U[3] / U[0]
    x = p.xmin + p.dx / 2 + np.arange(p.nx) * p.dx
    assert np.abs(vy - 0.01 * np.sin(2 * np.pi * x)).max() <= 1e-3 + 1e-12
    pr = (p.gamma0 - 1) * (U[1] - 0.5 * (U[2] ** 2 + U[3] ** 2) / U[0])
    assert np.allclose(pr, 2.5, rtol=1e-12)
    assert np.array_equal(U, serving.kh3d(j.params, 7))

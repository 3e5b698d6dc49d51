"""The smallest grids the solvers accept (ghostWidth interior cells on an axis: 2 for hydro,
3 for MHD) and other grids far below one tile, with every boundary type: the FP64 step is
bitwise the oracle's. Tiles are then almost empty and every halo read goes through the
boundary map, several times over the axis for periodic and reflecting sides. Initial states
are random (seeded) so that every variable varies from cell to cell."""
import numpy as np
import pytest

from euler_kokkos_amd.params import make_params

import test_mhd3d_gpu
import test_mhd_gpu
import test_parity_gpu

pytestmark = pytest.mark.gpu

ID, IP, IU, IV, IW, IA, IB, IC = range(8)


def random_state(oracle, op, mhd, seed):
    rng = np.random.default_rng(seed)
    U = np.zeros(oracle.ref_shape(op))
    shape = U.shape[:-1]
    rho = 1.0 + rng.uniform(0.0, 1.0, shape)
    vel = [rng.uniform(-0.5, 0.5, shape) for _ in range(3)]
    p = 1.0 + rng.uniform(0.0, 1.0, shape)
    U[..., ID] = rho
    ekin = 0.0
    nvel = 3 if (mhd or U.shape[-1] == 5) else 2
    for k, iv in enumerate((IU, IV, IW)[:nvel]):
        U[..., iv] = rho * vel[k]
        ekin = ekin + 0.5 * rho * vel[k] ** 2
    emag = 0.0
    if mhd:
        for iv in (IA, IB, IC):
            U[..., iv] = rng.uniform(-0.5, 0.5, shape)
            emag = emag + 0.5 * U[..., iv] ** 2
    U[..., IP] = p / (op.gamma0 - 1.0) + ekin + emag
    return U


@pytest.mark.parametrize("shape,bc", [((2, 2), (1, 2, 3, 3)), ((2, 5), (3, 3, 1, 2)), ((5, 2), (2, 1, 1, 3)),
                                      ((3, 3), (1, 1, 1, 1)), ((17, 2), (3, 3, 2, 2))])
def test_hydro2d_minimum_grids_bitwise(oracle, shape, bc):
    ekp = make_params("Hydro_Muscl_2D", *shape, riemann="hllc", slope_type="2", cfl="0.5", bc=bc + (3, 3))
    U0 = random_state(oracle, oracle.from_ek(ekp), False, 11)
    g, o = test_parity_gpu.run_pair(oracle, ekp, U0, 6)
    print(f"case=hydro2d {shape} bc={bc}: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("shape,bc", [((2, 2, 2), (3,) * 6), ((3, 2, 5), (1, 2, 3, 3, 2, 1)),
                                      ((2, 4, 3), (1, 1, 2, 2, 3, 3)), ((17, 9, 2), (2, 3, 1, 3, 1, 2))])
def test_hydro3d_minimum_grids_bitwise(oracle, shape, bc):
    ekp = make_params("Hydro_Muscl_3D", *shape, riemann="hllc", slope_type="2", cfl="0.5", bc=bc)
    U0 = random_state(oracle, oracle.from_ek(ekp), False, 12)
    g, o = test_parity_gpu.run_pair(oracle, ekp, U0, 6)
    print(f"case=hydro3d {shape} bc={bc}: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("shape,bc", [((3, 3), (3, 3, 3, 3)), ((3, 5), (1, 2, 3, 3)), ((4, 3), (2, 1, 1, 2)),
                                      ((15, 3), (1, 1, 3, 3))])
def test_mhd2d_minimum_grids_bitwise(oracle, shape, bc):
    ekp = test_mhd_gpu.mparams(*shape, bc=bc + (3, 3))
    R0 = random_state(oracle, oracle.from_ek(ekp), True, 13)
    g, o = test_mhd_gpu.run_pair(oracle, ekp, R0, 6)
    print(f"case=mhd2d {shape} bc={bc}: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("shape,bc", [((3, 3, 3), (3,) * 6), ((3, 4, 5), (1, 2, 3, 3, 2, 1)),
                                      ((5, 3, 3), (2, 2, 1, 1, 3, 3))])
def test_mhd3d_minimum_grids_bitwise(oracle, shape, bc):
    ekp = test_mhd3d_gpu.m3(*shape, bc=bc)
    R0 = random_state(oracle, oracle.from_ek(ekp), True, 14)
    g, o = test_mhd3d_gpu.run_pair(oracle, ekp, R0, 6)
    print(f"case=mhd3d {shape} bc={bc}: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)

"""Long FP64 runs on small grids, bitwise against the oracle: hundreds to thousands of steps,
so that shocks reflect off walls, interact and pass through periodic sides, and the rarely
taken branches (density and pressure floors, every HLLC / HLLD fan region, the two-shock
solver's iterations) are reached. Each run stops at its own step count; t and it must agree
exactly as well."""
import numpy as np
import pytest

from euler_kokkos_amd import native, problems
from euler_kokkos_amd.params import make_params

import test_mhd3d_gpu
import test_mhd_gpu
import test_parity_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rs", ["hllc", "approx", "llf", "hll"])
def test_hydro2d_implode_2000_steps_bitwise(oracle, rs):
    # strong implosion in a reflecting box: the shock focuses in the corner, reflects and
    # interacts with the contact
    ekp = make_params("Hydro_Muscl_2D", 40, 40, riemann=rs, slope_type="2", cfl="0.8", bc=(1,) * 6)
    U0 = oracle.implode(oracle.from_ek(ekp))
    g, o = test_parity_gpu.run_pair(oracle, ekp, U0, 2000)
    print(f"case=hydro2d implode 40^2 rs={rs} 2000 steps: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("rs,bc", [("hllc", (1, 1, 1, 1, 1, 1)), ("hllc", (3, 3, 2, 2, 1, 3)),
                                   ("approx", (1, 2, 3, 3, 1, 1)), ("llf", (3,) * 6)])
def test_hydro3d_blast_600_steps_bitwise(oracle, rs, bc):
    # a strong off-centre blast (pressure ratio 1e4, low background density near the floors)
    ekp = make_params("Hydro_Muscl_3D", 24, 20, 18, riemann=rs, slope_type="2", cfl="0.5", bc=bc)
    U0 = oracle.blast(oracle.from_ek(ekp), radius=0.2, center_x=0.35, center_y=0.55, center_z=0.45,
                      density_in=1.0, density_out=0.01, pressure_in=100.0, pressure_out=0.01)
    g, o = test_parity_gpu.run_pair(oracle, ekp, U0, 600)
    print(f"case=hydro3d blast 24x20x18 rs={rs} bc={bc} 600 steps: max|diff|={np.max(np.abs(g - o)):.3e}",
          flush=True)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("rs", ["hlld", "hll", "llf"])
def test_mhd2d_orszag_tang_1500_steps_bitwise(oracle, rs):
    # Orszag-Tang through the formation and interaction of its shocks
    ekp = test_mhd_gpu.mparams(48, 48, riemann=rs)
    R0 = oracle.orszag_tang(oracle.from_ek(ekp))
    g, o = test_mhd_gpu.run_pair(oracle, ekp, R0, 1500)
    print(f"case=mhd2d OT 48^2 rs={rs} 1500 steps: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)


@pytest.mark.parametrize("bc", [(3,) * 6, (1, 2, 1, 2, 3, 3)])
def test_mhd3d_blast_300_steps_bitwise(oracle, bc):
    # magnetised blast (strong field, low plasma beta outside the sphere)
    ekp = test_mhd3d_gpu.m3(20, 18, 16, bc=bc)
    U = problems.mhd_blast3d(ekp, B0=2.0, radius="0.25", center=("0.45", "0.5", "0.55"))
    op = oracle.from_ek(ekp)
    R0 = np.zeros(oracle.ref_shape(op))
    R0[3:-3, 3:-3, 3:-3, :] = U.transpose(3, 2, 1, 0)
    g, o = test_mhd3d_gpu.run_pair(oracle, ekp, R0, 300)
    print(f"case=mhd3d blast 20x18x16 bc={bc} 300 steps: max|diff|={np.max(np.abs(g - o)):.3e}", flush=True)
    assert np.array_equal(g, o)

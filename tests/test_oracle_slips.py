"""Slips planted in copies of the oracle, on the CPU, and the pins that catch them.

Each case copies one oracle source file, replaces one expression by a plausible
transcription slip (a wrong wave-speed bound, a swapped corner, a flipped sign, a
mismatched index), builds the copy into its own library and runs the closed-form pin of
tests/test_oracle_pins.py on it: the pin must report a misfit far above rounding. This is
the evidence that those pins constrain the functions they pin (the unmodified oracle passes
them to rounding).
"""
import os
import subprocess

import numpy as np
import pytest

import test_oracle_pins as pins
from oracle import pyoracle

ORACLE_DIR = os.path.dirname(os.path.abspath(pyoracle.__file__))
CFLAGS = ["-O1", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-std=c11", "-w"]
HYDRO, MHD = "ek_oracle.c", "ek_oracle_mhd.c"

# (name, file, expression in the oracle, slip, misfit of the pin that must catch it)
SLIPS = [
    ("llf_signal_speed", HYDRO, "double cmax = fmax(fabs(ul) + cl, fabs(ur) + cr);",
     "double cmax = fmax(fabs(ul), fabs(ur)) + fmax(cl, cr);", lambda o: pins.check_hydro_llf(o, 2)),
    ("llf_dissipation_sign", HYDRO,
     "flux[nVar] = HALF * (fleft[nVar] + fright[nVar] - cmax * (uright[nVar] - uleft[nVar]));",
     "flux[nVar] = HALF * (fleft[nVar] + fright[nVar] + cmax * (uright[nVar] - uleft[nVar]));",
     lambda o: pins.check_hydro_llf(o, 3)),
    ("hll_davis_bound", HYDRO, "double SL = fmin(fmin(ul, ur) - fmax(cl, cr), (double)ZERO);",
     "double SL = fmin(fmin(ul - cl, ur - cr), (double)ZERO);", lambda o: pins.check_hydro_hll(o)[0]),
    ("hll_unclamped_bound", HYDRO, "double SR = fmax(fmax(ul, ur) + fmax(cl, cr), (double)ZERO);",
     "double SR = fmax(ul, ur) + fmax(cl, cr);", lambda o: pins.check_hydro_hll(o)[0]),
    ("mhd_hll_davis_bound", MHD, "double sl = fmin(fmin(vl, vr) - fmax(cfl, cfr), 0.0);",
     "double sl = fmin(fmin(vl - cfl, vr - cfr), 0.0);", lambda o: pins.check_mhd_hll_llf(o, "hll")[0]),
    ("mhd_hll_one_sided_bn", MHD, "double bx_mean = 0.5 * (ql[IA] + qr[IA]);", "double bx_mean = ql[IA];",
     lambda o: pins.check_mhd_hll_llf(o, "hll")[0]),
    ("mhd_llf_wrong_axis_speed", MHD, "double vel_info = fmax(sil[0], sir[0]);",
     "double vel_info = fmax(sil[1], sir[1]);", lambda o: pins.check_mhd_hll_llf(o, "llf")[0]),
    ("fast_speed_wrong_component", MHD, "  if (dir == 0)\n    cf = sqrt(d2 + sqrt(d2 * d2 - c2 * a * a / d));",
     "  if (dir == 0)\n    cf = sqrt(d2 + sqrt(d2 * d2 - c2 * b * b / d));",
     lambda o: pins.check_mhd_hll_llf(o, "hll")[0]),
    ("emf_swapped_corners", MHD, "- SAL * SAT * EstarRL - SAR * SAB * EstarLR",
     "- SAL * SAT * EstarLR - SAR * SAB * EstarRL", lambda o: pins.check_emf(o)[0]),
    ("emf_dissipation_sign", MHD, "SAT * SAB / (SAT - SAB) * (AstarT - AstarB)",
     "SAT * SAB / (SAT - SAB) * (AstarB - AstarT)", lambda o: pins.check_emf(o)[0]),
    ("emf_alfven_speed_density", MHD, "fmax5(fabs(bLL) / sqrt(rstarLLy)", "fmax5(fabs(bLL) / sqrt(rstarLLx)",
     lambda o: pins.check_emf(o)[0]),
    ("emf_contact_pressure_sign", MHD, "(PtotLL - PtotLR + PtotRL - PtotRR)", "(PtotLL - PtotLR - PtotRL + PtotRR)",
     lambda o: pins.check_emf(o)[0]),
    ("emf_star_field_average", MHD, "double BstarR = (SAT * BstarRR - SAB * BstarRL) / (SAT - SAB);",
     "double BstarR = (SAT * BstarRL - SAB * BstarRR) / (SAT - SAB);", lambda o: pins.check_emf(o)[0]),
    ("emf_one_sided_branch_field", MHD, "(SAR * EstarLLx - SAL * EstarRLx + SAR * SAL * (bRL - bLL))",
     "(SAR * EstarLLx - SAL * EstarRLx + SAR * SAL * (bLL - bRL))", lambda o: pins.check_emf(o)[0]),
    ("emf_double_star_density", MHD,
     "double rstarLR = rLR * (SL - uLR) / (SL - ustar) * (ST - vLR) / (ST - vstar);",
     "double rstarLR = rLR * (SL - uLR) / (SL - ustar) * (SB - vLR) / (SB - vstar);",
     lambda o: pins.check_emf(o)[0]),
    ("implode_boundary_cells", HYDRO, "tmp = x + y > (xmin + xmax) / 2. + ymin;",
     "tmp = x + y >= (xmin + xmax) / 2. + ymin;", lambda o: pins.check_implode_ic(o)),
    ("blast_wrong_centre_axis", HYDRO, "(z - bp->center_z) * (z - bp->center_z)",
     "(z - bp->center_y) * (z - bp->center_y)",
     lambda o: pins.check_blast_ic(o, center=("0.3", "0.6", "0.45"), radius="0.2", extent=(1.5, 1.0, 0.8))),
]


@pytest.fixture(scope="module")
def objects(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_objects")
    for f in (HYDRO, MHD):
        subprocess.run(["gcc", *CFLAGS, "-I", ORACLE_DIR, "-c", "-o", str(d / (f + ".o")), os.path.join(ORACLE_DIR, f)],
                       check=True)
    return d


def build_variant(objects, tmp_path, fname, old, new):
    src = open(os.path.join(ORACLE_DIR, fname)).read()
    assert src.count(old) == 1, f"expression not found once in {fname}: {old!r}"
    mod = tmp_path / fname
    mod.write_text(src.replace(old, new))
    other = MHD if fname == HYDRO else HYDRO
    so = tmp_path / "libek_oracle_variant.so"
    subprocess.run(["gcc", *CFLAGS, "-I", ORACLE_DIR, "-shared", "-o", str(so), str(mod), str(objects / (other + ".o")),
                    "-lm"], check=True)
    return pyoracle.load(str(so))


@pytest.mark.parametrize("name,fname,old,new,check", SLIPS, ids=[s[0] for s in SLIPS])
def test_pin_catches_planted_slip(objects, tmp_path, name, fname, old, new, check):
    L = build_variant(objects, tmp_path, fname, old, new)
    with pyoracle.using(L):
        misfit = check(pyoracle)
    print(f"slip {name}: misfit {misfit}")
    assert misfit > 1e-9, f"{name} not caught (misfit {misfit})"


def test_unmodified_copy_passes_every_pin(objects, tmp_path):
    # the harness itself: the oracle rebuilt the same way passes the pins to rounding
    L = build_variant(objects, tmp_path, HYDRO, "/* riemann_llf:", "/* riemann_llf:")
    with pyoracle.using(L):
        assert pins.check_hydro_llf(pyoracle, 2) <= 1e-14
        assert pins.check_hydro_hll(pyoracle)[0] <= 1e-14
        assert pins.check_mhd_hll_llf(pyoracle, "hll")[0] <= 1e-13
        assert pins.check_emf(pyoracle)[0] <= 1e-12
        assert pins.check_implode_ic(pyoracle) == 0

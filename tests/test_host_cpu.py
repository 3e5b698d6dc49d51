"""The C++ host layer: ConfigMap parsing rules and HydroParams, through the C entry points
of include/ek_host.h. Expectations come from inih's documented rules as the reference uses
them (src/utils/config/inih/ini.cpp, INIReader.cpp, ConfigMap.cpp) and from
HydroParams' defaults (src/shared/HydroParams.cpp:37-180)."""
import ctypes as C
import os

import pytest

from euler_kokkos_amd import native
from euler_kokkos_amd.params import getfloat, make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def L():
    lib = native.lib()
    lib.ek_ini_get.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    lib.ek_ini_parse_error.argtypes = [C.c_char_p, C.c_int]
    lib.ek_ini_params.argtypes = [C.c_char_p, C.POINTER(native.EkParams)]
    return lib


def get(text, section, name, kind=0):
    b = text.encode()
    out = C.create_string_buffer(256)
    L().ek_ini_get(b, len(b), section.encode(), name.encode(), kind, out, 256)
    return out.value.decode()


INI = """
# comment at line start
; another
[Run]
Solver_Name = Hydro_Muscl_2D   ; inline comment after whitespace
tEnd=0.25;not a comment without whitespace
nStepMax = 0x10
label = first
   continued value
[hydro]
gamma0=1.4
flag = yes
off = off
empty =
"""


def test_keys_are_case_insensitive_and_comments_stripped():
    assert get(INI, "run", "solver_name") == "Hydro_Muscl_2D"
    assert get(INI, "RUN", "SOLVER_NAME") == "Hydro_Muscl_2D"
    assert get(INI, "run", "tend") == "0.25;not a comment without whitespace"
    assert get(INI, "hydro", "missing") == "<unset>"


def test_continuation_line_replaces_value():
    assert get(INI, "run", "label") == "continued value"


def test_typed_getters():
    assert get(INI, "run", "nstepmax", 1) == "16"            # strtol base 0
    assert float(get(INI, "hydro", "gamma0", 2)) == getfloat("1.4") == 1.399999976158142
    assert float(get(INI, "run", "tend", 2)) == getfloat("0.25")  # strtof stops at ';'
    assert get(INI, "hydro", "flag", 3) == "true"
    assert get(INI, "hydro", "off", 3) == "false"
    assert float(get(INI, "hydro", "empty", 2)) == -1.5     # empty -> default


def test_parse_error_line():
    bad = "[run]\nnx=3\nthis line has no equals\n[broken\n"
    b = bad.encode()
    assert L().ek_ini_parse_error(b, len(b)) == 3


def test_hydro_params_from_ini_match_python_mirror():
    p = native.EkParams()
    rc = L().ek_ini_params(os.path.join(ROOT, "settings", "implode_2D_512.ini").encode(), C.byref(p))
    assert rc == 0
    q = make_params("Hydro_Muscl_2D", 512, 512, gamma0="1.4", cfl="0.8", slope_type="2",
                    riemann="hllc")
    for f, _ in native.EkParams._fields_:
        if f in ("bc",):
            assert list(p.bc)[:4] == list(q.bc)[:4]
        elif f == "nz":
            assert p.nz == 1
        else:
            assert getattr(p, f) == getattr(q, f), f

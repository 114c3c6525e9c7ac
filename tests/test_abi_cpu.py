"""CPU-side checks of the C ABI library: it loads without a GPU and exports
every function the public headers declare (no compute calls)."""
import ctypes
import os

import pytest

from bubbles_amd import abi


def test_library_built_and_loads():
    assert os.path.exists(abi.LIB_PATH), "run python bubbles_amd/build.py"
    abi.load()


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    names = abi.header_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_symbols_have_ctypes_signatures():
    abi.load()
    for n in abi.header_symbols():
        f = getattr(abi.load(), n)
        assert f.argtypes is not None, n


def test_no_device_is_reported_cleanly():
    # device_count never fails; creating a solver without a device is an error code
    n = abi.device_count()
    if n == 0:
        from bubbles_amd import scenes

        sc = scenes.dam_break(4, 4, 4, spacing=0.01, tank=(0.2, 0.2, 0.2))
        with pytest.raises(abi.BbxError):
            abi.Solver(sc)


def test_default_params_match_reference_defaults():
    p = abi.Params()
    abi.check(abi.load().bbx_default_params(ctypes.byref(p)))
    # DefaultSphSolverData3 (src/solvers/sph_solver3.cpp:124-151)
    assert p.eos_exponent == 7.0
    assert p.negative_pressure_scale == 0.0
    assert p.viscosity == 0.04
    assert p.pseudo_viscosity == 10.0
    assert p.sound_speed == 100.0
    assert p.drag_coefficient == 0.0001
    assert p.gravity[1] == float(ctypes.c_float(-9.8).value)

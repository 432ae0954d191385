"""The C-ABI library (CPU-side checks): it loads, exports every function that
include/*.h declares, and its host logic rejects bad input before touching a
device.  No kernel runs here."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if fn.endswith(".h"):
            text = open(os.path.join(INCLUDE, fn)).read()
            names.update(re.findall(r"BCSIM_API\s+[\w\s\*]+?\b(bcsim_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_reference_interface():
    names = set(declared_functions())
    # one entry point per IAlgorithm member (src/core/LibBCSim.hpp:42-97)
    for n in ["bcsim_create", "bcsim_set_parameter", "bcsim_get_parameter", "bcsim_clear_fixed_scatterers",
              "bcsim_add_fixed_scatterers", "bcsim_clear_spline_scatterers", "bcsim_add_spline_scatterers",
              "bcsim_set_scan_sequence", "bcsim_set_excitation", "bcsim_set_analytical_profile",
              "bcsim_set_lut_profile", "bcsim_simulate_lines", "bcsim_get_debug_data",
              "bcsim_get_total_num_scatterers", "bcsim_set_logger",
              # set_* and create_simulator aliases of the same entry points
              "bcsim_create_simulator", "bcsim_set_fixed_scatterers", "bcsim_set_spline_scatterers"]:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    from openbcsim_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python signature table covers the header exactly
    assert set(_lib.SIGNATURES) | {"bcsim_last_error"} == set(declared_functions())


def test_create_rejects_unknown_type_without_device():
    from openbcsim_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    rc = lib.bcsim_create(b"cpu", C.byref(h))
    assert rc == 1 and b"Illegal algorithm type: cpu" in lib.bcsim_last_error()
    assert not h.value
    rc = lib.bcsim_set_parameter(None, b"verbose", b"1")
    assert rc == 1 and b"null simulator handle" in lib.bcsim_last_error()


def test_cxx_facade_compiles():
    """include/bcsim/LibBCSim.hpp is a header-only IAlgorithm facade over the C-ABI."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cxx", "facade_check.cpp")
    res = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I", INCLUDE, src], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_c_header_compiles_as_c99_and_links():
    """include/bcsim_c.h is a C header: a C99 caller (tests/cxx/c_caller.c) compiles
    with warnings as errors and links against the built library."""
    import shutil
    import subprocess
    import tempfile
    from openbcsim_amd import _lib
    from openbcsim_amd.build import C_CALLER_FLAGS
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    lib_dir = os.path.dirname(_lib.load()._name)
    src = os.path.join(ROOT, "tests", "cxx", "c_caller.c")
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "c_caller")
        res = subprocess.run(["gcc", *C_CALLER_FLAGS, "-I", INCLUDE, src, "-o", exe, "-L", lib_dir, "-lbcsim_hip",
                              "-lm"], capture_output=True, text=True)
        assert res.returncode == 0, res.stderr
        # the header alone, in every C dialect a binding generator may use
        for std in ("-std=c89", "-std=c11"):
            res = subprocess.run(["gcc", std, "-Wall", "-Werror", "-pedantic", "-fsyntax-only", "-x", "c", "-I",
                                  INCLUDE, "-"], input='#include "bcsim_c.h"\nint main(void) { return 0; }\n',
                                 capture_output=True, text=True)
            assert res.returncode == 0, (std, res.stderr)


def test_ceiling_probes_library_binds():
    """The bench's ceiling probes (libbcsim_ceilings.so) load and export every
    entry point the binding declares; nothing is launched without a device."""
    from openbcsim_amd import ceilings
    if not os.path.exists(ceilings.LIB_PATH):
        pytest.skip("libbcsim_ceilings.so is not built")
    lib = ceilings.load()
    for name in ("ceil_lds_add", "ceil_gather", "ceil_spline", "ceil_valu", "ceil_copy", "ceil_cu_count"):
        assert getattr(lib, name).restype is not None

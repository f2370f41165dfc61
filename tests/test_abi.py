"""The C ABI without a GPU: the library loads, exports every function that
include/m2c_amd.h declares, and the ctypes mirror of the header's structs and
constants has the C compiler's layout and values."""
import ctypes as C
import os
import re
import subprocess

import pytest

from m2c_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "m2c_amd.h")
LIB = os.path.join(ROOT, "m2c_amd", "libm2c_amd.so")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(m2cx_[a-z_0-9]+)\s*\(", txt, re.M)))


def test_header_declares_the_entry_points():
    names = _declared()
    for n in ("m2cx_create", "m2cx_compute_residual", "m2cx_compute_time_step", "m2cx_advance_one_time_step",
              "m2cx_levelset_residual", "m2cx_riemann_batch", "m2cx_numerical_flux_batch"):
        assert n in names
    assert len(names) >= 20


@pytest.mark.skipif(not os.path.exists(LIB), reason="libm2c_amd.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(_declared()) <= exported


def test_argument_errors_need_no_gpu():
    """Null handles are rejected before any device call."""
    if not os.path.exists(LIB):
        pytest.skip("libm2c_amd.so not built")
    lib = C.CDLL(LIB)
    lib.m2cx_run_steps.argtypes = [C.c_void_p, C.c_int]
    assert lib.m2cx_run_steps(None, 1) == abi.E_ARG
    lib.m2cx_destroy.argtypes = [C.c_void_p]
    assert lib.m2cx_destroy(None) == 0
    lib.m2cx_last_error.restype = C.c_char_p
    lib.m2cx_last_error.argtypes = [C.c_void_p]
    assert isinstance(lib.m2cx_last_error(None), bytes)


_PREFIXES = ("EOS_", "BC_", "LSBC_", "FLUX_", "REC_", "LIM_", "SLOPE_", "VAR_", "TI_", "MPFLUX_", "MPREC_", "GUESS_",
             "FIELD_", "E_", "GEOM_", "RIEMANN_", "REINIT_", "RNORMAL_", "PHASECHANGE_", "LSSOLVER_",
             "LSFLUX_", "LS_", "MAX_LEVELSETS")


def _consts():
    return sorted(n for n in dir(abi) if n.startswith(_PREFIXES) or n == "OK")


def _cname(n):
    return abi.C_NAMES.get(n, "M2CX_" + n)


def _c_layout(tmp_path):
    """sizeof/offsetof of the header's structs and the values of its enums,
    as the C compiler sees them."""
    checks = []
    for cname, py in (("m2cx_material", abi.Material), ("m2cx_grid", abi.Grid), ("m2cx_scheme", abi.Scheme),
                      ("m2cx_counters", abi.Counters), ("m2cx_geometry", abi.Geometry),
                      ("m2cx_levelset", abi.Levelset)):
        checks.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            checks.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    for n in _consts():
        checks.append(f'printf("const {n} %ld\\n", (long)({_cname(n)}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"m2c_amd.h\"\nint main(void){\n" +
                   "\n".join(checks) + "\nreturn 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return dict((l.rsplit(" ", 1)[0], int(l.rsplit(" ", 1)[1])) for l in out.splitlines())


def test_ctypes_mirror_matches_header(tmp_path):
    got = _c_layout(tmp_path)
    for cname, py in (("m2cx_material", abi.Material), ("m2cx_grid", abi.Grid), ("m2cx_scheme", abi.Scheme),
                      ("m2cx_counters", abi.Counters), ("m2cx_geometry", abi.Geometry),
                      ("m2cx_levelset", abi.Levelset)):
        assert got[f"{cname} size"] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[f"{cname}.{f}"] == getattr(py, f).offset, (cname, f)
    for n in _consts():
        assert got[f"const {n}"] == getattr(abi, n), n


def test_facades_compile_and_link():
    """The C++ façade header stands alone, its library exports the
    reference-signature methods, and the test driver built on it loads."""
    import ctypes
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = "#include \"m2c_amd_facade.hpp\"\nint main() { return 0; }\n"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-I" + os.path.join(root, "include"), "-"],
                       input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr
    lib = os.path.join(root, "m2c_amd", "libm2c_amd_facade.so")
    syms = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    for sig in ("m2c_amd::SpaceOperator::ComputeResidual(m2c_amd::SpaceVariable3D&, m2c_amd::SpaceVariable3D&, "
                "m2c_amd::SpaceVariable3D&, double, RiemannSolutions*",
                "m2c_amd::SpaceOperator::ComputeTimeStepSize(m2c_amd::SpaceVariable3D&, m2c_amd::SpaceVariable3D&, "
                "double&, double&, m2c_amd::SpaceVariable3D*)",
                "m2c_amd::TimeIntegratorBase::AdvanceOneTimeStep(",
                "m2c_amd::SpaceVariable3D::GetDataPointer()",
                "m2c_amd::SpaceVariable3D::RestoreDataPointerAndInsert()",
                "m2c_amd::LevelSetOperator::ComputeResidual("):
        assert sig in syms, sig
    d = ctypes.CDLL(os.path.join(root, "tests", "libfacade_driver.so"))
    assert hasattr(d, "facade_run")

"""The C-ABI boundary without a GPU: both product libraries load and export
every entry point that include/*.h declares (the reference's six C entry
points, cfd2lcs_c.c:67-142, their Fortran-callable names, cfd2lcs_c.c:16-58,
and the additive ones), the flag values match the reference header
(cfd2lcs_inc_double.h:27-65), and the Python binding fails loudly when a
library has not been built."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        for m in re.finditer(r"\b(cfd2lcs_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def header_macros():
    text = open(os.path.join(INCLUDE, "cfd2lcs_api.h")).read()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define\s+(\w+)\s+(-?\d+)\b", text)}


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref in ("cfd2lcs_init_c", "cfd2lcs_diagnostic_init_c", "cfd2lcs_update_c",
                "cfd2lcs_diagnostic_destroy_c", "cfd2lcs_set_option_c", "cfd2lcs_set_param_c",
                "cfd2lcs_init_", "cfd2lcs_diagnostic_init_", "cfd2lcs_update_",
                "cfd2lcs_diagnostic_destroy_", "cfd2lcs_set_option_", "cfd2lcs_set_param_"):
        assert ref in names, ref


def test_flag_values_match_reference_header():
    m = header_macros()
    expect = {"LCS_INTERNAL": 0, "LCS_INFLOW": 1, "LCS_OUTFLOW": 2, "LCS_WALL": 3, "LCS_SLIP": 4,
              "LCS_2D": 5, "LCS_MASK": 6, "FTLE_FWD": 0, "FTLE_BKWD": 1, "LP_TRACER": 2, "EULER": 0,
              "TRAPEZOIDAL": 1, "RK2": 2, "RK3": 3, "RK4": 4, "NEAREST_NBR": 0, "LINEAR": 1,
              "QUADRATIC": 2, "CUBIC": 3, "TSE": 5, "TSE_LIMIT": 6, "LCS_FALSE": 0, "LCS_TRUE": 1,
              "LCS_NAMELEN": 128}
    for k, v in expect.items():
        assert m.get(k) == v, k
    import libcfd2lcs_amd as L
    for k, v in expect.items():
        if hasattr(L, k):
            assert getattr(L, k) == v, k


@pytest.mark.parametrize("variant", ["dp", "sp"])
def test_library_exports_every_declared_symbol(variant):
    import libcfd2lcs_amd as L
    path = L.lib_path(variant)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run python -m libcfd2lcs_amd.build)")
    lib = C.CDLL(path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"{variant}: not exported: {missing}"
    for n in L.EXPORTED:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("variant", ["dp", "sp"])
def test_exported_names_are_c_linkage(variant):
    """No C++-mangled entry point: the boundary is extern "C" only."""
    import libcfd2lcs_amd as L
    out = subprocess.run(["nm", "-D", "--defined-only", L.lib_path(variant)], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for n in declared_functions():
        assert n in exported, n


def test_headers_compile_as_c_for_both_precisions(tmp_path):
    src = tmp_path / "t.c"
    for inc in ("cfd2lcs_inc_dp.h", "cfd2lcs_inc_sp.h"):
        src.write_text(f'#include "{inc}"\nint main(void) {{ return LCS_MASK + FTLE_BKWD + RK4 + CUBIC; }}\n')
        subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", INCLUDE, "-c", str(src), "-o",
                        str(tmp_path / "t.o")], check=True)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import libcfd2lcs_amd as L
    monkeypatch.setattr(L, "LIBDIR", str(tmp_path))
    monkeypatch.setattr(L, "_libs", {})
    with pytest.raises(RuntimeError, match="not built"):
        L.load("dp")


@pytest.mark.parametrize("variant", ["sp", "dp"])
def test_single_precision_channel_grid_is_accepted(variant, capfd):
    """A rectilinear grid with walls on x, slip faces on z and a periodic y
    axis, given in lcsdata_t: the periodic face ghosts take the spacing at one
    end of y and the wall edges keep the mirror spacing of the other end, so
    with float coordinates they differ by float rounding.  The grid set-up
    (host code, no GPU needed) must accept it; without a GPU, init then stops
    at the device check."""
    import flows as F
    import libcfd2lcs_amd as L
    try:
        lib = L.Cfd2lcs(variant)
    except RuntimeError as e:
        pytest.skip(str(e))
    X, Y, Z = F.cell_grid((24, 16, 12), (0, 0, 0), (2 * np.pi,) * 3)
    fl = np.full(X.shape, F.LCS_INTERNAL, dtype=np.int32)
    fl[0], fl[-1] = F.LCS_WALL, F.LCS_WALL
    fl[:, :, 0], fl[:, :, -1] = F.LCS_SLIP, F.LCS_SLIP
    lib.finalize()
    lib.init(X, Y, Z, fl)
    out = capfd.readouterr().out
    assert "not separable" not in out and "rectilinear" not in out, out
    assert lib.error == 0 or "no HIP device" in out, out
    lib.finalize()

"""CPU: the Fortran ISO_C_BINDING layer (gem_amd/fortran) builds with amdflang and its
bind(C) derived types have the byte sizes of the C structs of include/adz_hip.h (a
field missing, misplaced or of the wrong kind changes the size or the padding)."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "gem_amd", "fortran", "adz_f_driver")


def test_bind_c_type_sizes_match_c_abi():
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    from gem_amd import adz as A

    out = subprocess.run([DRIVER, "sizes"], capture_output=True, text=True, timeout=60)
    line = [ln for ln in out.stdout.splitlines() if ln.startswith("sizes")][0]
    sizes = [int(x) for x in line.split()[1:]]
    assert sizes == [C.sizeof(s) for s in (A.AdzGrid, A.AdzField, A.AdzItem, A.AdzTune, A.AdzBcTracer,
                                           A.AdzBcDomain, A.AdzFluxTracer, A.AdzFluxDomain,
                                           A.AdzAirmassFields)]


def test_fortran_layer_library_exports_its_entry_points():
    """libadz_gpu.so (module adz_gpu and its C-callable entry points) loads on a host
    without a GPU and exports every routine gem_amd.adz_fortran binds."""
    from gem_amd import adz_fortran as AF

    if not os.path.exists(AF.LIB_PATH):
        pytest.skip("Fortran layer not built (no amdflang)")
    F = AF.flib()
    for name in AF._SIGS:
        assert hasattr(F, name), name


@pytest.mark.parametrize("what", ["ilmc", "qv"])
def test_fp64_fields_refused_where_the_kernel_reads_real(what):
    """ILMC_LAM and adz_tricub_hlt(F_QV_L) run kernels on `real` arrays: with FP64 field
    storage the Fortran layer stops with a message before any device call (here: no GPU)."""
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    out = subprocess.run([DRIVER, "refuse_f64", what], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0
    assert "FP32 fields" in out.stderr + out.stdout
    assert "not refused" not in out.stdout

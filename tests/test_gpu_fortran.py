"""The Fortran ISO_C_BINDING shim on the GPU: padeops_amd/fortran/test_shim
(built by padeops_amd.build when amdflang is installed) drives every type of
the shim on device arrays (cd06, cf90, cd06stagg periodic and non-periodic
with its real and complex generics, PoissonPeriodic%init with a Get_ModKz
callback, derivatives, filters, spectral, Pade6stagg, padepoisson, cgrid and the
decomp_info / transpose_* module) and checks each against a closed form
(modified wavenumbers, single Fourier modes, constants) and the reference's
init codes."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "padeops_amd", "fortran",
                   "test_shim")


def test_fortran_shim(cuda):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (no amdflang)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "PASS" in r.stdout

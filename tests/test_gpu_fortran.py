"""The Fortran ISO_C_BINDING shim on the GPU: padeops_amd/fortran/test_shim
(built by padeops_amd.build when amdflang is installed) calls cd06%dd1 and
cf90%filter1 through the shim on device arrays and checks them against the
scheme's closed forms and the reference's init codes."""
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

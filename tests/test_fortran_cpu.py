"""The Fortran interface module compiles and links against the C-ABI library, and its error
path turns a nonzero status into a stop with the library's message (MOLSIM's Stop)."""
import os
import shutil

import pytest

from molsim_amd import systems
from fortran_io import DRIVER, run, write_input


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no amdflang)")
def test_driver_reports_library_errors(tmp_path):
    import torch  # noqa: F401  (only to ask whether a GPU is present)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present; the GPU test covers the driver")
    sp = systems.small_spme_system(n=20, L=30.0, nmesh=16, alpha=0.3, rcut=10.0)
    path = tmp_path / "in.txt"
    write_input(path, sp, [(0, sp.ro[0])])
    r = run(str(path))
    assert r.returncode != 0
    assert "molsim_gpu: msg_create failed with status" in r.stdout

"""The drop-in I_data/L_data/O_data (bsr_amd/fortran/bsr_gpu_dropin.f90)
compile against the reference's own modules cmdata and bsr_mat, as built
from src/bsr/BSR_MAT3/mod_cmdata.f90 and mod_param.f90 by oracle/ref.mk, so
Gen_matrix's calls (gen_matrix.f90:52-56) bind to them unchanged.  The GPU
side of the same subroutines runs in test_gpu_parity.py's Fortran driver
test, whose lists go through them."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFMOD = os.path.join(ROOT, "oracle", "_ref", "mod")
FORTRAN = os.path.join(ROOT, "bsr_amd", "fortran")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


@pytest.mark.skipif(not os.path.exists(os.path.join(REFMOD, "cmdata.mod")),
                    reason="reference modules not built (oracle/ref.mk)")
@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler")
def test_dropin_compiles_against_reference_modules(tmp_path):
    mod = tmp_path / "mod"
    mod.mkdir()
    # module bsr_gpu first, into a private module directory
    subprocess.run([FC, "-c", "-module-dir", str(mod), os.path.join(FORTRAN, "bsr_gpu.f90"),
                    "-o", str(tmp_path / "bsr_gpu.o")], check=True)
    obj = tmp_path / "dropin.o"
    subprocess.run([FC, "-c", "-I", REFMOD, "-I", str(mod), os.path.join(FORTRAN, "bsr_gpu_dropin.f90"),
                    "-o", str(obj)], check=True)
    syms = subprocess.run(["nm", str(obj)], check=True, capture_output=True, text=True).stdout
    defined = {ln.split()[-1] for ln in syms.splitlines() if " T " in ln}
    # the reference's own external names (gfortran/flang mangling: lower case + '_')
    assert {"i_data_", "l_data_", "o_data_"} <= defined

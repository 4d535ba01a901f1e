"""Host-side regression check of the K-matrix reverse pass (the same header
the gradient kernel uses, compiled for the CPU by build.sh): adjoints of the
K-matrix couplings, background terms and pole masses against central finite
differences of the forward K-matrix."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pwa-xib2jpsilambdak_amd", "_build", "km_adjoint_check")


@pytest.mark.skipif(not os.path.exists(EXE), reason="km_adjoint_check not built (run build.sh)")
def test_kmatrix_adjoint_matches_finite_differences():
    out = subprocess.run([EXE], capture_output=True, text=True, check=True).stdout
    worst = float(out.strip().splitlines()[-1].split()[-1])
    assert worst < 2e-6, out[-2000:]

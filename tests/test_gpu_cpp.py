"""C++ acceptance tests on the GPU.

* trl_amd/lib/test_lanczos_acceptance: the reference's test_lanczos_extend /
  test_lanczos_convergence checks (tests/test_lanczos_extend.hh,
  tests/test_lanczos_convergence.hh) on trl::hip with trl_amd::BlockLanczos.
* oracle/_ref/refdriver_hip: the same checks with the reference's UNCHANGED
  trl::BlockLanczos (compiled from the reference headers by
  oracle/Makefile.ref where they exist), with ModifiedGS and BlockCGS2.
"""

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
pytestmark = pytest.mark.gpu


def run(exe: Path) -> str:
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    return r.stdout


def test_acceptance_trl_amd_driver():
    exe = ROOT / "trl_amd" / "lib" / "test_lanczos_acceptance"
    assert exe.exists(), "build the C++ tests first (python -m trl_amd.build)"
    assert "ALL PASSED" in run(exe)


def test_acceptance_reference_driver_unchanged():
    exe = ROOT / "oracle" / "_ref" / "refdriver_hip"
    if not exe.exists():
        pytest.skip("oracle/_ref/refdriver_hip is built only where the reference headers exist (oracle/Makefile.ref)")
    assert "ALL PASSED" in run(exe)

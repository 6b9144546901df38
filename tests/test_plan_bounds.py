"""The update kernel's work plans on the host (CPU): decompose() covers the
j tiles with its chunks for every active count, and the j-split partials of
every plan fit the buffer the library allocates (partial_bound and
ensemble_partial_bound in gravidy_amd/csrc/h4_kernels.hpp).  An exhaustive
check over active counts, lane-group widths, work-item targets and split
caps, plus random ensembles, compiled from the library's own header."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_plans_cover_the_tiles_and_fit_the_partial_buffer(tmp_path):
    exe = tmp_path / "plan_bounds"
    subprocess.run([HIPCC, "-std=c++17", "-O2", "--offload-arch=gfx950",
                    "-I", os.path.join(ROOT, "gravidy_amd", "csrc"),
                    os.path.join(ROOT, "tests", "plan_bounds.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "plan bounds ok" in proc.stdout

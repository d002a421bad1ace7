"""CPU emulation of the index and address arithmetic of the semi-implicit
step's first sweep (tests/emulation/bounds_emulation.hip): the fused
right-hand side + z R2C kernel and the two-kernel path it replaces, for
every block, thread and x-plane, at (4,4,1024), (12,8,512), 512^3 and
ragged shapes.  Every global and LDS access must stay inside its
allocation and every output element must be written exactly once.  The
program includes the kernels' own index helpers; it is compiled for the
host with hipcc and never touches a GPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "emulation" / "bounds_emulation.hip"


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    return None


@pytest.mark.skipif(hipcc() is None, reason="hipcc is needed to build the host emulation")
def test_first_sweep_index_math_stays_in_bounds(tmp_path):
    exe = tmp_path / "bounds_emulation"
    cmd = [hipcc(), "-O2", "-std=c++17", "--offload-arch=gfx950", "-I", str(ROOT / "include"),
           "-I", str(ROOT / "evoxels_amd" / "csrc"), str(SRC), "-o", str(exe)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(run.stdout)
    assert run.returncode == 0, run.stdout + run.stderr
    lines = run.stdout.strip().splitlines()
    assert len(lines) == 15 and all(line.startswith("OK ") for line in lines)

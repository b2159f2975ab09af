"""The other mechanisms (RELACS_MX, CB07_TUV, RACM_TUV) on the GPU: the
library built for each (libbrams_chem_<mech>.so, from the same kernels and
the mechanism's generated code) against that mechanism's oracle — the RODAS3
step at both tolerance settings (equal attempts and sub-steps, mask, dt
history), the ROS2 step, the QSSA step and the TUV mapping with the
mechanism's J relations — on 16x16x10, several grids in one pass, and the
RODAS3 step at full depth on sampled blocks (50x50x40 with 65-cell blocks,
200x200x40).  A process holds one mechanism, so each runs
tests/mech_check.py in a child process with BRAMS_MECH set."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MECHS = ["relacs_mx", "cb07_tuv", "racm_tuv"]


def _run(mech, mode):
    for so in ("brams_amd/libbrams_chem_%s.so" % mech, "oracle/liboracle_%s.so" % mech):
        assert os.path.exists(os.path.join(ROOT, so)), "%s not built" % so
    print("mechanism %s: mech_check.py %s" % (mech, mode), flush=True)
    env = dict(os.environ, BRAMS_MECH=mech)
    out = subprocess.run([sys.executable, "-u", os.path.join(HERE, "mech_check.py"), mode],
                         env=env, capture_output=True, text=True, timeout=110)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed" in out.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("mech", MECHS)
def test_mechanism_on_gpu(mech):
    _run(mech, "--gpu")


@pytest.mark.gpu
@pytest.mark.parametrize("mech", MECHS)
def test_mechanism_full_depth_on_gpu(mech):
    """50x50x40 (65-cell blocks) and the 1.6M-cell grid on sampled blocks."""
    _run(mech, "--gpu-scale")

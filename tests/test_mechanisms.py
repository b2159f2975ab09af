"""The other mechanisms (RELACS_MX, CB07_TUV, RACM_TUV; src/brams/ccatt/<MECH>/)
on the CPU: their oracle against the reference's own kernels compiled from
that mechanism's directory (kinetic, fexchem, jacdchemdc, fexprod/fexloss
bit for bit), and the whole RODAS3 step against the reference's Sparse1.3
path.  A process holds one mechanism, as a BRAMS build does, so each runs
tests/mech_check.py --cpu in a child process with BRAMS_MECH set."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MECHS = ["relacs_mx", "cb07_tuv", "racm_tuv"]


@pytest.mark.ref
@pytest.mark.parametrize("mech", MECHS)
def test_mechanism_oracle_against_reference(mech):
    for so in ("oracle/liboracle_%s.so" % mech, "oracle/_ref/libbrams_ref_%s.so" % mech):
        if not os.path.exists(os.path.join(ROOT, so)):
            pytest.skip("%s not built" % so)
    env = dict(os.environ, BRAMS_MECH=mech)
    out = subprocess.run([sys.executable, os.path.join(HERE, "mech_check.py"), "--cpu"],
                         env=env, capture_output=True, text=True, timeout=900)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed" in out.stdout

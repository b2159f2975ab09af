"""mod_brams_chem_history (brams_amd/fortran/): the Fortran side of the dt
history record, which a BRAMS history writer and a restart call.  The test
driver writes two grids' records to a stream file, reads them back and
compares (host calls only, no GPU); the file's records are then checked
byte for byte against the ones the C-ABI packs from Python."""
import os
import struct
import subprocess

import numpy as np
import pytest

from brams_amd.chem import dt_history_pack
from brams_amd.workload import Grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "brams_amd", "fortran", "history_driver")


def test_fortran_history_round_trip(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("history_driver not built (flang absent)")
    path = str(tmp_path / "chem_dt.hist")
    r = subprocess.run([DRIVER, path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout + r.stderr
    data = open(path, "rb").read()
    want = [(Grid(16, 16, 10), 10.0 + 0.5 * np.arange(1, 41)),
            (Grid(40, 30, 12), 120.0 / np.arange(1, 226))]
    off = 0
    for grid, dt in want:
        (n,) = struct.unpack_from("<q", data, off)
        rec = data[off + 8:off + 8 + n]
        assert rec == dt_history_pack(grid, dt)
        off += 8 + n
    assert off == len(data)

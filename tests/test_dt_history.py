"""The dt history record (brams_chem_dt_history_pack/unpack): host code of
the library, called through the C-ABI without a GPU.  AllocIndex starts the
per-block step sizes at dtlt*N_DYN_CHEM (chem_spack_utils.f90:170-172) and
plans to keep them in the history file; the record carries them with the
grid, block size and mechanism they belong to.  The byte layout and the
checksum are pinned against a restatement here (struct + FNV-1a, itself
checked against the published FNV-1a test vectors), and every way a record
can be wrong for the grid it is restored into is refused."""
import struct

import numpy as np
import pytest

from brams_amd import abi
from brams_amd.chem import alloc_index, dt_history_pack, dt_history_unpack
from brams_amd.workload import Grid


def fnv1a64(data):
    h = 0xcbf29ce484222325
    for b in data:
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def record(mech, grid, block_size, dt, version=1):
    head = b"BRAMSDTH" + struct.pack("<II", version, 72) + mech.encode().ljust(16, b"\0")
    head += struct.pack("<8i", grid.m1, grid.m2, grid.m3, grid.ia, grid.iz, grid.ja, grid.jz, block_size)
    head += struct.pack("<ii", len(dt), 0)
    body = head + np.asarray(dt, "<f8").tobytes()
    return body + struct.pack("<Q", fnv1a64(body))


def test_fnv1a_known_answers():
    # published FNV-1a 64-bit test vectors
    assert fnv1a64(b"") == 0xcbf29ce484222325
    assert fnv1a64(b"a") == 0xaf63dc4c8601ec8c
    assert fnv1a64(b"foobar") == 0x85944171f73967e8


def test_mechanism_query():
    name, ns, nr, nphot = abi.mechanism()
    assert (name, ns, nr, nphot) == ("RELACS_TUV", 47, 128, 17)


@pytest.mark.parametrize("grid,block_size", [(Grid(16, 16, 10), 64), (Grid(40, 30, 12), 64), (Grid(8, 6, 5), 1000)])
def test_round_trip_and_layout(grid, block_size):
    nob = alloc_index(grid.m1, grid.ia, grid.iz, grid.ja, grid.jz, block_size).size
    dt = np.random.default_rng(nob).uniform(1.0, 180.0, nob)
    rec = dt_history_pack(grid, dt, block_size)
    assert len(rec) == 80 + 8 * nob == abi.lib().brams_chem_dt_history_bytes(nob)
    assert rec == record("RELACS_TUV", grid, block_size, dt)
    back = dt_history_unpack(rec, grid, block_size)
    assert back.tobytes() == dt.tobytes()


def _code(fn):
    with pytest.raises(abi.ChemError) as e:
        fn()
    return int(str(e.value).split("(")[1].split(")")[0])


def test_refusals():
    grid = Grid(16, 16, 10)
    nob = alloc_index(grid.m1, grid.ia, grid.iz, grid.ja, grid.jz).size
    dt = np.full(nob, 90.0)
    rec = dt_history_pack(grid, dt)
    # corrupt byte, truncation, another version, not a record
    bad = bytearray(rec)
    bad[72 + 8 * 3] ^= 0x01
    assert _code(lambda: dt_history_unpack(bytes(bad), grid)) == -7
    assert _code(lambda: dt_history_unpack(rec[:-1], grid)) == -7
    assert _code(lambda: dt_history_unpack(record("RELACS_TUV", grid, 64, dt, version=2), grid)) == -7
    assert _code(lambda: dt_history_unpack(b"\0" * len(rec), grid)) == -7
    # another grid, block size or mechanism
    assert _code(lambda: dt_history_unpack(rec, Grid(16, 16, 11))) == -8
    assert _code(lambda: dt_history_unpack(rec, grid, block_size=32)) == -8
    assert _code(lambda: dt_history_unpack(record("RACM_TUV", grid, 64, dt), grid)) == -8
    # a dt that no step leaves: zero, negative, NaN (pack and a forged record)
    for v in (0.0, -1.0, np.nan, np.inf):
        d = dt.copy()
        d[5] = v
        assert _code(lambda: dt_history_pack(grid, d)) == -7
        assert _code(lambda: dt_history_unpack(record("RELACS_TUV", grid, 64, d), grid)) == -7
    # nob that is not the grid's block count
    assert _code(lambda: dt_history_pack(grid, dt[:-1])) == -8

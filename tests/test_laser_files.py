"""The reader of the reference's intermediate-state operator files
(SCRATCH/dipole_<c>, SCRATCH/cap: opxvec_ext, source/propagation/
tdselib.f90:1031-1153) on a file written here in that layout: Fortran
sequential records with 4-byte length markers, buf_size, the diagonal, then
records (oij(buf_size), indxi(buf_size), indxj(buf_size), nlim) with 1-based
indices; the last record partly filled."""
import numpy as np

from adc_ottawa_amd import lib
from oracle import laser


def _record(payload: bytes) -> bytes:
    n = np.array([len(payload)], "<i4").tobytes()
    return n + payload + n


def test_read_is_operator_roundtrip(tmp_path):
    nd, bs = 37, 16
    rng = np.random.default_rng(0)
    diag = rng.normal(size=nd)
    row = rng.integers(0, nd, 40)
    col = (row + 1 + rng.integers(0, nd - 1, 40)) % nd
    val = rng.normal(size=40)
    out = _record(np.array([bs], "<i4").tobytes()) + _record(diag.astype("<f8").tobytes())
    for k0 in range(0, row.size, bs):
        k1 = min(row.size, k0 + bs)
        nlim = k1 - k0
        o = np.zeros(bs)
        i = np.zeros(bs, np.int32)
        j = np.zeros(bs, np.int32)
        o[:nlim], i[:nlim], j[:nlim] = val[k0:k1], row[k0:k1] + 1, col[k0:k1] + 1
        out += _record(o.astype("<f8").tobytes() + i.astype("<i4").tobytes() + j.astype("<i4").tobytes()
                       + np.array([nlim], "<i4").tobytes())
    path = tmp_path / "dipole_x"
    path.write_bytes(out)
    d, r, c, v = lib.read_is_operator(str(path), nd)
    assert np.array_equal(d, diag) and np.array_equal(r, row) and np.array_equal(c, col) and np.array_equal(v, val)
    D = laser.dense_dipole(d, r, c, v, nd)
    x = rng.normal(size=nd)
    # opxvec_ext: v2 = oii v1, then both (i, j) and (j, i) of each element
    want = diag * x
    for a, b, w in zip(row, col, val):
        want[a] += w * x[b]
        want[b] += w * x[a]
    assert np.abs(D @ x - want).max() < 1e-13

"""On-disk matrix files (adc2g_write_hmlt) against the oracle.

The files are read back here as rdham_on / rdham_off read them
(source/eigen/block_davidson.f90:600-686): Fortran sequential unformatted
records framed by 4-byte lengths.  The stored elements must be the upper
triangle of the literal restatement with the minc = 1e-12 filter
(oracle/elements.py, apply_minc=True): the same (i, j) set except elements
within a factor 2 of minc, where rounding may decide; values to 1e-13
relative; the diagonal record to 1e-14 relative.  Small record sizes exercise
the full records and the last, partial one; a record count that divides
evenly leaves an empty last record, as the reference writes it.
"""
import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system
from oracle import elements

pytestmark = pytest.mark.gpu

MINC = 1e-12


def _records(path):
    data = open(path, "rb").read()
    out, p = [], 0
    while p < len(data):
        n = int(np.frombuffer(data, np.int32, 1, p)[0])
        out.append(data[p + 4:p + 4 + n])
        assert int(np.frombuffer(data, np.int32, 1, p + 4 + n)[0]) == n
        p += n + 8
    return out


def _read(dia, off, ndim):
    rd = _records(dia)
    assert len(rd) == 2
    maxbl, nbuf = np.frombuffer(rd[0], np.int32)
    hii = np.frombuffer(rd[1], np.float64)
    assert hii.size == ndim
    ro = _records(off)
    assert len(ro) == nbuf
    vals, ii, jj = [], [], []
    for r in ro:
        assert len(r) == 16 * maxbl + 4
        buff = np.frombuffer(r, np.float64, maxbl, 0)
        oi = np.frombuffer(r, np.int32, maxbl, 8 * maxbl)
        oj = np.frombuffer(r, np.int32, maxbl, 12 * maxbl)
        nlim = int(np.frombuffer(r, np.int32, 1, 16 * maxbl)[0])
        vals.append(buff[:nlim])
        ii.append(oi[:nlim])
        jj.append(oj[:nlim])
    return int(maxbl), int(nbuf), hii, np.concatenate(vals), np.concatenate(ii), np.concatenate(jj)


@pytest.mark.parametrize("form", ("df", "int"))
@pytest.mark.parametrize("o,v,naux,seed,maxbl", [(3, 8, 10, 5, 1000), (4, 12, 20, 6, 4096)])
def test_hmlt_vs_oracle(o, v, naux, seed, maxbl, form, tmp_path):
    print(f"[case] hmlt o={o} v={v} naux={naux} maxbl={maxbl} coupling={form}", flush=True)
    sys = make_system(o, v, naux, seed=seed)
    Hm = elements.build_blocks(sys, apply_minc=True).dense()
    Hf = elements.build_blocks(sys, apply_minc=False).dense()
    n = Hm.shape[0]
    dia, off = tmp_path / "hmlt.diai", tmp_path / "hmlt.offi"
    with lib.Operator.from_system(sys, coupling=form) as op:
        noffd, nbuf = op.write_hmlt(dia, off, minc=MINC, maxbl=maxbl)
    mb, nb, hii, vals, ii, jj = _read(dia, off, n)
    assert mb == maxbl and nb == nbuf == noffd // maxbl + 1 and vals.size == noffd
    assert np.abs(hii - np.diag(Hf)).max() < 1e-14 * np.abs(np.diag(Hf)).max()
    assert (ii >= 1).all() and (ii < jj).all() and (jj <= n).all()
    # the stored pairs are unique
    key = (ii.astype(np.int64) - 1) * n + (jj - 1)
    assert np.unique(key).size == key.size
    got = np.zeros((n, n))
    got[ii - 1, jj - 1] = vals
    ref = np.triu(Hm, 1)
    near = np.abs(np.abs(np.triu(Hf, 1)) - MINC) < MINC  # within a factor 2 of minc
    diff_set = ((got != 0) != (ref != 0)) & ~near
    assert not diff_set.any()
    both = (got != 0) & (ref != 0)
    scale = np.abs(ref).max()
    assert np.abs(got[both] - ref[both]).max() < 1e-13 * scale


def test_hmlt_empty_last_record(tmp_path):
    """A record size equal to the element count gives one full record and an
    empty last one."""
    sys = make_system(3, 8, 10, seed=5)
    with lib.Operator.from_system(sys) as op:
        ndim = op.ndim
        noffd, nbuf = op.write_hmlt(tmp_path / "d", tmp_path / "o")
        assert nbuf == 1
        noffd2, nbuf2 = op.write_hmlt(tmp_path / "d2", tmp_path / "o2", maxbl=noffd)
    assert noffd2 == noffd and nbuf2 == 2
    *_, vals, _, _ = _read(tmp_path / "d2", tmp_path / "o2", ndim)
    assert vals.size == noffd
    with open(tmp_path / "o2", "rb") as f:
        data = f.read()
    # the last record's nlim is 0
    assert int(np.frombuffer(data, np.int32, 1, len(data) - 8)[0]) == 0


def test_hmlt_water_size_vs_stored_matrix(tmp_path):
    """At the water size (o=5, v=36, naux=164; ndim 16470) the element list
    equals the reference's stored matrix as oracle/coo_sigma.c builds it
    (ph-ph values from the device's block, every coupling evaluated from the
    factors, the same minc filter): the same (i, j) pairs away from minc,
    values to 1e-13 relative."""
    from oracle.cpu_baseline import COO
    print("[case] hmlt water size o=5 v=36 naux=164", flush=True)
    sys = make_system(5, 36, 164, seed=0)
    with lib.Operator.from_system(sys) as op:
        n = op.ndim
        h11 = op.phph()
        noffd, nbuf = op.write_hmlt(tmp_path / "hmlt.diai", tmp_path / "hmlt.offi")
    coo = COO(sys, h11=np.ascontiguousarray(h11))
    _, _, _, vals, ii, jj = _read(tmp_path / "hmlt.diai", tmp_path / "hmlt.offi", n)
    ka = (ii.astype(np.int64) - 1) * n + (jj - 1)
    kb = (coo.oi[:coo.nnz].astype(np.int64) - 1) * n + (coo.oj[:coo.nnz] - 1)
    va = dict(zip(ka.tolist(), vals.tolist()))
    vb = dict(zip(kb.tolist(), coo.val[:coo.nnz].tolist()))
    only = set(va) ^ set(vb)
    for k in only:
        x = va.get(k, vb.get(k))
        assert abs(abs(x) - MINC) < MINC, (k, x)
    common = np.array(sorted(set(va) & set(vb)))
    a = np.array([va[k] for k in common])
    b = np.array([vb[k] for k in common])
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
    print(f"[hmlt water] noffd {noffd} (stored matrix {coo.nnz}), records {nbuf}", flush=True)

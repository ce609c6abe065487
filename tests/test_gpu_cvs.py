"""CVS view (adc2g_create_cvs) against the oracle.

The oracle's CVS configuration map (oracle/csf.py cvs_indices, pinned by a
brute-force filter of the full table) selects the principal sub-matrix of
the literal restatement.  The view is checked against it:
  * dimensions and the diagonal (exact);
  * sigma-vectors to 1e-12 relative, both coupling forms;
  * Davidson eigenvalues to 1e-10 against eigh of the sub-matrix;
  * SIL autocorrelation to 1e-10 against the oracle propagation;
  * the hmlt files against the filtered sub-matrix;
  * the argument and lifetime errors.
"""
import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system
from oracle import csf, elements, sil

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", ("df", "int"))
@pytest.mark.parametrize("o,v,naux,seed,core", [(4, 8, 12, 3, [1]), (5, 7, 10, 4, [1, 2])])
def test_cvs_sigma_davidson_sil(o, v, naux, seed, core, form):
    print(f"[case] cvs o={o} v={v} naux={naux} core={core} coupling={form}", flush=True)
    sys = make_system(o, v, naux, seed=seed)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    idx = csf.cvs_indices(o, v, [c - 1 for c in core])
    Hc = H[np.ix_(idx, idx)]
    nc = len(core)
    with lib.Operator.from_system(sys, coupling=form) as op:
        cv = op.cvs(core)
        try:
            assert cv.ndim == idx.size and cv.n1 == nc * v
            assert cv.counts == [0, 0, nc * (o - nc) * v] + [nc * (o - nc) * v * (v - 1) // 2] * 2
            assert np.array_equal(cv.diagonal(), np.diag(H)[idx])
            X = np.random.default_rng(seed).standard_normal((cv.ndim, 5))
            W = cv.sigma(X)
            assert np.abs(W - Hc @ X).max() < 1e-12 * np.abs(Hc @ X).max()
            ev, vec, st = cv.davidson(3, blocksize=6, maxvec=36, tol=1e-8, maxiter=300)
            lam = np.linalg.eigvalsh(Hc)
            assert np.abs(ev - lam[:3]).max() < 1e-10 * np.abs(lam).max()
            f = np.zeros(cv.ndim)
            f[: cv.n1] = np.random.default_rng(1).normal(0, 0.1, cv.n1)
            auto, _ = cv.sil_autocorr(f, tout=0.1, tfinal=1.0, kdim=12, autotol=1e-10)
            _, ref = sil.propagate(lambda x: -1j * (Hc @ x), f, tout=0.1, tfinal=1.0, kdim=12, autotol=1e-10)
            assert np.abs(auto - ref).max() < 1e-10 * np.dot(f, f)
        finally:
            cv.close()


def _read(dia, off, n):
    """(maxbl, nbuf, hii, vals, oi, oj) of the hmlt records (Fortran
    sequential unformatted, 4-byte length markers)."""
    def records(path):
        data, out, p = open(path, "rb").read(), [], 0
        while p < len(data):
            m = int(np.frombuffer(data, np.int32, 1, p)[0])
            out.append(data[p + 4:p + 4 + m])
            p += m + 8
        return out
    rd = records(dia)
    maxbl, nbuf = (int(x) for x in np.frombuffer(rd[0], np.int32))
    hii = np.frombuffer(rd[1], np.float64)
    vals, ii, jj = [], [], []
    for r in records(off):
        nlim = int(np.frombuffer(r, np.int32, 1, 16 * maxbl)[0])
        vals.append(np.frombuffer(r, np.float64, maxbl, 0)[:nlim])
        ii.append(np.frombuffer(r, np.int32, maxbl, 8 * maxbl)[:nlim])
        jj.append(np.frombuffer(r, np.int32, maxbl, 12 * maxbl)[:nlim])
    assert hii.size == n
    return maxbl, nbuf, hii, np.concatenate(vals), np.concatenate(ii), np.concatenate(jj)


def test_cvs_hmlt(tmp_path):
    sys = make_system(4, 8, 12, seed=3)
    Hm = elements.build_blocks(sys, apply_minc=True).dense()
    H = elements.build_blocks(sys, apply_minc=False).dense()
    idx = csf.cvs_indices(4, 8, [0])
    with lib.Operator.from_system(sys) as op:
        cv = op.cvs([1])
        noffd, _ = cv.write_hmlt(tmp_path / "d", tmp_path / "o")
        n = cv.ndim
        cv.close()
    _, _, hii, vals, ii, jj = _read(tmp_path / "d", tmp_path / "o", n)
    ref = np.triu(Hm[np.ix_(idx, idx)], 1)
    got = np.zeros((n, n))
    got[ii - 1, jj - 1] = vals
    near = np.abs(np.abs(np.triu(H[np.ix_(idx, idx)], 1)) - 1e-12) < 1e-12
    assert not (((got != 0) != (ref != 0)) & ~near).any()
    both = (got != 0) & (ref != 0)
    assert np.abs(got[both] - ref[both]).max() < 1e-13 * np.abs(ref).max()
    assert np.array_equal(hii, np.diag(H)[idx])


def test_cvs_errors():
    sys = make_system(4, 8, 12, seed=3)
    op = lib.Operator.from_system(sys)
    for bad in ([0], [5], [1, 1], [1, 2, 3, 4], []):
        with pytest.raises(lib.Adc2gError) as e:
            op.cvs(bad)
        assert e.value.code == 10
    cv = op.cvs([2])
    with pytest.raises(lib.Adc2gError) as e:
        cv.phph()
    assert e.value.code == 10
    with pytest.raises(lib.Adc2gError) as e:
        op.close()                       # a view is still alive
    assert e.value.code == 15
    cv.close()
    op.close()

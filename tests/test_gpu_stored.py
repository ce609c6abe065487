"""Stored-matrix handles (adc2g_create_from_hmlt): the reference's hmlt.dia /
hmlt.off files loaded as rdham_on / rdham_off load them, with the product of
hxvec_incore.

* Files written here in the reference's record layout (wrtdg / wrtoffdg:
  Fortran sequential unformatted, 4-byte markers) from a random symmetric
  sparse matrix, with a small maxbl, a last partial record and an empty one:
  the sigma-vector equals the dense product (1e-13 relative), the diagonal is
  H_jj, the Davidson roots are numpy's eigenvalues, SIL equals the dense
  exponential.
* The round trip through adc2g_write_hmlt of an operator built from
  integrals: the stored handle's sigma-vector, Davidson roots and SIL
  autocorrelation equal the integral operator's (elements with |h| <= 1e-12
  dropped by the writer, as the reference drops them).
* Entry points that need the orbitals raise, corrupt files raise.
"""
import struct

import numpy as np
import pytest
import scipy.linalg

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system

pytestmark = pytest.mark.gpu


def _record(payload: bytes) -> bytes:
    m = struct.pack("<i", len(payload))
    return m + payload + m


def _write_reference_files(path, hii, rows, cols, vals, maxbl):
    """hmlt.dia (maxbl, nbuf) + H_jj and hmlt.off records (buff, oi, oj,
    nlim) as wrtdg / wrtoffdg write them; 1-based indices, i < j."""
    nz = vals.size
    nbuf = nz // maxbl + 1
    dia = _record(struct.pack("<ii", maxbl, nbuf)) + _record(np.asarray(hii, "<f8").tobytes())
    off = b""
    for b in range(nbuf):
        sl = slice(b * maxbl, min(nz, (b + 1) * maxbl))
        nlim = sl.stop - sl.start if sl.stop > sl.start else 0
        buf = np.zeros(maxbl, "<f8")
        oi = np.zeros(maxbl, "<i4")
        oj = np.zeros(maxbl, "<i4")
        buf[:nlim] = vals[sl]
        oi[:nlim] = rows[sl] + 1
        oj[:nlim] = cols[sl] + 1
        off += _record(buf.tobytes() + oi.tobytes() + oj.tobytes() + struct.pack("<i", nlim))
    (path / "hmlt.diai").write_bytes(dia)
    (path / "hmlt.offi").write_bytes(off)
    return path / "hmlt.diai", path / "hmlt.offi"


def _random_matrix(n, nnz, seed):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, n, nnz)
    cols = rng.integers(0, n, nnz)
    keep = rows < cols
    rows, cols = rows[keep], cols[keep]
    vals = rng.normal(0.0, 0.1, rows.size)
    hii = np.sort(rng.uniform(0.2, 3.0, n))
    A = np.diag(hii)
    np.add.at(A, (rows, cols), vals)
    np.add.at(A, (cols, rows), vals)
    return A, hii, rows, cols, vals


def test_stored_reference_layout(tmp_path):
    print("[case] stored matrix n=700 from reference-layout files", flush=True)
    n = 700
    A, hii, rows, cols, vals = _random_matrix(n, 20000, 1)
    maxbl = 997                      # several full records, one partial
    dia, off = _write_reference_files(tmp_path, hii, rows, cols, vals, maxbl)
    rng = np.random.default_rng(2)
    with lib.Operator.from_hmlt(dia, off) as op:
        assert op.ndim == n and op.coupling == "stored" and op.n1 == 0
        assert np.array_equal(op.diagonal(), hii)
        X = rng.normal(size=(n, 9))
        for prec in ("fp64", "fp32", "bf16"):    # computed in FP64 whatever the request
            W = op.sigma(X, precision=prec)
            assert np.abs(W - A @ X).max() < 1e-13 * np.abs(A @ X).max(), prec
        ev, vec, st = op.davidson(5, blocksize=8, maxvec=40, tol=1e-9, precon="olsen", maxiter=300)
        w = np.linalg.eigvalsh(A)[:5]
        assert np.abs(np.sort(ev) - w).max() < 1e-10
        for prec_name in ("dpr",):
            ev2, _, _ = op.davidson(5, blocksize=8, maxvec=40, tol=1e-9, precon=prec_name, maxiter=300,
                                    vectors=False, guess="subdiag", subdim=100)
            assert np.abs(np.sort(ev2) - w).max() < 1e-10
        for bad in ("coupled", "learned"):
            with pytest.raises(lib.Adc2gError):
                op.davidson(2, blocksize=2, maxvec=8, precon=bad, maxiter=2, vectors=False)
        with pytest.raises(lib.Adc2gError):
            op.mp2(density=False)
        with pytest.raises(lib.Adc2gError):
            lib.Operator.frozen_core(op, [1])
    # an empty last record (nz a multiple of maxbl) and a file with no elements
    nz = (vals.size // 100) * 100
    dia, off = _write_reference_files(tmp_path, hii, rows[:nz], cols[:nz], vals[:nz], 100)
    B = np.diag(hii)
    np.add.at(B, (rows[:nz], cols[:nz]), vals[:nz])
    np.add.at(B, (cols[:nz], rows[:nz]), vals[:nz])
    with lib.Operator.from_hmlt(dia, off) as op:
        X = rng.normal(size=(n, 3))
        assert np.abs(op.sigma(X) - B @ X).max() < 1e-13 * np.abs(B @ X).max()
    dia, off = _write_reference_files(tmp_path, hii, rows[:0], cols[:0], vals[:0], 16)
    with lib.Operator.from_hmlt(dia, off) as op:
        X = rng.normal(size=(n, 2))
        assert np.array_equal(op.sigma(X), hii[:, None] * X)


def test_stored_rejects_corrupt_files(tmp_path):
    print("[case] stored matrix corrupt files", flush=True)
    n = 50
    A, hii, rows, cols, vals = _random_matrix(n, 300, 3)
    dia, off = _write_reference_files(tmp_path, hii, rows, cols, vals, 40)
    good_off = off.read_bytes()
    off.write_bytes(good_off[:-10])                                   # truncated
    with pytest.raises(lib.Adc2gError):
        lib.Operator.from_hmlt(dia, off)
    bad = rows.copy()
    bad[0] = n + 5                                                    # index out of range
    dia, off = _write_reference_files(tmp_path, hii, bad, cols, vals, 40)
    with pytest.raises(lib.Adc2gError):
        lib.Operator.from_hmlt(dia, off)
    with pytest.raises(lib.Adc2gError):
        lib.Operator.from_hmlt(tmp_path / "missing", off)


def test_stored_round_trip_of_the_integral_operator(tmp_path):
    """write_hmlt of the integral operator, then the stored handle: same
    sigma-vector, roots and SIL autocorrelation."""
    print("[case] stored round trip o=4 v=10 naux=14", flush=True)
    sys_ = make_system(4, 10, 14, seed=4)
    rng = np.random.default_rng(4)
    with lib.Operator.from_system(sys_) as op:
        noffd, nbuf = op.write_hmlt(tmp_path / "hmlt.diai", tmp_path / "hmlt.offi")
        X = rng.normal(size=(op.ndim, 6))
        W = op.sigma(X)
        ev, _, _ = op.davidson(4, blocksize=4, maxvec=24, tol=1e-9, precon="olsen", maxiter=300, vectors=False)
        f = np.zeros(op.ndim)
        f[: op.n1] = rng.normal(size=op.n1)
        f /= np.linalg.norm(f)
        a_int = op.sil_autocorr(f, tout=0.5, tfinal=5.0, kdim=12, autotol=1e-10)[0]
    with lib.Operator.from_hmlt(tmp_path / "hmlt.diai", tmp_path / "hmlt.offi") as st:
        Ws = st.sigma(X)
        evs, _, _ = st.davidson(4, blocksize=4, maxvec=24, tol=1e-9, precon="olsen", maxiter=300, vectors=False)
        a_st = st.sil_autocorr(f, tout=0.5, tfinal=5.0, kdim=12, autotol=1e-10)[0]
    print(f"{noffd} stored elements in {nbuf} records; max |sigma diff| {np.abs(Ws - W).max():.2e}", flush=True)
    assert np.abs(Ws - W).max() < 1e-10 * np.abs(W).max()
    assert np.abs(np.sort(evs) - np.sort(ev)).max() < 1e-9
    assert np.abs(a_st - a_int).max() < 1e-8


def test_fortran_driver_stored_matrix(tmp_path):
    """The Fortran driver with dia_file / off_file: the Davidson on the
    reference's files, the roots printed equal numpy's, and an input that
    needs the orbitals is refused."""
    import os
    import subprocess
    driver = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "adc_ottawa_amd",
                          "adc2g_driver")
    if not os.path.exists(driver):
        pytest.skip("Fortran driver not built")
    print("[case] fortran driver stored matrix n=400", flush=True)
    n = 400
    A, hii, rows, cols, vals = _random_matrix(n, 9000, 5)
    _write_reference_files(tmp_path, hii, rows, cols, vals, 333)
    nml = ("&adc2g nstates=4 blocksize=6 maxvec=36 tol=1d-9 precon=2 davname='davstates' "
           "dia_file='hmlt.diai' off_file='hmlt.offi' /\n")
    r = subprocess.run([driver], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-1200:])
    assert r.returncode == 0, r.stdout + r.stderr
    got = []
    for ln in r.stdout.splitlines():
        t = ln.split()
        if len(t) == 3 and t[0].isdigit():
            got.append(float(t[1]))
    w = np.linalg.eigvalsh(A)[:4]
    assert len(got) == 4 and np.abs(np.array(got) - w).max() < 1e-10
    bad = nml.replace(" /", " adc2x=.true. /")
    r = subprocess.run([driver], input=bad, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode != 0 and "dia_file allows" in r.stdout


def test_stored_handle_writes_and_runs_lanczos(tmp_path):
    """A stored-matrix handle writes its own hmlt files back (the same
    diagonal and elements), and block Lanczos and relaxation on it equal
    those of the integral operator it came from."""
    print("[case] stored handle write-back, Lanczos, relaxation o=3 v=6", flush=True)
    sys_ = make_system(3, 6, 9, seed=9)
    rng = np.random.default_rng(9)
    with lib.Operator.from_system(sys_) as op:
        op.write_hmlt(tmp_path / "a.dia", tmp_path / "a.off", minc=0.0)
        q0 = np.linalg.qr(rng.normal(size=(op.ndim, 4)))[0]
        ev_i = op.block_lanczos(q0, 6)[0]
        e_i = op.relax(2, step=20.0)[0]
    with lib.Operator.from_hmlt(tmp_path / "a.dia", tmp_path / "a.off") as st:
        st.write_hmlt(tmp_path / "b.dia", tmp_path / "b.off", minc=0.0)
        ev_s = st.block_lanczos(q0, 6)[0]
        e_s = st.relax(2, step=20.0)[0]
    assert (tmp_path / "a.dia").read_bytes() == (tmp_path / "b.dia").read_bytes()

    def elements(path):
        data, out, p = open(path, "rb").read(), [], 0
        while p < len(data):
            m = int(np.frombuffer(data, np.int32, 1, p)[0])
            out.append(data[p + 4:p + 4 + m])
            p += m + 8
        return out
    maxbl = int(np.frombuffer(elements(tmp_path / "a.dia")[0], np.int32)[0])
    sets = []
    for name in ("a.off", "b.off"):
        trip = {}
        for r in elements(tmp_path / name):
            nlim = int(np.frombuffer(r, np.int32, 1, 16 * maxbl)[0])
            v = np.frombuffer(r, np.float64, maxbl, 0)[:nlim]
            i = np.frombuffer(r, np.int32, maxbl, 8 * maxbl)[:nlim]
            j = np.frombuffer(r, np.int32, maxbl, 12 * maxbl)[:nlim]
            trip.update({(int(a), int(b)): float(c) for a, b, c in zip(i, j, v)})
        sets.append(trip)
    assert sets[0].keys() == sets[1].keys()
    assert max(abs(sets[0][k] - sets[1][k]) for k in sets[0]) < 1e-13
    assert np.abs(np.sort(ev_s) - np.sort(ev_i)).max() < 1e-10
    assert np.abs(np.asarray(e_s) - np.asarray(e_i)).max() < 1e-8


def test_stored_handle_laser_operator(tmp_path):
    """The laser operator (adc2g_matxvec_treal_laser) on a stored-matrix
    handle equals the integral operator's: -i H from the stored elements,
    the IS dipole terms unchanged."""
    print("[case] stored handle laser operator o=3 v=5", flush=True)
    sys_ = make_system(3, 5, 8, seed=10)
    rng = np.random.default_rng(10)
    pulses = [{"envelope": "box", "envpar": (0.0, 10.0), "carrier": "cos", "freq": 0.3,
               "pulse_vec": (1.0, 0.0, 0.0), "strength": 0.1}]
    with lib.Operator.from_system(sys_) as op:
        nd = op.ndim
        row = rng.integers(0, nd, 2 * nd)
        col = (row + 1 + rng.integers(0, nd - 1, row.size)) % nd
        val = rng.normal(0.0, 0.05, row.size)
        diag = rng.normal(0.0, 0.2, nd)
        d0j = rng.normal(0.0, 0.1, nd)
        x = rng.normal(size=nd + 1) + 1j * rng.normal(size=nd + 1)
        op.write_hmlt(tmp_path / "h.dia", tmp_path / "h.off", minc=0.0)
        op.set_dipole_is(0, diag, row, col, val, 0.2, d0j)
        want = op.matxvec_treal_laser(1.5, pulses, x)
    with lib.Operator.from_hmlt(tmp_path / "h.dia", tmp_path / "h.off") as st:
        st.set_dipole_is(0, diag, row, col, val, 0.2, d0j)
        got = st.matxvec_treal_laser(1.5, pulses, x)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()

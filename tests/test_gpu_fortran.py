"""The Fortran host program (iso_c_binding drop-in) on the GPU: eigenpair file
in the reference's davstates record layout and the 'auto' file format, with
the values checked against the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

from adc_ottawa_amd.synth import make_system, unit_dipole
from oracle import dfsigma, fvec, sil

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "adc_ottawa_amd", "adc2g_driver")


def read_davstates(path, ndim):
    recs = []
    with open(path, "rb") as f:
        while True:
            head = f.read(4)
            if not head:
                break
            (n,) = struct.unpack("<i", head)
            body = f.read(n)
            (tail,) = struct.unpack("<i", f.read(4))
            assert tail == n == 4 + 8 + 8 * ndim
            i, e = struct.unpack("<id", body[:12])
            vec = np.frombuffer(body[12:], dtype="<f8")
            recs.append((i, e, vec))
    return recs


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver(tmp_path):
    print("[case] fortran driver o=3 v=7 naux=9", flush=True)
    o, v, naux = 3, 7, 9
    sys_ = make_system(o, v, naux, seed=11)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    df = dfsigma.DF(sys_)
    ndim = df.n1 + df.D2.size
    f = np.zeros(ndim)
    f[: df.n1] = unit_dipole(o, v, seed=2)
    f.astype("<f8").tofile(tmp_path / "f.bin")
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=3 "
           f"blocksize=6 maxvec=30 tol=1d-9 precon=1 davname='davstates' sil=.true. fvec_file='f.bin' "
           f"auto_file='auto' tout=0.1d0 tfinal=1.0d0 krydim=12 autotol=1d-10 nto_file='trd.bin' "
           f"fdtd_nsel=2 fdtd_nfbas=5 fdtd_iwfunc=2 fdtd_coef_file='fdt.bin' fdtd_ebound=0.5d0,1.5d0 /\n")
    fdt = np.random.default_rng(7).standard_normal((2, 5))
    np.asfortranarray(fdt).T.astype("<f8").tofile(tmp_path / "fdt.bin")     # column-major (2, 5)
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    H = df.sigma(np.eye(ndim))
    w = np.linalg.eigvalsh(H)[:3]
    recs = read_davstates(tmp_path / "davstates", ndim)
    assert [x[0] for x in recs] == [1, 2, 3]
    ev = np.array([x[1] for x in recs])
    assert np.abs(ev - w).max() < 1e-10
    X = np.stack([x[2] for x in recs], axis=1)
    assert np.linalg.norm(H @ X - X * ev, axis=0).max() < 1e-8
    # fdtd_*: the time-domain filter-diagonalisation states
    fdrec = read_davstates(tmp_path / "fdstates_td", ndim)
    E_td, V_td = sil.fdstates_td(lambda x: -1j * (H @ x), lambda x: H @ x, f, 0.1, 1.0, 12, 1e-10, (0.5, 1.5),
                                 fdt, 2)
    assert np.abs(np.array([x[1] for x in fdrec]) - E_td).max() < 1e-9
    assert np.abs(np.stack([x[2] for x in fdrec], axis=1) - V_td).max() < 1e-8
    # nto_file: the transition density of every state (adc2_nto_gs input)
    from oracle import nto
    trd = np.fromfile(tmp_path / "trd.bin", dtype="<f8").reshape(3, o, v).transpose(0, 2, 1)
    for s in range(3):
        ref = nto.trdens_vo(sys_, X[:, s])
        assert np.abs(trd[s] - ref).max() < 1e-12 * np.abs(ref).max()
    lines = open(tmp_path / "auto").read().splitlines()
    assert lines[0].startswith("#    time[fs]")
    data = np.array([[float(t) for t in ln.split()] for ln in lines[1:]])
    _, ref = sil.propagate(lambda x: -1j * (H @ x), f, tout=0.1, tfinal=1.0, kdim=12, autotol=1e-10)
    assert data.shape == (ref.size, 4)
    assert np.allclose(data[:, 0], np.arange(ref.size) * 0.1 / 41.34137333656, atol=5e-9)
    assert np.abs(data[:, 1] + 1j * data[:, 2] - ref).max() < 1e-10 * np.dot(f, f) + 1e-13


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_consumers(tmp_path):
    """The Fortran host program driving every other consumer through the C
    ABI.  It covers:
      * relaxation (solver = 2) against the exact lowest eigenpairs;
      * the block-Lanczos state file against the oracle;
      * SIL autocorrelation orders 0-2;
      * the Chebyshev file;
      * the hmlt matrix files."""
    from oracle import cheby, elements, lanczos
    print("[case] fortran driver consumers o=3 v=7 naux=9", flush=True)
    o, v, naux = 3, 7, 9
    sys_ = make_system(o, v, naux, seed=11)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    df = dfsigma.DF(sys_)
    n1 = df.n1
    ndim = n1 + df.D2.size
    f = np.zeros(ndim)
    f[:n1] = unit_dipole(o, v, seed=2)
    f.astype("<f8").tofile(tmp_path / "f.bin")
    H = df.sigma(np.eye(ndim))
    w, U = np.linalg.eigh(H)
    span = w[-1] - w[0]
    emin, emax = w[0] - 0.02 * span, w[-1] + 0.02 * span
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=2 "
           f"solver=2 stepsize=100d0 maxiter=300 tol=1d-7 davname='rlxstates' "
           f"sil=.true. fvec_file='f.bin' auto_file='auto' tout=0.1d0 tfinal=1.0d0 krydim=12 "
           f"autotol=1d-12 autoord=2 lanczos=.true. lmain=3 ncycles=4 lancname='lancstates' "
           f"chebyord=20 cheby_emin={emin:.17e} cheby_emax={emax:.17e} hmlt=.true. "
           f"fd_coef_file='fdc.bin' fd_nsel=2 /\n")
    nml = nml.replace("e-", "d-").replace("e+", "d+")
    fdc = np.random.default_rng(4).standard_normal((21, 2)) / (1.0 + np.arange(21))[:, None]
    np.asfortranarray(fdc).T.astype("<f8").tofile(tmp_path / "fdc.bin")   # column-major (0:20, 2)
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    # relaxation: the two lowest eigenpairs
    recs = read_davstates(tmp_path / "rlxstates", ndim)
    ev = np.array([x[1] for x in recs])
    assert np.abs(ev - w[:2]).max() < 1e-10 * np.abs(w).max()
    # block Lanczos states: the oracle from the same start block
    idx = np.argsort(-np.abs(f[:n1]), kind="stable")[:3]
    q0 = np.zeros((ndim, 3))
    q0[idx, np.arange(3)] = 1.0
    ref = lanczos.block_lanczos(lambda X: H @ X, q0, 4)
    recs = read_davstates(tmp_path / "lancstates", ndim)
    assert [x[0] for x in recs] == list(range(1, 13))
    le = np.array([x[1] for x in recs])
    assert np.abs(le - ref["eigval"]).max() < 1e-10 * np.abs(ref["eigval"]).max()
    R = np.stack([x[2] for x in recs], axis=1)
    d = np.minimum(np.abs(R - ref["ritz"]).max(axis=0), np.abs(R + ref["ritz"]).max(axis=0))
    assert d.max() < 1e-8
    # SIL orders 0..2
    _, autos = sil.propagate(lambda x: -1j * (H @ x), f, tout=0.1, tfinal=1.0, kdim=12, autotol=1e-12,
                             autoord=2)
    c = U.T @ f
    for k, name in enumerate(("auto", "auto1", "auto2")):
        lines = open(tmp_path / name).read().splitlines()[1:]
        data = np.array([[float(t) for t in ln.split()] for ln in lines])
        scale = np.sum(c * c * np.abs(w) ** k)
        assert np.abs(data[:, 1] + 1j * data[:, 2] - autos[k]).max() < 1e-10 * scale + 1e-13, name
    # Chebyshev autocorrelation
    cref = cheby.chebyshev_auto(lambda x: H @ x, f, (emin, emax), 20)
    vals = [ln.split() for ln in open(tmp_path / "chebyauto").read().splitlines()
            if ln.strip() and not ln.startswith("#")]
    ck = np.array([float(x[1]) for x in vals])
    assert ck.size == 41
    assert np.abs(ck - cref).max() < 1e-10 * cref[0]
    assert np.isclose(cref[0], np.dot(f, f))
    # matrix files: the upper triangle of the filtered restatement
    Hm = np.triu(elements.build_blocks(sys_, apply_minc=True).dense(), 1)
    data = open(tmp_path / "hmlt.offi", "rb").read()
    assert len(data) > 0 and (tmp_path / "hmlt.diai").stat().st_size == 8 + 8 + 8 * ndim + 8
    near = (np.abs(np.abs(np.triu(H, 1)) - 1e-12) < 1e-12).sum()
    nrec_elems = int(np.frombuffer(data, np.int32, 1, len(data) - 8)[0])
    assert abs(nrec_elems - np.count_nonzero(Hm)) <= near
    # Chebyshev filter-diagonalisation states (fd_coef_file, fd_nsel)
    recs = read_davstates(tmp_path / "fdstates", ndim)
    E_ref, V_ref = cheby.fdstates(lambda x: H @ x, f, (emin, emax), fdc)
    assert [x[0] for x in recs] == [1, 2]
    assert np.abs(np.array([x[1] for x in recs]) - E_ref).max() < 1e-10 * span
    assert np.abs(np.stack([x[2] for x in recs], axis=1) - V_ref).max() < 1e-10


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_fvector_initial_state(tmp_path):
    """dpl_file: the F-vector of the operator matrix gives the states'
    transition moments and oscillator strengths (adc2_ener) and is the SIL
    initial state (contract_dipole_initial_state, statenumber 0); both and the
    'auto' file equal the oracle's values for the oracle's F-vector."""
    print("[case] fortran driver F-vector initial state o=3 v=7 naux=9", flush=True)
    o, v, naux = 3, 7, 9
    sys_ = make_system(o, v, naux, seed=11)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    d = np.random.default_rng(4).normal(0.0, 0.3, (o + v, o + v))
    d = d + d.T
    np.asfortranarray(d).T.astype("<f8").tofile(tmp_path / "d.bin")  # column-major bytes
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=2 "
           f"blocksize=4 maxvec=20 tol=1d-10 precon=1 davname='davstates' sil=.true. dpl_file='d.bin' "
           f"auto_file='auto' tout=0.1d0 tfinal=1.0d0 krydim=12 autotol=1d-10 /\n")
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    f = fvec.fvector(sys_, d)
    assert "F-vector" in r.stdout
    df = dfsigma.DF(sys_)
    H = df.sigma(np.eye(f.size))
    # transition moments and oscillator strengths of adc2_ener (sign of tm
    # follows the eigenvector's sign, so compare |tm|)
    w, U = np.linalg.eigh(H)
    rows = [ln.split() for ln in r.stdout.splitlines() if len(ln.split()) == 5 and ln.split()[0] in ("1", "2")]
    assert len(rows) == 2
    for i, row in enumerate(rows):
        tm = U[:, i] @ f
        assert abs(float(row[1]) - w[i]) < 1e-9
        assert abs(abs(float(row[3])) - abs(tm)) < 1e-7 * np.linalg.norm(f)
        assert abs(float(row[4]) - 2.0 / 3.0 * w[i] * tm ** 2) < 1e-7 * np.dot(f, f) * max(abs(w[i]), 1.0)
    lines = open(tmp_path / "auto").read().splitlines()
    data = np.array([[float(t) for t in ln.split()] for ln in lines[1:]])
    _, ref = sil.propagate(lambda x: -1j * (H @ x), f, tout=0.1, tfinal=1.0, kdim=12, autotol=1e-10)
    assert data.shape == (ref.size, 4)
    assert np.abs(data[:, 1] + 1j * data[:, 2] - ref).max() < 1e-10 * np.dot(f, f) + 1e-13


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_lanczos_double_subspace(tmp_path):
    """lancguess=8 (init_vec_doublesubspace, 800 ISs) with orthotype=2 at the
    water size: the Lanczos-state energies equal the oracle's block Lanczos
    from the oracle's guess."""
    from oracle import lanczos
    print("[case] fortran driver lancguess 8 water o=5 v=36 naux=164", flush=True)
    o, v, naux = 5, 36, 164
    sys_ = make_system(o, v, naux, seed=0)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    df = dfsigma.DF(sys_)
    n = df.n1 + df.D2.size
    f = np.zeros(n)
    f[: df.n1] = unit_dipole(o, v, seed=2)
    f.astype("<f8").tofile(tmp_path / "f.bin")
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=1 "
           f"blocksize=4 maxvec=16 tol=1d-6 precon=2 davname='davstates' fvec_file='f.bin' "
           f"lanczos=.true. lmain=2 ncycles=6 lancguess=8 orthotype=2 lancname='lancstates' /\n")
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr

    def hsub(sel):
        X = np.zeros((n, sel.size))
        X[sel, np.arange(sel.size)] = 1.0
        return df.sigma(X)[sel]

    q0 = lanczos.subspace_guess(df.diag(), hsub, 2, kind=8, subdim=800)
    ref = lanczos.block_lanczos(df.sigma, q0, 6, orthotype=2)
    recs = read_davstates(tmp_path / "lancstates", n)
    ev = np.array([x[1] for x in recs])
    assert ev.size == 12
    assert np.abs(ev - ref["eigval"]).max() < 1e-9 * np.abs(ref["eigval"]).max()


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_cvs(tmp_path):
    """ncore/core: CVS on core orbital 1.  The Davidson eigenvalues equal the
    lowest eigenvalues of the CVS sub-matrix, and the printed transition
    moments and oscillator strengths come from the CVS F-vector (the
    oracle's F-vector at the CVS configurations)."""
    from oracle import csf
    print("[case] fortran driver cvs core=1 o=3 v=7 naux=9", flush=True)
    o, v, naux = 3, 7, 9
    sys_ = make_system(o, v, naux, seed=11)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    d = np.random.default_rng(4).normal(0.0, 0.3, (o + v, o + v))
    d = d + d.T
    np.asfortranarray(d).T.astype("<f8").tofile(tmp_path / "d.bin")
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=2 "
           f"blocksize=4 maxvec=20 tol=1d-10 precon=1 davname='davstates' dpl_file='d.bin' "
           f"ncore=1 core=1 /\n")
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    idx = csf.cvs_indices(o, v, [0])
    df = dfsigma.DF(sys_)
    Hf = df.sigma(np.eye(df.n1 + df.D2.size))
    Hc = Hf[np.ix_(idx, idx)]
    w, U = np.linalg.eigh(Hc)
    f = fvec.fvector(sys_, d)[idx]
    recs = read_davstates(tmp_path / "davstates", idx.size)
    ev = np.array([x[1] for x in recs])
    assert np.abs(ev - w[:2]).max() < 1e-10
    rows = [ln.split() for ln in r.stdout.splitlines() if len(ln.split()) == 5 and ln.split()[0] in ("1", "2")]
    assert len(rows) == 2
    for i, row in enumerate(rows):
        tm = U[:, i] @ f
        assert abs(abs(float(row[3])) - abs(tm)) < 1e-7 * np.linalg.norm(f)
        assert abs(float(row[4]) - 2.0 / 3.0 * w[i] * tm ** 2) < 1e-7 * np.dot(f, f) * max(abs(w[i]), 1.0)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_state_table(tmp_path):
    """'davstates.dat' as table2/wrstateinfo write it: the MP2 ground-state
    line, per state the total and excitation energies, the transition moment
    and oscillator strength, and the dominant configurations (|C| >= 0.05 among
    the 50 largest) whose printed MO indices name configurations with exactly
    the printed coefficients (oracle/csf.py order, dense eigenvectors)."""
    from oracle import csf, mp2
    print("[case] fortran driver state table o=3 v=7 naux=9", flush=True)
    o, v, naux = 3, 7, 9
    sys_ = make_system(o, v, naux, seed=11)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    d = np.random.default_rng(4).normal(0.0, 0.3, (o + v, o + v))
    d = d + d.T
    np.asfortranarray(d).T.astype("<f8").tofile(tmp_path / "d.bin")
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=3 "
           f"blocksize=6 maxvec=30 tol=1d-10 precon=1 davname='davstates' dpl_file='d.bin' ehf=-1.5d0 /\n")
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    text = open(tmp_path / "davstates.dat").read()
    print(text[:1500])
    df = dfsigma.DF(sys_)
    H = df.sigma(np.eye(df.n1 + df.D2.size))
    w, U = np.linalg.eigh(H)
    e_mp2 = mp2.energy(sys_)
    f = fvec.fvector(sys_, d)
    lines = text.splitlines()
    gs = [ln for ln in lines if "Ground state MP2 energy:" in ln]
    assert len(gs) == 1 and abs(float(gs[0].split()[-1]) - (-1.5 + e_mp2)) < 1e-7
    typ, a, b, k, l = csf.table(o, v)
    key = {}
    for j in range(typ.size):
        if typ[j] == 0:
            key[(k[j] + 1, -1, a[j] + o + 1, -1)] = j
        else:
            key[(k[j] + 1, l[j] + 1, a[j] + o + 1, b[j] + o + 1, "II" if typ[j] == 5 else "I")] = j
    states = text.split("State")[1:]
    assert len(states) == 3
    for i, blk in enumerate(states):
        bl = blk.splitlines()
        assert abs(float(bl[0].split()[-1]) - (w[i] - 1.5 + e_mp2)) < 1e-7
        ex = [x for x in bl if "Excitation Energy (eV):" in x][0]
        assert abs(float(ex.split()[-1]) - w[i] * 27.2113845) < 1e-4
        tm = U[:, i] @ f
        tdm = [x for x in bl if "Transition Dipole Moment:" in x][0]
        assert abs(abs(float(tdm.split(":")[1])) - abs(tm)) < 1e-5
        rows = [x.split() for x in bl if "->" in x and "C_jkab" not in x]
        assert rows, blk
        big = np.sort(np.abs(U[:, i])[np.abs(U[:, i]) >= 0.05])[::-1][:50]
        assert len(rows) == big.size
        sgn = None
        for row in rows:
            if len(row) == 4:                      # k -> a C
                j = key[(int(row[0]), -1, int(row[2]), -1)]
            elif len(row) == 7:                    # k l -> a b spin C
                j = key[(int(row[0]), int(row[1]), int(row[3]), int(row[4]), row[5])]
            else:                                  # k l -> a b C
                kk, ll, aa, bb = int(row[0]), int(row[1]), int(row[3]), int(row[4])
                j = key[(kk, ll, aa, bb, "I")]
            c = float(row[-1])
            sgn = sgn if sgn is not None else np.sign(c * U[j, i])
            assert abs(c - sgn * U[j, i]) < 2e-5


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_adc2x_frozen_core(tmp_path):
    """adc2x = .true. with nfrozen = 1: the Davidson eigenvalues equal the
    lowest eigenvalues of the oracle's ADC(2)-x matrix on the frozen-core
    configurations (the reference's method = adc2-x with frozen_core)."""
    from oracle import adcx, csf, elements
    print("[case] fortran driver adc2x nfrozen=1 o=3 v=6 naux=9", flush=True)
    o, v, naux = 3, 6, 9
    sys_ = make_system(o, v, naux, seed=12)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=3 "
           f"blocksize=5 maxvec=30 tol=1d-10 precon=2 davname='davstates' adc2x=.true. nfrozen=1 /\n")
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ADC(2)-x" in r.stdout
    blk = elements.build_blocks(sys_, apply_minc=False)
    n1 = blk.n1
    H = np.zeros((blk.ndim, blk.ndim))
    H[:n1, :n1] = blk.H11
    H[:n1, n1:] = blk.C
    H[n1:, :n1] = blk.C.T
    H[n1:, n1:] = np.diag(blk.D) + adcx.block22_dense(sys_)
    idx = csf.frozen_core_indices(o, v, [0])
    w = np.linalg.eigvalsh(H[np.ix_(idx, idx)])
    recs = read_davstates(tmp_path / "davstates", idx.size)
    ev = np.array([x[1] for x in recs])
    assert np.abs(ev - w[:3]).max() < 1e-10


def _write_is_operator(path, diag, row, col, val, bs=16):
    """The reference's SCRATCH/dipole_<c> / SCRATCH/cap layout (Fortran
    sequential records: buf_size, the diagonal, then (oij, indxi, indxj, nlim)
    with 1-based indices), the last record partly filled."""
    def rec(payload):
        n = struct.pack("<i", len(payload))
        return n + payload + n
    out = rec(struct.pack("<i", bs)) + rec(np.asarray(diag, "<f8").tobytes())
    for k0 in range(0, len(row), bs):
        k1 = min(len(row), k0 + bs)
        o = np.zeros(bs)
        i = np.zeros(bs, "<i4")
        j = np.zeros(bs, "<i4")
        o[: k1 - k0], i[: k1 - k0], j[: k1 - k0] = val[k0:k1], row[k0:k1] + 1, col[k0:k1] + 1
        out += rec(o.astype("<f8").tobytes() + i.tobytes() + j.tobytes() + struct.pack("<i", k1 - k0))
    with open(path, "wb") as f:
        f.write(out)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_laser_cap(tmp_path):
    """laser = .true.: the driver reads IS dipole and CAP operators in the
    reference's file layout, propagates the ground state under two pulses
    with the CAP (Lanczos-Arnoldi steps, ground state projected out of the
    CAP) and prints the norms; norms and final wavefunction equal the
    oracle's propagation (oracle/laser.py) of the same operators."""
    from oracle import laser
    print("[case] fortran driver laser + CAP o=3 v=5 naux=8", flush=True)
    o, v, naux = 3, 5, 8
    sys_ = make_system(o, v, naux, seed=21)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    df = dfsigma.DF(sys_)
    nd = df.n1 + df.D2.size
    H = df.sigma(np.eye(nd))
    H = 0.5 * (H + H.T)
    rng = np.random.default_rng(21)
    dips = [None, None, None]
    d00 = [0.3, 0.0, -0.2]
    for c in (0, 2):
        row = rng.integers(0, nd, 3 * nd)
        col = (row + 1 + rng.integers(0, nd - 1, row.size)) % nd
        val = rng.normal(0.0, 0.05, row.size)
        diag = rng.normal(0.0, 0.2, nd)
        d0j = rng.normal(0.0, 0.1, nd)
        _write_is_operator(tmp_path / f"dip{c}", diag, row, col, val)
        d0j.astype("<f8").tofile(tmp_path / f"d0j{c}.bin")
        dips[c] = (laser.dense_dipole(diag, row, col, val, nd), d00[c], d0j)
    B = rng.normal(0.0, 0.08, (nd + 1, 3))
    Wf = B @ B.T
    w00 = float(Wf[nd, nd])
    W = Wf[:nd, :nd] - w00 * np.eye(nd)
    iu = np.triu_indices(nd, 1)
    _write_is_operator(tmp_path / "cap", np.diag(W), iu[0], iu[1], W[iu], bs=1000)
    Wf[:nd, nd].astype("<f8").tofile(tmp_path / "w0j.bin")
    trow = rng.integers(0, nd, 2 * nd)
    tcol = (trow + 1 + rng.integers(0, nd - 1, trow.size)) % nd
    tval = rng.normal(0.0, 0.1, trow.size)
    tdiag = rng.uniform(0.0, 1.0, nd)
    theta0j = rng.normal(0.0, 0.1, nd)
    _write_is_operator(tmp_path / "theta", tdiag, trow, tcol, tval)
    theta0j.astype("<f8").tofile(tmp_path / "theta0j.bin")
    theta = (laser.dense_dipole(tdiag, trow, tcol, tval, nd), 0.15, theta0j)
    pulses = [{"envelope": "cos2", "envpar": (2.0, 2.0), "carrier": "cos", "freq": 0.7, "t0": 0.5, "phase": 0.2,
               "pulse_vec": (0.8, 0.0, 0.6), "strength": 0.2},
              {"envelope": "box", "envpar": (1.0, 3.0), "carrier": "sin", "freq": 1.3,
               "pulse_vec": (0.0, 0.0, 1.0), "strength": 0.05}]
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=2 blocksize=2 "
           f"maxvec=20 tol=1d-6 laser=.true. tout=0.5d0 tfinal=4d0 krydim=16 proptol=1d-9 "
           f"dipis_file='dip0','','dip2' d0j_file='d0j0.bin','','d0j2.bin' d00=0.3d0,0d0,-0.2d0 "
           f"cap_file='cap' w0j_file='w0j.bin' w00={w00!r} cap_ground=.true. npulse=2 "
           f"carrier=1,2 envelope=1,3 pulse_vec=0.8d0,0d0,0.6d0,0d0,0d0,1d0 strength=0.2d0,0.05d0 "
           f"freq=0.7d0,1.3d0 t0=0.5d0,0d0 phase=0.2d0,0d0 envpar=2d0,2d0,1d0,3d0 psi_file='psi.bin' "
           f"theta_file='theta' theta0j_file='theta0j.bin' theta00=0.15d0 trdens_file='trdens.bin' /\n")
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-1500:])
    assert r.returncode == 0, r.stdout + r.stderr
    norms = np.array([float(x.split()[3]) for x in r.stdout.splitlines() if x.strip().startswith("Time:")])
    psi = np.fromfile(tmp_path / "psi.bin", dtype="<c16")
    psi0 = np.zeros(nd + 1, dtype=complex)
    psi0[nd] = 1.0
    ref_psi, ref_norms, ref_flux = laser.propagate(lambda z: H @ z, dips, pulses, psi0, 0.5, 4.0, 16, 1e-9,
                                                   (W, w00, Wf[:nd, nd]), (None, True), theta)
    assert norms.size == ref_norms.size == 9
    assert np.abs(norms - ref_norms).max() < 2e-9     # printed with 9 decimals
    assert np.abs(psi - ref_psi).max() < 1e-8
    assert norms[-1] < 1.0 - 1e-6
    # flux.dat: three '#' header lines, then time and flux (ES15.8)
    rows = [x.split() for x in (tmp_path / "flux.dat").read_text().splitlines() if not x.startswith("#")]
    t, flux = np.array([float(a) for a, _ in rows]), np.array([float(b) for _, b in rows])
    assert np.allclose(t, 0.5 * np.arange(9))
    assert np.abs(flux - ref_flux).max() < 1e-9 + 1e-8 * np.abs(ref_flux).max()
    assert np.abs(flux).max() > 1e-6
    # trdens_file: the TD-NTO input (nvirt x nocc complex per record); the
    # last record is the density of the final wavefunction
    from oracle import nto
    trd = np.fromfile(tmp_path / "trdens.bin", dtype="<c16").reshape(9, o, v).transpose(0, 2, 1)
    Tr = nto.tdnto(sys_, ref_psi[:nd])[0]
    assert np.abs(trd[-1] - Tr).max() < 1e-8 * np.abs(Tr).max()
    assert np.abs(trd[0]).max() == 0.0                 # psi_IS = 0 at t = 0


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_laser_excited_state(tmp_path):
    """statenumber = 1 starts the laser propagation from the lowest Davidson
    state (initwf_exci_initstate) and cap_nproj = 2 projects the two lowest
    states out of the CAP (iprojcap); norms and final wavefunction equal the
    oracle's propagation from the eigenvectors the driver wrote."""
    from oracle import laser
    print("[case] fortran driver laser from an excited state o=3 v=5 naux=8", flush=True)
    o, v, naux = 3, 5, 8
    sys_ = make_system(o, v, naux, seed=22)
    sys_.eps.astype("<f8").tofile(tmp_path / "eps.bin")
    sys_.B.astype("<f8").tofile(tmp_path / "B.bin")
    df = dfsigma.DF(sys_)
    nd = df.n1 + df.D2.size
    H = df.sigma(np.eye(nd))
    H = 0.5 * (H + H.T)
    rng = np.random.default_rng(22)
    row = rng.integers(0, nd, 3 * nd)
    col = (row + 1 + rng.integers(0, nd - 1, row.size)) % nd
    val = rng.normal(0.0, 0.05, row.size)
    diag = rng.normal(0.0, 0.2, nd)
    d0j = rng.normal(0.0, 0.1, nd)
    _write_is_operator(tmp_path / "dipz", diag, row, col, val)
    d0j.astype("<f8").tofile(tmp_path / "d0jz.bin")
    dips = [None, None, (laser.dense_dipole(diag, row, col, val, nd), 0.1, d0j)]
    wd = np.abs(rng.normal(0.0, 0.05, nd))
    _write_is_operator(tmp_path / "cap", wd, np.zeros(0, int), np.zeros(0, int), np.zeros(0))
    np.zeros(nd).astype("<f8").tofile(tmp_path / "w0j.bin")
    pulses = [{"envelope": "sin2", "envpar": (0.0, 1.5), "carrier": "cos", "freq": 0.9,
               "pulse_vec": (0.0, 0.0, 1.0), "strength": 0.1}]
    nml = (f"&adc2g nocc={o} nvirt={v} naux={naux} eps_file='eps.bin' b_file='B.bin' nstates=3 blocksize=4 "
           f"maxvec=24 tol=1d-9 davname='davstates' laser=.true. tout=0.5d0 tfinal=3d0 krydim=16 proptol=1d-9 "
           f"dipis_file='','','dipz' d0j_file='','','d0jz.bin' d00=0d0,0d0,0.1d0 "
           f"cap_file='cap' w0j_file='w0j.bin' w00=0d0 cap_nproj=2 statenumber=1 npulse=1 "
           f"carrier=1 envelope=4 pulse_vec=0d0,0d0,1d0 strength=0.1d0 freq=0.9d0 envpar=0d0,1.5d0 "
           f"psi_file='psi.bin' /\n")
    r = subprocess.run([DRIVER], input=nml, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-1200:])
    assert r.returncode == 0, r.stdout + r.stderr
    recs = read_davstates(tmp_path / "davstates", nd)
    X = np.stack([x[2] for x in recs], axis=1)
    norms = np.array([float(x.split()[3]) for x in r.stdout.splitlines() if x.strip().startswith("Time:")])
    psi = np.fromfile(tmp_path / "psi.bin", dtype="<c16")
    psi0 = np.zeros(nd + 1, dtype=complex)
    psi0[:nd] = X[:, 0]
    ref_psi, ref_norms = laser.propagate(lambda z: H @ z, dips, pulses, psi0, 0.5, 3.0, 16, 1e-9,
                                         (np.diag(wd), 0.0, np.zeros(nd)), (X[:, :2], False))
    assert norms.size == ref_norms.size == 7
    assert np.abs(norms - ref_norms).max() < 2e-9
    assert np.abs(psi - ref_psi).max() < 1e-8

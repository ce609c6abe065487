"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests or fixtures for this path, so every restatement
in ``oracle/`` is checked here against an independent fact:

* configuration tables: per-type counts and bijectivity;
* first-order ph-ph block: the singlet eigenvalues must be a subset of an
  independently built spin-orbital CIS spectrum;
* zero-coupling limit: with all factors zero the spectrum is the union of the
  orbital-energy differences;
* full matrix: the literal element restatement (``elements``) and the
  density-fitted reformulation (``dfsigma``) agree element by element;
* Davidson: eigenvalues match ARPACK and the Lowdin-partitioned roots;
* SIL: the autocorrelation matches exp(-iHt) from a dense eigendecomposition;
* Chebyshev: the order-domain autocorrelation matches sum_n (v_n.q0)^2 T_k(x_n);
* block Lanczos: the three-term recurrence, orthonormality, exact block Gauss
  quadrature and the full-Krylov spectrum;
* relaxation: each timestep against exp(-H dtau) from scipy.linalg.expm, and
  the relaxed states against a dense eigendecomposition;
* network: analytic gradients match central finite differences;
* coupled preconditioner: the step solves the shifted system it stands for,
  and the Schur complement's inertia counts the eigenvalues of H below the
  shift.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from adc_ottawa_amd.synth import make_system
from oracle import cheby, coupled, csf, davidson, dfsigma, elements, fvec, lanczos, mlp, mp2, power, relax, schur, sil


def test_csf_counts_and_bijectivity():
    for o, v in [(2, 3), (3, 5), (5, 36)]:
        c = csf.counts(o, v)
        typ, a, b, k, l = csf.table(o, v)
        assert typ.size == c["ndim"]
        for t, name in zip(range(6), ["n1", "t1", "t2", "t3", "t4", "t5"]):
            assert np.sum(typ == t) == c[name]
        key = set(zip(typ.tolist(), a.tolist(), b.tolist(), k.tolist(), l.tolist()))
        assert len(key) == c["ndim"]
    c = csf.counts(16, 200)
    assert (c["n1"], c["t2"], c["t3"], c["t4"], c["n2"]) == (3200, 318400, 24000, 2388000, 5121600)


def spin_orbital_cis(sys):
    o, v = sys.nocc, sys.nvirt
    V = elements.eri(sys.B)
    e = sys.eps
    occ = [(p, s) for p in range(o) for s in (0, 1)]
    vir = [(p, s) for p in range(o, o + v) for s in (0, 1)]
    exc = [(i, a) for i in occ for a in vir]
    n = len(exc)
    A = np.zeros((n, n))
    for x, (i, a) in enumerate(exc):
        for y, (j, b) in enumerate(exc):
            val = 0.0
            if i == j and a == b:
                val += e[a[0]] - e[i[0]]
            if a[1] == i[1] and j[1] == b[1]:
                val += V[a[0], i[0], j[0], b[0]]
            if a[1] == b[1] and j[1] == i[1]:
                val -= V[a[0], b[0], j[0], i[0]]
            A[x, y] = val
    return np.linalg.eigvalsh(A)


def test_first_order_block_is_singlet_cis():
    sys = make_system(3, 4, 6, seed=3)
    blk = elements.build_blocks(sys, second_order=False)
    singlet = np.linalg.eigvalsh(blk.H11)
    so = spin_orbital_cis(sys)
    for w in singlet:
        assert np.min(np.abs(so - w)) < 1e-12


def test_zero_coupling_limit():
    sys = make_system(3, 5, 4, seed=1)
    sys.B[:] = 0.0
    blk = elements.build_blocks(sys)
    o, v = 3, 5
    e = sys.eps
    expect = np.concatenate([(e[o:, None] - e[None, :o]).ravel(), dfsigma.diag2(sys)])
    got = np.linalg.eigvalsh(blk.dense())
    assert np.allclose(np.sort(expect), got, atol=1e-14)
    df = dfsigma.DF(sys)
    assert np.allclose(df.diag(), blk.diag(), atol=1e-14)


@pytest.mark.parametrize("o,v,naux,seed", [(2, 3, 2, 0), (3, 5, 7, 1), (4, 7, 11, 2), (5, 9, 13, 3)])
def test_literal_vs_density_fitted(o, v, naux, seed):
    sys = make_system(o, v, naux, seed=seed)
    blk = elements.build_blocks(sys, apply_minc=False)
    df = dfsigma.DF(sys)
    scale = np.abs(blk.H11).max()
    assert np.abs(blk.H11 - df.H11).max() < 1e-13 * scale
    assert np.abs(blk.H11 - blk.H11.T).max() < 1e-13 * scale
    # coupling block column by column through the DF sigma
    n1 = o * v
    Hd = blk.dense()
    Hdf = df.sigma(np.eye(blk.ndim))
    assert np.abs(Hd - Hdf).max() < 1e-13 * np.abs(Hd).max()
    assert n1 == blk.n1


def test_oracle_eigenvalues_cfg_water():
    sys = make_system(5, 36, 164, seed=0)
    blk = elements.build_blocks(sys)
    op = sla.LinearOperator((blk.ndim, blk.ndim), matvec=blk.sigma, dtype=np.float64)
    w = np.sort(sla.eigsh(op, k=8, which="SA", tol=1e-13)[0])[:5]
    nb = schur.count_below(blk.H11, blk.C, blk.D, 5)
    assert nb >= 1
    ex = schur.roots(blk.H11, blk.C, blk.D, nb)
    assert np.abs(ex - w[:nb]).max() < 1e-12
    r = davidson.davidson(blk.sigma, blk.diag(), 5, blocksize=5, maxvec=40, tol=1e-7,
                          precon=1, deflate=True, maxiter=400)
    assert np.abs(r["eigval"] - w).max() < 1e-9


@pytest.mark.parametrize("precon,deflate", [(1, False), (2, True)])
def test_oracle_davidson_small(precon, deflate):
    sys = make_system(3, 8, 10, seed=4)
    df = dfsigma.DF(sys)
    H = df.sigma(np.eye(df.n1 + df.D2.size))
    w = np.linalg.eigvalsh(H)[:4]
    r = davidson.davidson(df.sigma, df.diag(), 4, tol=1e-8, precon=precon, deflate=deflate, maxiter=400)
    assert np.abs(r["eigval"] - w).max() < 1e-10
    # Ritz vectors are eigenvectors
    X = r["eigvec"]
    assert np.abs(H @ X - X * r["eigval"]).max() < 1e-7


def test_oracle_sil_exact_propagator():
    sys = make_system(3, 5, 6, seed=5)
    df = dfsigma.DF(sys)
    H = df.sigma(np.eye(df.n1 + df.D2.size))
    w, U = np.linalg.eigh(H)
    f = np.zeros(H.shape[0])
    f[: df.n1] = np.random.default_rng(0).normal(0, 0.1, df.n1)
    def func(x):
        return -1j * (H @ x)
    times, auto = sil.propagate(func, f, tout=0.1, tfinal=4.0, kdim=12, autotol=1e-12)
    c = U.T @ f
    exact = np.array([np.sum(c * c * np.exp(-1j * w * t)) for t in times])
    assert np.abs(auto - exact).max() < 1e-9 * np.sum(c * c)
    # t = 0 record is |F|^2
    assert abs(auto[0] - np.dot(f, f)) < 1e-14


def test_oracle_sil_higher_orders_exact():
    """Autocorrelation functions of orders 1 and 2 against the closed form
    sum_n c_n^2 E_n^k exp(-i E_n t) (the t/2 trick holds for H^k because H is
    real symmetric); a missing factor i, a conjugated bilinear form or a
    wrong power fails it."""
    sys = make_system(3, 5, 6, seed=5)
    df = dfsigma.DF(sys)
    H = df.sigma(np.eye(df.n1 + df.D2.size))
    w, U = np.linalg.eigh(H)
    f = np.zeros(H.shape[0])
    f[: df.n1] = np.random.default_rng(0).normal(0, 0.1, df.n1)
    times, autos = sil.propagate(lambda x: -1j * (H @ x), f, tout=0.1, tfinal=4.0, kdim=12,
                                 autotol=1e-12, autoord=2)
    c = U.T @ f
    for k in (0, 1, 2):
        exact = np.array([np.sum(c * c * w ** k * np.exp(-1j * w * t)) for t in times])
        assert np.abs(autos[k] - exact).max() < 1e-9 * np.sum(c * c * np.abs(w) ** k), k


def test_oracle_chebyshev_closed_form():
    """The Chebyshev recursion restatement against sum_n (v_n.q0)^2 T_k(x_n)
    from a dense eigendecomposition, on every order including the doubled ones
    (C_2k = 2 q_k.q_k - C_0); a wrong recursion coefficient, scaling or
    doubling formula fails it."""
    sys = make_system(3, 6, 9, seed=4)
    blk = elements.build_blocks(sys, apply_minc=False)
    H = blk.dense()
    lam = np.linalg.eigvalsh(H)
    bounds = (lam[0] - 0.05 * (lam[-1] - lam[0]), lam[-1] + 0.05 * (lam[-1] - lam[0]))
    q0 = np.random.default_rng(3).standard_normal(H.shape[0])
    for order in (1, 7, 20):
        got = cheby.chebyshev_auto(lambda x: H @ x, q0, bounds, order)
        ref = cheby.closed_form(H, q0, bounds, 2 * order + 1)
        assert np.abs(got - ref).max() < 1e-10 * np.abs(ref[0])
    with pytest.raises(ValueError):
        cheby.closed_form(H, q0, (lam[0], 0.5 * (lam[0] + lam[-1])), 3)


def _unit_block(n, idx):
    q = np.zeros((n, len(idx)))
    q[idx, np.arange(len(idx))] = 1.0
    return q


def test_oracle_block_lanczos_recurrence_and_quadrature():
    """The block Lanczos restatement against what the mathematics fixes:
    H L = L T + Q_{m+1} B_m E_m^T to rounding (a transposed B_{j-1}, a wrong
    sign or a wrong diagonal block breaks it); L^T L = I over a few cycles;
    exact block Gauss quadrature sum_n (f.r_n)^2 theta_n^k = f^T H^k f for
    f in span(Q_1) and k < 2m (a wrong Ritz normalisation or a wrong
    eigenvector breaks it)."""
    sys = make_system(4, 12, 20, seed=6)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    n = H.shape[0]
    b, m = 4, 3
    q0 = _unit_block(n, [0, 5, 17, 30])
    out = lanczos.block_lanczos(lambda X: H @ X, q0, m)
    L, T = out["L"], out["T"]
    E = np.zeros((b * m, b))
    E[-b:, :] = np.eye(b)
    resid = H @ L - L @ T - out["Qnext"] @ out["Bnext"] @ E.T
    hn = np.abs(H).max()
    assert np.abs(resid).max() < 1e-11 * hn * n ** 0.5
    assert np.abs(L.T @ L - np.eye(b * m)).max() < 1e-11
    assert np.allclose(np.tril(out["Bnext"], -1), 0.0)
    f = q0 @ np.array([0.3, -1.1, 0.7, 0.2])
    tm = lanczos.tranmom(out["ritz"], f)
    th = out["eigval"]
    hk = f.copy()
    for k in range(2 * m):
        lhs = (tm ** 2 * th ** k).sum()
        rhs = f @ hk
        assert abs(lhs - rhs) < 1e-9 * (np.abs(tm) ** 2 * np.abs(th) ** k).sum()
        hk = H @ hk


def test_oracle_block_lanczos_partial_reorthogonalisation():
    """orthotype 1 on a run long enough to lose orthogonality without it.
    Without it, max |L^T L - I| is O(1) and a ghost copy of a converged
    eigenvalue appears; with it, the stored vectors stay orthogonal to well
    below sqrt(eps), re-orthogonalisation happens, and every extreme
    eigenvalue appears once and matches eigh.  A wrong omega recurrence that
    never triggers, or a pro_mgs that misses a block, fails it."""
    sys = make_system(3, 8, 10, seed=5)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    n = H.shape[0]
    lam = np.linalg.eigvalsh(H)
    q0 = _unit_block(n, [0, 5, 11])
    out0 = lanczos.block_lanczos(lambda X: H @ X, q0, 40, orthotype=0)
    out1 = lanczos.block_lanczos(lambda X: H @ X, q0, 40, orthotype=1)
    L0, L1 = out0["L"], out1["L"]
    assert np.abs(L0.T @ L0 - np.eye(L0.shape[1])).max() > 1e-2
    assert out1["npro"] > 0
    assert np.abs(L1.T @ L1 - np.eye(L1.shape[1])).max() < 1e-10
    ev = out1["eigval"]
    assert np.abs(ev[:3] - lam[:3]).max() < 1e-10 and np.abs(ev[-3:] - lam[-3:]).max() < 1e-10
    assert (np.abs(ev - lam[-1]) < 1e-6).sum() == 1
    # orthotype 2 (mpro): re-orthogonalisations happen, the stored vectors stay
    # orthogonal below sqrt(eps) and the extreme eigenvalues appear once
    out2 = lanczos.block_lanczos(lambda X: H @ X, q0, 40, orthotype=2)
    L2, ev2 = out2["L"], out2["eigval"]
    assert out2["npro"] > 0
    assert np.abs(L2.T @ L2 - np.eye(L2.shape[1])).max() < 1e-10
    assert np.abs(ev2[:3] - lam[:3]).max() < 1e-10 and np.abs(ev2[-3:] - lam[-3:]).max() < 1e-10
    assert (np.abs(ev2 - lam[-1]) < 1e-6).sum() == 1


def test_oracle_mpro_bounds():
    """lowerbound/upperbound of mpro as the reference's loops run: the lower
    loop (do i=k-2,1) only runs for k = 3."""
    om = np.zeros((12, 12))
    j, eta = 9, lanczos.EPMACH ** 0.75
    om[j + 1, 1:9] = [1e-20, 1, 1, 1, 1, 1, 1e-20, 1]
    assert lanczos._lowerbound(om, j, 1, eta) == 1 and lanczos._lowerbound(om, j, 2, eta) == 1
    assert lanczos._lowerbound(om, j, 3, eta) == 2          # om(j+1,1) below eta
    om[j + 1, 1] = 1.0
    assert lanczos._lowerbound(om, j, 3, eta) == 1
    assert lanczos._lowerbound(om, j, 6, eta) == 5          # no loop for k >= 4
    assert lanczos._upperbound(om, j, 2, eta) == 6          # 3..6 above eta, 7 below
    assert lanczos._upperbound(om, j, 7, eta) == 8          # starts at k + 1 whatever om(j+1,k+1)


def test_oracle_lanczos_subspace_guess():
    """init_vec_doublesubspace: with subdim = ndim both subspaces are the whole
    space, so the columns are eigenvectors 1..bs/2 and bs/2+1..bs of H; with
    small disjoint subspaces the columns are orthonormal, supported on the
    selected ISs, and the Ritz values of the block lie in the spectrum."""
    rng = np.random.default_rng(3)
    n, b = 40, 4
    G = rng.standard_normal((n, n))
    H = G + G.T + np.diag(np.linspace(-20, 20, n))
    lam, U = np.linalg.eigh(H)
    hii = np.diag(H).copy()
    V = lanczos.subspace_guess(hii, lambda s: H[np.ix_(s, s)], b, kind=8, subdim=n)
    for c in range(b):
        assert min(np.abs(V[:, c] - U[:, c]).max(), np.abs(V[:, c] + U[:, c]).max()) < 1e-10
    V = lanczos.subspace_guess(hii, lambda s: H[np.ix_(s, s)], b, kind=8, subdim=10)
    assert np.abs(V.T @ V - np.eye(b)).max() < 1e-12
    lo, hi = np.argsort(hii)[:10], np.argsort(-hii)[:10]
    assert np.abs(np.delete(V[:, :2], lo, axis=0)).max() == 0.0
    assert np.abs(np.delete(V[:, 2:], hi, axis=0)).max() == 0.0
    # the second half holds eigenvectors 3 and 4 of the high block, not its lowest
    A = H[np.ix_(hi, hi)]
    w = np.linalg.eigvalsh(A)
    assert np.allclose(np.sort(np.diag(V[:, 2:].T @ H @ V[:, 2:])), w[2:4])
    V9 = lanczos.subspace_guess(hii, lambda s: H[np.ix_(s, s)], 2, kind=9, subdim=10)
    assert np.allclose(np.sort(np.diag(V9.T @ H @ V9)), w[:2])


def test_oracle_block_lanczos_full_krylov():
    """With blocksize * ncycles = ndim the Krylov space is the whole space and
    the spectrum of T is the spectrum of H."""
    rng = np.random.default_rng(2)
    n, b = 24, 3
    G = rng.standard_normal((n, n))
    H = G + G.T
    q0, _ = np.linalg.qr(rng.standard_normal((n, b)))
    out = lanczos.block_lanczos(lambda X: H @ X, q0, n // b)
    assert np.abs(out["eigval"] - np.linalg.eigvalsh(H)).max() < 1e-9 * np.abs(H).max()
    # the Ritz vectors are then eigenvectors of H
    R = out["ritz"]
    assert np.abs(H @ R - R * out["eigval"]).max() < 1e-8 * np.abs(H).max()


def test_oracle_cvs_indices():
    """The CVS loop restatement against a brute-force filter of the full
    configuration table: 1h1p with a core hole, t3/t4/t5 with exactly one
    core hole among i < j, nothing of t1/t2, in ascending full-space order
    (the CVS loops nest like the full ones).  Counts: n1 = nc v, t3 = nc (o-nc) v,
    t4 = t5 = nc (o-nc) v (v-1)/2."""
    for o, v, core in [(4, 5, [0]), (5, 6, [1, 3]), (3, 4, [0, 1])]:
        typ, a, b, k, l = csf.table(o, v)
        isc = np.zeros(o, dtype=bool)
        isc[core] = True
        keep = np.where(typ == 0, isc[k], False)
        two = (typ >= 3) & (isc[k].astype(int) + isc[l].astype(int) == 1)
        ref = np.flatnonzero(keep | two)
        got = csf.cvs_indices(o, v, core)
        assert np.array_equal(got, ref)
        nc = len(core)
        assert (typ[got] == 0).sum() == nc * v
        assert (typ[got] == 3).sum() == nc * (o - nc) * v
        assert (typ[got] == 4).sum() == (typ[got] == 5).sum() == nc * (o - nc) * v * (v - 1) // 2


def test_oracle_mp2():
    """MP2 energy against the spin-orbital formula 1/4 sum |<ij||ab>|^2 / D;
    d2 against the largest singular values of the amplitude matricisations;
    the density's oo/uu blocks against the closed-shell textbook forms and
    its trace (the second-order correction is traceless); the assembly
    doubles the off-diagonal oo elements as the reference's loop does."""
    sys = make_system(3, 5, 7, seed=2)
    o, v, e = sys.nocc, sys.nvirt, sys.eps
    V = mp2.integrals(sys)
    occ = [(p, s) for p in range(o) for s in (0, 1)]
    vir = [(p, s) for p in range(o, o + v) for s in (0, 1)]

    def asym(p, q, r, s):
        a = V[p[0], r[0], q[0], s[0]] if (p[1] == r[1] and q[1] == s[1]) else 0.0
        b = V[p[0], s[0], q[0], r[0]] if (p[1] == s[1] and q[1] == r[1]) else 0.0
        return a - b

    E = sum(0.25 * asym(i, j, a, b) ** 2 / (e[i[0]] + e[j[0]] - e[a[0]] - e[b[0]])
            for i in occ for j in occ for a in vir for b in vir)
    assert abs(mp2.energy(sys, V) - E) < 1e-14 * abs(E)
    t = mp2.amplitudes(sys, V)
    so = np.linalg.svd(t.reshape(o, -1), compute_uv=False)[0]
    sv = np.linalg.svd(t.transpose(2, 0, 1, 3).reshape(v, -1), compute_uv=False)[0]
    assert abs(mp2.d2(sys, V) - max(so, sv)) < 1e-13
    rho, blk = mp2.density(sys, V)
    Poo = -2.0 * np.einsum("ikab,jkab->ij", t, 2.0 * t - t.transpose(0, 1, 3, 2))
    Puu = 2.0 * np.einsum("ijac,ijbc->ab", t, 2.0 * t - t.transpose(0, 1, 3, 2))
    assert np.abs(blk["oo"] - Poo).max() < 1e-13
    assert np.abs(blk["uu"] - Puu).max() < 1e-13
    assert abs(np.trace(blk["oo"]) + np.trace(blk["uu"])) < 1e-14
    off = ~np.eye(o, dtype=bool)
    assert np.abs(rho[:o, :o][off] - 2.0 * blk["oo"][off]).max() < 1e-15
    assert np.abs(np.diag(rho)[:o] - 2.0 - np.diag(blk["oo"])).max() < 1e-15


def test_oracle_power_method():
    """The power-method restatement converges to the largest eigenvalue of the
    dense matrix (o=3, v=8 with a 20-IS guess: about 200 iterations); with
    the whole space as the guess sub-space it starts on the exact top
    eigenvector and stops at the first iteration."""
    sys = make_system(3, 8, 10, seed=5)
    Hm = elements.build_blocks(sys, apply_minc=True).dense()
    H = elements.build_blocks(sys, apply_minc=False).dense()
    lam = np.linalg.eigvalsh(H)
    hs = lambda s: Hm[np.ix_(s, s)]
    e, v, it, ok = power.power_method(lambda x: H @ x, power.guess(np.diag(H), hs, subdim=20))
    assert ok and 100 < it < 350 and abs(e - lam[-1]) < 1e-12 * abs(lam[-1])
    e, v, it, ok = power.power_method(lambda x: H @ x, power.guess(np.diag(H), hs, subdim=H.shape[0]))
    assert ok and it == 1


def test_oracle_subdiag_guess():
    """initvec_subdiag: with subdim = ndim the guess vectors are the exact
    lowest eigenvectors, so the Davidson converges in the first iteration;
    a tie at the boundary extends the selection over the whole degenerate
    group; the vectors live on the ISs of lowest H_jj."""
    sys = make_system(3, 8, 10, seed=4)
    H = elements.build_blocks(sys, apply_minc=True).dense()
    hii = np.diag(H).copy()
    n = H.shape[0]
    hs = lambda sel: H[np.ix_(sel, sel)]
    lam, U = np.linalg.eigh(H)
    V = davidson.subdiag_guess(hii, hs, 4, subdim=n)
    for c in range(4):
        assert min(np.abs(V[:, c] - U[:, c]).max(), np.abs(V[:, c] + U[:, c]).max()) < 1e-10
    out = davidson.davidson(lambda X: H @ X, hii, 4, blocksize=4, tol=1e-8, guess=V)
    assert out["niter"] == 1 and np.abs(out["eigval"] - lam[:4]).max() < 1e-12
    # support: the 30 lowest H_jj
    V = davidson.subdiag_guess(hii, hs, 4, subdim=30)
    sel = np.argsort(hii, kind="stable")[:30]
    assert np.abs(np.delete(V, sel, axis=0)).max() == 0.0
    # a degenerate group at the boundary is kept whole
    h2 = hii.copy()
    o = np.argsort(h2, kind="stable")
    h2[o[30]] = h2[o[29]] + 1e-8
    V = davidson.subdiag_guess(h2, hs, 4, subdim=30)
    assert np.abs(V[o[30]]).max() > 0.0
    assert np.abs(np.delete(V, o[:31], axis=0)).max() == 0.0


def test_oracle_relax_steps_are_imaginary_time_propagation():
    """Each XSIL timestep propagates psi to exp(-H dtau) psi (normalised), up to
    the Krylov truncation error, which is negligible here (dtau * spread ~ 1,
    12 Lanczos vectors).  Steps 2-4 use the inter-timestep matrix elements
    and step 3 follows a subspace collapse, so a wrong H_ij or S_ij, a wrong
    canonical orthogonalisation or a wrong collapse fails it."""
    from scipy.linalg import expm
    sys = make_system(3, 8, 10, seed=5)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    lam = np.linalg.eigvalsh(H)
    dtau = 1.0 / (lam[-1] - lam[0])
    g = relax.ondiag_guess(np.diag(H), 1)
    out = relax.relax(lambda x: H @ x, g, krydim=12, maxiter=4, step=dtau, eps=0.0, trace=True)
    P = expm(-dtau * H)
    psi = g[:, 0]
    for k, got in enumerate(out["trace"][0]["psi"]):
        ref = P @ psi
        ref /= np.linalg.norm(ref)
        assert np.abs(got - ref).max() < 1e-10, k
        psi = got


def test_oracle_relax_converges_to_lowest_states():
    """The relaxed states are the lowest eigenpairs (each state's first
    Lanczos vector is orthogonalised against the states already converged)."""
    sys = make_system(3, 8, 10, seed=5)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    lam, U = np.linalg.eigh(H)
    g = relax.ondiag_guess(np.diag(H), 3)
    out = relax.relax(lambda x: H @ x, g, krydim=10, maxiter=300, step=30.0, eps=1e-7)
    assert (out["residuals"] <= 1e-7).all()
    # energy error ~ residual^2 / gap, vector error ~ residual / gap
    assert np.abs(out["energies"] - lam[:3]).max() < 1e-12 * np.abs(lam).max()
    for s in range(3):
        v = out["vectors"][:, s]
        assert min(np.abs(v - U[:, s]).max(), np.abs(v + U[:, s]).max()) < 2e-5


def test_oracle_mlp_gradients():
    rng = np.random.default_rng(1)
    p = mlp.init_params(0)
    p["w3"] = rng.normal(0, 0.1, mlp.H)
    n = 37
    x = mlp.transform(rng.normal(0, 2, n), rng.normal(1, 0.2, n), rng.normal(0, 1, n))
    cls = rng.integers(0, mlp.NCLASS, n)
    dn = rng.normal(0, 1, n)
    en = rng.normal(0, 1, n)
    g = mlp.backward(p, x, cls, dn, en)
    h = 1e-6
    for k in p:
        flat = p[k].reshape(-1)
        for idx in rng.choice(flat.size, size=min(5, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = mlp.loss(p, x, cls, dn, en)
            flat[idx] = old - h
            lm = mlp.loss(p, x, cls, dn, en)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k].reshape(-1)[idx]) < 1e-6 * max(1.0, abs(fd)), k


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159], dtype=np.float32)
    r = mlp.bf16_round(x)
    # ties go to the even significand; bf16 keeps 8 significant bits
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0 + 2 ** -7
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


@pytest.mark.parametrize("o,v,naux,seed", [(3, 8, 10, 5), (4, 6, 12, 2)])
def test_oracle_fvector(o, v, naux, seed):
    """F-vector: the literal restatement of every term function equals the
    compact amplitude form (symmetric and non-symmetric dpl); dpl = identity
    (the number operator) gives f = 0; the density-like intermediates equal
    the MP2 density blocks (tauA = -uu^T/4, tauB = oo^T/4, tau21 + tau22 =
    -ou^T/2)."""
    sys = make_system(o, v, naux, seed=seed)
    nb = o + v
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(nb, nb))
    for dpl in (d + d.T, d):
        L = fvec.literal_terms(sys, dpl)
        C = fvec.compact_terms(sys, dpl)
        for k in ("F0", "tauA", "FA", "tauB", "FB", "tau21", "F21", "tau22", "F22", "tau23", "F23", "tau24",
                  "F24", "tauC", "FC", "F27", "F28", "F29", "F210", "FI", "FII", "FIII", "FIV1", "FIV2"):
            assert np.abs(L[k] - C[k]).max() <= 1e-13 * max(np.abs(L[k]).max(), 1e-300), k
        assert np.abs(L["F25"] + L["F26"] - C["F25+F26"]).max() < 1e-13 * np.abs(L["F25"]).max()
        f = fvec.assemble(sys, dpl, L)
        assert np.abs(f - fvec.fvector_compact(sys, dpl)).max() < 1e-13 * np.abs(f).max()
        # the zeroth-order 1h1p term dominates and the 2h2p block is first order
        n1 = o * v
        assert np.abs(f[:n1] + np.sqrt(2.0) * dpl[o:, :o].reshape(-1)).max() < 0.2 * np.abs(f[:n1]).max()
    eye = np.eye(nb)
    assert np.abs(fvec.fvector(sys, eye)).max() < 1e-15
    assert np.abs(fvec.fvector_compact(sys, eye)).max() < 1e-15
    _, parts = mp2.density(sys)
    L = fvec.literal_terms(sys, d)
    sc = np.abs(parts["uu"]).max()
    assert np.abs(L["tauA"] + parts["uu"].T / 4.0).max() < 1e-13 * sc
    assert np.abs(L["tauB"] - parts["oo"].T / 4.0).max() < 1e-13 * np.abs(parts["oo"]).max()
    assert np.abs(L["tau21"] + L["tau22"] + parts["ou"].T / 2.0).max() < 1e-13 * np.abs(parts["ou"]).max()


def test_oracle_bulirsch_stoer_relaxation():
    """bs_rlx: one Bulirsch-Stoer timestep equals exp(-H dtau) psi from
    scipy.linalg.expm to 1e-8 (tolerance 1e-10), and the relaxation reaches
    the lowest eigenpair; the second state reaches it too, since the
    reference's orthogonalisation loop projects on vec_conv(:,s), which is
    zero."""
    from scipy.linalg import expm
    sys = make_system(3, 8, 10, seed=5)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    lam, U = np.linalg.eigh(H)
    g = relax.ondiag_guess(np.diag(H), 2)
    bs = relax.BSStepper()
    f = lambda x: -(H @ x)
    psi, t, order, nxt, inttime, step = g[:, 0].copy(), 0.0, 16, 0.0, 0.0, 2.0
    while True:
        ss = step if inttime == 0.0 else (step - inttime if inttime + nxt > step else nxt)
        psi, order, tr, nxt, code = bs.step(f, psi, f(psi), step, t, order, ss, 1e-10)
        assert code == 0
        t += tr
        inttime += tr
        if not abs(step - inttime) > 1e-9 * step:
            break
    ex = expm(-H * step) @ g[:, 0]
    assert np.abs(psi - ex).max() < 1e-8 * np.abs(ex).max()
    out = relax.bs_relax(lambda x: H @ x, g, maxiter=300, step=2.0, eps=1e-6, toler=1e-8)
    assert np.abs(out["eigval"] - lam[0]).max() < 1e-9 * abs(lam[0])
    for s in range(2):
        v = out["eigvec"][:, s]  # residual 1e-6 over a gap of 0.023
        assert min(np.abs(v - U[:, 0]).max(), np.abs(v + U[:, 0]).max()) < 1e-4


def test_oracle_rkf45_relaxation():
    """rkf45_rlx: one r8_rkf45 call to tout = 2 equals exp(-2H) psi from
    scipy.linalg.expm to 1e-9 (tolerances 1e-10), the relaxation reaches the
    lowest eigenpair for both states (the orthogonalisation loop is a no-op as
    in bs_rlx), and a run that needs more than 30000 derivative evaluations
    stops with the reference's error code 4."""
    from scipy.linalg import expm
    sys = make_system(3, 8, 10, seed=5)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    lam, U = np.linalg.eigh(H)
    g = relax.ondiag_guess(np.diag(H), 2)
    rk = relax.RKF45()
    y, _, t, flag = rk.step(lambda x: -(H @ x), g[:, 0].copy(), np.zeros(H.shape[0]), 0.0, 2.0, 1e-10, 1e-10, 1)
    ex = expm(-2.0 * H) @ g[:, 0]
    assert flag == 2 and t == 2.0
    assert np.abs(y - ex).max() < 1e-9 * np.abs(ex).max()
    out = relax.rkf45_relax(lambda x: H @ x, g, maxiter=300, step=5.0, eps=1e-6, toler=1e-7)
    assert np.abs(out["eigval"] - lam[0]).max() < 1e-9 * abs(lam[0])
    for s in range(2):
        v = out["eigvec"][:, s]
        assert min(np.abs(v - U[:, 0]).max(), np.abs(v + U[:, 0]).max()) < 1e-4
    with pytest.raises(RuntimeError, match="Error code: 4"):
        relax.rkf45_relax(lambda x: H @ x, g, maxiter=300, step=2.0, eps=1e-6, toler=1e-8)


def test_oracle_coupled():
    """The coupled step solves M(theta) t = r, M = [[H11 - theta, C], [C^T, D - sigma]],
    and #eig(A0(sigma)) < sigma equals #eig(H) < sigma for several sigma below
    min D (Sylvester's law of inertia), against a dense eigh of H.  A dropped
    P, a transposed C or a wrong sign in A0 fails the inertia count."""
    sys_ = make_system(3, 8, 10, seed=4)
    df = dfsigma.DF(sys_)
    C = coupled.coupling_matrix(df)
    H11, D = df.H11, df.D2
    n1 = df.n1
    H = np.block([[H11, C], [C.T, np.diag(D)]])
    assert np.abs(H - df.sigma(np.eye(H.shape[0]))).max() < 1e-12 * np.abs(H).max()
    w = np.linalg.eigvalsh(H)
    dmin = D.min()
    counts = []
    for sig in np.linspace(w[0] - 0.5, dmin - 1e-3, 9):
        A0 = coupled.schur_complement(H11, C, D, sig)
        assert coupled.count_below(A0, sig) == int(np.sum(w < sig))
        counts.append(int(np.sum(w < sig)))
    assert max(counts) > 1          # the shifts do separate roots
    rng = np.random.default_rng(2)
    R = rng.standard_normal((H.shape[0], 3))
    theta = np.array([w[0] + 0.1, w[1] - 0.05, w[2]])
    sig = dmin - 0.2
    T = coupled.coupled_step(H11, C, D, sig, theta, R)
    for i in range(3):
        M = H - theta[i] * np.eye(H.shape[0])
        M[n1:, n1:] += (theta[i] - sig) * np.eye(H.shape[0] - n1)
        assert np.abs(M @ T[:, i] - R[:, i]).max() < 1e-10 * np.abs(R[:, i]).max()
    with pytest.raises(ValueError):
        coupled.schur_complement(H11, C, D, dmin + 1e-3)


@pytest.mark.parametrize("o,v,naux,seed", [(2, 3, 6, 3), (2, 4, 5, 8), (3, 3, 6, 1)])
def test_oracle_mp2_density_brute_force(o, v, naux, seed):
    """The MP2 energy and every block of the second-order density -- the oo,
    uu and ou blocks of rho_mp2 (source/adclib/mp2.f90:421-596) -- against
    exact Rayleigh-Schroedinger perturbation theory in the determinant space
    of the same Hamiltonian (tests/detspace.py: Psi1 = R V Phi0,
    Psi2 = R (V - E1) Psi1, rho2 from <Psi|g|Psi>), which shares nothing with
    the restated formulas.  The ou block comes from the singles of Psi2, so a
    wrong denominator, sign or exchange term in rhogs2_ou fails here."""
    from tests.detspace import mp_second_order
    sys_ = make_system(o, v, naux, seed=seed)
    V = mp2.integrals(sys_)
    E2, rho0, rho2 = mp_second_order(sys_.eps, V, o)
    assert abs(E2 - mp2.energy(sys_, V)) < 1e-13 * abs(E2)
    assert np.abs(rho0 - np.diag([2.0] * o + [0.0] * v)).max() == 0.0
    _, blk = mp2.density(sys_, V)
    scale = np.abs(rho2).max()
    assert np.abs(rho2[:o, :o] - blk["oo"]).max() < 1e-12 * scale
    assert np.abs(rho2[o:, o:] - blk["uu"]).max() < 1e-12 * scale
    assert np.abs(rho2[:o, o:] - blk["ou"]).max() < 1e-12 * scale
    assert np.abs(rho2[o:, :o] - blk["ou"].T).max() < 1e-12 * scale
    assert np.abs(blk["ou"]).max() > 1e-3 * scale     # the ou block is not trivially zero


def _scaled(sys_, lam):
    """The same system with every two-electron integral scaled by lam (the
    fluctuation potential of the Moller-Plesset partition, orbital energies
    unchanged)."""
    from adc_ottawa_amd.synth import System
    return System(sys_.nocc, sys_.nvirt, sys_.naux, sys_.eps.copy(), sys_.B * np.sqrt(lam))


def _orders(fn, sys_):
    """Coefficients of lam^0, lam^1, lam^2 of fn(system with integrals x lam);
    exact for the ADC(2) quantities, which are polynomials of degree <= 2 in
    the integrals."""
    lams = np.array([0.5, 1.0, 2.0])
    vals = np.array([np.asarray(fn(_scaled(sys_, x)), dtype=np.float64) for x in lams])
    c = np.linalg.solve(np.vander(lams, 3, increasing=True), vals.reshape(3, -1))
    return c.reshape((3,) + vals.shape[1:])


@pytest.mark.parametrize("o,v,naux,seed", [(2, 3, 6, 3), (2, 4, 5, 8), (3, 3, 6, 1)])
def test_oracle_isr_brute_force(o, v, naux, seed):
    """The ADC(2)-s matrix and f-vector of the oracle against the
    intermediate-state representation built exactly in the determinant space
    (tests/detspace.py: exact ground state of H0 + lam V, precursor states
    C_J Psi0, orthogonalised to Psi0 and the lower class, Loewdin within a
    class; strict perturbation orders read off a polynomial in lam), which
    shares nothing with the restated formulas:
    * 1h1p/1h1p block H11 (adc_ph.f90 K, C1, CA, CB, CC) order by order
      through second order, element by element;
    * 1h1p f-vector (get_moment.f90:153-442: F0, F21..F210, FA, FB, FC) order
      by order through second order, element by element, the sign of the
      reference's 1h1p configurations being -1 against C_ak = g_ak/sqrt(2);
    * coupling C1 (adc_ph.f90:974-1073), 2h2p diagonal and 2h2p f-vector
      (FI..FIV2) through the invariants of the 2h2p basis: C C^T, C D C^T,
      C f2 (same sign convention) and |f2|^2."""
    from tests.detspace import isr_orders
    sys_ = make_system(o, v, naux, seed=seed)
    rng = np.random.default_rng(100 + seed)
    d = rng.standard_normal((o + v, o + v))
    d = d + d.T
    bf = isr_orders(sys_.eps, mp2.integrals(sys_), o, d)
    n1 = o * v
    typ, a, _, k, _ = csf.table(o, v)
    ph = np.where(typ == 0)[0]
    assert np.array_equal(a[ph] * o + k[ph], np.arange(n1))  # 1h1p configurations virtual-major
    h11 = _orders(lambda s: elements.build_blocks(s, apply_minc=False).H11, sys_)
    f1 = _orders(lambda s: fvec.fvector(s, d)[:n1], sys_)
    tol = (1e-12, 1e-9, 1e-6)   # the fit of the brute-force series limits order 2 to ~1e-8
    for n in range(3):
        ref_m, ref_f = bf["M11"][n], -bf["f1"][n].reshape(n1)
        assert np.abs(h11[n] - ref_m).max() < tol[n] * np.abs(ref_m).max(), f"H11 order {n}"
        assert np.abs(f1[n] - ref_f).max() < tol[n] * np.abs(ref_f).max(), f"f(1h1p) order {n}"
    assert np.abs(bf["M11"][2]).max() > 1e-4 and np.abs(bf["f1"][2]).max() > 1e-5  # second order is not trivial
    blk = elements.build_blocks(sys_, apply_minc=False)
    C, D2 = blk.C, blk.D
    f2 = fvec.fvector(sys_, d)[n1:]
    checks = {"CC": C @ C.T, "CDC": (C * D2[None, :]) @ C.T, "Cf": -(C @ f2), "ff": np.array([f2 @ f2])}
    for key, ref in checks.items():
        got = bf[key]
        scale = np.abs(ref).max()
        assert np.abs(got[2] - ref).max() < 1e-9 * scale, key
        assert max(np.abs(got[0]).max(), np.abs(got[1]).max()) < 1e-12 * scale, key


@pytest.mark.parametrize("o,v,naux,seed", [(2, 3, 6, 3), (2, 4, 5, 8), (3, 3, 6, 1)])
def test_oracle_csf_states_and_adc2x_block(o, v, naux, seed):
    """The configuration basis itself and the ADC(2)-x block, against the
    determinant space (tests/detspace.py csf_first_order): each 2h2p CSF is
    the normalised state sum_cbkl Z_J(c,b,k,l) g_ck g_bl Phi0 of its pair
    tensor (oracle/dfsigma.py pair_tensor), each 1h1p CSF g_ak/sqrt(2) Phi0.
    * The 2h2p states are orthonormal, and the first-order coupling
      <Phi_ak|V|Phi_J> equals the reference's element functions C1..C5
      (oracle/elements.py) element by element: this fixes the order, the
      phases and the two a!=b, k!=l spin couplings of the reference's CSFs.
    * The first-order 2h2p block <Phi_J|V - E1|Phi_J'> (the ADC(2)-x block)
      equals the pair-space form of oracle/adcx.py element by element, and
      the reference's element functions restated literally (the diagonal of
      t1, t2, t3, t5 and the (t5,t1), (t3,t2) blocks) equal both."""
    from oracle import adcx
    from tests.detspace import csf_first_order
    sys_ = make_system(o, v, naux, seed=seed)
    C1, M1, S = csf_first_order(sys_.eps, mp2.integrals(sys_), o, dfsigma.pair_tensor)
    assert np.abs(S - np.eye(S.shape[0])).max() < 1e-14
    blk = elements.build_blocks(sys_, apply_minc=False)
    assert np.abs(C1 - blk.C).max() < 1e-13 * np.abs(blk.C).max()
    scale = np.abs(M1).max()
    assert np.abs(adcx.block22_dense(sys_) - M1).max() < 1e-13 * scale
    d = adcx.diagonal_first_order(sys_)
    m = ~np.isnan(d)
    assert np.abs(d[m] - np.diag(M1)[m]).max() < 1e-13 * scale
    for r, c, val in adcx.offdiagonal_blocks(sys_).values():
        assert np.abs(M1[np.ix_(r, c)] - val).max() < 1e-13 * scale


def _laser_inputs(o=2, v=4, naux=6, seed=3):
    """A small ADC(2)-s matrix and random IS dipole operators in the
    reference's file layout (diagonal + each off-diagonal element once)."""
    from oracle import laser
    sys_ = make_system(o, v, naux, seed=seed)
    df = dfsigma.DF(sys_)
    nd = df.n1 + df.D2.size
    H = df.sigma(np.eye(nd))
    H = 0.5 * (H + H.T)
    rng = np.random.default_rng(seed)
    dips, raw = [], []
    for c in range(3):
        row = rng.integers(0, nd, 4 * nd)
        col = rng.integers(0, nd, 4 * nd)
        keep = row != col
        row, col = row[keep], col[keep]
        val = rng.normal(0.0, 0.05, row.size)
        diag = rng.normal(0.0, 0.2, nd)
        d0j = rng.normal(0.0, 0.1, nd)
        d00 = float(rng.normal(0.0, 0.5))
        dips.append((laser.dense_dipole(diag, row, col, val, nd), d00, d0j))
        raw.append((diag, row, col, val, d00, d0j))
    return sys_, H, dips, raw


def test_oracle_laser_envelopes():
    """efield: the cos^2 and sin^2 envelopes have half maximum at +-fwhm/2
    about their peak (so fwhm is their full width at half maximum) and vanish
    outside their support; the sin^2 ramp is 1/2 at tau/2 and 1 from tau on;
    the box is 1 on [ti, tf] only; the carrier peaks (cos) or vanishes (sin)
    where freq (t - t0) = phase; pulses add along their polarisation."""
    from oracle import laser
    c2 = {"envelope": "cos2", "envpar": (50.0, 20.0), "pulse_vec": (1, 0, 0), "strength": 1.0}
    assert laser.envelope(c2, 50.0) == 1.0
    assert abs(laser.envelope(c2, 40.0) - 0.5) < 1e-15 and abs(laser.envelope(c2, 60.0) - 0.5) < 1e-15
    assert laser.envelope(c2, 30.0) == 0.0 and laser.envelope(c2, 70.0) == 0.0
    s2 = {"envelope": "sin2", "envpar": (10.0, 20.0)}
    assert abs(laser.envelope(s2, 30.0) - 1.0) < 1e-15
    assert abs(laser.envelope(s2, 20.0) - 0.5) < 1e-15 and abs(laser.envelope(s2, 40.0) - 0.5) < 1e-15
    assert laser.envelope(s2, 10.0) == 0.0 and laser.envelope(s2, 50.0) == 0.0
    r = {"envelope": "sin2ramp", "envpar": (30.0, 0.0)}
    assert laser.envelope(r, -1.0) == 0.0 and abs(laser.envelope(r, 15.0) - 0.5) < 1e-15
    assert laser.envelope(r, 30.0) == 1.0 and laser.envelope(r, 100.0) == 1.0
    b = {"envelope": "box", "envpar": (5.0, 9.0)}
    assert [laser.envelope(b, t) for t in (4.99, 5.0, 7.0, 9.0, 9.01)] == [0.0, 1.0, 1.0, 1.0, 0.0]
    w, t0, ph = 0.7, 3.0, 0.4
    pc = {"envelope": "box", "envpar": (0.0, 100.0), "carrier": "cos", "freq": w, "t0": t0, "phase": ph,
          "pulse_vec": (0.6, 0.8, 0.0), "strength": 0.05}
    ps = dict(pc, carrier="sin", pulse_vec=(0.0, 0.0, 1.0))
    tp = t0 + ph / w
    assert np.allclose(laser.efield([pc], tp), [0.03, 0.04, 0.0], atol=1e-15)
    assert np.allclose(laser.efield([ps], tp), 0.0, atol=1e-15)
    assert np.allclose(laser.efield([pc, ps], 7.3), laser.efield([pc], 7.3) + laser.efield([ps], 7.3), atol=1e-16)


def test_oracle_laser_constant_field_exact():
    """A box pulse of zero frequency is a constant field E: from the ground
    state the propagation equals exp(-i (H - E.D) t) psi0 on the ndim + 1
    states (dense exponential), the IS dipole entering with the reference's
    sign (-mu.E), d00 on both blocks and d0j symmetric; the norm stays 1."""
    import scipy.linalg
    from oracle import laser
    sys_, H, dips, _ = _laser_inputs()
    nd = H.shape[0]
    pulse = {"envelope": "box", "envpar": (0.0, 100.0), "carrier": "cos", "freq": 0.0,
             "pulse_vec": (0.48, 0.6, 0.64), "strength": 0.3}
    E = laser.efield([pulse], 1.0)
    psi0 = np.zeros(nd + 1, dtype=complex)
    psi0[nd] = 1.0
    hmul = (lambda x: H @ x)  # noqa: E731
    psi, norms = laser.propagate(hmul, dips, [pulse], psi0, 0.5, 3.0, 20, 1e-12)
    want = scipy.linalg.expm(-1j * 3.0 * laser.full_matrix(H, dips, E)) @ psi0
    assert np.abs(psi - want).max() < 1e-9
    assert np.abs(norms - 1.0).max() < 1e-10
    # the same generator through matxvec_treal_laser
    v = np.random.default_rng(1).normal(size=nd + 1) + 1j * np.random.default_rng(2).normal(size=nd + 1)
    assert np.abs(laser.matxvec_treal_laser(hmul, dips, E, v) + 1j * (laser.full_matrix(H, dips, E) @ v)).max() \
        < 1e-12 * np.abs(v).max() * np.abs(H).max()


def test_oracle_laser_time_dependent_norm():
    """Under a cos^2 pulse (Hermitian H(t)) the norm is conserved, the
    wavefunction leaves the ground state, and a period before the pulse
    (E = 0 on the ground state) is skipped unchanged."""
    from oracle import laser
    sys_, H, dips, _ = _laser_inputs(seed=4)
    nd = H.shape[0]
    pulse = {"envelope": "cos2", "envpar": (3.0, 2.0), "carrier": "cos", "freq": 0.5,
             "pulse_vec": (1.0, 0.0, 0.0), "strength": 0.2}
    psi0 = np.zeros(nd + 1, dtype=complex)
    psi0[nd] = 1.0
    psi, norms = laser.propagate(lambda x: H @ x, dips, [pulse], psi0, 0.25, 5.0, 16, 1e-10)
    assert np.abs(norms - 1.0).max() < 1e-8
    assert np.linalg.norm(psi[:nd]) > 1e-3
    psi1, _ = laser.propagate(lambda x: H @ x, dips, [pulse], psi0, 0.25, 1.0, 16, 1e-10)
    assert np.array_equal(psi1, psi0)


def _cap_inputs(nd, seed):
    """A positive semi-definite CAP on the ndim + 1 states (W_full = B B^T)
    split into the reference's W (IS block without w00), w00 and w0j, and a
    projector onto the complement of two orthonormal IS vectors."""
    rng = np.random.default_rng(seed)
    B = rng.normal(0.0, 0.08, (nd + 1, 4))
    Wf = B @ B.T
    w00 = Wf[nd, nd]
    cap = (Wf[:nd, :nd] - w00 * np.eye(nd), w00, Wf[:nd, nd].copy())
    R, _ = np.linalg.qr(rng.normal(size=(nd, 2)))
    return cap, Wf, R


@pytest.mark.parametrize("ground", [False, True])
def test_oracle_csil_cap_constant_field_exact(ground):
    """With a CAP the generator is H - E.D - i P W P (non-Hermitian): the
    Lanczos-Arnoldi propagation under a constant field equals the dense
    exp(-i (H - E.D - i P W P) t) psi0, and the norm never grows (P W P is
    positive semi-definite)."""
    import scipy.linalg
    from oracle import laser
    sys_, H, dips, _ = _laser_inputs(seed=6)
    nd = H.shape[0]
    cap, Wf, R = _cap_inputs(nd, 6)
    proj = (R, ground)
    pulse = {"envelope": "box", "envpar": (0.0, 100.0), "carrier": "cos", "freq": 0.0,
             "pulse_vec": (0.0, 0.6, 0.8), "strength": 0.2}
    E = laser.efield([pulse], 1.0)
    rng = np.random.default_rng(3)
    psi0 = rng.normal(size=nd + 1) + 1j * rng.normal(size=nd + 1)
    psi0 /= np.linalg.norm(psi0)
    psi, norms = laser.propagate(lambda x: H @ x, dips, [pulse], psi0, 0.5, 3.0, 24, 1e-12, cap, proj)
    G = laser.full_matrix(H, dips, E) - 1j * laser.cap_matrix(cap, proj, nd)
    want = scipy.linalg.expm(-1j * 3.0 * G) @ psi0
    assert np.abs(psi - want).max() < 1e-9
    assert np.all(np.diff(norms) <= 1e-12) and norms[-1] < norms[0] - 1e-4
    # the generator through matxvec_treal_laser
    v = rng.normal(size=nd + 1) + 1j * rng.normal(size=nd + 1)
    got = laser.matxvec_treal_laser(lambda x: H @ x, dips, E, v, cap, proj)
    assert np.abs(got + 1j * (G @ v)).max() < 1e-12 * np.abs(G).max() * np.abs(v).max() * 10


def _exact_path(G, psi0, t, h):
    """psi(t - h), psi(t), psi(t + h) of psi' = -i G psi (dense exponential)."""
    import scipy.linalg
    return [scipy.linalg.expm(-1j * s * G) @ psi0 for s in (t - h, t, t + h)]


@pytest.mark.parametrize("ground", [False, True])
def test_oracle_flux_is_rate_of_change(ground):
    """With theta00 = w00 = 0 the flux is d/dt [<psi|P Theta P|psi> -
    <psi|psi>] along the exact constant-field propagation with the CAP:
    2 Re <psi'|P Theta P|psi> is the rate of the population in the absorbing
    region and 2 <psi|P W P|psi> the loss of norm.  Central differences of the
    dense exponential (step 1e-4, error ~1e-9) check both terms, the
    projector on both sides, the conjugated bra and the ground-state rows."""
    from oracle import laser
    sys_, H, dips, _ = _laser_inputs(seed=6)
    nd = H.shape[0]
    cap, Wf, R = _cap_inputs(nd, 7)
    cap = (cap[0], 0.0, cap[2])
    proj = (R, ground)
    rng = np.random.default_rng(8)
    T = rng.normal(0.0, 0.3, (nd, nd))
    theta = (0.5 * (T + T.T), 0.0, rng.normal(0.0, 0.3, nd))
    Tf = laser.cap_matrix(theta, proj, nd)
    pulse = {"envelope": "box", "envpar": (0.0, 100.0), "carrier": "cos", "freq": 0.0,
             "pulse_vec": (0.0, 0.6, 0.8), "strength": 0.2}
    E = laser.efield([pulse], 1.0)
    G = laser.full_matrix(H, dips, E) - 1j * laser.cap_matrix(cap, proj, nd)
    psi0 = rng.normal(size=nd + 1) + 1j * rng.normal(size=nd + 1)
    psi0 /= np.linalg.norm(psi0)
    h = 1e-4
    pm, p0, pp = _exact_path(G, psi0, 1.3, h)
    pop = [np.vdot(x, Tf @ x).real - np.vdot(x, x).real for x in (pm, pp)]
    want = (pop[1] - pop[0]) / (2.0 * h)
    dtpsi = laser.matxvec_treal_laser(lambda x: H @ x, dips, E, p0, cap, proj)
    got = laser.flux(theta, cap, proj, p0, dtpsi)
    assert abs(got - want) < 1e-7 * max(1.0, abs(want))
    loss = -(np.vdot(pp, pp).real - np.vdot(pm, pm).real) / (2.0 * h)
    assert abs(laser.flux((0.0 * theta[0], 0.0, 0.0 * theta[2]), cap, proj, p0, dtpsi) - loss) < 1e-7
    assert loss > 0.0


def test_oracle_flux_theta00_w00_placement():
    """Theta = 2 and W = c on the intermediate states, no off-diagonal or
    ground-IS elements, no projector: the reference's (1 + theta00) Theta u
    gives 2 (1 + theta00) d/dt |psi_IS|^2 + theta00 d/dt |psi_0|^2 for the
    first term and 2 (1 + w00) c |psi_IS|^2 + 2 w00 |psi_0|^2 for the
    second, which differ from (Theta + theta00) u when Theta != 1.  The
    rates come from central differences of the exact propagation."""
    from oracle import laser
    sys_, H, dips, _ = _laser_inputs(seed=9)
    nd = H.shape[0]
    cap, Wf, R = _cap_inputs(nd, 9)
    a, b, c = 0.3, 0.2, 0.7
    theta = (2.0 * np.eye(nd), a, np.zeros(nd))
    flux_cap = (c * np.eye(nd), b, np.zeros(nd))
    pulse = {"envelope": "box", "envpar": (0.0, 100.0), "carrier": "cos", "freq": 0.0,
             "pulse_vec": (0.48, 0.6, 0.64), "strength": 0.3}
    E = laser.efield([pulse], 1.0)
    G = laser.full_matrix(H, dips, E) - 1j * laser.cap_matrix(cap, (None, False), nd)
    rng = np.random.default_rng(10)
    psi0 = rng.normal(size=nd + 1) + 1j * rng.normal(size=nd + 1)
    psi0 /= np.linalg.norm(psi0)
    h = 1e-4
    pm, p0, pp = _exact_path(G, psi0, 0.9, h)
    rate_is = (np.linalg.norm(pp[:nd]) ** 2 - np.linalg.norm(pm[:nd]) ** 2) / (2.0 * h)
    rate_0 = (abs(pp[nd]) ** 2 - abs(pm[nd]) ** 2) / (2.0 * h)
    dtpsi = laser.matxvec_treal_laser(lambda x: H @ x, dips, E, p0, cap, (None, False))
    want = (2.0 * (1.0 + a) * rate_is + a * rate_0
            + 2.0 * (1.0 + b) * c * np.linalg.norm(p0[:nd]) ** 2 + 2.0 * b * abs(p0[nd]) ** 2)
    got = laser.flux(theta, flux_cap, (None, False), p0, dtpsi)
    assert abs(got - want) < 1e-7 * max(1.0, abs(want))


@pytest.mark.parametrize("o,v,naux,seed", [(2, 3, 5, 1), (3, 4, 6, 2)])
def test_oracle_trdens_is_fvector_transpose(o, v, naux, seed):
    """The transition density of adc2_trdens_gs_virt_occ is the transpose of
    the F-vector's dpl_vo terms: with rho replaced by the F-vector's own MP2
    intermediates, trdens_vo(x)[a, i] = x . f(E_ai) for every (a, i), where
    f is the 1h1p block of the F-vector oracle (pinned against the exact
    intermediate-state representation) of the operator whose only element is
    dpl(a, i) = 1.  A dropped or doubled term, a sign, a transposed
    integral or a denominator of the wrong pair fails it."""
    from oracle import fvec, mp2, nto
    sys = make_system(o, v, naux, seed=seed)
    V = mp2.integrals(sys)
    nb = o + v
    terms = fvec.literal_terms(sys, np.zeros((nb, nb)))
    rho = np.zeros((nb, nb))
    rho[o:, o:] = -4.0 * terms["tauA"].T
    rho[:o, :o] = 4.0 * terms["tauB"].T
    rng = np.random.default_rng(seed)
    xs = rng.normal(size=(3, o * v))
    got = np.stack([nto.trdens_vo(sys, x, rho, V) for x in xs])
    want = np.zeros_like(got)
    for a in range(v):
        for i in range(o):
            dpl = np.zeros((nb, nb))
            dpl[o + a, i] = 1.0
            f1 = fvec.fvector(sys, dpl)[: o * v]
            want[:, a, i] = xs @ f1
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # with rho_mp2 as the reference assembles it only the MP2-density terms change
    full = nto.trdens_vo(sys, xs[0], None, V)
    R = mp2.density(sys, V)[0] - rho
    X = xs[0].reshape(v, o)
    extra = 0.25 * np.sqrt(2.0) * (R[o:, o:] @ X - X @ R[:o, :o].T)
    assert np.abs(full - got[0] - extra).max() < 1e-12 * np.abs(full).max()


def test_oracle_tdnto():
    """tdadc2_nto: the SVD reconstructs the normalised complex density
    (U diag(sigma) V^H = trdens^T), the density is that of the normalised IS
    part (scaling wf leaves it unchanged), and the pair count follows
    sigma^2 >= 0.001; a pure 1h1p zeroth-order-dominated state has one
    dominant pair."""
    from oracle import nto
    sys = make_system(3, 5, 6, seed=7)
    o, v = sys.nocc, sys.nvirt
    rng = np.random.default_rng(7)
    nd = o * v + 10
    wf = rng.normal(size=nd) + 1j * rng.normal(size=nd)
    T, sigma, U, Vm, npair = nto.tdnto(sys, wf)
    S = np.zeros((o, v))
    S[:o, :o] = np.diag(sigma)
    assert np.abs(U @ S @ Vm.conj().T - T.T).max() < 1e-12
    T2 = nto.tdnto(sys, 3.7j * wf)[0]
    assert np.abs(T2 - 1j * T).max() < 1e-12          # the phase of wf carries through
    assert npair == int(np.sum(sigma ** 2 >= 1e-3))
    e = np.zeros(nd, dtype=complex)
    e[2 * o + 1] = 1.0                                 # configuration (a=2, i=1)
    T1, s1, U1, V1, n1 = nto.tdnto(sys, e)
    assert s1[0] > 0.9 * np.sqrt(2.0) and s1[1] < 0.2 * s1[0]
    assert abs(abs(U1[1, 0]) - 1.0) < 0.05 and abs(abs(V1[2, 0]) - 1.0) < 0.05


def test_oracle_mp2_density_frozen_core():
    """The frozen-core MP2 density: with no frozen orbital it is the full
    density; frozen rows and columns hold only the 2 on the diagonal; the
    second-order correction stays traceless (trace 2 nocc); the active block
    differs from the full density's (the frozen orbitals no longer
    correlate)."""
    from oracle import mp2
    sys = make_system(4, 7, 9, seed=12)
    full = mp2.density(sys)[0]
    assert np.abs(mp2.density_frozen(sys, []) - full).max() < 1e-14
    rho = mp2.density_frozen(sys, [0, 2])
    for f in (0, 2):
        row = rho[f].copy()
        assert row[f] == 2.0
        row[f] = 0.0
        assert not row.any() and not np.delete(rho[:, f], f).any()
    assert abs(np.trace(rho) - 2 * sys.nocc) < 1e-12
    assert np.abs(rho[np.ix_([1, 3], [1, 3])] - full[np.ix_([1, 3], [1, 3])]).max() > 1e-6


def test_oracle_cheby_fdstates():
    """calc_eigenstates of fdstates_cfd: the accumulated states equal
    sum_k coef[k, i] V cos(k arccos x) V^T q0 from the dense
    eigendecomposition (normalised), and a filter built from the Chebyshev
    expansion of a narrow Gaussian around one eigenvalue returns that
    eigenvector and eigenvalue."""
    from oracle import cheby, dfsigma
    sys = make_system(2, 4, 6, seed=3)
    df = dfsigma.DF(sys)
    nd = df.n1 + df.D2.size
    H = df.sigma(np.eye(nd))
    H = 0.5 * (H + H.T)
    lam, U = np.linalg.eigh(H)
    bounds = (lam[0] - 0.1, lam[-1] + 0.1)
    rng = np.random.default_rng(3)
    q0 = rng.normal(size=nd)
    K = 60
    coef = rng.normal(size=(K + 1, 3)) / (1.0 + np.arange(K + 1))[:, None]
    E, V = cheby.fdstates(lambda x: H @ x, q0, bounds, coef)
    x = (2.0 * lam - bounds[0] - bounds[1]) / (bounds[1] - bounds[0])
    T = np.cos(np.arange(K + 1)[:, None] * np.arccos(x)[None, :])          # T_k(x_n)
    want = U @ ((T.T @ coef) * (U.T @ q0)[:, None])
    want /= np.linalg.norm(want, axis=0)
    assert np.abs(V - want).max() < 1e-10
    assert np.abs(E - np.einsum("ni,ni->i", want, H @ want)).max() < 1e-10
    # a Gaussian filter at one eigenvalue (its Chebyshev coefficients from Gauss-Chebyshev quadrature)
    gaps = [np.min(np.abs(np.delete(x, i) - x[i])) for i in range(nd)]
    m = int(np.argmax(gaps))                 # the most isolated eigenvalue (gap 0.078 in x)
    K = 800
    th = np.pi * (np.arange(4000) + 0.5) / 4000
    g = np.exp(-((np.cos(th) - x[m]) / 0.01) ** 2)
    c = np.array([(2.0 - (k == 0)) * np.mean(g * np.cos(k * th)) for k in range(K + 1)])
    E1, V1 = cheby.fdstates(lambda v: H @ v, U[:, m] + 0.3 * q0 / np.linalg.norm(q0), bounds, c[:, None])
    assert gaps[m] > 0.07
    assert abs(abs(V1[:, 0] @ U[:, m]) - 1.0) < 1e-8 and abs(E1[0] - lam[m]) < 1e-8


def test_oracle_sil_fdstates_td():
    """calc_fdstates_tdfd: the SIL propagation's filter states equal those
    built from the exact psi(t) = U exp(-i lam t) U^T psi0 (1e-9), and with
    a single filter energy at an isolated eigenvalue and a long window the
    state is that eigenvector (energy to 1e-4 of the gap)."""
    from oracle import dfsigma, sil
    sys = make_system(2, 4, 6, seed=5)
    df = dfsigma.DF(sys)
    nd = df.n1 + df.D2.size
    H = df.sigma(np.eye(nd))
    H = 0.5 * (H + H.T)
    lam, U = np.linalg.eigh(H)
    rng = np.random.default_rng(5)
    f = rng.normal(size=nd)
    psi0 = f / np.linalg.norm(f)
    exact = lambda t: U @ (np.exp(-1j * lam * t) * (U.T @ psi0))  # noqa: E731
    func = lambda x: -1j * (H @ x)  # noqa: E731
    F = rng.normal(size=(3, 7))
    eb = (lam[0], lam[-1])
    E1, V1 = sil.fdstates_td(func, lambda x: H @ x, f, 0.5, 20.0, 20, 1e-12, eb, F, 1)
    E2, V2 = sil.fdstates_td(func, lambda x: H @ x, f, 0.5, 20.0, 20, 1e-12, eb, F, 1, states=exact)
    assert np.abs(V1 - V2).max() < 1e-9 and np.abs(E1 - E2).max() < 1e-9
    gaps = [np.min(np.abs(np.delete(lam, i) - lam[i])) for i in range(nd)]
    m = int(np.argmax(gaps))
    Fm = np.array([[1.0, 0.0]])
    E3, V3 = sil.fdstates_td(func, lambda x: H @ x, f, 1.0, 400.0, 20, 1e-12, (lam[m], lam[m] + 1.0), Fm, 2,
                             states=exact)
    assert abs(E3[0] - lam[m]) < 1e-4 * gaps[m] and abs(abs(V3[:, 0] @ U[:, m]) - 1.0) < 1e-3

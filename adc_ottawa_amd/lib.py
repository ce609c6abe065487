"""ctypes binding of the C ABI in include/adc2g.h.

This is the Python-side mirror of the Fortran ``iso_c_binding`` module
(``adc_ottawa_amd/fortran/adc2g_mod.f90``): the same entry points, argument
meaning and error behaviour.  It loads the in-tree ``libadc2g.so`` and fails
loudly when it is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libadc2g.so")

ADC2G_OK = 0
ERRORS = {1: "SIL illegal integration order", 2: "SIL stepsize underflow",
          3: "SIL diagonalisation failed", 4: "SIL illegal initial stepsize",
          10: "bad argument", 11: "HIP failure", 12: "not all vectors have converged",
          13: "Rayleigh-Ritz eigensolve failed", 14: "device allocation failed",
          15: "call out of order", 16: "block Lanczos breakdown"}

PREC = {"fp64": 0, "fp32": 1, "bf16": 2}
PRECON = {"dpr": 1, "olsen": 2, "learned": 3, "coupled": 4, "learned_coupled": 5}


class Adc2gError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"adc2g error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("build_phph", C.c_int32), ("keep_lowprec", C.c_int32),
                ("coupling", C.c_int32)]


COUPLING = {"auto": 0, "df": 1, "int": 2}


class DavParams(C.Structure):
    _fields_ = [("nstates", C.c_int32), ("blocksize", C.c_int32), ("maxvec", C.c_int32),
                ("maxiter", C.c_int32), ("tol", C.c_double), ("precon", C.c_int32),
                ("deflate", C.c_int32), ("precision", C.c_int32), ("refine", C.c_int32),
                ("lowprec_tol", C.c_double), ("guess", C.c_int32), ("verbose", C.c_int32),
                ("subdim", C.c_int32), ("reserved", C.c_int32)]


GUESS = {"ondiag": 0, "subdiag": 1, "vectors": 2}


class DavStats(C.Structure):
    _fields_ = [("niter", C.c_int32), ("nsigma", C.c_int32), ("nmult_ref", C.c_int64),
                ("niter_lowprec", C.c_int32), ("nfallback", C.c_int32), ("time_total", C.c_double),
                ("time_sigma", C.c_double), ("resnorm_max", C.c_double), ("time_precon", C.c_double),
                ("nprecon_builds", C.c_int32), ("reserved", C.c_int32), ("precon_sigma", C.c_double)]


class SilParams(C.Structure):
    _fields_ = [("tout", C.c_double), ("tfinal", C.c_double), ("kdim", C.c_int32),
                ("precision", C.c_int32), ("autotol", C.c_double), ("verbose", C.c_int32),
                ("reserved", C.c_int32)]


class SilStats(C.Structure):
    _fields_ = [("nsteps", C.c_int64), ("nsigma", C.c_int64), ("max_order", C.c_int32),
                ("reserved", C.c_int32), ("time_total", C.c_double)]


class Pulse(C.Structure):
    """adc2g_pulse: one laser pulse of efield (tdselib.f90:850-1029)."""
    _fields_ = [("carrier", C.c_int32), ("envelope", C.c_int32), ("pulse_vec", C.c_double * 3),
                ("strength", C.c_double), ("freq", C.c_double), ("t0", C.c_double), ("phase", C.c_double),
                ("envpar", C.c_double * 2)]


CARRIER = {"cos": 1, "sin": 2}
ENVELOPE = {"cos2": 1, "sin2ramp": 2, "box": 3, "sin2": 4}


def pulses_array(pulses):
    """ctypes array of adc2g_pulse from dicts with keys carrier ('cos'/'sin'),
    envelope ('cos2', 'sin2ramp', 'box', 'sin2'), pulse_vec (3), strength,
    freq, t0, phase and envpar (2)."""
    arr = (Pulse * max(len(pulses), 1))()
    for i, q in enumerate(pulses):
        arr[i].carrier = CARRIER[q.get("carrier", "cos")]
        arr[i].envelope = ENVELOPE[q["envelope"]]
        for c in range(3):
            arr[i].pulse_vec[c] = float(q["pulse_vec"][c])
        arr[i].strength = float(q["strength"])
        arr[i].freq = float(q.get("freq", 0.0))
        arr[i].t0 = float(q.get("t0", 0.0))
        arr[i].phase = float(q.get("phase", 0.0))
        ep = list(q.get("envpar", (0.0, 0.0))) + [0.0, 0.0]
        arr[i].envpar[0], arr[i].envpar[1] = float(ep[0]), float(ep[1])
    return arr


class LaserParams(C.Structure):
    _fields_ = [("tout", C.c_double), ("tfinal", C.c_double), ("kdim", C.c_int32), ("precision", C.c_int32),
                ("proptol", C.c_double), ("verbose", C.c_int32), ("reserved", C.c_int32)]


def read_is_operator(path, ndim):
    """An intermediate-state operator file as the reference writes them for
    SCRATCH/dipole_<c> and SCRATCH/cap and opxvec_ext reads them
    (source/propagation/tdselib.f90:1031-1153): Fortran sequential records
    (4-byte length markers) buf_size; the ndim diagonal elements; then
    records (oij(buf_size), indxi(buf_size), indxj(buf_size), nlim) with
    1-based indices, each off-diagonal element once.  Returns (diag, row,
    col, val) with 0-based indices, for Operator.set_dipole_is."""
    raw = open(path, "rb").read()
    recs, pos = [], 0
    while pos < len(raw):
        ln = int(np.frombuffer(raw, "<i4", 1, pos)[0])
        recs.append(raw[pos + 4: pos + 4 + ln])
        pos += 8 + ln
    bs = int(np.frombuffer(recs[0], "<i4", 1)[0])
    diag = np.frombuffer(recs[1], "<f8").copy()
    if diag.size != ndim:
        raise ValueError(f"{path}: {diag.size} diagonal elements, expected {ndim}")
    rows, cols, vals = [], [], []
    for r in recs[2:]:
        nlim = int(np.frombuffer(r, "<i4", 1, 16 * bs)[0])
        vals.append(np.frombuffer(r, "<f8", nlim, 0))
        rows.append(np.frombuffer(r, "<i4", nlim, 8 * bs).astype(np.int64) - 1)
        cols.append(np.frombuffer(r, "<i4", nlim, 12 * bs).astype(np.int64) - 1)
    cat = (lambda x, t: np.concatenate(x).astype(t) if x else np.zeros(0, t))  # noqa: E731
    return diag, cat(rows, np.int64), cat(cols, np.int64), cat(vals, np.float64)


def nto(trdens, c_occ=None, c_virt=None, thresh=1e-3):
    """Natural transition orbitals of an occupied-virtual transition density
    trdens[a, i] (nvirt x nocc, real or complex), as adc2_nto_gs and
    tdadc2_nto (nto.f90:183-236, 486-531) form them: the SVD
    trdens^T = U diag(sigma) V^H of the nocc x nvirt matrix (nocc <= nvirt),
    the pairs with sigma^2 >= thresh kept.  Returns (sigma (all nocc values),
    hole (nocc x npair, or nao x npair with the occupied MO coefficients
    c_occ), particle (nvirt x npair, or nao x npair with c_virt))."""
    T = np.asarray(trdens)
    nv, no = T.shape
    if no > nv:
        raise ValueError("the NTOs need nocc <= nvirt")
    U, sigma, Vh = np.linalg.svd(T.T, full_matrices=True)
    npair = int(np.count_nonzero(sigma ** 2 >= thresh))
    hole = U[:, :npair]
    particle = Vh.conj().T[:, :npair]
    if c_occ is not None:
        hole = np.asarray(c_occ) @ hole
    if c_virt is not None:
        particle = np.asarray(c_virt) @ particle
    return sigma, hole, particle


class LanczosParams(C.Structure):
    _fields_ = [("blocksize", C.c_int32), ("ncycles", C.c_int32), ("precision", C.c_int32),
                ("check_orth", C.c_int32), ("orthotype", C.c_int32), ("reserved", C.c_int32)]


class LanczosStats(C.Structure):
    _fields_ = [("nvec", C.c_int32), ("nsigma", C.c_int32), ("npro", C.c_int32), ("reserved", C.c_int32),
                ("time_sigma", C.c_double),
                ("time_tdiag", C.c_double), ("time_ritz", C.c_double), ("time_total", C.c_double),
                ("orth_loss", C.c_double)]


class RelaxParams(C.Structure):
    _fields_ = [("nstates", C.c_int32), ("krydim", C.c_int32), ("maxdim", C.c_int32),
                ("maxiter", C.c_int32), ("step", C.c_double), ("tol", C.c_double),
                ("precision", C.c_int32), ("verbose", C.c_int32), ("integrator", C.c_int32),
                ("reserved", C.c_int32), ("toler", C.c_double)]


class RelaxStats(C.Structure):
    _fields_ = [("nconv", C.c_int32), ("reserved", C.c_int32), ("nsigma", C.c_int64),
                ("time_sigma", C.c_double), ("time_total", C.c_double)]


# every symbol declared in include/adc2g.h
SYMBOLS = [
    "adc2g_create", "adc2g_destroy", "adc2g_dims", "adc2g_diagonal", "adc2g_phph",
    "adc2g_sigma", "adc2g_sigma_device", "adc2g_matxvec_treal", "adc2g_davidson",
    "adc2g_sil_autocorr", "adc2g_precon_harvest", "adc2g_precon_train", "adc2g_precon_nparams",
    "adc2g_precon_get", "adc2g_precon_set", "adc2g_precon_eval", "adc2g_last_error",
    "adc2g_sigma_profile", "adc2g_precon_absorb", "adc2g_coupling_form", "adc2g_cheby_auto",
    "adc2g_block_lanczos", "adc2g_relax", "adc2g_write_hmlt", "adc2g_davidson_vectors",
    "adc2g_sil_autocorr_orders", "adc2g_create_cvs", "adc2g_create_frozen_core", "adc2g_set_adc2x", "adc2g_mp2", "adc2g_fvector", "adc2g_lanczos_guess",
    "adc2g_spectral_bounds", "adc2g_power", "adc2g_coupled_schur", "adc2g_sigma_device_noalloc",
    "adc2g_set_dipole_is", "adc2g_matxvec_treal_laser", "adc2g_sil_laser", "adc2g_set_cap_is",
    "adc2g_set_cap_projector", "adc2g_set_flux_theta", "adc2g_trdens_vo", "adc2g_create_from_hmlt",
    "adc2g_cheby_fdstates", "adc2g_sil_fdstates",
]

_lib = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def load(path: str = LIBPATH):
    """Load libadc2g.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python -m adc_ottawa_amd.build`")
    lib = C.CDLL(path)
    vp = C.c_void_p
    i32, i64 = C.c_int32, C.c_int64
    lib.adc2g_create.argtypes = [i32, i32, i32, vp, vp, C.POINTER(Opts), C.POINTER(vp)]
    lib.adc2g_destroy.argtypes = [vp]
    lib.adc2g_dims.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
    lib.adc2g_diagonal.argtypes = [vp, vp]
    lib.adc2g_phph.argtypes = [vp, vp]
    lib.adc2g_coupled_schur.argtypes = [vp, C.c_double, i32, vp]
    lib.adc2g_sigma.argtypes = [vp, i32, vp, vp, i32]
    lib.adc2g_sigma_device.argtypes = [vp, i32, vp, vp, i32, vp]
    lib.adc2g_sigma_device_noalloc.argtypes = [vp, i32, vp, vp, i32, vp]
    lib.adc2g_matxvec_treal.argtypes = [vp, vp, vp, i32]
    lib.adc2g_davidson.argtypes = [vp, C.POINTER(DavParams), vp, vp, C.POINTER(DavStats)]
    lib.adc2g_sil_autocorr.argtypes = [vp, C.POINTER(SilParams), vp, vp, i64, C.POINTER(i64),
                                       C.POINTER(SilStats)]
    lib.adc2g_precon_harvest.argtypes = [vp, C.POINTER(DavParams), i64, C.POINTER(i64)]
    lib.adc2g_precon_train.argtypes = [vp, vp, i32, i32, C.c_double, C.c_uint64, vp]
    lib.adc2g_precon_nparams.argtypes = []
    lib.adc2g_precon_get.argtypes = [vp, vp]
    lib.adc2g_precon_set.argtypes = [vp, vp]
    lib.adc2g_precon_eval.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp]
    lib.adc2g_last_error.argtypes = [C.c_char_p, i32]
    lib.adc2g_sigma_profile.argtypes = [vp, i32, vp]
    lib.adc2g_coupling_form.argtypes = [vp, vp]
    lib.adc2g_cheby_auto.argtypes = [vp, vp, vp, i32, i32, vp]
    lib.adc2g_cheby_fdstates.argtypes = [vp, vp, vp, i32, i32, vp, i32, vp, vp]
    lib.adc2g_sil_fdstates.argtypes = [vp, C.POINTER(SilParams), vp, i32, vp, i32, i32, vp, vp, vp,
                                       C.POINTER(SilStats)]
    lib.adc2g_precon_absorb.argtypes = [vp, vp, C.POINTER(i64)]
    lib.adc2g_write_hmlt.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_double, i32, C.POINTER(i64),
                                     C.POINTER(i32)]
    lib.adc2g_davidson_vectors.argtypes = [vp, C.POINTER(DavParams), vp, vp, vp, C.POINTER(DavStats)]
    lib.adc2g_create_cvs.argtypes = [vp, i32, vp, C.POINTER(vp)]
    lib.adc2g_create_frozen_core.argtypes = [vp, i32, vp, C.POINTER(vp)]
    lib.adc2g_set_adc2x.argtypes = [vp, i32]
    lib.adc2g_mp2.argtypes = [vp, vp, vp, vp]
    lib.adc2g_fvector.argtypes = [vp, vp, vp]
    lib.adc2g_lanczos_guess.argtypes = [vp, i32, i32, i32, vp]
    lib.adc2g_spectral_bounds.argtypes = [vp, i32, i32, vp]
    lib.adc2g_power.argtypes = [vp, i32, i32, C.c_double, vp, vp, C.POINTER(i32)]
    lib.adc2g_sil_autocorr_orders.argtypes = [vp, C.POINTER(SilParams), vp, i32, vp, vp, vp, i64,
                                              C.POINTER(i64), C.POINTER(SilStats)]
    lib.adc2g_relax.argtypes = [vp, C.POINTER(RelaxParams), vp, vp, vp, vp, vp, C.POINTER(RelaxStats)]
    lib.adc2g_block_lanczos.argtypes = [vp, C.POINTER(LanczosParams), vp, vp, vp, vp, vp, vp,
                                        C.POINTER(LanczosStats)]
    lib.adc2g_set_dipole_is.argtypes = [vp, i32, vp, i64, vp, vp, vp, C.c_double, vp]
    lib.adc2g_matxvec_treal_laser.argtypes = [vp, C.c_double, i32, vp, vp, vp]
    lib.adc2g_set_cap_is.argtypes = [vp, vp, i64, vp, vp, vp, C.c_double, vp]
    lib.adc2g_set_cap_projector.argtypes = [vp, i32, vp, i32]
    lib.adc2g_sil_laser.argtypes = [vp, C.POINTER(LaserParams), i32, vp, vp, vp, vp, vp, i64, C.POINTER(i64),
                                    C.POINTER(SilStats)]
    lib.adc2g_trdens_vo.argtypes = [vp, i32, vp, i64, vp]
    lib.adc2g_create_from_hmlt.argtypes = [C.c_char_p, C.c_char_p, i32, C.POINTER(C.c_void_p)]
    lib.adc2g_set_flux_theta.argtypes = [vp, vp, i64, vp, vp, vp, C.c_double, vp]
    for s in SYMBOLS:
        getattr(lib, s).restype = C.c_int32
    _lib = lib
    return lib


def check(rc: int):
    if rc != ADC2G_OK:
        buf = C.create_string_buffer(1024)
        _lib.adc2g_last_error(buf, 1024)
        raise Adc2gError(rc, buf.value.decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Operator:
    """Device-resident ADC(2)-s operator of one system (``adc2g_create``)."""

    def __init__(self, nocc, nvirt, naux, eps, B, device=0, second_order=True, keep_lowprec=False,
                 coupling="auto"):
        lib = load()
        self._h = C.c_void_p()
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        nbas = nocc + nvirt
        if eps.size != nbas or B.shape != (naux, nbas, nbas):
            raise ValueError("eps/B shapes do not match nocc, nvirt, naux")
        o = Opts(device, 1 if second_order else 0, 1 if keep_lowprec else 0, COUPLING[coupling])
        check(lib.adc2g_create(nocc, nvirt, naux, _ptr(eps), _ptr(B), C.byref(o), C.byref(self._h)))
        nd, n1 = C.c_int64(), C.c_int64()
        cnt = (C.c_int64 * 5)()
        check(lib.adc2g_dims(self._h, C.byref(nd), C.byref(n1), cnt))
        self.ndim, self.n1, self.counts = nd.value, n1.value, list(cnt)
        self.nocc, self.nvirt, self.naux = nocc, nvirt, naux
        form = C.c_int32()
        check(lib.adc2g_coupling_form(self._h, C.byref(form)))
        self.coupling = {1: "df", 2: "int"}[form.value]

    @classmethod
    def from_system(cls, sys, **kw):
        return cls(sys.nocc, sys.nvirt, sys.naux, sys.eps, sys.B, **kw)

    @classmethod
    def from_hmlt(cls, dia_path, off_path, device=0):
        """A stored-matrix operator from the reference's hmlt.dia / hmlt.off
        files (adc2g_create_from_hmlt): the solvers, sigma, diagonal and
        write_hmlt work on it; what needs the orbitals raises Adc2gError."""
        lib = load()
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        check(lib.adc2g_create_from_hmlt(os.fsencode(dia_path), os.fsencode(off_path), int(device),
                                         C.byref(self._h)))
        nd, n1 = C.c_int64(), C.c_int64()
        cnt = (C.c_int64 * 5)()
        check(lib.adc2g_dims(self._h, C.byref(nd), C.byref(n1), cnt))
        self.ndim, self.n1, self.counts = nd.value, n1.value, list(cnt)
        self.nocc = self.nvirt = self.naux = 0
        self.coupling = "stored"
        return self

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            check(_lib.adc2g_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def diagonal(self):
        out = np.empty(self.ndim)
        check(_lib.adc2g_diagonal(self._h, _ptr(out)))
        return out

    def phph(self):
        out = np.empty((self.n1, self.n1))
        check(_lib.adc2g_phph(self._h, _ptr(out)))
        return out

    def coupled_schur(self, sigma, precision="fp64"):
        """A0 = H11 - C (D - sigma)^-1 C^T (n1 x n1) of the coupled preconditioner."""
        out = np.empty((self.n1, self.n1), order="F")
        check(_lib.adc2g_coupled_schur(self._h, float(sigma), PREC[precision], _ptr(out)))
        return out

    def sigma(self, X, precision="fp64"):
        X = np.asarray(X, dtype=np.float64)
        one = X.ndim == 1
        Xf = np.asfortranarray(X.reshape(self.ndim, -1))
        W = np.empty_like(Xf, order="F")
        check(_lib.adc2g_sigma(self._h, Xf.shape[1], _ptr(Xf), _ptr(W), PREC[precision]))
        return W[:, 0] if one else W

    def matxvec_treal(self, v1, precision="fp64"):
        v1 = np.ascontiguousarray(v1, dtype=np.complex128)
        v2 = np.empty_like(v1)
        check(_lib.adc2g_matxvec_treal(self._h, _ptr(v1), _ptr(v2), PREC[precision]))
        return v2

    def davidson(self, nstates, blocksize=0, maxvec=-1, maxiter=100, tol=1e-7, precon="dpr",
                 deflate=False, precision="fp64", refine=True, lowprec_tol=1e-4, vectors=True,
                 verbose=False, guess="ondiag", subdim=800):
        """Block Davidson.  guess: "ondiag", "subdiag" (with subdim) or an
        (ndim, blocksize) array of orthonormal initial vectors."""
        g = None
        if not isinstance(guess, str):
            g = np.asfortranarray(np.asarray(guess, dtype=np.float64))
            if g.shape[0] != self.ndim:
                raise ValueError("guess must have ndim rows")
            blocksize = g.shape[1]
            guess = "vectors"
        p = DavParams(nstates=nstates, blocksize=blocksize, maxvec=maxvec, maxiter=maxiter, tol=tol,
                      precon=PRECON[precon], deflate=1 if deflate else 0, precision=PREC[precision],
                      refine=1 if refine else 0, lowprec_tol=lowprec_tol, guess=GUESS[guess],
                      verbose=1 if verbose else 0, subdim=subdim)
        ev = np.empty(nstates)
        vec = np.empty((nstates, self.ndim)) if vectors else None
        st = DavStats()
        if g is not None:
            check(_lib.adc2g_davidson_vectors(self._h, C.byref(p), _ptr(g), _ptr(ev),
                                              _ptr(vec) if vectors else None, C.byref(st)))
        else:
            check(_lib.adc2g_davidson(self._h, C.byref(p), _ptr(ev),
                                      _ptr(vec) if vectors else None, C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in DavStats._fields_}
        return ev, (vec.T if vectors else None), stats

    def sil_autocorr(self, fvec, tout, tfinal, kdim, autotol, precision="fp64", verbose=False):
        fvec = np.ascontiguousarray(fvec, dtype=np.float64)
        nrec = int(tfinal / (0.5 * tout))
        out = np.empty(nrec + 1, dtype=np.complex128)
        p = SilParams(tout, tfinal, kdim, PREC[precision], autotol, 1 if verbose else 0, 0)
        n = C.c_int64()
        st = SilStats()
        check(_lib.adc2g_sil_autocorr(self._h, C.byref(p), _ptr(fvec), _ptr(out), out.size,
                                      C.byref(n), C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in SilStats._fields_}
        return out[: n.value], stats

    def sil_autocorr_orders(self, fvec, tout, tfinal, kdim, autotol, autoord, precision="fp64",
                            verbose=False):
        """SIL autocorrelation functions of orders 0..autoord; returns
        ([auto0, auto1, ...], stats)."""
        fvec = np.ascontiguousarray(fvec, dtype=np.float64)
        nrec = int(tfinal / (0.5 * tout))
        outs = [np.empty(nrec + 1, dtype=np.complex128) for _ in range(autoord + 1)]
        ptrs = [_ptr(o) for o in outs] + [None] * (2 - autoord)
        p = SilParams(tout, tfinal, kdim, PREC[precision], autotol, 1 if verbose else 0, 0)
        n = C.c_int64()
        st = SilStats()
        check(_lib.adc2g_sil_autocorr_orders(self._h, C.byref(p), _ptr(fvec), int(autoord), ptrs[0], ptrs[1],
                                             ptrs[2], nrec + 1, C.byref(n), C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in SilStats._fields_}
        return [o[: n.value] for o in outs], stats

    def cheby_auto(self, q0, bounds, order, precision="fp64"):
        """Chebyshev order-domain autocorrelation C_0..C_2K (K = order) of q0
        with the spectrum scaled by bounds = (Emin, Emax)."""
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        if q0.size != self.ndim:
            raise ValueError("q0 must have ndim entries")
        b = np.ascontiguousarray(bounds, dtype=np.float64)
        out = np.empty(2 * order + 1)
        check(_lib.adc2g_cheby_auto(self._h, _ptr(b), _ptr(q0), int(order), PREC[precision], _ptr(out)))
        return out

    def cheby_fdstates(self, q0, bounds, coef, precision="fp64"):
        """Chebyshev filter-diagonalisation eigenstates (calc_fdstates_cfd):
        coef[k, i] (order + 1 rows, one column per selected state) weights
        T_k(H_norm) q0; returns (energies, vectors (ndim x nsel), normalised)."""
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        if q0.size != self.ndim:
            raise ValueError("q0 must have ndim entries")
        c = np.asfortranarray(np.asarray(coef, dtype=np.float64).reshape(len(coef), -1))
        nk, nsel = c.shape
        b = np.ascontiguousarray(bounds, dtype=np.float64)
        vecs = np.empty((self.ndim, nsel), order="F")
        e = np.empty(nsel)
        check(_lib.adc2g_cheby_fdstates(self._h, _ptr(b), _ptr(q0), nk - 1, nsel, _ptr(c), PREC[precision],
                                        _ptr(vecs), _ptr(e)))
        return e, vecs

    def sil_fdstates(self, fvec, tout, tfinal, kdim, autotol, ebound, fbas2eig, iwfunc=1, precision="fp64"):
        """Time-domain filter-diagonalisation states (calc_fdstates_tdfd):
        fbas2eig (nsel x nfbas) combines the filter states of nfbas energies
        on ebound; returns (energies, vectors (ndim x nsel), stats)."""
        fvec = np.ascontiguousarray(fvec, dtype=np.float64)
        if fvec.size != self.ndim:
            raise ValueError("fvec must have ndim entries")
        F = np.asfortranarray(np.atleast_2d(np.asarray(fbas2eig, dtype=np.float64)))
        nsel, nfbas = F.shape
        eb = np.ascontiguousarray(ebound, dtype=np.float64)
        vecs = np.empty((self.ndim, nsel), order="F")
        e = np.empty(nsel)
        p = SilParams(tout, tfinal, kdim, PREC[precision], autotol, 0, 0)
        st = SilStats()
        check(_lib.adc2g_sil_fdstates(self._h, C.byref(p), _ptr(fvec), nfbas, _ptr(eb), int(iwfunc), nsel, _ptr(F),
                                      _ptr(vecs), _ptr(e), C.byref(st)))
        return e, vecs, {f: getattr(st, f) for f, _ in SilStats._fields_}

    def block_lanczos(self, q0, ncycles, precision="fp64", travec=None, tmat=False, ritz=False,
                      check_orth=False, orthotype=0):
        """Block Lanczos pseudospectrum from the orthonormal block q0 (ndim x b).

        Returns (eigval, extras, stats); extras holds "tranmom" (travec given),
        "tmat" and "ritz" (ndim x nvec) when asked for."""
        q0 = np.asfortranarray(np.asarray(q0, dtype=np.float64).reshape(self.ndim, -1))
        b = q0.shape[1]
        nvec = b * int(ncycles)
        p = LanczosParams(b, int(ncycles), PREC[precision], 1 if check_orth else 0, int(orthotype), 0)
        ev = np.empty(nvec)
        extras = {}
        tm_out = f = None
        if travec is not None:
            f = np.ascontiguousarray(travec, dtype=np.float64)
            if f.size != self.ndim:
                raise ValueError("travec must have ndim entries")
            tm_out = extras["tranmom"] = np.empty(nvec)
        t_out = extras["tmat"] = np.empty((nvec, nvec), order="F") if tmat else None
        r_out = extras["ritz"] = np.empty((self.ndim, nvec), order="F") if ritz else None
        st = LanczosStats()
        check(_lib.adc2g_block_lanczos(
            self._h, C.byref(p), _ptr(q0), _ptr(ev), _ptr(t_out) if tmat else None,
            _ptr(f) if f is not None else None, _ptr(tm_out) if f is not None else None,
            _ptr(r_out) if ritz else None, C.byref(st)))
        extras = {k: v for k, v in extras.items() if v is not None}
        stats = {k: getattr(st, k) for k, _ in LanczosStats._fields_}
        return ev, extras, stats

    def relax(self, nstates, guess=None, krydim=10, maxdim=-1, maxiter=100, step=10.0, tol=1e-7,
              precision="fp64", vectors=True, verbose=False, integrator="xsil", toler=1e-5):
        """Imaginary-time relaxation of the nstates lowest states, with the XSIL,
        the Bulirsch-Stoer ('bs') or the RKF45 ('rkf45') integrator (error
        tolerance toler for the last two).

        guess: (ndim, nstates) or None (unit vectors on the smallest |H_jj|).
        Returns (energies, vectors or None, residuals, iterations, stats)."""
        g = None
        if guess is not None:
            g = np.asfortranarray(np.asarray(guess, dtype=np.float64).reshape(self.ndim, nstates))
        p = RelaxParams(nstates, krydim, maxdim, maxiter, step, tol, PREC[precision], 1 if verbose else 0,
                        {"xsil": 1, "bs": 2, "rkf45": 3}[integrator], 0, toler)
        ev = np.empty(nstates)
        vec = np.empty((self.ndim, nstates), order="F") if vectors else None
        res = np.empty(nstates)
        nit = np.empty(nstates, dtype=np.int32)
        st = RelaxStats()
        check(_lib.adc2g_relax(self._h, C.byref(p), _ptr(g) if g is not None else None, _ptr(ev),
                               _ptr(vec) if vectors else None, _ptr(res), _ptr(nit), C.byref(st)))
        stats = {k: getattr(st, k) for k, _ in RelaxStats._fields_}
        return ev, vec, res, nit, stats

    def write_hmlt(self, dia_path, off_path, minc=1e-12, maxbl=0):
        """hmlt.dia / hmlt.off files of the matrix; returns (noffd, nbuf)."""
        nz, nb = C.c_int64(), C.c_int32()
        check(_lib.adc2g_write_hmlt(self._h, os.fsencode(dia_path), os.fsencode(off_path), float(minc),
                                    int(maxbl), C.byref(nz), C.byref(nb)))
        return nz.value, nb.value

    def mp2(self, density=True):
        """(E_MP2, D2, rho or None) of mp2_ener, mp2_d2 and rho_mp2."""
        e, d = C.c_double(), C.c_double()
        nb = self.nocc + self.nvirt
        rho = np.empty((nb, nb), order="F") if density else None
        check(_lib.adc2g_mp2(self._h, C.byref(e), C.byref(d), _ptr(rho) if density else None))
        return e.value, d.value, rho

    def fvector(self, dpl):
        """f_J = <Psi_J|D|Psi_0> (ndim) of get_modifiedtm_adc2 for the nbas x nbas operator matrix dpl."""
        nb = self.nocc + self.nvirt
        d = np.asfortranarray(dpl, dtype=np.float64)
        if d.shape != (nb, nb):
            raise ValueError(f"dpl must be {nb} x {nb}")
        f = np.empty(self.ndim)
        check(_lib.adc2g_fvector(self._h, _ptr(d), _ptr(f)))
        return f

    def spectral_bounds(self, method="davlanc", ncycles=0):
        """(Emin, Emax) as spectral_bounds estimates them ('davlanc', 'davpower' or 'lanczos')."""
        b = np.empty(2)
        check(_lib.adc2g_spectral_bounds(self._h, {"davlanc": 1, "davpower": 2, "lanczos": 3}[method],
                                         int(ncycles), _ptr(b)))
        return float(b[0]), float(b[1])

    def lanczos_guess(self, kind, blocksize, subdim=0):
        """Block Lanczos start block: kind 8 init_vec_doublesubspace, 9 init_vec_singlesubspace."""
        q0 = np.empty((self.ndim, blocksize), order="F")
        check(_lib.adc2g_lanczos_guess(self._h, int(kind), int(subdim), int(blocksize), _ptr(q0)))
        return q0

    def power(self, subdim=0, maxit=0, thrsh=0.0, vector=False):
        """Highest eigenpair by the power method: (E, v or None, iterations)."""
        e = C.c_double()
        it = C.c_int32()
        vec = np.empty(self.ndim) if vector else None
        check(_lib.adc2g_power(self._h, int(subdim), int(maxit), float(thrsh), C.byref(e),
                               _ptr(vec) if vector else None, C.byref(it)))
        return e.value, vec, it.value

    def cvs(self, core):
        """CVS view on the core orbitals ``core`` (1-based occupied indices);
        close it before this operator."""
        return CvsOperator(self, core)

    def set_adc2x(self, enable=True):
        """Switch the operator to ADC(2)-x (first-order 2h2p-2h2p block in the
        sigma-vector and the diagonal) or back to ADC(2)-s."""
        check(_lib.adc2g_set_adc2x(self._h, 1 if enable else 0))

    def set_dipole_is(self, component, diag, row, col, val, d00, d0j):
        """Intermediate-state representation of dipole component 0/1/2 (x/y/z):
        diagonal (ndim), off-diagonal elements once each (0-based row, col,
        val), ground-state element d00 and column d0j (ndim).  diag = d0j =
        None with no elements and d00 = 0 clears the component."""
        row = np.ascontiguousarray(row, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int64)
        val = np.ascontiguousarray(val, dtype=np.float64)
        if not (row.size == col.size == val.size):
            raise ValueError("row, col and val must have the same length")
        if diag is None and d0j is None:
            if row.size or d00 != 0.0:
                raise ValueError("diag and d0j are required unless the component is cleared")
            check(_lib.adc2g_set_dipole_is(self._h, int(component), None, 0, None, None, None, 0.0, None))
            return
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        d0j = np.ascontiguousarray(d0j, dtype=np.float64)
        if diag.shape != (self.ndim,) or d0j.shape != (self.ndim,):
            raise ValueError("diag and d0j must have ndim elements")
        check(_lib.adc2g_set_dipole_is(self._h, int(component), _ptr(diag), row.size, _ptr(row), _ptr(col),
                                       _ptr(val), float(d00), _ptr(d0j)))

    def set_cap_is(self, diag, row, col, val, w00, w0j):
        """The CAP's intermediate-state representation (as set_dipole_is);
        diag = w0j = None with no elements and w00 = 0 removes the CAP."""
        row = np.ascontiguousarray(row, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int64)
        val = np.ascontiguousarray(val, dtype=np.float64)
        if not (row.size == col.size == val.size):
            raise ValueError("row, col and val must have the same length")
        if diag is None and w0j is None:
            if row.size or w00 != 0.0:
                raise ValueError("diag and w0j are required unless the CAP is removed")
            check(_lib.adc2g_set_cap_is(self._h, None, 0, None, None, None, 0.0, None))
            return
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        w0j = np.ascontiguousarray(w0j, dtype=np.float64)
        if diag.shape != (self.ndim,) or w0j.shape != (self.ndim,):
            raise ValueError("diag and w0j must have ndim elements")
        check(_lib.adc2g_set_cap_is(self._h, _ptr(diag), row.size, _ptr(row), _ptr(col), _ptr(val), float(w00),
                                    _ptr(w0j)))

    def set_cap_projector(self, vecs=None, ground=False):
        """Project the CAP term onto the complement of the orthonormal IS
        vectors vecs (ndim x nvec) and, with ground, of the ground state."""
        if vecs is None or np.size(vecs) == 0:
            check(_lib.adc2g_set_cap_projector(self._h, 0, None, 1 if ground else 0))
            return
        V = np.asfortranarray(np.asarray(vecs, dtype=np.float64).reshape(self.ndim, -1))
        check(_lib.adc2g_set_cap_projector(self._h, V.shape[1], _ptr(V), 1 if ground else 0))

    def matxvec_treal_laser(self, t, pulses, v1):
        """-i H(t) v1 for complex v1 of ndim + 1 components (ground state last)."""
        v1 = np.ascontiguousarray(v1, dtype=np.complex128)
        if v1.shape != (self.ndim + 1,):
            raise ValueError("v1 must have ndim + 1 components")
        v2 = np.empty_like(v1)
        arr = pulses_array(pulses)
        check(_lib.adc2g_matxvec_treal_laser(self._h, float(t), len(pulses), C.cast(arr, C.c_void_p), _ptr(v1),
                                             _ptr(v2)))
        return v2

    def set_flux_theta(self, diag, row, col, val, theta00, theta0j):
        """The projector Theta onto the absorbing region for the CAP flux (as
        set_dipole_is); diag = theta0j = None with no elements clears it."""
        row = np.ascontiguousarray(row, dtype=np.int64)
        col = np.ascontiguousarray(col, dtype=np.int64)
        val = np.ascontiguousarray(val, dtype=np.float64)
        if not (row.size == col.size == val.size):
            raise ValueError("row, col and val must have the same length")
        if diag is None and theta0j is None:
            if row.size or theta00 != 0.0:
                raise ValueError("diag and theta0j are required unless Theta is cleared")
            check(_lib.adc2g_set_flux_theta(self._h, None, 0, None, None, None, 0.0, None))
            return
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        theta0j = np.ascontiguousarray(theta0j, dtype=np.float64)
        if diag.shape != (self.ndim,) or theta0j.shape != (self.ndim,):
            raise ValueError("diag and theta0j must have ndim elements")
        check(_lib.adc2g_set_flux_theta(self._h, _ptr(diag), row.size, _ptr(row), _ptr(col), _ptr(val),
                                        float(theta00), _ptr(theta0j)))

    def sil_laser(self, psi0, pulses, tout, tfinal, kdim=20, proptol=1e-8, precision="fp64", verbose=False,
                  flux=False, trdens=False):
        """SIL propagation of psi0 (ndim + 1 complex components) under the
        pulses; returns (psi(tfinal), norms at i * tout, stats), with after
        the norms the CAP flux at the same times when flux is set and the
        transition densities of psi_IS / |psi_IS| (records x nvirt x nocc,
        complex; the TD-NTO input) when trdens is set."""
        psi = np.array(psi0, dtype=np.complex128)
        if psi.shape != (self.ndim + 1,):
            raise ValueError("psi0 must have ndim + 1 components")
        nrec = int(tfinal / tout)
        norms = np.empty(nrec + 1)
        fl = np.empty(nrec + 1) if flux else None
        n1 = self.nocc * self.nvirt
        td = np.empty((nrec + 1) * n1, dtype=np.complex128) if trdens else None
        p = LaserParams(tout, tfinal, kdim, PREC[precision], proptol, 1 if verbose else 0, 0)
        arr = pulses_array(pulses)
        n = C.c_int64()
        st = SilStats()
        check(_lib.adc2g_sil_laser(self._h, C.byref(p), len(pulses), C.cast(arr, C.c_void_p), _ptr(psi),
                                   _ptr(norms), _ptr(fl) if flux else None, _ptr(td) if trdens else None,
                                   norms.size, C.byref(n), C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in SilStats._fields_}
        out = [psi, norms[: n.value]]
        if flux:
            out.append(fl[: n.value])
        if trdens:
            out.append(td.reshape(nrec + 1, self.nocc, self.nvirt).transpose(0, 2, 1)[: n.value].copy())
        return tuple(out) + (stats,)

    def trdens_vo(self, x):
        """The occupied-virtual ground-to-excited transition density
        trdens[a, i] (nvirt x nocc) of adc2_trdens_gs_virt_occ for x of
        ndim components, or (nvec, nvirt, nocc) for x of shape (ndim, nvec)."""
        x = np.asarray(x, dtype=np.float64)
        one = x.ndim == 1
        X = np.asfortranarray(x.reshape(self.ndim, -1))
        if X.shape[0] != self.ndim:
            raise ValueError("x must have ndim rows")
        k = X.shape[1]
        out = np.empty(k * self.nocc * self.nvirt)
        check(_lib.adc2g_trdens_vo(self._h, k, _ptr(X), self.ndim, _ptr(out)))
        T = out.reshape(k, self.nocc, self.nvirt).transpose(0, 2, 1)
        return T[0].copy() if one else T.copy()

    def frozen_core(self, frozen):
        """Frozen-core view: the configurations with no hole in the orbitals
        ``frozen`` (1-based occupied indices); close it before this operator."""
        return FrozenCoreOperator(self, frozen)

    def sigma_device(self, nvec, dX, dW, precision="fp64", stream=None):
        """W = H X on device pointers (ints), asynchronous on ``stream``."""
        check(_lib.adc2g_sigma_device(self._h, nvec, C.c_void_p(dX), C.c_void_p(dW), PREC[precision],
                                      C.c_void_p(stream) if stream else None))

    def profile(self, enable=True):
        """Collect the per-kernel timing recorded so far and (re)arm it.
        Returns {class: (ms, launches, work)} for coupling/pack/unpack."""
        out = np.zeros(12)
        check(_lib.adc2g_sigma_profile(self._h, 1 if enable else 0, _ptr(out)))
        names = ("coupling_gemm", "csf_to_pair", "pair_to_csf_diag", "coupling_minor")
        return {n: tuple(out[3 * i:3 * i + 3]) for i, n in enumerate(names)}

    # ---- learned preconditioner
    def precon_harvest(self, nstates, max_records=1 << 40, base="dpr", **kw):
        """Training records with the network's base step: "dpr" or "coupled"."""
        p = DavParams(nstates=nstates, blocksize=kw.get("blocksize", 0), maxvec=kw.get("maxvec", -1),
                      maxiter=kw.get("maxiter", 200), tol=kw.get("tol", 1e-7), precon=PRECON[base])
        added = C.c_int64()
        check(_lib.adc2g_precon_harvest(self._h, C.byref(p), max_records, C.byref(added)))
        return added.value

    def absorb(self, other):
        """Append other's harvested records to this handle's; returns the count."""
        n = C.c_int64()
        check(_lib.adc2g_precon_absorb(self._h, other.handle, C.byref(n)))
        return n.value

    def precon_train(self, data=None, epochs=10, batch=8192, lr=1e-3, seed=0):
        data = self if data is None else data
        loss = np.empty(max(1, epochs))
        check(_lib.adc2g_precon_train(self._h, data.handle, epochs, batch, lr, seed, _ptr(loss)))
        return loss[:epochs]

    def precon_params(self):
        out = np.empty(_lib.adc2g_precon_nparams(), dtype=np.float32)
        check(_lib.adc2g_precon_get(self._h, _ptr(out)))
        return out

    def set_precon_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float32)
        check(_lib.adc2g_precon_set(self._h, _ptr(params)))

    def precon_eval(self, feat, cls, dn, en):
        n = feat.shape[0]
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        cls = np.ascontiguousarray(cls, dtype=np.int32)
        dn = np.ascontiguousarray(dn, dtype=np.float32)
        en = np.ascontiguousarray(en, dtype=np.float32)
        m = np.empty(n, dtype=np.float32)
        g = np.empty(_lib.adc2g_precon_nparams(), dtype=np.float32)
        loss = C.c_double()
        check(_lib.adc2g_precon_eval(self._h, n, _ptr(feat), _ptr(cls), _ptr(dn), _ptr(en), _ptr(m),
                                     _ptr(g), C.byref(loss)))
        return m, g, loss.value


class _View(Operator):
    """A configuration-subspace view of an Operator: the principal sub-matrix
    on the view's configurations, with the same solver methods."""

    _create = None

    def __init__(self, parent, orbitals):
        lib = load()
        self._h = C.c_void_p()
        orbs = np.ascontiguousarray(orbitals, dtype=np.int32)
        check(getattr(lib, self._create)(parent.handle, orbs.size, _ptr(orbs), C.byref(self._h)))
        self._parent = parent
        nd, n1 = C.c_int64(), C.c_int64()
        cnt = (C.c_int64 * 5)()
        check(lib.adc2g_dims(self._h, C.byref(nd), C.byref(n1), cnt))
        self.ndim, self.n1, self.counts = nd.value, n1.value, list(cnt)
        self.nocc, self.nvirt, self.naux = parent.nocc, parent.nvirt, parent.naux
        self.coupling = parent.coupling
        self.orbitals = [int(c) for c in orbs]


class CvsOperator(_View):
    """CVS view (``adc2g_create_cvs``) on the core orbitals (1-based occupied
    indices): the core-hole configurations."""

    _create = "adc2g_create_cvs"

    @property
    def core(self):
        return self.orbitals


class FrozenCoreOperator(_View):
    """Frozen-core view (``adc2g_create_frozen_core``) of the frozen orbitals
    (1-based occupied indices): the configurations with no hole in them."""

    _create = "adc2g_create_frozen_core"

"""ctypes wrapper of oracle/coo_sigma.c: the reference's stored-matrix CPU
sigma-vector (COO build + hxvec_incore), used as the checker of small cases
and as the CPU baseline timing.  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libcoo_sigma.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        lib.coo_ndim.restype = i64
        lib.coo_ndim.argtypes = [i32, i32]
        lib.coo_capacity.restype = i64
        lib.coo_capacity.argtypes = [i32, i32, i32, i32]
        lib.coo_row_coupling_count.restype = i64
        lib.coo_row_coupling_count.argtypes = [i32, i32]
        lib.coo_build.restype = i64
        lib.coo_build.argtypes = [i32, i32, i32, vp, vp, i32, i32, vp, vp, vp]
        lib.coo_hxvec.restype = C.c_double
        lib.coo_hxvec.argtypes = [i64, i32, vp, i64, vp, vp, vp, vp, vp, C.POINTER(C.c_double)]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class COO:
    """COO records of the 1h1p rows a in [a0, a1) (all k)."""

    def __init__(self, sys, h11=None, a0=0, a1=None):
        lib = load()
        o, v = sys.nocc, sys.nvirt
        a1 = v if a1 is None else a1
        cap = lib.coo_capacity(o, v, a0, a1)
        self.val = np.empty(cap)
        self.oi = np.empty(cap, dtype=np.int32)
        self.oj = np.empty(cap, dtype=np.int32)
        B = np.ascontiguousarray(sys.B)
        h = None if h11 is None else np.ascontiguousarray(h11)
        nnz = lib.coo_build(o, v, sys.naux, _p(B), None if h is None else _p(h), a0, a1,
                            _p(self.val), _p(self.oi), _p(self.oj))
        self.nnz = nnz
        self.ndim = lib.coo_ndim(o, v)
        self.rows = (a1 - a0) * o
        self.full_nnz = o * v * (o * v - 1) // 2 + o * v * lib.coo_row_coupling_count(o, v)

    def hxvec(self, hii, V):
        """W = H V through hxvec_incore; returns (W, t_offdiag, t_diag)."""
        V = np.asfortranarray(V.reshape(self.ndim, -1))
        W = np.empty_like(V, order="F")
        td = C.c_double()
        t = _lib.coo_hxvec(self.ndim, V.shape[1], _p(np.ascontiguousarray(hii)), self.nnz, _p(self.val),
                           _p(self.oi), _p(self.oj), _p(V), _p(W), C.byref(td))
        return W, t, td.value

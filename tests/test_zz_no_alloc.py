"""adc2g_sigma_device_noalloc (the sigma-vector that allocates nothing, all
workspaces in the library's static scratch) against adc2g_sigma_device on the
same device buffers, through the C ABI: two calls in a row, the second with
fewer vectors; FP64 to 1e-12 relative, FP32/BF16 to 1e-5 of the largest
element (the passes differ only in how many vectors share a GEMM, which can
change split-K summation orders).  Error paths: a CVS view and a precision
whose operands are not resident are rejected without work."""
import ctypes as C

import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import SIZES, make_system

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]

    def malloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def h2d(self, dst, arr):
        assert self.h.hipMemcpy(C.c_void_p(dst), arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0

    def d2h(self, arr, src):
        assert self.h.hipMemcpy(arr.ctypes.data_as(C.c_void_p), C.c_void_p(src), arr.nbytes, 2) == 0

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self, p):
        self.h.hipFree(C.c_void_p(p))


def run_pair(op, hip, X, prec):
    """W from adc2g_sigma_device and from adc2g_sigma_device_noalloc."""
    L = lib.load()
    ndim, nvec = X.shape
    Xf = np.asfortranarray(X)
    dX, dA, dB = hip.malloc(Xf.nbytes), hip.malloc(Xf.nbytes), hip.malloc(Xf.nbytes)
    try:
        hip.h2d(dX, Xf)
        lib.check(L.adc2g_sigma_device(op.handle, nvec, C.c_void_p(dX), C.c_void_p(dA), lib.PREC[prec], None))
        lib.check(L.adc2g_sigma_device_noalloc(op.handle, nvec, C.c_void_p(dX), C.c_void_p(dB), lib.PREC[prec],
                                               None))
        hip.sync()
        A = np.empty_like(Xf)
        B = np.empty_like(Xf)
        hip.d2h(A, dA)
        hip.d2h(B, dB)
    finally:
        for p in (dX, dA, dB):
            hip.free(p)
    return A, B


@pytest.mark.parametrize("form", ["df", "int"])
def test_noalloc_small(form):
    print(f"[case] noalloc small o=4 v=12 naux=20 coupling={form}", flush=True)
    hip = Hip()
    sys_ = make_system(4, 12, 20, seed=3)
    rng = np.random.default_rng(5)
    with lib.Operator.from_system(sys_, coupling=form, keep_lowprec=True) as op:
        for prec, tol in (("fp64", 1e-12), ("fp32", 1e-5), ("bf16", 1e-5)):
            for nvec in (7, 3):
                X = rng.standard_normal((op.ndim, nvec))
                A, B = run_pair(op, hip, X, prec)
                err = np.abs(A - B).max() / np.abs(A).max()
                print(f"{prec} nvec {nvec}: max |W_noalloc - W| / max |W| = {err:.2e}")
                assert err <= tol


def test_noalloc_benzene_and_errors():
    sz = SIZES["benzene"]
    print(f"[case] noalloc benzene {sz}", flush=True)
    hip = Hip()
    sys_ = make_system(seed=0, **sz)
    rng = np.random.default_rng(6)
    L = lib.load()
    with lib.Operator.from_system(sys_) as op:
        # FP32 operands are not resident yet: rejected before any work
        X = rng.standard_normal((op.ndim, 1))
        dX = hip.malloc(X.nbytes)
        dW = hip.malloc(X.nbytes)
        rc = L.adc2g_sigma_device_noalloc(op.handle, 1, C.c_void_p(dX), C.c_void_p(dW), lib.PREC["fp32"], None)
        hip.free(dX)
        hip.free(dW)
        assert rc == 10
        # FP64: two vectors per pass at this size; 5 vectors then 2
        for nvec in (5, 2):
            X = rng.standard_normal((op.ndim, nvec))
            A, B = run_pair(op, hip, X, "fp64")
            err = np.abs(A - B).max() / np.abs(A).max()
            print(f"fp64 nvec {nvec}: max |W_noalloc - W| / max |W| = {err:.2e}")
            assert err <= 1e-12
        with op.cvs([1]) as view:
            Xv = np.zeros((view.ndim, 1))
            dX = hip.malloc(Xv.nbytes)
            dW = hip.malloc(Xv.nbytes)
            rc = L.adc2g_sigma_device_noalloc(view.handle, 1, C.c_void_p(dX), C.c_void_p(dW), lib.PREC["fp64"],
                                              None)
            hip.free(dX)
            hip.free(dW)
            assert rc == 10

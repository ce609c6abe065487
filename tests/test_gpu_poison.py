"""No kernel reads device memory that no kernel has written.

With ADC2G_POISON=1 the library fills every new floating-point device buffer
with NaN bytes, so a read of unwritten memory (even one multiplied by zero)
turns the results into NaN.  Reused device memory can hold such bytes in
ordinary runs, where a read of this kind failed now and then (the inverse
iteration of the small symmetric eigensolver read the last row's
super-diagonals, times zero, before writing them).  One child process with
the variable set runs the Davidson with each correction step, the ADC(2)-x
block and a frozen-core view, and compares the roots with those of this
process (no poison), whose roots other tests compare with the oracle.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system
out = {}
sys_ = make_system(5, 36, 164, seed=0)
with lib.Operator.from_system(sys_) as op:
    for pre in ("dpr", "olsen", "coupled"):
        ev, _, st = op.davidson(5, blocksize=5, maxvec=40, tol=1e-7, precon=pre, maxiter=500, vectors=False)
        out[pre] = [float(x) for x in ev]
    with op.frozen_core([1]) as view:
        ev, _, _ = view.davidson(4, blocksize=8, maxvec=40, tol=1e-7, precon="olsen", maxiter=500, vectors=False)
        out["frozen"] = [float(x) for x in ev]
    op.set_adc2x(True)
    ev, _, _ = op.davidson(4, blocksize=8, maxvec=40, tol=1e-7, precon="olsen", maxiter=500, vectors=False)
    out["adcx"] = [float(x) for x in ev]
print("RESULT " + json.dumps(out), flush=True)
"""


def test_poisoned_buffers_give_the_same_roots():
    print("[case] poisoned buffers o=5 v=36 naux=164", flush=True)
    env = dict(os.environ, ADC2G_POISON="1")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = json.loads(next(x for x in r.stdout.splitlines() if x.startswith("RESULT "))[7:])
    sys_ = make_system(5, 36, 164, seed=0)
    with lib.Operator.from_system(sys_) as op:
        ev, _, _ = op.davidson(5, blocksize=5, maxvec=40, tol=1e-7, precon="dpr", maxiter=500, vectors=False)
        with op.frozen_core([1]) as view:
            ev_f, _, _ = view.davidson(4, blocksize=8, maxvec=40, tol=1e-7, precon="olsen", maxiter=500,
                                       vectors=False)
        op.set_adc2x(True)
        ev_x, _, _ = op.davidson(4, blocksize=8, maxvec=40, tol=1e-7, precon="olsen", maxiter=500, vectors=False)
    for pre in ("dpr", "olsen", "coupled"):
        assert np.abs(np.array(got[pre]) - ev).max() < 1e-9, pre
    assert np.abs(np.array(got["frozen"]) - ev_f).max() < 1e-9
    assert np.abs(np.array(got["adcx"]) - ev_x).max() < 1e-9

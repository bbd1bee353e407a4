"""pt_eval_noalloc -- the evaluation that allocates nothing (its per-trial staging lives in
the library's fixed device scratch) -- against pt_eval(PT_FORM_BATCHED), the entry point
whose kernels it shares: bit for bit, on two calls in a row (the second smaller), then
the limits it refuses.  Device buffers come from the HIP runtime the library itself links
(hipMalloc through ctypes), outside the entry point under test."""
import ctypes

import numpy as np
import pytest

from powder6_amd import problems as PR
from powder6_amd import pt

pytestmark = pytest.mark.gpu


def _hip():
    pt.lib()                                   # loads the library and with it the HIP runtime it links
    h = ctypes.CDLL('libamdhip64.so.7')
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, hip, host):
        self.hip, self.n = hip, host.size
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), host.nbytes) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc(plan, hip, X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    dx, dc = DeviceArray(hip, X), DeviceArray(hip, np.full(len(X), -1.0))
    try:
        L = pt.lib()
        L.pt_eval_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        rc = L.pt_eval_noalloc(plan.h, dx.ptr, len(X), dc.ptr)
        if rc < 0:
            raise pt.PTError(L.pt_last_error().decode())
        return dc.get()
    finally:
        dx.free()
        dc.free()


def _trials(pb, n, seed):
    rng = np.random.default_rng(seed)
    x0 = pb.start_vector()
    X = x0 + rng.uniform(-1, 1, size=(n, pb.nv)) * np.array(pb.scale)
    return np.minimum(np.array(pb.hi), np.maximum(np.array(pb.lo), X))


def test_noalloc_equals_batched_on_two_calls():
    pb = PR.pbso4_pnma(simulate=pt.gpu_simulate, chains=16)
    hip = _hip()
    with pt.Plan(pb) as plan:
        for n, seed in ((300, 1), (37, 2)):
            X = _trials(pb, n, seed)
            want = plan.eval(X, pt.FORM_BATCHED)
            got = _noalloc(plan, hip, X)
            assert np.all(np.isfinite(got)) and np.all(got < 3e38)
            assert np.array_equal(got, want), np.max(np.abs(got - want))


def test_noalloc_refuses_what_the_scratch_cannot_hold():
    pb = PR.pbso4_pnma(simulate=pt.gpu_simulate, chains=16)
    hip = _hip()
    with pt.Plan(pb) as plan:
        with pytest.raises(pt.PTError, match='at most'):
            _noalloc(plan, hip, _trials(pb, 4000, 3))
        # the refusal leaves the scratch usable
        X = _trials(pb, 8, 4)
        assert np.array_equal(_noalloc(plan, hip, X), plan.eval(X, pt.FORM_BATCHED))


def test_noalloc_refuses_multi_dataset_and_free_intensity_plans():
    hip = _hip()
    pbs = PR.xray_neutron_pair(simulate=pt.gpu_simulate)
    with pt.Plan(pbs) as plan:
        with pytest.raises(pt.PTError, match='one dataset'):
            _noalloc(plan, hip, _trials(pbs[0], 4, 5))
    pb = PR.lebail_probe(simulate=pt.gpu_simulate)
    with pt.Plan(pb) as plan:
        with pytest.raises(pt.PTError, match='free intensities'):
            _noalloc(plan, hip, _trials(pb, 4, 6))

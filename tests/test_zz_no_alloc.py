"""dh_c2p_grad_noalloc on the GPU: the loss gradient of the learned guess
computed in the library's fixed device array (no allocation) equals the
allocating dh_c2p_grad bit for bit, on two calls in a row (the second with
fewer rows, so it runs over what the first left in the array), and a call
with more rows than the array holds is refused.  Device buffers come from
the HIP runtime the library itself uses."""
import ctypes as C

import numpy as np
import pytest

from disco_amd import c2p_model, capi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class _Hip:
    def __init__(self):
        capi.lib()  # loads libdisco_hip.so and with it its HIP runtime
        self.h = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed ({rc})"

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 4))), "hipMalloc")
        self.bufs.append(p)
        self.check(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D),
                   "hipMemcpy")
        return p

    def get(self, p, n):
        out = np.empty(n, np.float32)
        self.check(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self.check(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), D2H),
                   "hipMemcpy")
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _problem(n, seed):
    from test_c2p_train import _problem as prob
    return prob(n=n, seed=seed)


def test_noalloc_gradient_equals_dh_c2p_grad():
    lib = capi.lib()
    hip = _Hip()
    try:
        for n, seed in ((4000, 3), (1500, 4)):
            print(f"case dh_c2p_grad_noalloc n={n}", flush=True)
            p, X, Y = _problem(n, seed)
            g_ref, l_ref = c2p_model.grad(p, X, Y)
            dp, dX, dY = hip.put(p), hip.put(X), hip.put(Y)
            dg, dl = hip.put(np.full(c2p_model.N_PARAM, np.nan)), hip.put(np.full(1, np.nan))
            rc = lib.dh_c2p_grad_noalloc(dp, dX, dY, n, dg, dl, None)
            assert rc == 0, lib.dh_last_error().decode()
            g, loss = hip.get(dg, c2p_model.N_PARAM), hip.get(dl, 1)[0]
            print(f"  loss {loss:.6e} (allocating call {l_ref:.6e})", flush=True)
            assert np.array_equal(g, g_ref)
            assert loss == np.float32(l_ref)
        rc = lib.dh_c2p_grad_noalloc(dp, dX, dY, 4097, dg, dl, None)
        assert rc != 0 and b"rows" in lib.dh_last_error()
    finally:
        hip.free()

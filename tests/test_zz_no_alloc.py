"""tohip_compliance_noalloc (the entry point that allocates nothing) against
tohip_compliance on the same data, through the C ABI: two calls in a row on
one GPU, the second on a smaller grid, so both share the library's fixed
scratch.  The test's own device buffers come from hipMalloc of the HIP
runtime the library links (the entry point itself allocates nothing)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamCreate.restype = C.c_int
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.restype = C.c_int
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamDestroy.restype = C.c_int
    return h


class DevBuf:
    def __init__(self, h, nbytes):
        self.h, self.ptr = h, C.c_void_p()
        assert h.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.nbytes == self.nbytes
        assert self.h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n):
        a = np.empty(n)
        assert self.h.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.h.hipFree(self.ptr)


def one_case(T, h, stream, nels, seed):
    from topopt_jl_amd._lib import dptr
    prob = T.PointLoadCantilever(nels)
    s = T.FEASolver(T.HIPMatrixFreeSolver, prob, xmin=1e-9, penalty=T.PowerPenaltyFun(3.0))
    g = np.random.default_rng(seed)
    x = g.uniform(0.01, 1.0, prob.ncells)
    u = g.standard_normal(prob.ndofs)
    u[prob.fixed.astype(bool)] = 0.0
    s.set_densities(x)
    # the existing entry point (host arrays, the context's own partials)
    obj = C.c_double()
    ce = np.empty(prob.ncells)
    gr = np.empty(prob.ncells)
    assert s.lib.tohip_compliance(s.ctx.ptr, dptr(u), C.byref(obj), dptr(ce), dptr(gr)) == 0
    # the no-allocation entry point on device buffers (node-major u)
    u_nm = u[(np.asarray(prob.node_dofs) - 1).T.reshape(-1)]
    bufs = [DevBuf(h, 8 * prob.ndofs), DevBuf(h, 8 * prob.ncells), DevBuf(h, 8 * prob.ncells),
            DevBuf(h, 8 * prob.ncells), DevBuf(h, 8)]
    try:
        bufs[0].put(u_nm)
        bufs[1].put(x)
        rc = s.lib.tohip_compliance_noalloc(s.ctx.ptr, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                                            bufs[3].ptr, bufs[4].ptr, stream)
        assert rc == 0, s.lib.tohip_last_error(s.ctx.ptr)
        assert h.hipStreamSynchronize(stream) == 0
        ce2, gr2, obj2 = bufs[2].get(prob.ncells), bufs[3].get(prob.ncells), bufs[4].get(1)[0]
        # same kernel and grid, so the same bits
        assert np.array_equal(ce2, ce) and np.array_equal(gr2, gr)
        assert obj2 == obj.value
        # a missing buffer is an argument error, not a launch
        assert s.lib.tohip_compliance_noalloc(s.ctx.ptr, None, bufs[1].ptr, bufs[2].ptr,
                                              bufs[3].ptr, bufs[4].ptr, stream) == -1
    finally:
        for b in bufs:
            b.free()
        s.close()
    return prob.ncells


def test_noalloc_compliance_matches_tohip_compliance(T, gpu):
    h = hip_runtime()
    stream = C.c_void_p()
    assert h.hipStreamCreate(C.byref(stream)) == 0
    try:
        n1 = one_case(T, h, stream, (40, 20, 16), 1)
        n2 = one_case(T, h, stream, (12, 6, 6), 2)  # smaller, after the larger one
        assert n2 < n1
    finally:
        h.hipStreamDestroy(stream)

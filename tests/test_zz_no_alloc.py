"""clb_batch_solve_noalloc against clb_batch_solve, through the C ABI.

The no-allocation entry point runs the batched FP64 Newton speciation on device buffers the
caller owns, with its work-queue head in the library's fixed per-device scratch. Its results
must equal those of the allocating path (load + solve + fetch) bit for bit: each voxel's
solve is the same sequence of operations whichever lane takes it. Two calls in a row, the
second with fewer voxels, so the scratch's queue head is reused from a nonzero value. The
device buffers come straight from the HIP runtime (hipMalloc / hipMemcpy through ctypes)."""
import ctypes

import numpy as np
import pytest

HIP_H2D, HIP_D2H = 1, 2
CLB_ERR_ARG = -1
CLB_NOALLOC_MAX_SOLVES = 1 << 30  # include/clb.h


def hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class DevBuf:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), max(nbytes, 8)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, HIP_H2D) == 0

    def get(self, dtype, count):
        a = np.zeros(count, dtype=dtype)
        assert self.h.hipMemcpy(a.ctypes.data, self.p, a.nbytes, HIP_D2H) == 0
        return a

    def free(self):
        self.h.hipFree(self.p)


@pytest.mark.gpu
def test_noalloc_matches_batch_solve(clb):
    print("case: clb_batch_solve_noalloc vs clb_batch_solve, 8/24 tableau, n = 20000 then 3001", flush=True)
    h = hip()
    lib = clb.load()
    lib.clb_batch_solve_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.clb_batch_solve_noalloc.restype = ctypes.c_int
    S, lk = clb.gen_tableau(8, 24, seed=3)
    totals_all = clb.gen_totals_field(32, 32, 32, 8, sigma=4.0, seed=4)
    lat = clb.Lattice(4, 1, 1)
    lat.set_tableau(S, lk, list(range(8)), tol=1e-10, max_iter=200)
    for n in (20000, 3001):
        T = np.ascontiguousarray(totals_all[-n:])
        lat.batch_load(T)
        lat.batch_solve(clb.WARM_REF)
        x_ref, it_ref, r_ref, cv_ref = lat.batch_fetch()
        bt, bx = DevBuf(h, T.nbytes), DevBuf(h, T.nbytes)
        bi, br, bc = DevBuf(h, 4 * n), DevBuf(h, 8 * n), DevBuf(h, n)
        bt.put(T)
        bx.put(np.full(T.shape, np.nan))
        rc = lib.clb_batch_solve_noalloc(lat.h, n, bt.p, 0, None, None, clb.WARM_REF, bx.p, bi.p, br.p, bc.p)
        assert rc == 0, lib.clb_last_error(lat.h)
        lat.sync()
        x = bx.get(np.float64, T.size).reshape(T.shape)
        it, r, cv = bi.get(np.int32, n), br.get(np.float64, n), bc.get(np.uint8, n)
        for b in (bt, bx, bi, br, bc):
            b.free()
        assert cv.all() and cv_ref.all()
        assert np.array_equal(x, x_ref)
        assert np.array_equal(it, it_ref)
        assert np.array_equal(r, r_ref)
    # sizes beyond the maximum are refused before anything is enqueued
    assert lib.clb_batch_solve_noalloc(lat.h, CLB_NOALLOC_MAX_SOLVES + 1, None, 0, None, None, clb.WARM_REF, None,
                                       None, None, None) == CLB_ERR_ARG
    assert b"must be in" in lib.clb_last_error(lat.h)
    lat.close()

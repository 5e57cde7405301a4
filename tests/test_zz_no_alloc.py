"""mba_evaluate_noalloc (device buffers in, nothing allocated) against
mba_evaluate (host buffers) through the C ABI: two calls in a row on one
device, the second on a smaller problem, in both precisions, with and
without the Huber loss.  Residuals and Jacobians must be bit-identical (the
same kernel code on the same values), the cost within 1e-12 relative (a
device sum instead of a host sum).  Named to be collected last."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ba_problem


class _Hip:
    """Device buffers through the HIP runtime libmba.so is bound to (the
    test allocates; the entry point under test does not).  torch is not
    used: its wheel bundles a second HIP runtime, which cannot initialise
    in a process where libmba.so's runtime already has."""

    def __init__(self, mba):
        mba.load_library()
        try:
            self.lib = ctypes.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | ctypes.RTLD_GLOBAL)
        except OSError:
            self.lib = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what} failed: hip error {rc}"

    def put(self, a, dtype):
        a = np.ascontiguousarray(a, dtype)
        p = ctypes.c_void_p()
        self._ok(self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))), "hipMalloc")
        self.bufs.append(p)
        self._ok(self.lib.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1),
                 "hipMemcpy H2D")
        return p.value

    def get(self, ptr, shape, dtype):
        out = np.empty(shape, dtype)
        self._ok(self.lib.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr),
                                    ctypes.c_size_t(out.nbytes), 2), "hipMemcpy D2H")
        return out

    def stream(self):
        st = ctypes.c_void_p()
        self._ok(self.lib.hipStreamCreate(ctypes.byref(st)), "hipStreamCreate")
        return st.value

    def sync(self, st):
        self._ok(self.lib.hipStreamSynchronize(ctypes.c_void_p(st)), "hipStreamSynchronize")

    def close(self, st=None):
        if st:
            self.lib.hipStreamDestroy(ctypes.c_void_p(st))
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def _noalloc(mba, d, precision, huber):
    n, nc, npt = d["obs"].shape[0], d["cam_init"].shape[0], d["pt_init"].shape[0]
    h = _Hip(mba)
    st = h.stream()
    try:
        cams, pts, obs = h.put(d["cam_init"], np.float64), h.put(d["pt_init"], np.float64), h.put(d["obs"], np.float64)
        ci, pi = h.put(d["cam_idx"], np.int32), h.put(d["pt_idx"], np.int32)
        res = h.put(np.full((n, 2), -1.0), np.float64)
        jc = h.put(np.full((n, 2, 9), -1.0), np.float64)
        jp = h.put(np.full((n, 2, 3), -1.0), np.float64)
        cost = h.put(np.full(1, -1.0), np.float64)
        mba.evaluate_noalloc(nc, npt, n, cams, pts, obs, ci, pi, res, jc, jp, cost, precision=precision,
                             huber_a=huber, stream=st)
        h.sync(st)
        return (h.get(res, (n, 2), np.float64), h.get(jc, (n, 2, 9), np.float64), h.get(jp, (n, 2, 3), np.float64),
                float(h.get(cost, 1, np.float64)[0]))
    finally:
        h.close(st)


def _smaller(d, ncam):
    """The observations of cameras < ncam, cameras cut to ncam (points kept)."""
    keep = d["cam_idx"] < ncam
    return dict(cam_init=d["cam_init"][:ncam].copy(), pt_init=d["pt_init"], obs=d["obs"][keep].copy(),
                cam_idx=d["cam_idx"][keep].copy(), pt_idx=d["pt_idx"][keep].copy())


@pytest.mark.gpu
@pytest.mark.parametrize("precision,huber", [(0, 0.0), (0, 4.0), (1, 0.0), (1, 4.0)])
def test_noalloc_matches_evaluate_twice_second_smaller(mba, c1, precision, huber):
    for d in (c1, _smaller(c1, 7)):
        print(f"noalloc: precision {precision} huber {huber} cameras {d['cam_init'].shape[0]} "
              f"observations {d['obs'].shape[0]}", flush=True)
        r, jc, jp, cost = _noalloc(mba, d, precision, huber)
        rr, jcr, jpr, costr = mba.evaluate(ba_problem(mba, d), precision, huber)
        np.testing.assert_array_equal(r, rr)
        np.testing.assert_array_equal(jc, jcr)
        np.testing.assert_array_equal(jp, jpr)
        assert abs(cost - costr) <= 1e-12 * costr, (cost, costr)


@pytest.mark.gpu
def test_noalloc_out_of_range_index_gives_nan_rows(mba, c1):
    """An out-of-range camera index yields NaN for that observation only."""
    d = _smaller(c1, 5)
    bad = dict(d)
    bad["cam_idx"] = d["cam_idx"].copy()
    bad["cam_idx"][3] = d["cam_init"].shape[0]
    print(f"noalloc: out-of-range index, {d['obs'].shape[0]} observations", flush=True)
    r, jc, jp, cost = _noalloc(mba, bad, 0, 0.0)
    rr, jcr, jpr, _ = mba.evaluate(ba_problem(mba, d), 0, 0.0)
    assert np.isnan(r[3]).all() and np.isnan(jc[3]).all() and np.isnan(jp[3]).all() and np.isnan(cost)
    ok = np.arange(r.shape[0]) != 3
    np.testing.assert_array_equal(r[ok], rr[ok])
    np.testing.assert_array_equal(jc[ok], jcr[ok])


def test_noalloc_rejects_more_cameras_than_its_scratch(mba):
    """The scratch holds the tables of at most 16384 cameras: a larger problem
    is MBA_ERR_INVALID_ARGUMENT, decided on the host before any device call
    (no device needed)."""
    with pytest.raises(mba.MbaError) as e:
        mba.evaluate_noalloc(16385, 10, 10, 1, 1, 1, 1, 1, residuals=1)
    assert e.value.status == mba.ERR_INVALID_ARGUMENT
    assert "16384" in str(e.value)

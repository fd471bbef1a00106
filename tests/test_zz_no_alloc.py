"""flip_gpu_sort_cells_noalloc (no device allocation; scratch in a fixed
per-GPU array) against flip_gpu_sorted_cells, through the C ABI: two calls in
a row, the second with fewer particles and a smaller grid, each bit-exact
(counts and ids, NaN keys and a bucket of thousands included)."""
import ctypes

import numpy as np
import pytest

import flip3d_amd as fa

pytestmark = pytest.mark.gpu


def _positions(rs, m, crowd):
    p = rs.uniform(-0.05, 1.05, size=(m, 3)).astype(np.float32)
    p[:crowd] = (0.37 + 1e-4 * rs.uniform(size=(crowd, 3))).astype(np.float32)  # one crowded cell
    p[crowd:crowd + 40, 1] = np.nan  # NaN keys land in cell n-1 on that axis
    return p


def _reference_path(n, p):
    with fa.FlipContext(n) as g:
        g.upload(p, np.zeros_like(p), np.zeros_like(p), np.full(len(p), 1, np.int8),
                 dens=np.ones(len(p), np.float32))
        return g.sorted_cells()


class _Hip:
    """Device buffers through the HIP runtime the library itself links
    (libamdhip64.so.7 of /opt/rocm; a second runtime in the process, such as
    one bundled with another package, would not share its devices)."""

    def __init__(self):
        fa.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, f"{what}: HIP error {rc}"

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self._ok(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 4))), "hipMalloc")
        self.bufs.append(p)
        self._ok(self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy")
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        self._ok(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self._ok(self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2), "hipMemcpy")
        return out

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def test_noalloc_sort_matches_context_sort_twice():
    lib = fa.lib()
    rs = np.random.RandomState(11)
    hip = _Hip()
    try:
        for n, m, crowd in ((96, 900000, 5000), (40, 150000, 700)):
            p = _positions(rs, m, crowd)
            want_counts, want_ids = _reference_path(n, p)
            counts0 = np.full(n ** 3, -7, np.int32)
            ids0 = np.full(m, -7, np.int32)
            dp, dc, di = hip.put(p.reshape(-1)), hip.put(counts0), hip.put(ids0)
            rc = lib.flip_gpu_sort_cells_noalloc(dp, m, n, dc, di, None)
            assert rc == 0, lib.flip_gpu_last_error()
            counts, ids = hip.get(dc, counts0), hip.get(di, ids0)
            assert counts.max() >= crowd
            assert np.array_equal(counts, want_counts), n
            assert np.array_equal(ids, want_ids), n
    finally:
        hip.close()


def test_noalloc_sort_rejects_sizes_beyond_the_scratch():
    lib = fa.lib()
    assert lib.flip_gpu_sort_cells_noalloc(ctypes.c_void_p(8), (1 << 20) + 1, 8, ctypes.c_void_p(8),
                                           ctypes.c_void_p(8), None) == -1
    assert b"scratch limit" in lib.flip_gpu_last_error()
    assert lib.flip_gpu_sort_cells_noalloc(ctypes.c_void_p(8), 10, 102, ctypes.c_void_p(8), ctypes.c_void_p(8),
                                           None) == -1
    assert b"scratch limit" in lib.flip_gpu_last_error()

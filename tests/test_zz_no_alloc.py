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


def test_noalloc_sort_matches_context_sort_twice():
    import torch
    lib = fa.lib()
    rs = np.random.RandomState(11)
    for n, m, crowd in ((96, 900000, 5000), (40, 150000, 700)):
        p = _positions(rs, m, crowd)
        want_counts, want_ids = _reference_path(n, p)
        dp = torch.from_numpy(p.reshape(-1)).cuda()
        dc = torch.full((n ** 3,), -7, dtype=torch.int32, device="cuda")
        di = torch.full((m,), -7, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        rc = lib.flip_gpu_sort_cells_noalloc(ctypes.c_void_p(dp.data_ptr()), m, n, ctypes.c_void_p(dc.data_ptr()),
                                             ctypes.c_void_p(di.data_ptr()), None)
        assert rc == 0, lib.flip_gpu_last_error()
        torch.cuda.synchronize()
        counts, ids = dc.cpu().numpy(), di.cpu().numpy()
        assert counts.max() >= crowd
        assert np.array_equal(counts, want_counts), n
        assert np.array_equal(ids, want_ids), n


def test_noalloc_sort_rejects_sizes_beyond_the_scratch():
    lib = fa.lib()
    assert lib.flip_gpu_sort_cells_noalloc(ctypes.c_void_p(8), (1 << 20) + 1, 8, ctypes.c_void_p(8),
                                           ctypes.c_void_p(8), None) == -1
    assert b"scratch limit" in lib.flip_gpu_last_error()
    assert lib.flip_gpu_sort_cells_noalloc(ctypes.c_void_p(8), 10, 102, ctypes.c_void_p(8), ctypes.c_void_p(8),
                                           None) == -1
    assert b"scratch limit" in lib.flip_gpu_last_error()

"""The learned coarse solver's U-Net: HIP BF16-MFMA forward/backward against
the numpy reference, the reference pinned by finite differences, and
training on solver-generated systems."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import unet_ref as ur


def random_params(seed, scale=1.0):
    rs = np.random.RandomState(seed)
    offs, hw, hb, n = ur.layout()
    p = np.zeros(n)
    for (w0, b0), (cin, cout, _) in zip(offs, ur.CONVS):
        p[w0:b0] = rs.normal(0, scale * np.sqrt(2.0 / (27 * cin)), b0 - w0)
        p[b0:b0 + cout] = rs.normal(0, 0.05, cout)
    p[hw:hb] = rs.normal(0, 0.3, 16)
    p[hb] = 0.1
    return p


def random_inputs(seed, S, R):
    rs = np.random.RandomState(seed)
    F = np.zeros((S, R, R, R, 8))
    F[..., :5] = rs.normal(size=(S, R, R, R, 5))
    mask = (rs.uniform(size=(S, R, R, R)) < 0.7).astype(np.float64)
    return F, mask


def unit_operator(A, L):
    """Unit-spacing 7-point operator of the solver on the FLUID cells of one
    R^3 system (solver.cpp:19-41 with subcell ghost fluid): sparse matrix and
    the FLUID index map."""
    R = A.shape[0]
    fl = np.argwhere(A == 1)
    idx = -np.ones(A.shape, np.int64)
    idx[A == 1] = np.arange(len(fl))
    rows, cols, vals = [], [], []
    diag = np.full(len(fl), 6.0)
    Ls = L[A == 1]
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        q = fl + np.array(d)
        out = ((q < 0) | (q >= R)).any(1)
        qc = np.clip(q, 0, R - 1)
        Aq = A[qc[:, 0], qc[:, 1], qc[:, 2]]
        nbr = ~out & (Aq == 1)
        rows.append(np.flatnonzero(nbr))
        cols.append(idx[qc[nbr, 0], qc[nbr, 1], qc[nbr, 2]])
        vals.append(-np.ones(nbr.sum()))
        diag[out | (Aq == 2)] -= 1.0
        air = ~out & (Aq == 0)
        diag[air] -= L[qc[air, 0], qc[air, 1], qc[air, 2]].astype(np.float32) / np.minimum(
            np.float32(1e-6), Ls[air].astype(np.float32))
    rows.append(np.arange(len(fl)))
    cols.append(np.arange(len(fl)))
    vals.append(diag)
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(len(fl),) * 2)
    return M, A == 1


def box_system(seed, R):
    """A pressure system of a random liquid box: walls on the boundary,
    AIR above a random height, level set from the box, random divergence."""
    rs = np.random.RandomState(seed)
    A = np.zeros((R, R, R), np.int8)
    lo = rs.randint(1, R // 3, 3)
    hi = rs.randint(R // 2, R - 1, 3)
    A[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = 1
    A[0, :, :] = A[-1, :, :] = A[:, 0, :] = A[:, -1, :] = A[:, :, 0] = A[:, :, -1] = 2
    L = np.where(A == 1, -0.5, 0.4 + 0.4 * rs.uniform(size=A.shape)).astype(np.float32)
    L[A == 2] = 1.0
    M, fl = unit_operator(A, L)
    b = np.zeros(A.shape, np.float32)
    b[fl] = rs.normal(size=fl.sum())
    x = np.zeros(A.shape, np.float32)
    x[fl] = spla.spsolve(M.tocsc(), b[fl].astype(np.float64)) / R ** 2  # spacing 1/R
    return A, L, b, x


# ---------------------------------------------------------------- CPU
def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, 3.14159265, -2.5e-3, 0.0], np.float64)
    r = ur.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] - 3.140625) < 1e-12
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)


def test_reference_backward_matches_finite_differences():
    """Directional derivatives of sum(g * out) against central differences
    (float64, no rounding); a wrong transpose, tap flip or skip would fail.
    The step is small enough that no ReLU input changes sign."""
    R, S = 8, 2
    p = random_params(1)
    F, mask = random_inputs(2, S, R)
    g = np.random.RandomState(3).normal(size=mask.shape)
    net = ur.Net(p, R)
    net.forward(F, mask)
    grad = net.backward(g)
    rs = np.random.RandomState(4)
    offs, hw, hb, _ = ur.layout()
    blocks = [(w0, b0) for w0, b0 in offs] + [(b0, b0 + c[1]) for (w0, b0), c in zip(offs, ur.CONVS)] + [(hw, hb + 1)]
    for lo, hi in blocks:
        d = np.zeros_like(p)
        d[lo:hi] = rs.normal(size=hi - lo)
        d /= np.linalg.norm(d)
        eps = 1e-7
        fp = (g * ur.Net(p + eps * d, R).forward(F, mask)).sum()
        fm = (g * ur.Net(p - eps * d, R).forward(F, mask)).sum()
        fd = (fp - fm) / (2 * eps)
        an = grad @ d
        assert abs(fd - an) <= 1e-4 * max(1.0, abs(an)), (lo, hi, fd, an)


def test_operator_matches_oracle_laplacian():
    """The loss operator of the tests is the solver's stencil: compare with
    the oracle's compute_Ax on a box system (unit spacing = n^2 scaled)."""
    from oracle import pyoracle as po
    R = 12
    A, L, b, x = box_system(5, R)
    M, fl = unit_operator(A, L)
    v = np.random.RandomState(6).normal(size=A.shape)
    ax = po.compute_ax(R, A, L, v, f64=True)
    assert np.allclose(M @ v[fl], ax[fl] / R ** 2, rtol=1e-6, atol=1e-9)


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_forward_matches_reference():
    import flip3d_amd as fa
    R, S = 16, 2
    with fa.UNet(R, capacity=S, seed=7) as net:
        p = random_params(8).astype(np.float32)
        net.set_params(p)
        F, mask = random_inputs(9, S, R)
        out = net.forward_features(F, mask)
        ref = ur.Net(p.astype(np.float64), R, bf=True).forward(F, mask)
        err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
        assert err < 5e-3, err


@pytest.mark.gpu
def test_backward_matches_reference():
    import flip3d_amd as fa
    R, S = 16, 2
    with fa.UNet(R, capacity=S, seed=7) as net:
        p = random_params(10).astype(np.float32)
        net.set_params(p)
        F, mask = random_inputs(11, S, R)
        g = np.random.RandomState(12).normal(size=mask.shape).astype(np.float32)
        net.forward_features(F, mask)
        grad = net.backward(g)
        ref_net = ur.Net(p.astype(np.float64), R, bf=True)
        ref_net.forward(F, mask)
        ref = ref_net.backward(g)
        offs, hw, hb, _ = ur.layout()
        for i, (w0, b0) in enumerate(offs):
            for lo, hi, what in ((w0, b0, "w"), (b0, b0 + ur.CONVS[i][1], "b")):
                e = np.linalg.norm(grad[lo:hi] - ref[lo:hi]) / max(np.linalg.norm(ref[lo:hi]), 1e-30)
                assert e < 2e-2, (i, what, e)
        e = np.linalg.norm(grad[hw:hb + 1] - ref[hw:hb + 1]) / np.linalg.norm(ref[hw:hb + 1])
        assert e < 1e-3, e


@pytest.mark.gpu
def test_training_fits_solver_systems():
    """Adam on a fixed batch of box systems lowers the relative A-norm
    error well below the untrained network's, and predict() returns x in
    the system's unit."""
    import flip3d_amd as fa
    R, S = 16, 4
    sys_ = [box_system(100 + i, R) for i in range(S)]
    A = np.stack([s[0] for s in sys_])
    L = np.stack([s[1] for s in sys_])
    b = np.stack([s[2] for s in sys_])
    x = np.stack([s[3] for s in sys_])
    with fa.UNet(R, capacity=S, seed=3) as net:
        first = net.train_step(A, L, b, x, 1e-3)
        for _ in range(300):
            last = net.train_step(A, L, b, x, 1e-3)
        assert last < 0.5 * first, (first, last)
        pred = net.predict(A, L, b)
        for s in range(S):
            M, fl = unit_operator(A[s], L[s])
            e = (pred[s] - x[s])[fl].astype(np.float64)
            t = x[s][fl].astype(np.float64)
            rel = (e @ (M @ e)) / (t @ (M @ t))
            assert rel < 0.6, (s, rel)


@pytest.mark.gpu
@pytest.mark.parametrize("R", [16, 24, 32, 40, 48])
def test_layer_kernels_equal_separate_kernels(R):
    """The solver's seven layer kernels (each brick's halo and weights staged
    in LDS; features, poolings and head fused) return the separate kernels'
    result bit for bit: shipped weights at R=32, random weights at the other
    coarse-level sides the solver uses (16-48; at 24, 40 and 48 bricks reach
    past the cube)."""
    import flip3d_amd as fa
    A, L, b, _ = box_system(11 + R, R)
    with fa.UNet(R, capacity=1, seed=R) as net:
        if R == 32:
            net.load(os.path.join(os.path.dirname(fa.__file__), "weights", "unet_coarse.bin"))
        else:
            net.set_params(random_params(R).astype(np.float32))
        ref = net.solve_system(A, L, b, path=0)
        got = net.solve_system(A, L, b, path=1)
        assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), (R, np.abs(ref - got).max())
        assert np.abs(ref).max() > 0 and np.all(got[A != 1] == 0)
        # predict() forms sigma from other partial sums: equal up to its rounding
        pr = net.predict(A[None], L[None], b[None])[0]
        assert np.abs(pr - got).max() <= 1e-5 * np.abs(pr).max(), (R, np.abs(pr - got).max())


@pytest.mark.gpu
def test_fused_inference_equals_predict():
    """The solver's inference path (upsample-adds and head fused into the
    convolutions) returns predict()'s result bit for bit, with the shipped
    weights and with random ones, at the coarse level's side 32."""
    import flip3d_amd as fa
    R = 32
    for seed in (7, 8):
        A, L, b, _ = box_system(seed, R)
        with fa.UNet(R, capacity=1, seed=seed) as net:
            if seed == 7:
                net.load(os.path.join(os.path.dirname(fa.__file__), "weights", "unet_coarse.bin"))
            else:
                net.set_params(random_params(seed).astype(np.float32))
            ref = net.predict(A[None], L[None], b[None])[0]
            got = net.predict_fused(A, L, b)
            assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), (seed, np.abs(ref - got).max())
            assert np.abs(ref).max() > 0

"""CPU reference of the learned coarse solver's U-Net (numpy, float64).

TEST INFRASTRUCTURE ONLY (tests/ import it to check the HIP forward and
backward kernels).  Same architecture, parameter layout and channels-last
activation layout as flip3d_amd/csrc/unet.hip:

    act = leaky ReLU, slope LEAK
    e0 = act(conv(F0: 8->16))           R       p0 = pool(e0)
    e1 = act(conv(p0: 16->32))         R/2     p1 = pool(e1)
    e2 = act(conv(p1: 32->64))         R/4     p2 = pool(e2)
    e3 = act(conv(p2: 64->64))         R/8
    d2 = act(conv(up(e3) + e2: 64->32))
    d1 = act(conv(up(d2) + e1: 32->16))
    d0 = act(conv(up(d1) + e0: 16->16))
    out = mask * (w . d0 + b)

conv is 3x3x3 with zero padding, weights [27][cout][cin] with tap =
(dz+1)*9 + (dy+1)*3 + (dx+1) over (z, y, x) = the grid's (i, j, k).  With
bf16=True every tensor the device stores in bf16 (weights, input features,
activations, gradients) is rounded to bf16 at the same points, so the
comparison isolates accumulation order.
"""
import numpy as np

LEAK = 0.1
CONVS = [(8, 16, 0), (16, 32, 1), (32, 64, 2), (64, 64, 3), (64, 32, 2), (32, 16, 1), (16, 16, 0)]


def layout():
    """Parameter offsets: per conv (w_off, b_off), then head w (16), head b."""
    off, out = 0, []
    for cin, cout, _ in CONVS:
        w = off
        off += 27 * cin * cout
        out.append((w, off))
        off += cout
    hw = off
    return out, hw, hw + 16, hw + 17


def bf16(a):
    """Round to bfloat16 (nearest even) and back."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3(x, W, b=None):
    """x [S,R,R,R,cin], W [27,cout,cin] -> [S,R,R,R,cout]"""
    S, R = x.shape[0], x.shape[1]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros(x.shape[:4] + (W.shape[1],))
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        y += xp[:, dz:dz + R, dy:dy + R, dx:dx + R, :] @ W[t].T
    if b is not None:
        y += b
    return y


def conv3_back(x, W, dy, need_dx=True):
    """Gradients of y = conv3(x, W) + b: (dx, dW, db)."""
    R = x.shape[1]
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dW = np.zeros_like(W)
    dxp = np.zeros_like(xp) if need_dx else None
    for t in range(27):
        dz, dyy, dx = t // 9, (t // 3) % 3, t % 3
        xs = xp[:, dz:dz + R, dyy:dyy + R, dx:dx + R, :]
        dW[t] = np.einsum("szyxo,szyxi->oi", dy, xs)
        if need_dx:
            dxp[:, dz:dz + R, dyy:dyy + R, dx:dx + R, :] += dy @ W[t]
    db = dy.sum(axis=(0, 1, 2, 3))
    dxo = dxp[:, 1:-1, 1:-1, 1:-1, :] if need_dx else None
    return dxo, dW, db


def pool(x):
    S, R, C = x.shape[0], x.shape[1], x.shape[-1]
    h = R // 2
    return x.reshape(S, h, 2, h, 2, h, 2, C).mean(axis=(2, 4, 6))


def pool_back(d):
    return np.repeat(np.repeat(np.repeat(d, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_back(d):
    S, R, C = d.shape[0], d.shape[1], d.shape[-1]
    h = R // 2
    return d.reshape(S, h, 2, h, 2, h, 2, C).sum(axis=(2, 4, 6))


class Net:
    def __init__(self, params, R, bf=False):
        self.p = np.asarray(params, np.float64)
        self.R = R
        self.bf = bf
        self.offs, self.hw, self.hb, self.n = layout()
        assert self.p.size == self.n

    def _r(self, a):
        return bf16(a) if self.bf else a

    def W(self, i):
        cin, cout, _ = CONVS[i]
        w0, b0 = self.offs[i]
        return self._r(self.p[w0:b0].reshape(27, cout, cin)), self.p[b0:b0 + cout]

    def forward(self, F, mask):
        """F [S,R,R,R,8], mask [S,R,R,R] -> out [S,R,R,R]"""
        r, c = self._r, {}
        c["F0"] = r(F)
        relu = lambda a: np.where(a > 0, a, LEAK * a)

        def layer(i, x):
            W, b = self.W(i)
            return r(relu(conv3(x, W, b)))

        c["e0"] = layer(0, c["F0"])
        c["p0"] = r(pool(c["e0"]))
        c["e1"] = layer(1, c["p0"])
        c["p1"] = r(pool(c["e1"]))
        c["e2"] = layer(2, c["p1"])
        c["p2"] = r(pool(c["e2"]))
        c["e3"] = layer(3, c["p2"])
        c["u2"] = r(up(c["e3"]) + c["e2"])
        c["d2"] = layer(4, c["u2"])
        c["u1"] = r(up(c["d2"]) + c["e1"])
        c["d1"] = layer(5, c["u1"])
        c["u0"] = r(up(c["d1"]) + c["e0"])
        c["d0"] = layer(6, c["u0"])
        hw = self.p[self.hw:self.hb]
        c["mask"] = mask
        out = mask * (c["d0"] @ hw + self.p[self.hb])
        self.cache = c
        return out

    def backward(self, gout):
        """dL/dparams for dL/dout = gout (after forward)."""
        c, r = self.cache, self._r
        g = np.zeros_like(self.p)
        hw = self.p[self.hw:self.hb]
        gm = gout * c["mask"]
        g[self.hw:self.hb] = np.einsum("szyx,szyxc->c", gm, c["d0"])
        g[self.hb] = gm.sum()

        def conv_back(i, x, dy, need_dx=True):
            W, _ = self.W(i)
            dx, dW, db = conv3_back(x, W, dy, need_dx)
            w0, b0 = self.offs[i]
            g[w0:b0] = dW.reshape(-1)
            g[b0:b0 + CONVS[i][1]] = db
            return None if dx is None else r(dx)

        act = lambda y: np.where(y > 0, 1.0, LEAK)
        dd0 = r(r(gm[..., None] * hw) * act(c["d0"]))
        du0 = conv_back(6, c["u0"], dd0)
        dd1 = r(r(up_back(du0)) * act(c["d1"]))
        du1 = conv_back(5, c["u1"], dd1)
        dd2 = r(r(up_back(du1)) * act(c["d2"]))
        du2 = conv_back(4, c["u2"], dd2)
        de3 = r(r(up_back(du2)) * act(c["e3"]))
        dp2 = conv_back(3, c["p2"], de3)
        de2 = r(r(du2 + r(pool_back(dp2))) * act(c["e2"]))
        dp1 = conv_back(2, c["p1"], de2)
        de1 = r(r(du1 + r(pool_back(dp1))) * act(c["e1"]))
        dp0 = conv_back(1, c["p0"], de1)
        de0 = r(r(du0 + r(pool_back(dp0))) * act(c["e0"]))
        conv_back(0, c["F0"], de0, need_dx=False)
        return g


def features(A, L, b):
    """Input features of one or more R^3 systems (flags A, level set L, rhs b)."""
    A = np.asarray(A)
    fl = (A == 1)
    sig = np.sqrt((b[fl].astype(np.float64) ** 2).mean()) if fl.any() else 1.0
    sig = sig if sig > 0 else 1.0
    F = np.zeros(A.shape + (8,))
    F[..., 0] = np.where(fl, b / sig, 0.0)
    F[..., 1] = fl
    F[..., 2] = A == 0
    F[..., 3] = A == 2
    F[..., 4] = np.where(A != 2, np.clip(L, -2, 2), 0.0)
    return F, fl.astype(np.float64), sig

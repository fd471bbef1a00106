"""CPU restatement of the reference's output writers (the oracle for them).

TEST INFRASTRUCTURE ONLY: the checker for the library's file writers.  Only
tests/ may import it.

  - write_obj:      src/exporter.cpp:127-164 (write_obj)
  - write_sunflow:  src/exporter.cpp:72-125 (writeSunflow; its template text
                    is the reference's data, so it is checked by digest only)
  - render_preview: src/flip3D.cpp:1105-1192 (buildRotationMatrix, transform,
                    write_frame's point splats) and write_bmp.cpp:52-118
Float arithmetic follows the reference's FLOAT/double mixing: numpy float32
for FLOAT operands, float64 where a double literal takes part.
"""
import math

import numpy as np

F32 = np.float32


def obj_bytes(v, n, f, n_grid):
    """write_obj for frame data: vertices scaled by 1/(1 - 2 wall_thick)
    after subtracting wall_thick, y and z swapped; normals swapped; faces
    1-based with normal indices."""
    wt = float(F32(1.0 / n_grid))
    s = 1.0 / (1.0 - 2.0 * wt)
    out = []
    for x, y, z in v:
        out.append("v %f %f %f\n" % (s * (x - wt), s * (z - wt), s * (y - wt)))
    for x, y, z in n:
        out.append("vn %f %f %f\n" % (x, z, y))
    for a, b, c in f:
        out.append("f %d//%d %d//%d %d//%d\n" % (a + 1, a + 1, b + 1, b + 1, c + 1, c + 1))
    return "".join(out).encode()


def sunflow_bytes(v, n, f, n_grid):
    """writeSunflow's scene file: positions scaled to the 100-unit box and
    clamped 0.001 inside it, y and z swapped."""
    wt = float(F32(1.0 / n_grid))
    s = 100.0 * 1.0 / (1.0 - 2.0 * wt)
    h = float(F32(0.001))
    out = ["include template.sc\n\n",
           "object { \n shader Liquid \n type mesh \n name polySurfac \n %d %d \n" % (len(v), len(f))]
    for (x, y, z), (a, b, c) in zip(v, n):
        out.append("v %f %f %f %f %f %f 0 0\n" % (s * min(1.0 - h, max(h, x - wt)), s * min(1.0 - h, max(h, z - wt)),
                                                  s * min(1.0 - h, max(h, y - wt)), a, c, b))
    for a, b, c in f:
        out.append("t %d %d %d\n" % (a, b, c))
    out.append("}\n")
    return "".join(out).encode()


def _libm():
    import ctypes
    m = ctypes.CDLL("libm.so.6")
    for fn in ("cosf", "sinf"):
        getattr(m, fn).argtypes = [ctypes.c_float]
        getattr(m, fn).restype = ctypes.c_float
    return m


def _rotation():
    """buildRotationMatrix(-0.2, 0.2): Rx then Rx*Ry in FLOAT (the float
    overloads of cos and sin, i.e. cosf/sinf)."""
    m = _libm()
    xr, yr = F32(-0.2), F32(0.2)
    cx, sx = F32(m.cosf(float(xr))), F32(m.sinf(float(xr)))
    cy, sy = F32(m.cosf(float(yr))), F32(m.sinf(float(yr)))
    Rx = [[F32(1), F32(0), F32(0)], [F32(0), cx, -sx], [F32(0), sx, cx]]
    Ry = [[cy, F32(0), sy], [F32(0), F32(1), F32(0)], [-sy, F32(0), cy]]
    R = [[F32(0)] * 3 for _ in range(3)]
    for i in range(3):
        for j in range(3):
            acc = F32(0.0)
            for k in range(3):
                acc = F32(acc + F32(Rx[i][k] * Ry[k][j]))
            R[i][j] = acc
    return R


def render_preview(p, t, n_grid, density=0.5, width=1024, height=700):
    """write_frame: FLUID particles, centred, rotated and projected, splat a
    2x2 pixel block each, blending 0.3/0.5/0.8 at alpha exp(-DENSITY N/23)
    over black; returns the RGBA bytes (height rows of width pixels)."""
    R = _rotation()
    winrate = F32(F32(height) / F32(width))
    alpha = F32(math.exp(-density * n_grid / 23.0))
    color = [F32(0.3), F32(0.5), F32(0.8)]
    eye, offset = F32(2.0), F32(0.3)
    fl = t == 1
    q = p[fl].astype(np.float64)
    p0 = (q - 0.5).astype(np.float32)
    pr = []
    for i in range(3):
        acc = np.zeros(len(p0), np.float32)
        for k in range(3):
            acc = (acc + R[i][k] * p0[:, k]).astype(np.float32)
        pr.append(acc)
    z = ((float(offset) + 0.5) + pr[2].astype(np.float64)).astype(np.float32)
    fac = (eye / (eye + z)).astype(np.float32)
    x = (fac.astype(np.float64) * (pr[0].astype(np.float64) + 0.05)).astype(np.float32)
    y = (fac.astype(np.float64) * (pr[1].astype(np.float64) + 0.25)).astype(np.float32)
    xd, yd, wr = x.astype(np.float64), y.astype(np.float64), float(winrate)
    ok = (xd > -0.5) & (xd < 0.5) & (yd > -wr / 2.0) & (yd < wr / 2.0)
    pi = np.trunc(width * (xd[ok] + 0.5)).astype(np.int64)
    pj = np.trunc(width * (yd[ok] + wr / 2.0)).astype(np.int64)
    count = np.zeros((height, width), np.int64)
    for di in (-1, 0):
        for dj in (-1, 0):
            ni, nj = pi + di, pj + dj
            m = (ni > 0) & (ni < width) & (nj > 0) & (nj < height)
            np.add.at(count, (nj[m], ni[m]), 1)
    img = np.zeros((height, width, 4), np.float32)
    img[:, :, 3] = 1.0
    # every splat on a pixel is the same blend, so the value depends only on
    # how many splats landed there
    for c in range(3):
        v = np.zeros((height, width), np.float32)
        a_c = float(F32(alpha * color[c]))
        keep = 1.0 - float(alpha)
        for k in range(int(count.max()) if count.size else 0):
            nv = (a_c + keep * v.astype(np.float64)).astype(np.float32)
            v = np.where(count > k, nv, v)
        img[:, :, c] = v
    return (255.0 * np.minimum(1.0, img.astype(np.float64))).astype(np.uint8)


def bmp_bytes(rgba, invert_y=True):
    """write_bmp: 54-byte header, rows of BGR bytes padded to
    (3 (width + 1) / 4) * 4 bytes, rows in buffer order when invertY."""
    height, width = rgba.shape[:2]
    bpl = (3 * (width + 1) // 4) * 4
    import struct
    hdr = b"BM" + struct.pack("<iiiiiihhiiiiii", 54 + bpl * height, 0, 54, 40, width, height, 1, 24, 0,
                              bpl * height, 0, 0, 0, 0)
    lines = np.zeros((height, bpl), np.uint8)
    for v in range(height):
        i = v if invert_y else height - v - 1
        lines[i, 0:3 * width:3] = rgba[i, :, 2]
        lines[i, 1:3 * width:3] = rgba[i, :, 1]
        lines[i, 2:3 * width:3] = rgba[i, :, 0]
    return hdr + lines.tobytes()

"""Field view of a render/save/<step>_scene.dat checkpoint (flip3D.cpp:453-513):
int count, float max_dens, int pourTime, float pourPos[2], float pourRad,
count 72-byte particle records (src/common.h:34-45), n^3 pressure floats,
4 n^3 wall-normal floats."""
import hashlib

import numpy as np

DIGEST_FIELDS = ("p", "u_fluid", "n", "type", "visible", "m", "dens", "press", "wall_normal")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def checkpoint_fields(raw, n):
    """Split a checkpoint image into its fields.  The records' remove and tmp
    bytes are not initialised by the reference (flip3D.cpp:628-643), so they
    are left out."""
    hdr = np.frombuffer(raw[:24], np.uint8)
    count = int(hdr[:4].view(np.int32)[0])
    rec = np.frombuffer(raw[24:24 + 72 * count], np.uint8).reshape(count, 72)
    C = n ** 3
    tail = np.frombuffer(raw[24 + 72 * count:], np.float32)
    return {
        "count": count,
        "max_dens": hdr[4:8].view(np.float32)[0],
        "pour_time": hdr[8:12].view(np.int32)[0],
        "pour": hdr[12:24].view(np.float32),
        "p": rec[:, 0:12].copy().view(np.float32),
        "u": rec[:, 12:24].copy().view(np.float32),
        "n": rec[:, 24:36].copy().view(np.float32),
        "type": rec[:, 36].copy(),
        "visible": rec[:, 37].copy(),
        "m": rec[:, 64:68].copy().view(np.float32),
        "dens": rec[:, 68:72].copy().view(np.float32),
        "press": tail[:C],
        "wall_normal": tail[C:],
        "size": len(raw),
    }


def field_digests(f):
    """SHA-256 of each compared field.  Velocities count for fluid particles
    only: wall-particle velocities come from an out-of-bounds read in the
    reference (mapper.cpp:102-106) and are not defined."""
    fields = dict(f)
    fields["u_fluid"] = f["u"][f["type"] == 1]
    return {k: digest(fields[k]) for k in DIGEST_FIELDS}

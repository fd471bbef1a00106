"""Checkpoint files in the reference's layout, and the reference-signature
C++ shim (flip3D / mapper / solver / corrector namespaces) over the C ABI."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import flip3d_amd as fa
from oracle import pyoracle as po
from tests.checkpoint_format import checkpoint_fields, digest, field_digests

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "checkpoint.npz")
SHIM = os.path.join(ROOT, "flip3d_amd", "libflip3d_shim.so")
SHIM_CHECK = os.path.join(ROOT, "tests", "shim", "shim_check")

SHIM_SYMBOLS = [
    "flip3D::init(int)",
    "flip3D::simulateStep()",
    "mapper::mapP2G(sorter*, std::vector<_particle*, std::allocator<_particle*> >&, float****, int)",
    "mapper::mapG2P(std::vector<_particle*, std::allocator<_particle*> >&, float****, int)",
    "mapper::fetchVelocity(float*, float*, float****, int)",
    "solver::setSubcell(char)",
    "solver::solve(char***, float***, float***, float***, int)",
    "corrector::correct(sorter*, std::vector<_particle*, std::allocator<_particle*> >&, float, float)",
]


def _defined(lib):
    out = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {line.split(" ", 2)[2] for line in out.splitlines() if re.match(r"^[0-9a-f]+ [TW] ", line)}


def test_checkpoint_golden_matches_oracle_state():
    """The reference's saved state after 5 steps of its dam-break scene holds
    the particles the oracle computes for the same 5 steps."""
    g = np.load(GOLD, allow_pickle=False)
    s = po.Sim(int(g["n"]))
    s.scene_reference()
    s.srand(1)
    for _ in range(5):
        s.step()
    st = s.particles()
    assert s.num_particles() == int(g["count"])
    assert np.float32(s.max_dens) == g["max_dens"]
    fluid = st["type"] == 1
    for k, a in (("p", st["p"]), ("u_fluid", st["u"][fluid]), ("n", st["n"]), ("dens", st["dens"])):
        assert digest(a.astype(np.float32)) == str(g["sha_" + k]), k
    assert digest(st["type"].astype(np.uint8)) == str(g["sha_type"])
    assert digest(s.grid(po.G_PRESS).astype(np.float32)) == str(g["sha_press"])
    assert int(g["size"]) == 24 + 72 * int(g["count"]) + 20 * int(g["n"]) ** 3


def test_shim_exports_reference_signatures():
    """libflip3d_shim.so defines the reference's entry points with the
    reference's C++ signatures (demangled names)."""
    if not os.path.exists(SHIM):
        pytest.fail(f"{SHIM} not built; run __graft_entry__.build()")
    have = _defined(SHIM)
    missing = [s for s in SHIM_SYMBOLS if s not in have]
    assert not missing, missing


def test_shim_symbols_equal_the_reference_build():
    """Where the reference's own build of these modules is present (compiled
    here from its sources), each shim symbol is one it defines."""
    ref = po.REF_PATH
    if not os.path.exists(ref):
        pytest.skip("reference build not present")
    have = _defined(ref)
    for s in SHIM_SYMBOLS:
        assert s in have, s


@pytest.mark.gpu
def test_checkpoint_matches_reference_save():
    """Save after 5 steps of the reference scene: every initialised field of
    the file equals the reference's saveState output."""
    g = np.load(GOLD, allow_pickle=False)
    n = int(g["n"])
    d = tempfile.mkdtemp()
    path = os.path.join(d, "5_scene.dat")
    with fa.FlipContext(n) as c:
        c.build_scene(fa.SCENE_REFERENCE)
        c.srand(1)
        for _ in range(5):
            c.step()
        c.save_state(path)
    raw = open(path, "rb").read()
    f = checkpoint_fields(raw, n)
    assert f["size"] == int(g["size"]) and f["count"] == int(g["count"])
    assert f["max_dens"] == g["max_dens"] and f["pour_time"] == g["pour_time"]
    assert np.array_equal(f["pour"], g["pour"])
    bad = [k for k, v in field_digests(f).items() if v != str(g["sha_" + k])]
    assert not bad, bad


@pytest.mark.gpu
def test_checkpoint_round_trip_continues_bit_identically():
    d = tempfile.mkdtemp()
    path = os.path.join(d, "s.dat")
    with fa.FlipContext(32) as a, fa.FlipContext(32) as b:
        a.build_scene(fa.SCENE_BENCHMARK)
        a.srand(1)
        for _ in range(4):
            a.step()
        a.save_state(path)
        b.load_state(path)
        assert b.max_dens == a.max_dens
        assert np.array_equal(b.grid(fa.GRID_PRESS), a.grid(fa.GRID_PRESS))
        a.srand(9)
        b.srand(9)
        for _ in range(4):
            a.step()
            b.step()
        pa, pb = a.download(), b.download()
        for k in ("p", "u", "dens", "type", "n"):
            assert np.array_equal(pa[k], pb[k]), k
    with fa.FlipContext(16) as c:
        with pytest.raises(fa.FlipError, match="does not match the grid size"):
            c.load_state(path)


@pytest.mark.gpu
def test_shim_matches_c_abi():
    """tests/shim/shim_check: solver::solve, mapper::mapP2G/mapG2P/
    fetchVelocity, flip3D::init/simulateStep and a checkpoint reload through
    flip3D::init(step), each bit-identical to the flat C ABI."""
    if not os.path.exists(SHIM_CHECK):
        pytest.fail(f"{SHIM_CHECK} not built; run __graft_entry__.build()")
    d = tempfile.mkdtemp()
    r = subprocess.run([SHIM_CHECK], cwd=d, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr


# objects.dat (flip3D.cpp:498-508): per object a line "type shape material
# visible" and a line "r c0 c1 c2 p00 p01 p02 p10 p11 p12".  Each reference
# scene function fills one reused stack Object, so only the fields it sets
# (or carries from an earlier object of the same function) are defined; the
# masks below list them (1 = compared), from placeWalls (:300-362) and the
# scene functions (:203-297; the benchmark objects of ref_harness.cpp).
_WALL = ([1, 1, 1, 1], [0, 0, 0, 0, 1, 1, 1, 1, 1, 1])
_OBJ_MASKS = {
    0: [([1, 1, 0, 1], [0, 0, 0, 0, 1, 1, 1, 1, 1, 1])] * 2,                      # damBreakTest: two boxes
    1: [([1, 1, 1, 1], [1, 1, 1, 1, 0, 0, 0, 0, 0, 0]),                         # spherePourTest: wall sphere,
        ([1, 1, 1, 1], [1, 1, 1, 1, 1, 1, 1, 1, 1, 1])],                        #   box (material, r, c carried)
    2: [([1, 1, 1, 1], [0, 0, 0, 0, 1, 1, 1, 1, 1, 1])] * 2,                      # cliffPourTest: wall box, box
    3: [([1, 1, 0, 0], [0, 0, 0, 0, 1, 1, 1, 1, 1, 1]),                         # waterDropTest: box,
        ([1, 1, 0, 0], [1, 1, 1, 1, 1, 1, 1, 1, 1, 1])],                        #   sphere (p carried)
    4: [([1, 1, 0, 1], [0, 0, 0, 0, 1, 1, 1, 1, 1, 1]),                         # benchmark: box,
        ([1, 1, 0, 1], [1, 1, 1, 1, 1, 1, 1, 1, 1, 1])],                        #   sphere (p carried)
}
_SCENE_OF = {0: fa.SCENE_REFERENCE, 1: fa.SCENE_SPHERE_POUR, 2: fa.SCENE_CLIFF_POUR, 3: fa.SCENE_WATER_DROP,
             4: fa.SCENE_BENCHMARK}


def _objects(text):
    lines = text.strip().split("\n")
    count = int(lines[0])
    assert len(lines) == 1 + 2 * count
    return [(lines[1 + 2 * t].split(), lines[2 + 2 * t].split()) for t in range(count)]


@pytest.mark.gpu
@pytest.mark.parametrize("which", [0, 1, 2, 3, 4])
def test_objects_dat_matches_reference(which, tmp_path):
    """The first save of a context writes objects.dat beside the checkpoint:
    the reference's eight objects (six walls, two scene objects) with every
    field the reference defines equal to the reference's own file, token for
    token; a context that loads the pair and saves again writes the same
    file."""
    ref = _objects(open(os.path.join(ROOT, "tests", "golden", f"objects_{which}.txt")).read())
    masks = [_WALL] * 6 + _OBJ_MASKS[which]
    d1, d2 = tmp_path / "a", tmp_path / "b"
    d1.mkdir()
    d2.mkdir()
    with fa.FlipContext(32) as g:
        g.build_scene(_SCENE_OF[which])
        g.save_state(str(d1 / "1_scene.dat"))
    mine_text = open(d1 / "objects.dat").read()
    mine = _objects(mine_text)
    assert len(mine) == len(ref) == 8
    for t, ((a1, a2), (b1, b2), (m1, m2)) in enumerate(zip(mine, ref, masks)):
        for u in range(4):
            if m1[u]:
                assert a1[u] == b1[u], (which, t, u, a1, b1)
        for u in range(10):
            if m2[u]:
                assert a2[u] == b2[u], (which, t, u, a2, b2)
    with fa.FlipContext(32) as g:
        g.load_state(str(d1 / "1_scene.dat"))
        g.save_state(str(d2 / "1_scene.dat"))
    assert open(d2 / "objects.dat").read() == mine_text

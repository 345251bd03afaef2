"""Particle files (rx_ps.cpp:581-644) and scene files (rx_sph_config.h) on the host: no GPU involved."""
import math
import struct

import numpy as np
import pytest

from fltk_sph_ice_amd import _abi, scene


def _pv(n, seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 4), np.float32)
    v = np.zeros((n, 4), np.float32)
    p[:, :3] = rng.normal(size=(n, 3))
    v[:, :3] = rng.normal(size=(n, 3))
    return p, v


def test_particle_file_layout_and_round_trip(tmp_path):
    p, v = _pv(5)
    f = tmp_path / "sph_0001.dat"
    scene.write_particles(f, p, v)
    raw = f.read_bytes()
    # uint32 count, then the positions (x y z float32 each), then the velocities
    expect = struct.pack("<I", 5) + p[:, :3].astype("<f4").tobytes() + v[:, :3].astype("<f4").tobytes()
    assert raw == expect
    rp, rv = scene.read_particles(f)
    assert np.array_equal(rp[:, :3], p[:, :3]) and np.array_equal(rv[:, :3], v[:, :3])
    assert np.all(rp[:, 3] == 0) and np.all(rv[:, 3] == 0)


def test_positions_only_and_broken_files(tmp_path):
    p, _ = _pv(4, 1)
    f = tmp_path / "pos_only.dat"
    f.write_bytes(struct.pack("<I", 4) + p[:, :3].astype("<f4").tobytes())
    rp, rv = scene.read_particles(f)
    assert np.array_equal(rp[:, :3], p[:, :3]) and np.all(rv == 0)
    g = tmp_path / "short.dat"
    g.write_bytes(struct.pack("<I", 4) + p[:3, :3].astype("<f4").tobytes())  # ends inside the positions
    lib = _abi.load()
    lib.CuClearLastError()
    with pytest.raises(RuntimeError):
        scene.read_particles(g)
    with pytest.raises(RuntimeError):
        scene.read_particles(tmp_path / "missing.dat")
    # the count alone
    assert lib.rxParticlesRead(str(f).encode(), None, None, 0) == 4


def test_scene_file_space_and_liquid_boxes(tmp_path):
    cfg = tmp_path / "sph_scene_1.cfg"
    cfg.write_text(
        "// a dam break\n"
        "[space]\n"
        "cen = 0.0, 0.0, 0.0\n"
        "ext = 1.0, 0.5, 0.5\n"
        "density = 998.29\n"
        "mass = 0.0002\n"
        "kernel_particles = 20\n"
        "dt = 0.004\n"
        "max_iterations = 3\n"
        "\n"
        "[liquid box (r)]\n"
        "cen = -0.5, 0.0, 0.0   # relative to the space\n"
        "ext = 0.3, 0.6, 0.6\n"
        "[liquid box]\n"
        "cen = 0.5, 0.2, 0.0\n"
        "ext = 0.1, 0.1, 0.1\n"
        "vel = 0.0, -1.0, 0.0\n")
    sc = scene.scene_from_config(cfg, seed=3)
    h = (3 * 20 * 0.0002 / 998.29 / (4 * math.pi)) ** (1 / 3)  # rxPBDSPH_GPU::Initialize
    r = (math.pi / (6 * 20)) ** (1 / 3) * h
    assert sc.h == pytest.approx(h) and sc.spacing == pytest.approx(2 * r)
    assert sc.box_min == (-1.0, -0.5, -0.5) and sc.box_max == (1.0, 0.5, 0.5)
    assert sc.dt == 0.004 and sc.iterations == 3 and sc.mass == 0.0002
    (b1, o1, v1), (b2, o2, v2) = sc.extra["blocks"]
    # first box: centre (-0.5, 0, 0), half extent (0.3, 0.3, 0.3) -> 2 int(0.3 / 2r) - 1 particles per axis
    n1 = 2 * int(0.3 / (2 * r)) - 1
    assert b1 == (n1, n1, n1) and v1 == (0.0, 0.0, 0.0)
    n2 = 2 * int(0.1 / (2 * r)) - 1
    assert b2 == (n2, n2, n2) and v2 == (0.0, -1.0, 0.0)
    pos, vel = sc.particles()
    assert len(pos) == sc.n == n1 ** 3 + n2 ** 3
    first = pos[:n1 ** 3, :3]
    np.testing.assert_allclose(first.mean(0), [-0.5, 0.0, 0.0], atol=2e-3 * r)
    np.testing.assert_allclose(first.max(0) - first.min(0), (n1 - 1) * 2 * r, atol=5e-3 * r)  # jitter 0.001 of 2r per end
    assert np.all(vel[n1 ** 3:, 1] == -1.0) and np.all(vel[:n1 ** 3] == 0)
    sp = sc.params()
    assert sp.Mass == pytest.approx(0.0002) and sp.EffectiveRadius == pytest.approx(h, rel=1e-6)

"""The CPU marching-cubes path (oracle/mc.cpp) against what fixes it independently of the library:
the reference's case tables (digests), closed-surface topology, exact iso-planes and hand-worked voxels."""
import hashlib
import json
from collections import Counter
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden" / "mc_tables_digest.json"


def _tables():
    lib = oracle.load()
    edge = np.zeros(256, np.uint16)
    nv = np.zeros(256, np.uint8)
    tri = np.zeros((256, 16), np.uint8)
    lib.orc_mc_tables(oracle.ptr(edge), oracle.ptr(nv), oracle.ptr(tri))
    return edge, nv, tri


def test_case_tables_match_the_reference_digests():
    # rx_mc_tables.h:19-580: edgeTable, triTable (rows up to their terminator), numVertsTable
    edge, nv, tri = _tables()
    rows = [[int(e) for e in tri[c][: nv[c]]] for c in range(256)]
    assert all(int(v) == 0xFF for c in range(256) for v in tri[c][nv[c]:])
    canon = {
        "edge": ",".join(str(int(m)) for m in edge),
        "tri": ";".join(",".join(map(str, r)) for r in rows),
        "nverts": ",".join(str(int(v)) for v in nv),
    }
    ref = json.loads(GOLDEN.read_text())
    for k, v in canon.items():
        assert hashlib.sha256(v.encode()).hexdigest() == ref[k], k


def test_case_tables_cover_exactly_the_crossing_edges():
    # independent of the digests: each case's triangles use every edge whose corners straddle the level, and
    # no other; complementary cases use the same edges.
    edge, nv, tri = _tables()
    ends = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    for c in range(256):
        crossing = sum(1 << e for e, (a, b) in enumerate(ends) if ((c >> a) & 1) != ((c >> b) & 1))
        used = 0
        for e in tri[c][: nv[c]]:
            used |= 1 << int(e)
        assert nv[c] % 3 == 0 and used == crossing == edge[c], c
        assert edge[c] == edge[255 - c]


def _field(n, h, lo, f):
    z, y, x = np.meshgrid(*[lo[a] + np.arange(n[a], dtype=np.float64) * h[a] for a in (2, 1, 0)], indexing="ij")
    return f(x, y, z).astype(np.float32)


def _topology(r):
    T = r["tris"]
    e = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    directed = Counter(map(tuple, e))
    undirected = Counter(tuple(sorted(k)) for k in directed)
    return directed, undirected


@pytest.mark.parametrize("shape", ["sphere", "torus"])
def test_closed_surfaces_are_watertight_and_oriented(shape):
    n, h, lo = (24, 22, 20), (0.1, 0.1, 0.1), (-1.2, -1.1, -1.0)
    if shape == "sphere":
        f = lambda x, y, z: np.sqrt(x * x + y * y + z * z) - 0.713
        chi = 2
    else:
        f = lambda x, y, z: np.sqrt((np.sqrt(x * x + y * y) - 0.61) ** 2 + z * z) - 0.297
        chi = 0
    r = oracle.mc_mesh(_field(n, h, lo, f), n, h, lo, 0.0)
    assert r["ntris"] > 100
    directed, undirected = _topology(r)
    # every mesh edge is used by exactly two triangles, once in each direction
    assert max(directed.values()) == 1 and set(undirected.values()) == {2}
    # every vertex is used, Euler characteristic of the surface
    assert len(np.unique(r["tris"])) == r["nvrts"]
    assert r["nvrts"] - len(undirected) + r["ntris"] == chi
    V = r["vrts"][:, :3].astype(np.float64)
    assert np.abs(f(V[:, 0], V[:, 1], V[:, 2])).max() < 0.02
    # triangles and vertex normals face the side below the threshold (the inside of these fields)
    T = r["tris"]
    tn = np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]])
    grad = lambda p: np.stack([(f(*(p + d)) - f(*(p - d))) for d in np.eye(3)[:, :, None] * 1e-4], -1)
    c = V[T].mean(1)
    area = np.linalg.norm(tn, axis=1)
    solid = area > 1e-3 * np.median(area)  # slivers where the level passes through a sample point
    assert solid.mean() > 0.9
    assert (np.sum(tn * grad(c.T), 1) < 0)[solid].all()
    N = r["nrms"].astype(np.float64)
    assert np.allclose(np.linalg.norm(N, axis=1), 1.0, atol=1e-6)
    assert (np.sum(N * grad(V.T), 1) < 0).all()
    # counts of the first stage agree with the scans and the outputs
    assert r["nverts_vox"] == 3 * r["ntris"] and r["active"] == int(r["occ"].sum())


def test_linear_field_vertices_lie_on_the_iso_plane():
    # linear interpolation of a linear field is exact up to rounding
    n, h, lo = (9, 8, 7), (0.25, 0.2, 0.3), (0.0, -0.5, 0.1)
    a, b, c, d = 0.7, -0.4, 0.5, 0.05
    vol = _field(n, h, lo, lambda x, y, z: a * x + b * y + c * z + d)
    thr = 0.3
    r = oracle.mc_mesh(vol, n, h, lo, thr)
    assert r["nvrts"] > 20
    V = r["vrts"][:, :3].astype(np.float64)
    # edges of the outer layer (index n along an axis) read the clamped last sample: their vertices sit one
    # spacing outside the sampled box, off the plane, as in the reference.  Inside the box all are exact.
    lo3, hi3 = np.array(lo), np.array(lo) + (np.array(n) - 1) * np.array(h)
    inside = np.all((V >= lo3 - 1e-6) & (V <= hi3 + 1e-6), axis=1)
    assert inside.sum() > 20 and (~inside).sum() > 0
    assert np.abs(a * V[:, 0] + b * V[:, 1] + c * V[:, 2] + d - thr)[inside].max() < 1e-5
    assert np.all(r["vrts"][:, 3] == 1.0)
    # a plane: vertex normals away from the outer layer equal the plane's normal, towards the lower values
    core = np.all((V >= lo3 + np.array(h) - 1e-6) & (V <= hi3 - np.array(h) + 1e-6), axis=1)
    N = r["nrms"]
    g = np.array([a, b, c]) / np.linalg.norm([a, b, c])
    assert core.sum() > 5 and np.abs(N + g)[core].max() < 1e-5


def test_one_voxel_cases_by_hand():
    # 2x2x2 samples: one inner voxel at (0,0,0); the others read clamped samples and are degenerate.
    n, h, lo = (2, 2, 2), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    vol = np.ones(8, np.float32)
    vol[0] = 0.0  # corner 0 below 0.5: case 1, edges (0, 8, 3) at their midpoints
    r = oracle.mc_mesh(vol, n, h, lo, 0.5)
    assert r["voxbit"][0] == 1 and r["vnum"][0] == 3 and r["tnum"][0] == 1
    # the clamped voxels (1,0,0), (0,1,0), (0,0,1), ... see corner 0 at several corners
    assert r["active"] == int(r["occ"].sum())
    V = r["vrts"][:, :3]
    # vertices: x-edge (0,0,0), y-edge (0,0,0), z-edge (0,0,0) and the clamped copies
    mids = {tuple(v) for v in V.tolist()}
    assert {(0.5, 0.0, 0.0), (0.0, 0.5, 0.0), (0.0, 0.0, 0.5)} <= mids
    tri0 = r["tris"][r["tscan"][0]]
    pts = [tuple(V[k]) for k in tri0]
    # triangle (e0, e8, e3) written as (V(e8), V(e3), V(e0))
    assert pts == [(0.0, 0.0, 0.5), (0.0, 0.5, 0.0), (0.5, 0.0, 0.0)]


def test_empty_and_full_volumes():
    n, h, lo = (5, 4, 3), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    for v in (0.0, 2.0):
        r = oracle.mc_mesh(np.full(60, v, np.float32), n, h, lo, 1.0)
        assert r["active"] == 0 and r["nvrts"] == 0 and r["ntris"] == 0
    # the level exactly at a sample value: "below" is strict
    vol = np.full(60, 1.0, np.float32)
    r = oracle.mc_mesh(vol, n, h, lo, 1.0)
    assert r["active"] == 0

"""Step drivers written against the extern "C" ABI alone, in the call order of the reference host classes.

They restate rxSPH_GPU::Update (rx_sph_gpu.cpp:756-923) and rxPBDSPH_GPU::Update (rx_sph_pbd_gpu.cpp:790-957)
one ABI call at a time, so a test can check the drop-in boundary end to end.  The fused solvers of the
library (include/rx_sph_host.h) are the fast path; these drivers are the reference call sequence.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import check, float3, load, rxParticleCell, uint3
from .device import DevBuf


class GridBuffers:
    """Device arrays of rxParticleCell plus the per-particle hash/index (rxSPH_GPU::Allocate, rx_sph_gpu.cpp:391-472)."""

    def __init__(self, n: int, ncells: int, lib=None):
        self.lib = lib or load()
        self.n, self.ncells = n, ncells
        self.hash = DevBuf(4 * n)
        self.index = DevBuf(4 * n)
        self.cstart = DevBuf(4 * ncells)
        self.cend = DevBuf(4 * ncells)
        self.spos = DevBuf(16 * n)
        self.svel = DevBuf(16 * n)

    def cell(self) -> rxParticleCell:
        c = rxParticleCell()
        c.dSortedPos, c.dSortedVel = self.spos.ptr, self.svel.ptr
        c.dSortedIndex, c.dGridParticleHash = self.index.ptr, self.hash.ptr
        c.dCellStart, c.dCellEnd = self.cstart.ptr, self.cend.ptr
        c.uNumParticles, c.uNumCells = self.n, self.ncells
        return c

    def set_object_to_cell(self, dpos: int, dvel: int, dattr: int | None = None) -> rxParticleCell:
        """rxSPH_GPU::setObjectToCell: hash, sort, reorder + cell bounds."""
        lib = self.lib
        lib.CuCalcHash(self.hash.ptr, self.index.ptr, dpos, dattr, self.n)
        lib.CuSort(self.hash.ptr, self.index.ptr, self.n)
        c = self.cell()
        lib.CuReorderDataAndFindCellStart(c, dpos, dvel)
        check(lib)
        return c


def set_boundary_grid(sp, box_min, box_max, h: float) -> None:
    """Boundary-particle grid geometry of rxSPH_GPU::InitBoundary (rx_sph_gpu.cpp:495-511): the box expanded by
    4 particle radii, cell = h (the build's cell rule), GridSizeB = int(extent / h) + 1 per axis."""
    r = np.float32(sp.ParticleRadius)
    cells = 1
    for k in range(3):
        lo = np.float32(box_min[k]) - np.float32(4.0) * r
        hi = np.float32(box_max[k]) + np.float32(4.0) * r
        n = int((hi - lo) / np.float32(h)) + 1
        sp.WorldOriginB[k], sp.WorldMaxB[k], sp.CellWidthB[k] = float(lo), float(hi), float(np.float32(h))
        sp.GridSizeB[k] = n
        cells *= n
    sp.NumCellsB = cells


class BoundaryGrid:
    """Boundary particles through the ABI in the order of rxSPH_GPU::InitBoundary (rx_sph_gpu.cpp:482-572): hash
    (CuCalcHashB), sort, reorder + cell bounds (CuReorderDataAndFindCellStartB), then the volumes
    (CuSphBoundaryVolume).  `sp` must carry the boundary grid (set_boundary_grid) and be the parameters set
    with CuSetParameters."""

    def __init__(self, sp, bpos4: np.ndarray, lib=None):
        self.lib = lib = lib or load()
        self.n = n = bpos4.shape[0]
        self.g = GridBuffers(n, sp.NumCellsB, lib)
        self.dpos = DevBuf.from_array(np.ascontiguousarray(bpos4, np.float32))
        self.vol = DevBuf(4 * n)
        self.scl = DevBuf(4 * n)
        lib.CuCalcHashB(self.g.hash.ptr, self.g.index.ptr, self.dpos.ptr, float3(*sp.WorldOriginB), float3(*sp.CellWidthB),
                        uint3(*sp.GridSizeB), n)
        lib.CuSort(self.g.hash.ptr, self.g.index.ptr, n)
        self.cell = self.g.cell()
        lib.CuReorderDataAndFindCellStartB(self.cell, self.dpos.ptr)
        lib.CuSphBoundaryVolume(self.vol.ptr, sp.Mass, self.cell)
        check(lib)


class PolygonCells:
    """The triangles of a mesh listed per cell of the particle grid, on the device, for the *WithPolygon
    integrators (rxParticleCell::dSortedPolyIdx / dGridPolyHash / dPolyCellStart / dPolyCellEnd).  A triangle is
    listed in every cell its bounding box overlaps (rxSPH_GPU::setPolysToCell, rx_sph_gpu.cpp:2301-2420, keeps
    only the cells the triangle itself cuts; a superset lists more candidates and finds the same hits).
    vrts: (nv, 3) float32, tris: (nt, 3) int32."""

    def __init__(self, sp, vrts: np.ndarray, tris: np.ndarray, lib=None):
        self.lib = lib or load()
        self.vrts = np.ascontiguousarray(vrts, np.float32)
        self.tris = np.ascontiguousarray(tris, np.int32)
        org = np.array([sp.WorldOrigin[k] for k in range(3)], np.float32)
        cw = np.array([sp.CellWidth[k] for k in range(3)], np.float32)
        gs = np.array([sp.GridSize[k] for k in range(3)], np.int64)
        tv = self.vrts[self.tris]  # (nt, 3, 3)
        lo = np.clip(np.floor((tv.min(1) - org) / cw).astype(np.int64), 0, gs - 1)
        hi = np.clip(np.floor((tv.max(1) - org) / cw).astype(np.int64), 0, gs - 1)
        hashes, idx = [], []
        for t in range(len(self.tris)):
            xs, ys, zs = (np.arange(lo[t, k], hi[t, k] + 1) for k in range(3))
            z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
            h = ((z * gs[1] + y) * gs[0] + x).ravel()
            hashes.append(h)
            idx.append(np.full(len(h), t))
        h = np.concatenate(hashes).astype(np.uint32)
        ti = np.concatenate(idx).astype(np.uint32)
        order = np.argsort(h, kind="stable")
        self.hash, self.index = h[order], ti[order]
        ncells = int(np.prod(gs))
        self.start = np.full(ncells, 0xFFFFFFFF, np.uint32)
        self.end = np.full(ncells, 0xFFFFFFFF, np.uint32)
        first = np.r_[True, self.hash[1:] != self.hash[:-1]]
        last = np.r_[self.hash[1:] != self.hash[:-1], True]
        self.start[self.hash[first]] = np.nonzero(first)[0]
        self.end[self.hash[last]] = np.nonzero(last)[0] + 1
        self.d_vrts, self.d_tris = DevBuf.from_array(self.vrts), DevBuf.from_array(self.tris)
        self.d_index, self.d_hash = DevBuf.from_array(self.index), DevBuf.from_array(self.hash)
        self.d_start, self.d_end = DevBuf.from_array(self.start), DevBuf.from_array(self.end)

    def attach(self, cell: rxParticleCell) -> rxParticleCell:
        cell.dSortedPolyIdx, cell.dGridPolyHash = self.d_index.ptr, self.d_hash.ptr
        cell.dPolyCellStart, cell.dPolyCellEnd = self.d_start.ptr, self.d_end.ptr
        cell.uNumPolyHash = len(self.hash)
        return cell


def sph_step(scene, pos4: np.ndarray, vel4: np.ndarray, steps: int = 1):
    """WCSPH steps through the ABI: hash -> sort -> reorder -> density -> forces -> integrate."""
    lib = load()
    sp = scene.params()
    n = pos4.shape[0]
    lib.CuSetParameters(C.byref(sp))
    g = GridBuffers(n, sp.NumCells, lib)
    dpos, dvel = DevBuf.from_array(pos4), DevBuf.from_array(vel4)
    dens, pres, frc = DevBuf(4 * n), DevBuf(4 * n), DevBuf(16 * n)
    for _ in range(steps):
        c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
        lib.CuSphDensity(dens.ptr, pres.ptr, c)
        lib.CuSphForces(dens.ptr, pres.ptr, frc.ptr, c, scene.dt)
        lib.CuSphIntegrate(dpos.ptr, dvel.ptr, frc.ptr, dens.ptr, None, scene.dt, n)
        check(lib)
    return (dpos.download(np.float32).reshape(n, 4), dvel.download(np.float32).reshape(n, 4),
            dens.download(np.float32))


class PbfAbi:
    """PBF steps through the ABI in the order of rxPBDSPH_GPU::Update with a fixed iteration count and the
    XSPH pass of the workload.  Buffers are allocated once; `step()` issues the calls of one step."""

    def __init__(self, scene, pos4: np.ndarray, vel4: np.ndarray):
        self.lib = lib = load()
        self.scene = scene
        self.sp = sp = scene.params()
        self.n = n = pos4.shape[0]
        lib.CuSetParameters(C.byref(sp))
        self.g = GridBuffers(n, sp.NumCells, lib)
        self.dpos, self.dvel = DevBuf.from_array(pos4), DevBuf.from_array(vel4)
        self.dens, self.frc = DevBuf(4 * n), DevBuf(16 * n)
        self.ppos, self.pvel = DevBuf(16 * n), DevBuf(16 * n)
        self.scl, self.dp = DevBuf(4 * n), DevBuf(16 * n)
        self.errs: list[float] = []

    def step(self, stats: bool = True, timer=None):
        """One step.  With `stats` the maximum density error of every iteration is read back to the host as the
        reference loop does (CuPbdSphDensityErrorStats synchronises); without it the iterations run unjudged.
        `timer`, when given, is called as timer(name, fn, *args) for the lambda and position-correction calls
        (to bracket them with events on the library stream)."""
        run = timer or (lambda _name, fn, *a: fn(*a))
        lib, sc, g, n = self.lib, self.scene, self.g, self.n
        dpos, dvel, dens, frc = self.dpos, self.dvel, self.dens, self.frc
        ppos, pvel, scl, dp = self.ppos, self.pvel, self.scl, self.dp
        mx, mean = C.c_float(), C.c_float()
        self.errs = []
        c = g.set_object_to_cell(dpos.ptr, dvel.ptr)
        lib.CuPbdSphDensity(dens.ptr, c)
        lib.CuPbdSphExternalForces(dens.ptr, frc.ptr, c, sc.dt)
        lib.CuPbdSphIntegrate(dpos.ptr, dvel.ptr, frc.ptr, None, ppos.ptr, pvel.ptr, sc.dt, n)
        for _it in range(sc.iterations):
            c = g.set_object_to_cell(ppos.ptr, pvel.ptr)
            run("lambda", lib.CuPbdSphScalingFactor, ppos.ptr, dens.ptr, scl.ptr, sc.epsilon, c)
            if stats:
                lib.CuPbdSphDensityErrorStats(dens.ptr, self.sp.Density, n, C.byref(mx), C.byref(mean))
                self.errs.append(mx.value)
            run("dp", lib.CuPbdSphPositionCorrection, ppos.ptr, scl.ptr, dp.ptr, c)
            lib.CuPbdSphCorrectPosition(ppos.ptr, dp.ptr, n)
            pvel.fill(0)
            frc.fill(0)
            lib.CuPbdSphIntegrate(ppos.ptr, pvel.ptr, frc.ptr, None, ppos.ptr, pvel.ptr, sc.dt, n)
        lib.CuPbdSphUpdateVelocity(ppos.ptr, dpos.ptr, pvel.ptr, sc.dt, n)
        lib.CuCopyArrayD2D(dpos.ptr, ppos.ptr, 16 * n)
        if sc.xsph_c != 0.0:
            c = g.set_object_to_cell(dpos.ptr, pvel.ptr)
            lib.CuXSphViscosity(dpos.ptr, pvel.ptr, dvel.ptr, dens.ptr, sc.xsph_c, c)
        else:
            lib.CuCopyArrayD2D(dvel.ptr, pvel.ptr, 16 * n)
        check(lib)

    def arrays(self):
        n = self.n
        return (self.dpos.download(np.float32).reshape(n, 4), self.dvel.download(np.float32).reshape(n, 4),
                self.dens.download(np.float32))


def pbf_step(scene, pos4: np.ndarray, vel4: np.ndarray, steps: int = 1):
    """PBF steps through the ABI (PbfAbi); returns (pos4, vel4, dens, max_err per iteration of the last step)."""
    s = PbfAbi(scene, pos4, vel4)
    for _ in range(steps):
        s.step()
    return (*s.arrays(), s.errs)

"""Surface extraction through the C ABI: the marching-cubes host object of the reference (rxMCMeshGPU,
rx_mc_gpu.cpp:104-366) over the entry points of include/rx_cu_mesh.h.

`MCMesh(n, lo, h)` allocates the device arrays `rxMCMeshGPU::Set` allocates for a lattice of n = (nx, ny, nz)
samples with minimum corner lo and spacing h; `create_mesh(volume, threshold)` runs the four stages in the order
of `rxMCMeshGPU::CreateMeshV` and returns the vertex and triangle counts; `vertices()`, `triangles()` and
`normals()` download the mesh.  The volume is a device pointer (a DevBuf or an int address) or a host array.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import check, float3, load, uint3
from .device import DevBuf


class MCMesh:
    """Device arrays and stage calls of one marching-cubes lattice (RX_CUMC_USE_GEOMETRY form)."""

    def __init__(self, n, lo, h, lib=None):
        self.lib = lib or load()
        self.n = tuple(int(v) for v in n)
        nx, ny, nz = self.n
        if min(self.n) < 1:
            raise ValueError(f"lattice {self.n} is empty")
        self.lo = tuple(float(v) for v in lo)
        self.h = tuple(float(v) for v in h)
        self.nvox = nx * ny * nz
        # rx_mc_gpu.cpp:126-133
        self.esize = (uint3 * 3)(uint3(nx, ny + 1, nz + 1), uint3(nx + 1, ny, nz + 1), uint3(nx + 1, ny + 1, nz))
        ne = [nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz]
        self.nedge = (C.c_uint * 4)(*ne, sum(ne))
        if self.nvox >= 2**32 or sum(ne) >= 2**32:
            raise ValueError("lattice too large for 32-bit voxel and edge indices")
        v4, e4 = 4 * self.nvox, 4 * sum(ne)
        b = lambda k: DevBuf(k, self.lib)
        self.vox_bit, self.vox_vnum, self.vox_vscan = b(v4), b(v4), b(v4)
        self.vox_tnum, self.vox_tscan = b(v4), b(v4)
        self.vox_occ, self.vox_occscan, self.vox_compact = b(v4), b(v4), b(v4)
        self.edge_occ, self.edge_occscan = b(e4), b(e4)
        self.edge_vrts, self.vrts = b(16 * sum(ne)), b(16 * sum(ne))
        self.nrms = b(12 * sum(ne))
        self.tris: DevBuf | None = None
        self.volume: DevBuf | None = None
        self.active = self.nvrts = self.ntris = 0

    def _volume_ptr(self, volume) -> int:
        if isinstance(volume, DevBuf):
            return volume.ptr
        if isinstance(volume, int):
            return volume
        a = np.ascontiguousarray(volume, np.float32).reshape(-1)
        if a.size != self.nvox:
            raise ValueError(f"volume has {a.size} samples, the lattice {self.nvox}")
        if self.volume is None:
            self.volume = DevBuf(4 * self.nvox, self.lib)
        self.volume.upload(a)
        return self.volume.ptr

    def create_mesh(self, volume, threshold: float) -> tuple[int, int]:
        """rxMCMeshGPU::CreateMeshV (rx_mc_gpu.cpp:336-366): returns (vertices, triangles)."""
        lib = self.lib
        vol = self._volume_ptr(volume)
        n3 = uint3(*self.n)
        h3 = float3(*self.h)
        active, nv, nt = C.c_uint(), C.c_uint(), C.c_uint()
        lib.CuMCCalTriNum(vol, self.vox_bit.ptr, self.vox_vnum.ptr, self.vox_vscan.ptr, self.vox_tnum.ptr, self.vox_tscan.ptr,
                          self.vox_occ.ptr, self.vox_occscan.ptr, self.vox_compact.ptr, n3, self.nvox, h3, threshold,
                          C.byref(active), C.byref(nv), C.byref(nt))
        check(lib)
        self.active, ntris = active.value, nt.value
        lib.CuMCCalEdgeVrts(vol, self.edge_vrts.ptr, self.vrts.ptr, self.edge_occ.ptr, self.edge_occscan.ptr, self.esize,
                            self.nedge, n3, self.nvox, h3, float3(*self.lo), threshold, C.byref(nv))
        check(lib)
        self.nvrts = nv.value
        if self.nvrts:
            if self.tris is None or self.tris.nbytes < 12 * max(ntris, 1):
                self.tris = DevBuf(12 * max(ntris, 1), lib)
            lib.CuMCCalTri(self.tris.ptr, self.vox_bit.ptr, self.vox_tscan.ptr, self.vox_compact.ptr, self.edge_occscan.ptr,
                           self.esize, self.nedge, n3, self.nvox, h3, threshold, self.active, self.nvrts, ntris)
            check(lib)
            lib.CuMCCalNrm(self.nrms.ptr, self.tris.ptr, self.vrts.ptr, self.nvrts, ntris)
            check(lib)
        else:
            ntris = 0
        self.ntris = ntris
        return self.nvrts, self.ntris

    def vertices(self) -> np.ndarray:
        return self.vrts.download(np.float32, 4 * self.nvrts).reshape(-1, 4)

    def triangles(self) -> np.ndarray:
        if not self.ntris:
            return np.zeros((0, 3), np.uint32)
        return self.tris.download(np.uint32, 3 * self.ntris).reshape(-1, 3)

    def normals(self) -> np.ndarray:
        return self.nrms.download(np.float32, 3 * self.nvrts).reshape(-1, 3)

    def stage_arrays(self) -> dict:
        """Every array the stages wrote, downloaded (for checks against the CPU path)."""
        u = lambda b, k=None: b.download(np.uint32, k)
        r = {
            "voxbit": u(self.vox_bit), "vnum": u(self.vox_vnum), "tnum": u(self.vox_tnum), "occ": u(self.vox_occ),
            "occscan": u(self.vox_occscan), "compacted": u(self.vox_compact, self.active) if self.active else np.zeros(0, np.uint32),
            "vscan": u(self.vox_vscan), "tscan": u(self.vox_tscan), "eocc": u(self.edge_occ), "escan": u(self.edge_occscan),
            "edge_vrts": self.edge_vrts.download(np.float32).reshape(-1, 4),
        }
        return r


class AnisoSurface:
    """Surface of a particle set through the ABI in the reference's order: rxSPH_GPU::CalAnisotropicKernel
    (rx_sph_gpu.cpp:2036-2095: grid on the positions, smoothed centres, grid on the centres, covariance,
    decomposition, transform) then the lattice and the mesh of rxFlGLWindow::calMeshSPH_GPU
    (rx_fltk_glcanvas.cpp:2086-2130: grid on the positions, anisotropic lattice density, marching cubes).

    `sp` is the rxSimParams set with CuSetParameters; `n_lat`, `lo`, `d` the lattice; `h` the anisotropic
    neighbourhood radius (the reference uses twice the effective radius), `lam` the smoothing weight (0.9)."""

    def __init__(self, sp, dpos: DevBuf, dvel: DevBuf, n: int, n_lat, lo, d: float, lib=None, h=None, lam=0.9):
        from .abi_driver import GridBuffers

        self.lib = lib = lib or load()
        self.sp, self.n = sp, n
        self.dpos, self.dvel = dpos, dvel
        self.h = float(h if h is not None else 2.0 * sp.EffectiveRadius)
        self.lam = float(lam)
        self.grid = GridBuffers(n, sp.NumCells, lib)
        self.up, self.wpos, self.posw = DevBuf(16 * n, lib), DevBuf(16 * n, lib), DevBuf(16 * n, lib)
        self.cmat, self.eig, self.rmat, self.gmat = DevBuf(36 * n, lib), DevBuf(12 * n, lib), DevBuf(36 * n, lib), DevBuf(36 * n, lib)
        self.n_lat = tuple(int(v) for v in n_lat)
        self.lo = tuple(float(v) for v in lo)
        self.d = float(d)
        self.vol = DevBuf(4 * int(np.prod(self.n_lat)), lib)
        self.mesh = MCMesh(self.n_lat, self.lo, (self.d,) * 3, lib)

    def run(self, threshold: float, timer=None) -> tuple[int, int]:
        """One extraction; returns (vertices, triangles).  `timer(name, fn, *args)` wraps each stage call."""
        lib, g = self.lib, self.grid
        t = timer or (lambda _name, fn, *a: fn(*a))
        c = t("grid", g.set_object_to_cell, self.dpos.ptr, self.dvel.ptr)
        t("smooth", lib.CuSphCalUpdatedPosition, self.up.ptr, self.wpos.ptr, self.lam, self.h, c)
        c = t("grid", g.set_object_to_cell, self.up.ptr, self.dvel.ptr)
        t("covariance", lib.CuSphCalCovarianceMatrix, self.posw.ptr, self.cmat.ptr, self.h, c)
        t("decomposition", lib.CuSphSVDecomposition, self.cmat.ptr, self.posw.ptr, self.eig.ptr, self.rmat.ptr, self.n)
        t("transform", lib.CuSphCalTransformMatrix, self.eig.ptr, self.rmat.ptr, self.gmat.ptr, self.n)
        c = t("grid", g.set_object_to_cell, self.dpos.ptr, self.dvel.ptr)
        nx, ny, nz = self.n_lat
        t("lattice", lib.CuSphGridDensityAniso, self.vol.ptr, self.gmat.ptr, 0.0, c, nx, ny, nz, *self.lo, self.d, self.d, self.d)
        check(lib)
        return t("mesh", self.mesh.create_mesh, self.vol, threshold)

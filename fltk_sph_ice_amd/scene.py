"""Scene set-up: the seeded workloads of the particle path and their rxSimParams.

Each workload is a dam-break style block of particles on a jittered lattice inside a box (the scene of
rxParticleSystemBase::AddBox, rx_ps.cpp:195-261, with a fixed seed).  Gravity acts along -y as in the
reference application (rx_fltk_glcanvas.cpp:193).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._abi import check, load, rxSimParams, rxSphEnv


@dataclass
class Scene:
    name: str
    box_min: tuple
    box_max: tuple
    h: float
    spacing: float
    block: tuple            # (nx, ny, nz)
    block_origin: tuple
    jitter: float           # absolute, per axis
    vel_sigma: float
    seed: int
    dt: float
    rest_dens: float = 998.29
    gas_k: float = 3.0
    viscosity: float = 1.0e-3
    gravity: tuple = (0.0, -9.8, 0.0)
    # PBF
    epsilon: float = 600.0
    iterations: int = 4
    use_ap: int = 1
    ap_k: float = 0.1
    ap_n: float = 4.0
    ap_q: float = 0.2
    xsph_c: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(np.prod(self.block))

    def env(self) -> rxSphEnv:
        lib = load()
        e = rxSphEnv()
        lib.rxSphEnvDefault(C.byref(e))
        e.max_particles = self.n
        for k in range(3):
            e.boundary_cen[k] = 0.5 * (self.box_min[k] + self.box_max[k])
            e.boundary_ext[k] = 0.5 * (self.box_max[k] - self.box_min[k])
            e.gravity[k] = self.gravity[k]
        e.h = self.h
        e.rest_dens = self.rest_dens
        e.mass = 0.0
        e.spacing = self.spacing
        e.particle_radius = 0.0
        e.dt = self.dt
        e.viscosity = self.viscosity
        e.gas_k = self.gas_k
        e.epsilon = self.epsilon
        e.min_iter = self.iterations
        e.max_iter = self.iterations
        e.use_ap = self.use_ap
        e.ap_k, e.ap_n, e.ap_q = self.ap_k, self.ap_n, self.ap_q
        e.xsph_c = self.xsph_c
        e.cell_width = 0.0
        e.expansion = 0.01
        return e

    def params(self) -> rxSimParams:
        lib = load()
        p = rxSimParams()
        e = self.env()
        if lib.rxSetupSimParams(C.byref(e), C.byref(p)) != 0:
            check(lib)
        p.NumBodies = self.n
        return p

    def particles(self) -> tuple[np.ndarray, np.ndarray]:
        """(pos4, vel4) float32 arrays of shape (n, 4)."""
        lib = load()
        n = self.n
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        org = (C.c_float * 3)(*self.block_origin)
        got = lib.rxSceneBox(pos.ctypes.data, vel.ctypes.data, n, org, *self.block, self.spacing, self.jitter,
                             self.vel_sigma, self.seed)
        assert got == n
        if "v0" in self.extra:
            vel[:, :3] += np.asarray(self.extra["v0"], np.float32)
        return pos, vel


# --------------------------------------------------------------------------------------------- workloads
def wcsph_dam_break(nx: int = 32, ny: int = 32, nz: int = 32, seed: int = 1) -> Scene:
    """WCSPH dam break: 32^3 block at spacing 0.5h, jitter 0.01h, unit box, h = 0.02, k = 3, mu = 3.5,
    dt = 1e-3 (32,768 particles at the default size)."""
    h = 0.02
    return Scene("wcsph_dam_break", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), h, 0.5 * h, (nx, ny, nz), (0.0, 0.0, 0.0),
                 0.01 * h, 0.0, seed, 1.0e-3, gas_k=3.0, viscosity=3.5)


def pbf_dam_break(nx: int = 64, ny: int = 64, nz: int = 64, seed: int = 2, box: tuple | None = None) -> Scene:
    """Plain PBF: block at spacing 0.5h (h = 0.1) with jitter 0.01h in a 2:1:1 box, 4 Jacobi iterations,
    eps = 600, s_corr k = 0.1 n = 4, XSPH c = 0.01, dt = 1/120, velocities N(0, 0.05)
    (262,144 particles at the default size)."""
    h = 0.1
    s = 0.5 * h
    if box is None:
        ly = max(ny, nz) * s
        box = (2.0 * ly, ly, ly)
    return Scene("pbf_dam_break", (0.0, 0.0, 0.0), box, h, s, (nx, ny, nz), (0.0, 0.0, 0.0), 0.01 * h, 0.05, seed,
                 1.0 / 120.0, epsilon=600.0, iterations=4, use_ap=1, ap_k=0.1, ap_n=4.0, ap_q=0.2, xsph_c=0.01)


def pbf_1m(seed: int = 5) -> Scene:
    """Learned-warm-start PBF workload: 128 x 128 x 64 = 1,048,576 particles, plain-PBF parameters."""
    sc = pbf_dam_break(128, 128, 64, seed=seed, box=(12.8, 6.4, 6.4))
    sc.name = "pbf_1m"
    return sc


@dataclass
class IceScene:
    """An ice block for batched shape matching: jittered lattice at rest, one cluster per particle."""

    name: str
    block: tuple
    spacing: float
    origin: tuple
    box_min: tuple
    box_max: tuple
    jitter: float
    impulse_sigma: float
    seed_pos: int
    seed_vel: int
    alpha: float = 0.9
    beta: float = 0.3
    dt: float = 1.0 / 120.0
    iterations: int = 2
    radius_spacings: float = 2.5
    gravity: tuple = (0.0, -9.8, 0.0)

    @property
    def n(self) -> int:
        return int(np.prod(self.block))

    @property
    def cluster_radius(self) -> float:
        return self.radius_spacings * self.spacing

    def particles(self):
        lib = load()
        n = self.n
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        dummy = np.zeros((n, 4), np.float32)
        org = (C.c_float * 3)(*self.origin)
        lib.rxSceneBox(pos.ctypes.data, None, n, org, *self.block, self.spacing, self.jitter, 0.0, self.seed_pos)
        lib.rxSceneBox(dummy.ctypes.data, vel.ctypes.data, n, org, *self.block, self.spacing, 0.0, self.impulse_sigma,
                       self.seed_vel)
        return pos, vel


def ice_block(nx: int = 64, ny: int = 64, nz: int = 32, seed_pos: int = 3, seed_vel: int = 4) -> IceScene:
    """Batched ice shape matching: 64 x 64 x 32 = 131,072 particles on a jittered lattice (spacing 0.05,
    jitter 1% of the spacing) 0.5 above the floor of a box; clusters from the neighbours within 2.5
    spacings; alpha = 0.9, beta = 0.3, 2 shape-matching iterations, gravity plus an N(0, 0.1) impulse."""
    s = 0.05
    ext = (nx * s, ny * s, nz * s)
    box_max = (2.0 * ext[0] + 1.0, ext[1] + 1.5, 2.0 * ext[2] + 1.0)
    origin = (0.5 * (box_max[0] - ext[0]), 0.5, 0.5 * (box_max[2] - ext[2]))
    return IceScene("ice_block", (nx, ny, nz), s, origin, (0.0, 0.0, 0.0), box_max, 0.01 * s, 0.1, seed_pos, seed_vel)


def pbf_splash(nx: int = 48, ny: int = 32, nz: int = 48, seed: int = 6, box: tuple = (6.4, 3.2, 3.2)) -> Scene:
    """PBF splash: a block released from mid-height with a downward velocity of 2 (plus N(0, 0.05) noise) into
    the same 2:1:1 box, plain-PBF parameters."""
    sc = pbf_dam_break(nx, ny, nz, seed=seed, box=box)
    sc.name = "pbf_splash"
    s = sc.spacing
    sc.block_origin = (0.5 * (box[0] - nx * s), box[1] - ny * s - 0.1, 0.5 * (box[2] - nz * s))
    sc.extra["v0"] = (0.0, -2.0, 0.0)
    return sc


@dataclass
class CoupledScene:
    """Fluid + melting ice: a PBF fluid block and an ice block (shape-matching clusters) in one box, with a
    temperature per particle.  Temperatures use the reference's scale (melting point 250 = 0 deg C):
    the ice starts at 250 (0 C), the fluid at 270 +- U(1) (20 C +- 1)."""

    name: str
    fluid: Scene             # PBF parameters, box and the fluid block
    ice_block: tuple         # (nx, ny, nz) at the fluid spacing
    ice_origin: tuple
    seed_ice: int
    seed_temp: int
    alpha: float = 0.9
    beta: float = 0.3
    sm_iterations: int = 2
    radius_spacings: float = 2.5
    heat_td: float = 40.0

    @property
    def n_fluid(self) -> int:
        return self.fluid.n

    @property
    def n_ice(self) -> int:
        return int(np.prod(self.ice_block))

    @property
    def n(self) -> int:
        return self.n_fluid + self.n_ice

    def particles(self):
        """(pos4, vel4, temps, phase): fluid first, then ice."""
        lib = load()
        fp, fv = self.fluid.particles()
        ni = self.n_ice
        ip = np.zeros((ni, 4), np.float32)
        org = (C.c_float * 3)(*self.ice_origin)
        s = self.fluid.spacing
        lib.rxSceneBox(ip.ctypes.data, None, ni, org, *self.ice_block, s, 0.01 * s, 0.0, self.seed_ice)
        pos = np.concatenate([fp, ip]).astype(np.float32)
        vel = np.concatenate([fv, np.zeros_like(ip)]).astype(np.float32)
        rng = np.random.default_rng(self.seed_temp)
        temps = np.concatenate([270.0 + rng.uniform(-1.0, 1.0, self.n_fluid), np.full(ni, 250.0)]).astype(np.float32)
        phase = np.concatenate([np.full(self.n_fluid, 2), np.full(ni, -2)]).astype(np.int32)
        return pos, vel, temps, phase


def coupled_melt(fluid=(128, 128, 96), ice=(64, 128, 64), seed: int = 9) -> CoupledScene:
    """Fluid + melting ice: 128 x 128 x 96 = 1,572,864 fluid particles (dam break at spacing h/2, h = 0.1,
    plain-PBF parameters) beside a 64 x 128 x 64 = 524,288-particle ice block, 2,097,152 in all, in a
    16 x 8 x 8 box.  Clusters from the ice neighbours within 2.5 spacings (uneven sizes up to 81 at the
    block faces and corners), alpha 0.9, beta 0.3, 2 shape-matching iterations."""
    s = 0.05
    box = (2.5 * fluid[0] * s, 1.25 * max(fluid[1], ice[1]) * s, 1.5 * max(fluid[2], ice[2]) * s)
    fl = pbf_dam_break(*fluid, seed=seed, box=box)
    fl.name = "coupled_fluid"
    ice_origin = (fluid[0] * s + 0.3 * box[0], 0.0 + s, 0.5 * (box[2] - ice[2] * s))
    return CoupledScene("coupled_melt", fl, tuple(ice), ice_origin, seed + 1, seed + 2)


# --------------------------------------------------------------------------------------------- boundary particles
def box_surface_particles(box_min, box_max, spacing: float) -> np.ndarray:
    """Boundary particles on the six faces of the box, a square lattice of the given spacing per face with the
    shared edges and corners listed once; float32 (n, 4).  (The reference samples the container surface with
    marching cubes and a relaxation, rxSolid::GenerateParticlesOnSurf, rx_sph_solid.cpp:148-216; the ABI
    takes any positions, so a lattice stands in for that host step.)"""
    lo = np.asarray(box_min, np.float64)
    hi = np.asarray(box_max, np.float64)
    cnt = np.maximum(np.round((hi - lo) / spacing).astype(int), 1)
    axes = [lo[k] + (hi[k] - lo[k]) * np.arange(cnt[k] + 1) / cnt[k] for k in range(3)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    on = np.zeros(len(g), bool)
    for k in range(3):
        on |= np.isclose(g[:, k], lo[k]) | np.isclose(g[:, k], hi[k])
    out = np.zeros((int(on.sum()), 4), np.float32)
    out[:, :3] = g[on]
    return out


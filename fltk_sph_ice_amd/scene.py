"""Scene set-up: the seeded workloads of the particle path and their rxSimParams.

Each workload is a dam-break style block of particles on a jittered lattice inside a box (the scene of
rxParticleSystemBase::AddBox, rx_ps.cpp:195-261, with a fixed seed).  Gravity acts along -y as in the
reference application (rx_fltk_glcanvas.cpp:193).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re
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
    grad_mass: int = 0       # PBF: 1 = constraint gradients carry the mass (RX_PBF_OPT_GRAD_MASS)
    mass: float = 0.0        # 0: solved from the lattice sum at `spacing` so that the lattice has rest density
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        if "blocks" in self.extra:
            return int(sum(np.prod(b[0]) for b in self.extra["blocks"]))
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
        e.mass = self.mass
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
        if "blocks" in self.extra:  # several boxes (scene files): one seeded lattice each, velocity per box
            parts = [_box(blk, org, self.spacing, self.jitter, self.vel_sigma, self.seed + t, v0)
                     for t, (blk, org, v0) in enumerate(self.extra["blocks"])]
            return np.concatenate([p for p, _ in parts]), np.concatenate([v for _, v in parts])
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
def wcsph_dam_break(nx: int = 32, ny: int = 32, nz: int = 32, seed: int = 1, box: tuple = (1.0, 1.0, 1.0)) -> Scene:
    """WCSPH dam break: block at spacing 0.5h, jitter 0.01h, h = 0.02, k = 3, mu = 3.5, dt = 1e-3; 32^3 = 32,768
    particles in the unit box at the default size."""
    h = 0.02
    return Scene("wcsph_dam_break", (0.0, 0.0, 0.0), tuple(box), h, 0.5 * h, (nx, ny, nz), (0.0, 0.0, 0.0),
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
    """Learned-warm-start PBF workload: 128 x 64 x 128 = 1,048,576 particles (64 layers deep, the depth of the
    64^3 plain-PBF block) in a 12.8 x 6.4 x 6.4 box, plain-PBF parameters.  (A 128-layer column compresses past
    what 4 corrections with eps = 600 hold and leaves the grid within 200 steps, plain or learned.)"""
    sc = pbf_dam_break(128, 64, 128, seed=seed, box=(12.8, 6.4, 6.4))
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
    radius_spacings: float = 1.6
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
    jitter 1% of the spacing) 0.5 above the floor of a box; clusters from the neighbours within 1.6
    spacings (the centre, its 6 face and 12 edge neighbours: 19 inside the block, down to 7 at its corners);
    alpha = 0.9, beta = 0.3, 2 shape-matching iterations, gravity plus an N(0, 0.1) impulse."""
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


def coupled_melt(fluid=(128, 64, 192), ice=(64, 128, 64), seed: int = 9) -> CoupledScene:
    """Fluid + melting ice: 128 x 64 x 192 = 1,572,864 fluid particles (a dam break 64 layers deep at spacing h/2,
    h = 0.1, plain-PBF parameters) beside a 64 x 128 x 64 = 524,288-particle ice block, 2,097,152 in all, in a
    16 x 8 x 14.4 box.  Clusters from the ice neighbours within 2.5 spacings (uneven sizes up to 81 at the
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


def _box(block, origin, spacing, jitter, vel_sigma, seed, v0):
    lib = load()
    n = int(np.prod(block))
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    org = (C.c_float * 3)(*origin)
    got = lib.rxSceneBox(pos.ctypes.data, vel.ctypes.data, n, org, *block, spacing, jitter, vel_sigma, seed)
    assert got == n
    vel[:, :3] += np.asarray(v0, np.float32)
    return pos, vel


# --------------------------------------------------------------------------------------------- scene files
_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


def read_config(path) -> list[tuple[str, dict]]:
    """Sections of a scene file in the reference's INI form (rxINI, read by rxSPHConfig, rx_sph_config.h:428-470):
    `[header]` lines open a section, `name = value` lines fill it; `#` and `//` start comments."""
    out: list[tuple[str, dict]] = []
    for raw in open(path, encoding="utf-8", errors="replace"):
        line = raw.split("//")[0].split("#")[0].strip()
        if not line:
            continue
        if line.startswith("[") and line.endswith("]"):
            out.append((line[1:-1].strip(), {}))
        elif "=" in line and out:
            k, v = line.split("=", 1)
            out[-1][1][k.strip()] = v.strip()
    return out


def _nums(v: str) -> list[float]:
    return [float(x) for x in _NUM.findall(v)]


def scene_from_config(path, seed: int = 1, name: str | None = None) -> Scene:
    """A Scene from a scene file: the `[space]` section (rxSPHConfig::SetSphSpace, rx_sph_config.h:193-262) and
    the `[liquid box]` / `[liquid box (r)]` sections (SetLiquidBox, :266-279; `(r)`: centre and extent relative
    to the space, in units of its half extent).

    Parameters follow rxPBDSPH_GPU::Initialize (rx_sph_pbd_gpu.cpp:166-200): h = (3 k m / (4 pi rho))^(1/3) for
    k kernel particles of mass m at density rho, particle radius r = (pi / (6 k))^(1/3) h.  A liquid box is the
    lattice of rxParticleSystemBase::AddBox (rx_ps.cpp:195-261) at spacing 2r: 2 int(ext / 2r) - 1 particles per
    axis centred on the box, jitter 0.001 of the spacing (seeded here), the box's velocity.  Defaults are those of
    rxSPHEnviroment (rx_sph.h:146-180)."""
    sections = read_config(path)
    space = next((kv for hd, kv in sections if hd == "space"), None)
    if space is None:
        raise ValueError(f"{path}: no [space] section")
    cen = np.array(_nums(space.get("cen", "0 0 0")) or [0, 0, 0], np.float64)
    ext = np.array(_nums(space.get("ext", "2.0 0.8 0.8")), np.float64)
    rho = float(space.get("density", 998.29))
    m = float(space.get("mass", 0.04))
    kp = float(space.get("kernel_particles", 20.0))
    h = (3.0 * kp * m / rho / (4.0 * math.pi)) ** (1.0 / 3.0)
    r = (math.pi / (6.0 * kp)) ** (1.0 / 3.0) * h
    spacing = 2.0 * r
    blocks = []
    for hd, kv in sections:
        if hd not in ("liquid box", "liquid box (r)"):
            continue
        bc = np.array(_nums(kv.get("cen", "0 0 0")), np.float64)
        be = np.array(_nums(kv.get("ext", "0 0 0")), np.float64)
        if hd.endswith("(r)"):
            bc, be = cen + bc * ext, be * ext
        v0 = tuple(_nums(kv.get("vel", "0 0 0")) or [0.0, 0.0, 0.0])
        sx = [int(be[k] / spacing) - 1 for k in range(3)]
        if min(sx) < 0:
            continue
        blk = tuple(2 * v + 1 for v in sx)
        org = tuple(float(bc[k] - (sx[k] + 0.5) * spacing) for k in range(3))
        blocks.append((blk, org, v0))
    if not blocks:
        raise ValueError(f"{path}: no liquid box with particles")
    sc = Scene(name or os.path.splitext(os.path.basename(str(path)))[0], tuple(cen - ext), tuple(cen + ext), h, spacing,
               blocks[0][0], blocks[0][1], 0.001 * spacing, 0.0, seed, float(space.get("dt", 0.01)), rest_dens=rho,
               gas_k=float(space.get("gas_stiffness", 3.0)), viscosity=float(space.get("viscosity", 1.0e-3)),
               epsilon=float(space.get("epsilon", 0.001)), iterations=int(space.get("max_iterations", 10)),
               use_ap=int(space.get("use_artificial_pressure", 1)), ap_k=float(space.get("ap_k", 0.1)),
               ap_n=float(space.get("ap_n", 4.0)), ap_q=float(space.get("ap_q", 0.2)), mass=m)
    sc.extra["blocks"] = blocks
    return sc


# --------------------------------------------------------------------------------------------- particle files
def write_particles(path, pos4: np.ndarray, vel4: np.ndarray | None = None) -> None:
    """rxParticleSystemBase::OutputParticles format (rxParticlesWrite)."""
    lib = load()
    p = np.ascontiguousarray(pos4, np.float32)
    v = None if vel4 is None else np.ascontiguousarray(vel4, np.float32)
    if lib.rxParticlesWrite(str(path).encode(), p.ctypes.data, None if v is None else v.ctypes.data, len(p)):
        check(lib)


def read_particles(path) -> tuple[np.ndarray, np.ndarray]:
    """rxParticleSystemBase::InputParticles format (rxParticlesRead): (pos4, vel4)."""
    lib = load()
    n = lib.rxParticlesRead(str(path).encode(), None, None, 0)
    if n < 0:
        check(lib)
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    if lib.rxParticlesRead(str(path).encode(), pos.ctypes.data, vel.ctypes.data, n) != n:
        check(lib)
    return pos, vel


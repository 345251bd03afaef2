"""ctypes bindings of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

The HIP path is compared with the oracle: tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py use it; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class OrcParams(C.Structure):
    _fields_ = [
        ("h", C.c_double), ("mass", C.c_double), ("rest_dens", C.c_double), ("gas_k", C.c_double),
        ("viscosity", C.c_double), ("gravity", C.c_double * 3), ("radius", C.c_double),
        ("bmin", C.c_double * 3), ("bmax", C.c_double * 3), ("origin", C.c_double * 3), ("cell", C.c_double * 3),
        ("gsize", C.c_int * 3), ("eps", C.c_double), ("use_ap", C.c_int), ("ap_k", C.c_double),
        ("ap_n", C.c_double), ("ap_q", C.c_double), ("xsph_c", C.c_double), ("iterations", C.c_int),
        ("dt", C.c_double), ("grad_mass", C.c_int), ("first_scale", C.c_double), ("first_dmax", C.c_double),
    ]


class OrcPbfStats(C.Structure):
    _fields_ = [("max_err", C.c_double * 16), ("mean_abs", C.c_double * 16), ("iterations", C.c_int)]


class OrcSmParams(C.Structure):
    _fields_ = [
        ("alpha", C.c_double), ("beta", C.c_double), ("dt", C.c_double), ("gravity", C.c_double * 3),
        ("bmin", C.c_double * 3), ("bmax", C.c_double * 3), ("iterations", C.c_int),
        ("volume_conservation", C.c_int), ("restitution", C.c_double),
    ]


_lib = None
P = C.c_void_p
D = C.c_double
I = C.c_int
F = C.c_float

_SIGS = {
    "orc_kernel_coef": (D, [I, D, I]),
    "orc_kernel_poly6": (D, [D, D, D]),
    "orc_kernel_spiky_grad": (D, [D, D, D]),
    "orc_kernel_visc_lap": (D, [D, D, D]),
    "orc_rest_density": (D, [D, D, D]),
    "orc_calc_hash_f": (None, [P, P, I, P, P, P]),
    "orc_stable_sort": (None, [P, P, I]),
    "orc_cell_bounds": (None, [P, I, I, P, P]),
    "orc_neighbors_brute_f": (C.c_long, [P, I, F, I, P, P]),
    "orc_sph_density_f": (None, [P, I, P, P, P]),
    "orc_sph_force_f": (None, [P, P, P, P, I, P, P]),
    "orc_sph_step_f": (None, [P, P, I, P, F, P, P, P]),
    "orc_sph_step_d": (None, [P, P, I, P, D, P]),
    "orc_pbf_density_f": (None, [P, I, P, P]),
    "orc_integrate_polygon_f": (None, [P, P, P, P, I, P, P, P, I, P, P, P, D, F, P, P]),
    "orc_grid_density_f": (None, [P, I, P, I, I, I, F, F, F, F, F, F, P, P]),
    "orc_boundary_volume_f": (None, [P, I, P, P]),
    "orc_boundary_density_f": (None, [P, I, P, P, I, P, P, P]),
    "orc_boundary_force_f": (None, [P, I, P, P, I, P, P, P, P]),
    "orc_pbf_scaling_factor_wb_f": (None, [P, I, P, P, I, P, P, P, P]),
    "orc_pbf_position_correction_wb_f": (None, [P, P, I, P, P, P, I, P, F, P]),
    "orc_pbf_ext_force_f": (None, [P, P, P, I, P, P]),
    "orc_pbf_scaling_factor_f": (None, [P, I, P, P, P]),
    "orc_pbf_position_correction_f": (None, [P, P, I, P, F, P]),
    "orc_pbf_xsph_f": (None, [P, P, P, I, P, F, P]),
    "orc_pbf_step_f": (None, [P, P, I, P, P, P, P]),
    "orc_sph_force_scale_f": (None, [P, P, P, P, I, P, P]),
    "orc_pbf_ext_force_scale_f": (None, [P, P, P, I, P, P]),
    "orc_boundary_force_scale_f": (None, [P, I, P, P, I, P, P, P, P]),
    "orc_pbf_lambda_scale_f": (None, [P, I, P, P, I, P, P, P]),
    "orc_pbf_dp_scale_f": (None, [P, P, I, P, P, P, I, P, F, P]),
    "orc_pbf_step_d": (None, [P, P, I, P, P, P]),
    "orc_jacobi_eigen3": (I, [P, P, D, I]),
    "orc_polar_decomposition": (None, [P, P, P]),
    "orc_sm_step_f": (None, [P, P, I, P, P, I, P, P]),
    "orc_sm_step_d": (None, [P, P, I, P, P, I, P, P]),
    "orc_sm_step_masked_f": (None, [P, P, I, P, P, I, P, P, P]),
    "orc_legacy_cluster_f": (None, [I, P, P, P, P, P, P, P, P, P, P, F, P, I]),
    "orc_legacy_cluster_d": (None, [I, P, P, P, P, P, P, P, P, P, P, D, P, I]),
    "orc_legacy_average_f": (None, [I, P, P, P, P, P, P, P, I, I]),
    "orc_heat_diffuse_f": (None, [P, P, P, I, P, F, F, P]),
    "orc_heat_transfer_gpu_f": (None, [P, P, P, P, P, P, P, P, P, F, F, F, F, F, I]),
    "orc_heat_cpu_f": (None, [P, P, P, I, P, F, F]),
    "orc_detect_neighbors_f": (None, [P, I, P, P]),
    "orc_detect_surface_f": (None, [P, P, P, I, F, F]),
    "orc_mlp_forward_d": (None, [P, I, I, P, P, I, P, P, I, P, P, P]),
    "orc_mlp_backward_d": (None, [P, P, I, I, P, P, I, P, P, I, P, P, P, P, P, P, P, P, P]),
    "orc_mc_tables": (None, [P, P, P]),
    "orc_aniso_smooth": (None, [P, P, P, P, C.c_uint, P, P, F, F, P, P]),
    "orc_aniso_cov": (None, [P, P, P, P, C.c_uint, P, P, F, P, P]),
    "orc_sym_eigen3": (None, [P, P]),
    "orc_aniso_svd": (None, [P, P, C.c_uint, P, P]),
    "orc_aniso_transform": (None, [P, P, C.c_uint, P]),
    "orc_aniso_grid_density": (None, [P, P, P, P, C.c_uint, P, P, F, F, F, P, I, I, I, F, F, F, F, F, F, P]),
    "orc_mc_tri_num": (None, [P, P, C.c_uint, F, P, P, P, P, P, P, P, P, P]),
    "orc_mc_edge_vrts": (C.c_uint, [P, P, P, P, P, P, F, P, P, P, P]),
    "orc_mc_tri": (None, [P, P, P, P, P, P, P, P, C.c_uint, C.c_uint, C.c_uint]),
    "orc_mc_nrm": (None, [P, P, P, C.c_uint, C.c_uint]),
}


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def params_from(sp, scene=None) -> OrcParams:
    """OrcParams carrying exactly the float32 values of an rxSimParams (plus the PBF knobs of a Scene)."""
    o = OrcParams()
    o.h = sp.EffectiveRadius
    o.mass = sp.Mass
    o.rest_dens = sp.Density
    o.gas_k = sp.GasStiffness
    o.viscosity = sp.Viscosity
    for k in range(3):
        o.gravity[k] = sp.Gravity[k]
        o.bmin[k] = sp.BoundaryMin[k]
        o.bmax[k] = sp.BoundaryMax[k]
        o.origin[k] = sp.WorldOrigin[k]
        o.cell[k] = sp.CellWidth[k]
        o.gsize[k] = sp.GridSize[k]
    o.radius = sp.ParticleRadius
    o.use_ap = sp.AP
    o.ap_k = sp.AP_K
    o.ap_n = sp.AP_N
    o.ap_q = sp.AP_Q
    o.dt = sp.Dt
    if scene is not None:
        o.eps = scene.epsilon
        o.xsph_c = scene.xsph_c
        o.iterations = scene.iterations
        o.grad_mass = int(getattr(scene, "grad_mass", 0))
    return o


def sph_step(pos4: np.ndarray, vel4: np.ndarray, P: OrcParams, dt: float, double: bool = False):
    """One WCSPH step of the CPU path; returns (pos4, vel4, dens)."""
    lib = load()
    n = pos4.shape[0]
    if double:
        pos = np.ascontiguousarray(pos4, np.float64).copy()
        vel = np.ascontiguousarray(vel4, np.float64).copy()
        dens = np.zeros(n, np.float64)
        lib.orc_sph_step_d(ptr(pos), ptr(vel), n, C.byref(P), dt, ptr(dens))
        return pos, vel, dens
    pos = np.ascontiguousarray(pos4, np.float32).copy()
    vel = np.ascontiguousarray(vel4, np.float32).copy()
    dens = np.zeros(n, np.float32)
    pres = np.zeros(n, np.float32)
    acc = np.zeros((n, 4), np.float32)
    lib.orc_sph_step_f(ptr(pos), ptr(vel), n, C.byref(P), dt, ptr(dens), ptr(pres), ptr(acc))
    return pos, vel, dens


def pbf_step(pos4: np.ndarray, vel4: np.ndarray, P: OrcParams, double: bool = False, want_lambda: bool = False):
    """One PBF step of the CPU path; returns (pos4, vel4, dens, stats[, lambda_sum])."""
    lib = load()
    n = pos4.shape[0]
    st = OrcPbfStats()
    if double:
        pos = np.ascontiguousarray(pos4, np.float64).copy()
        vel = np.ascontiguousarray(vel4, np.float64).copy()
        dens = np.zeros(n, np.float64)
        lib.orc_pbf_step_d(ptr(pos), ptr(vel), n, C.byref(P), ptr(dens), C.byref(st))
        return pos, vel, dens, st
    pos = np.ascontiguousarray(pos4, np.float32).copy()
    vel = np.ascontiguousarray(vel4, np.float32).copy()
    dens = np.zeros(n, np.float32)
    lam = np.zeros(n, np.float32)
    lib.orc_pbf_step_f(ptr(pos), ptr(vel), n, C.byref(P), ptr(dens), ptr(lam), C.byref(st))
    if want_lambda:
        return pos, vel, dens, st, lam
    return pos, vel, dens, st


def mc_layout(n):
    """Edge sizes and counts of a lattice of n = (nx, ny, nz) samples (rx_mc_gpu.cpp:126-133)."""
    nx, ny, nz = map(int, n)
    esize = np.array([[nx, ny + 1, nz + 1], [nx + 1, ny, nz + 1], [nx + 1, ny + 1, nz]], np.uint32)
    ne = [int(np.prod(e)) for e in esize]
    return esize, np.array(ne + [sum(ne)], np.uint32)


def mc_mesh(vol: np.ndarray, n, h, lo, thr: float) -> dict:
    """The four marching-cubes stages of the CPU path on one volume (x fastest, shape nz, ny, nx or flat).

    Returns every array the stages write (as the reference host keeps them) plus the counts."""
    lib = load()
    n3 = np.array(n, np.uint32)
    nvox = int(np.prod(n3))
    vol = np.ascontiguousarray(vol, np.float32).reshape(-1)
    assert vol.size == nvox
    u = lambda k: np.zeros(k, np.uint32)
    r = {k: u(nvox) for k in ("voxbit", "vnum", "vscan", "tnum", "tscan", "occ", "occscan", "compacted")}
    out3 = u(3)
    lib.orc_mc_tri_num(ptr(vol), ptr(n3), nvox, thr, ptr(r["voxbit"]), ptr(r["vnum"]), ptr(r["vscan"]), ptr(r["tnum"]),
                       ptr(r["tscan"]), ptr(r["occ"]), ptr(r["occscan"]), ptr(r["compacted"]), ptr(out3))
    r["active"], r["nverts_vox"], r["ntris"] = map(int, out3)
    esize, ne = mc_layout(n3)
    h3 = np.array(h, np.float32)
    lo3 = np.array(lo, np.float32)
    r["edge_vrts"] = np.zeros((int(ne[3]), 4), np.float32)
    r["vrts"] = np.zeros((int(ne[3]), 4), np.float32)
    r["eocc"] = u(int(ne[3]))
    r["escan"] = u(int(ne[3]))
    nv = lib.orc_mc_edge_vrts(ptr(vol), ptr(esize), ptr(ne), ptr(n3), ptr(h3), ptr(lo3), thr, ptr(r["edge_vrts"]), ptr(r["vrts"]),
                              ptr(r["eocc"]), ptr(r["escan"]))
    r["nvrts"] = int(nv)
    r["vrts"] = r["vrts"][:nv]
    nt = r["ntris"] if nv else 0
    r["tris"] = np.zeros((nt, 3), np.uint32)
    r["nrms"] = np.zeros((nv, 3), np.float32)
    if nv and nt:
        lib.orc_mc_tri(ptr(r["tris"]), ptr(r["voxbit"]), ptr(r["tscan"]), ptr(r["compacted"]), ptr(r["escan"]), ptr(esize), ptr(ne),
                       ptr(n3), r["active"], nv, nt)
        lib.orc_mc_nrm(ptr(r["nrms"]), ptr(r["tris"]), ptr(r["vrts"]), nv, nt)
    r["ntris"] = nt
    return r


def aniso_pipeline(spos4, sidx, cs, ce, geo, gsize, lam: float, h: float, h_eff: float, mass: float, wpoly6: float, lattice=None):
    """The five anisotropic-kernel stages of the CPU path on sorted particle data (positions in slot order,
    slot -> particle index, cell start / end with 0xFFFFFFFF for empty cells).  geo = origin xyz + cell width xyz.
    Returns the stage outputs in particle order, and the lattice density when lattice = (n3, lo3, d3) is given."""
    lib = load()
    n = len(sidx)
    spos4 = np.ascontiguousarray(spos4, np.float32)
    sidx = np.ascontiguousarray(sidx, np.uint32)
    cs = np.ascontiguousarray(cs, np.uint32)
    ce = np.ascontiguousarray(ce, np.uint32)
    geo = np.ascontiguousarray(geo, np.float32)
    gsize = np.ascontiguousarray(gsize, np.int32)
    r = {"up": np.zeros((n, 4), np.float32), "wpos": np.zeros((n, 4), np.float32), "posw": np.zeros((n, 4), np.float32),
         "cmat": np.zeros((n, 9), np.float32), "eigen": np.zeros((n, 3), np.float32), "rmat": np.zeros((n, 9), np.float32),
         "gmat": np.zeros((n, 9), np.float32)}
    lib.orc_aniso_smooth(ptr(spos4), ptr(sidx), ptr(cs), ptr(ce), n, ptr(geo), ptr(gsize), lam, h, ptr(r["up"]), ptr(r["wpos"]))
    lib.orc_aniso_cov(ptr(spos4), ptr(sidx), ptr(cs), ptr(ce), n, ptr(geo), ptr(gsize), h, ptr(r["posw"]), ptr(r["cmat"]))
    lib.orc_aniso_svd(ptr(r["cmat"]), ptr(r["posw"]), n, ptr(r["eigen"]), ptr(r["rmat"]))
    lib.orc_aniso_transform(ptr(r["eigen"]), ptr(r["rmat"]), n, ptr(r["gmat"]))
    if lattice is not None:
        (nx, ny, nz), lo, d = lattice
        out = np.zeros(nx * ny * nz, np.float32)
        lib.orc_aniso_grid_density(ptr(spos4), ptr(sidx), ptr(cs), ptr(ce), n, ptr(geo), ptr(gsize), h_eff, mass, wpoly6,
                                   ptr(r["gmat"]), nx, ny, nz, *map(float, lo), *map(float, d), ptr(out))
        r["grid"] = out
    return r


# ---------------------------------------------------------------------------------------------- per-particle scales
# The magnitudes FP32 results are judged against: for a sum of terms that nearly cancel (pressure forces in a
# uniform lattice, the position correction) the rounding of any evaluation order is bounded by a small multiple
# of sum_j |term_ij|, not by |sum|.
def _nb_args(b4, vol, lam_b):
    if b4 is None:
        return None, None, None, 0
    return ptr(b4), ptr(vol), (ptr(lam_b) if lam_b is not None else None), len(b4)


def sph_force_scale(pos4, vel4, dens, pres, P: OrcParams) -> np.ndarray:
    n = pos4.shape[0]
    out = np.zeros(n, np.float64)
    load().orc_sph_force_scale_f(ptr(pos4), ptr(vel4), ptr(dens), ptr(pres), n, C.byref(P), ptr(out))
    return out


def pbf_ext_force_scale(pos4, vel4, dens, P: OrcParams) -> np.ndarray:
    n = pos4.shape[0]
    out = np.zeros(n, np.float64)
    load().orc_pbf_ext_force_scale_f(ptr(pos4), ptr(vel4), ptr(dens), n, C.byref(P), ptr(out))
    return out


def boundary_force_scale(pos4, b4, vol, dens, pres, P: OrcParams, base=None) -> np.ndarray:
    """sum_b |boundary term| added to `base` (the fluid scale) or to zeros."""
    n = pos4.shape[0]
    out = np.zeros(n, np.float64) if base is None else np.array(base, np.float64)
    load().orc_boundary_force_scale_f(ptr(pos4), n, ptr(b4), ptr(vol), len(b4), C.byref(P), ptr(dens), ptr(pres), ptr(out))
    return out


def pbf_lambda_scale(pos4, P: OrcParams, b4=None, vol=None):
    """(rho_i / rho0 + |C_i|) / (sd_i + eps) per particle (and per boundary particle when b4 is given)."""
    n = pos4.shape[0]
    out = np.zeros(n, np.float64)
    outb = np.zeros(len(b4) if b4 is not None else 1, np.float64)
    pb, pv, _, nb = _nb_args(b4, vol, None)
    load().orc_pbf_lambda_scale_f(ptr(pos4), n, pb, pv, nb, C.byref(P), ptr(out), ptr(outb))
    return (out, outb) if b4 is not None else out


def pbf_dp_scale(pos4, lam, P: OrcParams, dt, b4=None, vol=None, lam_b=None) -> np.ndarray:
    n = pos4.shape[0]
    out = np.zeros(n, np.float64)
    pb, pv, pl, nb = _nb_args(b4, vol, lam_b)
    load().orc_pbf_dp_scale_f(ptr(pos4), ptr(lam), n, pb, pv, pl, nb, C.byref(P), np.float32(dt), ptr(out))
    return out


def rel_err_vs_scale(got, ref, scale) -> float:
    """max_i |got_i - ref_i| / scale_i (vectors: the Euclidean norm of the difference)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    d = np.abs(got - ref) if got.ndim == 1 else np.linalg.norm(got - ref, axis=1)
    return float((d / np.maximum(scale, 1e-300)).max())

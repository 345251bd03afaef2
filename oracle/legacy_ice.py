"""Reference-layout tables of the legacy ice launches (Ice_SM.h:23-53, IceObject.h:370-395), built from a
cluster CSR, and the oracle's versions of the launches.  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import OrcSmParams, load, ptr


def layout(rest4: np.ndarray, cs: np.ndarray, mem: np.ndarray, ptoc_param_size: int = 3) -> dict:
    """Slot tables of clusters c with members mem[cs[c]:cs[c+1]] (one slot per membership, cluster by cluster):
    indxSet (first, last slot), pIndxes, orgPos (3 per slot), orgCm (3 per cluster), and the particle -> cluster
    tables PtoCIndx / PtoC (entries (cluster, member number), PtoCMax per particle and layer, -1 padded)."""
    nc = len(cs) - 1
    n = len(rest4)
    pidx = np.ascontiguousarray(mem, np.int32)
    indx = np.stack([cs[:-1], cs[1:] - 1], 1).astype(np.int32).ravel()
    dt = np.float64 if rest4.dtype == np.float64 else np.float32  # float32 tables for float32 rest positions
    org = np.ascontiguousarray(rest4[pidx, :3], dt).ravel()
    ocm = np.stack([rest4[mem[cs[c]:cs[c + 1]], :3].astype(np.float64).mean(0) for c in range(nc)]).astype(dt).ravel()
    lists = [[] for _ in range(n)]
    for c in range(nc):
        for o, p in enumerate(mem[cs[c]:cs[c + 1]]):
            lists[p].append((c, o))
    pmax = max(1, max(len(l) for l in lists))
    ptoc = np.full((n, ptoc_param_size, pmax), -1, np.int32)
    cnt = np.zeros(n, np.int32)
    for p, l in enumerate(lists):
        cnt[p] = len(l)
        for j, (c, o) in enumerate(l):
            ptoc[p, 0, j], ptoc[p, 1, j] = c, o
    return dict(nc=nc, n=n, nslots=len(pidx), indx=indx, pidx=pidx, org=org, ocm=ocm, ptoc_indx=cnt, ptoc=ptoc.ravel().copy(),
                ptoc_max=pmax, ptoc_param_size=ptoc_param_size)


def sm_params(dt, gravity=(0.0, -9.8, 0.0), lo=(-5.0,) * 3, hi=(5.0,) * 3, restitution=0.9) -> OrcSmParams:
    S = OrcSmParams()
    S.alpha, S.beta, S.dt = 1.0, 0.0, dt
    for a in range(3):
        S.gravity[a], S.bmin[a], S.bmax[a] = gravity[a], lo[a], hi[a]
    S.iterations, S.volume_conservation, S.restitution = 1, 0, restitution
    return S


def cluster_pass(L, prt_pos4, prt_vel4, sld3, cur, ccm, vel, dt, S, itr: bool, double: bool = False):
    """orc_legacy_cluster_*: updates cur (3 per slot), ccm (3 per cluster) and, for the first pass, vel in place."""
    lib = load()
    T = np.float64 if double else np.float32
    f = lib.orc_legacy_cluster_d if double else lib.orc_legacy_cluster_f
    org, ocm = np.ascontiguousarray(L["org"], T), np.ascontiguousarray(L["ocm"], T)
    f(L["nc"], ptr(prt_pos4), ptr(prt_vel4), ptr(sld3), ptr(org), ptr(cur), ptr(ocm), ptr(ccm), ptr(vel), ptr(L["pidx"]),
      ptr(L["indx"]), dt, C.byref(S), 1 if itr else 0)


def average(L, sld_pos, sld_vel, sm_pos, sm_vel):
    load().orc_legacy_average_f(L["n"], ptr(sld_pos), ptr(sld_vel), ptr(sm_pos), ptr(sm_vel), ptr(L["indx"]), ptr(L["ptoc_indx"]),
                                ptr(L["ptoc"]), L["ptoc_max"], L["ptoc_param_size"])

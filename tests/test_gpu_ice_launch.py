"""The reference's GPU ice step through the C++-linkage launches of Ice_SM.h:23-53 and IceObject.h:370-395, in the
order of IceObject::StepObjCalcWidhIteration (IceObject.cpp:1000-1017) and StepInterPolationNormal
(IceObject.cpp:1019-1046): LaunchShapeMatchingGPU, LaunchCalcAverageGPU, then (LaunchShapeMatchingIterationGPU,
LaunchCalcAverageGPU) per further iteration, and LaunchInterPolationGPU into the fluid arrays.  Every launch is
compared with the oracle's restatement on overlapping clusters (one per particle, radius neighbourhoods) in the
reference's slot and PtoC layouts."""
import ctypes as C

import numpy as np
import pytest

from fltk_sph_ice_amd import scene
from fltk_sph_ice_amd.device import DevBuf
from oracle import legacy_ice as LI

pytestmark = pytest.mark.gpu
TOL = 2e-6  # FP32 storage of positions near 1; both sides compute in double


def _setup(lib, dims=(7, 6, 5), radius_spacings=1.8, lo=(-0.3, 0.09, -0.3), hi=(1.0, 1.0, 1.0)):
    sc = scene.ice_block(*dims)
    pos, vel = sc.particles()
    rest = pos.copy()
    pos[:, :3] -= np.float32(sc.origin) - np.float32([0.0, 0.05, 0.0])  # block near the floor at y = lo[1]
    rest[:, :3] = pos[:, :3]
    rng = np.random.default_rng(3)
    pos[:, :3] += rng.normal(0, 0.004, pos[:, :3].shape).astype(np.float32)
    vel[:, :3] = rng.normal(0, 0.6, vel[:, :3].shape).astype(np.float32)
    d = np.linalg.norm(rest[:, None, :3].astype(np.float64) - rest[None, :, :3], axis=-1)
    cs, mem = [0], []
    for i in range(len(rest)):
        m = np.nonzero(d[i] <= radius_spacings * sc.spacing)[0]
        mem += m.tolist()
        cs.append(len(mem))
    cs, mem = np.array(cs, np.int32), np.array(mem, np.int32)
    L = LI.layout(rest, cs, mem)
    sp = scene.wcsph_dam_break(4, 4, 4).params()
    for a in range(3):
        sp.BoundaryMin[a], sp.BoundaryMax[a] = lo[a], hi[a]
        sp.Gravity[a] = (0.0, -9.8, 0.0)[a]
    lib.CuSetParameters(C.byref(sp))
    S = LI.sm_params(0.01, lo=lo, hi=hi)
    return pos, vel, L, S


def test_step_with_iterations_matches_oracle(lib):
    pos, vel, L, S = _setup(lib)
    n, ns, nc = L["n"], L["nslots"], L["nc"]
    assert L["ptoc_max"] > 10 and ns > 10 * n  # overlapping clusters of more than 10 members
    dt = np.float32(0.01)
    z3 = lambda k: np.zeros(3 * k, np.float32)
    # device state of the reference's ice object
    g = {k: DevBuf.from_array(a) for k, a in dict(pos=pos, vel=vel, org=L["org"], cur=z3(ns), ocm=L["ocm"], ccm=z3(nc), v=z3(ns),
                                                  pidx=L["pidx"], indx=L["indx"], sld=z3(n), sldv=z3(n), ptocn=L["ptoc_indx"],
                                                  ptoc=L["ptoc"], sph=np.zeros((n, 4), np.float32),
                                                  sphv=np.zeros((n, 4), np.float32)).items()}
    # oracle state
    o = dict(cur=z3(ns), ccm=z3(nc), v=z3(ns), sld=z3(n), sldv=z3(n))
    iters = 3
    for it in range(iters):
        if it == 0:
            lib.LaunchShapeMatchingGPU(nc, g["pos"].ptr, g["vel"].ptr, g["org"].ptr, g["cur"].ptr, g["ocm"].ptr, g["ccm"].ptr, g["v"].ptr,
                                       g["pidx"].ptr, g["indx"].ptr, dt)
            LI.cluster_pass(L, pos, vel, o["sld"], o["cur"], o["ccm"], o["v"], dt, S, itr=False)
        else:
            lib.LaunchShapeMatchingIterationGPU(nc, g["pos"].ptr, g["vel"].ptr, g["sld"].ptr, g["sldv"].ptr, g["org"].ptr, g["cur"].ptr,
                                                g["ocm"].ptr, g["ccm"].ptr, g["v"].ptr, g["pidx"].ptr, g["indx"].ptr, dt)
            LI.cluster_pass(L, pos, vel, o["sld"], o["cur"], o["ccm"], o["v"], dt, S, itr=True)
        assert lib.CuGetLastErrorCode() == 0
        cur = g["cur"].download(np.float32)
        assert np.abs(cur - o["cur"]).max() <= TOL, (it, np.abs(cur - o["cur"]).max())
        np.testing.assert_allclose(g["ccm"].download(np.float32), o["ccm"], atol=TOL)
        np.testing.assert_allclose(g["v"].download(np.float32), o["v"], atol=TOL / float(dt))
        lib.LaunchCalcAverageGPU(n, g["sld"].ptr, g["sldv"].ptr, g["sph"].ptr, g["sphv"].ptr, g["cur"].ptr, g["v"].ptr, g["indx"].ptr,
                                 g["ptocn"].ptr, g["ptoc"].ptr, n, L["ptoc_max"], L["ptoc_param_size"])
        LI.average(L, o["sld"], o["sldv"], o["cur"], o["v"])
        assert lib.CuGetLastErrorCode() == 0
        np.testing.assert_allclose(g["sld"].download(np.float32), o["sld"], atol=TOL)
        np.testing.assert_allclose(g["sldv"].download(np.float32), o["sldv"], atol=TOL / float(dt))
    # the floor at y = 0.09 cuts through the lowest layer: the iteration passes clamp goals to it
    assert (o["cur"].reshape(-1, 3)[:, 1] <= 0.09 + 1e-7).any()
    # the solid result goes into the fluid arrays (blend 1)
    lib.LaunchInterPolationGPU(n, g["sld"].ptr, g["sldv"].ptr, g["sph"].ptr, g["sphv"].ptr)
    sph = g["sph"].download(np.float32).reshape(n, 4)
    sphv = g["sphv"].download(np.float32).reshape(n, 4)
    assert np.array_equal(sph[:, :3], g["sld"].download(np.float32).reshape(n, 3))
    assert np.array_equal(sphv[:, :3], g["sldv"].download(np.float32).reshape(n, 3))


def test_average_skips_empty_entries_and_particles(lib):
    pos, vel, L, S = _setup(lib, dims=(5, 4, 4))
    n = L["n"]
    rng = np.random.default_rng(9)
    ptoc = L["ptoc"].reshape(n, L["ptoc_param_size"], L["ptoc_max"]).copy()
    cnt = L["ptoc_indx"].copy()
    cnt[0] = 0  # particle 0 in no cluster
    ptoc[3, 0, 0] = -1  # an empty entry
    ptoc[4, 1, 1] = -1
    L2 = dict(L, ptoc=ptoc.ravel().copy(), ptoc_indx=cnt)
    smp = rng.normal(0, 1, 3 * L["nslots"]).astype(np.float32)
    smv = rng.normal(0, 1, 3 * L["nslots"]).astype(np.float32)
    sld0 = rng.normal(0, 1, 3 * n).astype(np.float32)
    g = {k: DevBuf.from_array(a) for k, a in dict(sld=sld0, sldv=-sld0, smp=smp, smv=smv, indx=L["indx"], cnt=cnt, ptoc=L2["ptoc"]).items()}
    lib.LaunchCalcAverageGPU(n, g["sld"].ptr, g["sldv"].ptr, None, None, g["smp"].ptr, g["smv"].ptr, g["indx"].ptr, g["cnt"].ptr, g["ptoc"].ptr,
                             n, L["ptoc_max"], L["ptoc_param_size"])
    assert lib.CuGetLastErrorCode() == 0
    rs, rv = sld0.copy(), -sld0
    LI.average(L2, rs, rv, smp, smv)
    np.testing.assert_allclose(g["sld"].download(np.float32), rs, atol=TOL)
    np.testing.assert_allclose(g["sldv"].download(np.float32), rv, atol=TOL)
    assert np.array_equal(g["sld"].download(np.float32)[:3], sld0[:3])

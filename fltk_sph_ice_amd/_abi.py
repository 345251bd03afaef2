"""ctypes view of the C ABI in include/*.h (librxcu.so).

The library is loaded from the package directory only; a missing library raises instead of falling back to
any other implementation.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "librxcu.so"

f3 = C.c_float * 3
u3 = C.c_uint * 3


class float3(C.Structure):
    """float3 passed by value."""

    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class uint3(C.Structure):
    """uint3 passed by value."""

    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint)]

MAX_POLY_NUM = MAX_BOX_NUM = MAX_SPHERE_NUM = 10


class rxSimParams(C.Structure):
    """rx_cu_common.cuh:142-237 (include/rx_cu_common.h)."""

    _fields_ = [
        ("Gravity", f3), ("GlobalDamping", C.c_float), ("ParticleRadius", C.c_float),
        ("BoundaryMax", f3), ("BoundaryMin", f3),
        ("GridSize", u3), ("NumCells", C.c_uint), ("WorldOrigin", f3), ("WorldMax", f3), ("CellWidth", f3),
        ("NumBodies", C.c_uint), ("MaxParticlesPerCell", C.c_uint),
        ("GridSizeB", u3), ("NumCellsB", C.c_uint), ("WorldOriginB", f3), ("WorldMaxB", f3), ("CellWidthB", f3),
        ("NumBodiesB", C.c_uint),
        ("EffectiveRadius", C.c_float), ("Mass", C.c_float), ("VorticityConfinement", C.c_float),
        ("Buoyancy", C.c_float), ("Density", C.c_float), ("Pressure", C.c_float),
        ("Tension", C.c_float), ("Viscosity", C.c_float), ("GasStiffness", C.c_float),
        ("Volume", C.c_float), ("KernelParticles", C.c_float), ("Restitution", C.c_float),
        ("Threshold", C.c_float), ("InitDensity", C.c_float), ("Dt", C.c_float),
        ("Wpoly6", C.c_float), ("GWpoly6", C.c_float), ("LWpoly6", C.c_float),
        ("Wspiky", C.c_float), ("GWspiky", C.c_float), ("LWspiky", C.c_float),
        ("Wvisc", C.c_float), ("GWvisc", C.c_float), ("LWvisc", C.c_float),
        ("Wd2", C.c_float), ("Wd3", C.c_float), ("GWd2", C.c_float), ("GWd3", C.c_float), ("Wd1", C.c_float),
        ("NoiseScl", C.c_float), ("NoiseEthr", C.c_float), ("NoiseMag", C.c_float),
        ("AP", C.c_int), ("AP_K", C.c_float), ("AP_N", C.c_float), ("AP_Q", C.c_float), ("AP_WQ", C.c_float),
        ("PolyNum", C.c_uint), ("PolyVel", f3 * MAX_POLY_NUM),
        ("BoxNum", C.c_uint), ("BoxCen", f3 * MAX_BOX_NUM), ("BoxExt", f3 * MAX_BOX_NUM),
        ("BoxRot", (f3 * 3) * MAX_BOX_NUM), ("BoxInvRot", (f3 * 3) * MAX_BOX_NUM), ("BoxFlg", C.c_uint * MAX_BOX_NUM),
        ("SphereNum", C.c_uint), ("SphereCen", f3 * MAX_SPHERE_NUM), ("SphereRad", C.c_float * MAX_SPHERE_NUM),
        ("SphereFlg", C.c_uint * MAX_SPHERE_NUM),
    ]


class rxParticleCell(C.Structure):
    """rx_cu_common.cuh:241-259; passed by value."""

    _fields_ = [
        ("dSortedPos", C.c_void_p), ("dSortedVel", C.c_void_p), ("dSortedIndex", C.c_void_p),
        ("dGridParticleHash", C.c_void_p), ("dCellStart", C.c_void_p), ("dCellEnd", C.c_void_p),
        ("uNumParticles", C.c_uint), ("uNumCells", C.c_uint), ("uNumArdGrid", C.c_uint),
        ("dSortedPolyIdx", C.c_void_p), ("dGridPolyHash", C.c_void_p), ("dPolyCellStart", C.c_void_p),
        ("dPolyCellEnd", C.c_void_p), ("uNumPolyHash", C.c_uint),
    ]


class rxIceEnv(C.Structure):
    """include/rx_ice_funcs.h."""

    _fields_ = [
        ("alpha", C.c_float), ("beta", C.c_float), ("dt", C.c_float), ("gravity", f3), ("box_min", f3),
        ("box_max", f3), ("iterations", C.c_int), ("volume_conservation", C.c_int), ("restitution", C.c_float),
        ("cluster_radius", C.c_float),
    ]


class rxSphEnv(C.Structure):
    """include/rx_sph_host.h."""

    _fields_ = [
        ("max_particles", C.c_int), ("boundary_cen", f3), ("boundary_ext", f3), ("h", C.c_float),
        ("rest_dens", C.c_float), ("mass", C.c_float), ("spacing", C.c_float), ("particle_radius", C.c_float),
        ("dt", C.c_float), ("viscosity", C.c_float), ("gas_k", C.c_float), ("gravity", f3),
        ("epsilon", C.c_float), ("eta", C.c_float), ("min_iter", C.c_int), ("max_iter", C.c_int),
        ("use_ap", C.c_int), ("ap_k", C.c_float), ("ap_n", C.c_float), ("ap_q", C.c_float),
        ("xsph_c", C.c_float), ("cell_width", C.c_float), ("expansion", C.c_float),
    ]


vp = C.c_void_p
fp = C.POINTER(C.c_float)
u32 = C.c_uint
i32 = C.c_int
flt = C.c_float

# name -> (restype, argtypes).  Every symbol declared in include/*.h is listed here.
SIGNATURES: dict[str, tuple] = {
    # rx_cu_funcs.h
    "CuInit": (None, [i32, vp]),
    "CuSetDevice": (None, [i32]),
    "CuDeviceProp": (None, []),
    "CuAllocateArray": (None, [C.POINTER(vp), i32]),
    "CuSetArrayValue": (None, [vp, i32, C.c_size_t]),
    "CuFreeArray": (None, [vp]),
    "CuCopyArrayD2D": (None, [vp, vp, i32]),
    "CuThreadSync": (None, []),
    "CuCopyArrayFromDevice": (None, [vp, vp, vp, i32]),
    "CuCopyArrayToDevice": (None, [vp, vp, i32, i32]),
    "CuRegisterGLBufferObject": (None, [u32, vp]),
    "CuUnregisterGLBufferObject": (None, [vp]),
    "CuMapGLBufferObject": (vp, [vp]),
    "CuUnmapGLBufferObject": (None, [vp]),
    "CuScan": (None, [vp, vp, u32]),
    "CuSort": (None, [vp, vp, u32]),
    "CuSortNoAlloc": (None, [vp, vp, u32]),
    "CuScanf": (None, [vp, vp, u32]),
    "CuSPHInit": (None, [i32]),
    "CuSPHClean": (None, []),
    "CuSetParameters": (None, [C.POINTER(rxSimParams)]),
    "CuClearData": (None, []),
    "CuCalcHash": (None, [vp, vp, vp, vp, i32]),
    "CuReorderDataAndFindCellStart": (None, [rxParticleCell, vp, vp]),
    "CuCalcHashB": (None, [vp, vp, vp, float3, float3, uint3, i32]),
    "CuReorderDataAndFindCellStartB": (None, [rxParticleCell, vp]),
    "CuSphDensity": (None, [vp, vp, rxParticleCell]),
    "CuSphForces": (None, [vp, vp, vp, rxParticleCell, flt]),
    "CuSphNormal": (None, [vp, vp, rxParticleCell]),
    "CuSphIntegrate": (None, [vp, vp, vp, vp, vp, flt, u32]),
    "CuSphCheckDelete": (None, [vp, vp, vp, fp, fp, fp, u32]),
    "CuSphCheckDeleteX": (None, [vp, vp, vp, flt, fp, u32]),
    "CuSphIntegrateWithPolygon": (None, [vp, vp, vp, vp, vp, vp, vp, i32, flt, rxParticleCell]),
    "CuPbdSphIntegrateWithPolygon": (None, [vp, vp, vp, vp, vp, vp, vp, vp, i32, flt, rxParticleCell]),
    "CuSphGridDensity": (None, [vp, rxParticleCell, i32, i32, i32, flt, flt, flt, flt, flt, flt]),
    "CuPbdSphGridDensity": (None, [vp, rxParticleCell, i32, i32, i32, flt, flt, flt, flt, flt, flt]),
    "CuIceMeshMake": (None, [vp, rxParticleCell, i32, i32, i32, flt, flt, flt, flt, flt, flt, vp]),
    "CuSphBoundaryVolume": (None, [vp, flt, rxParticleCell]),
    "CuSphBoundaryDensity": (None, [vp, vp, vp, vp, rxParticleCell, u32]),
    "CuSphBoundaryForces": (None, [vp, vp, vp, vp, vp, rxParticleCell, u32]),
    "CuPbdSphBoundaryDensity": (None, [vp, vp, vp, rxParticleCell, u32]),
    "CuPbdSphScalingFactorWithBoundary": (None, [vp, vp, vp, flt, rxParticleCell, vp, vp, rxParticleCell]),
    "CuPbdSphPositionCorrectionWithBoundary": (None, [vp, vp, vp, rxParticleCell, vp, vp, rxParticleCell]),
    "CuSphDetectNeighborParticles": (None, [vp, rxParticleCell, flt, i32]),
    "CuSphDetectSurfaceParticles": (None, [vp, vp, rxParticleCell, vp, flt]),
    "CuPbdSphDensity": (None, [vp, rxParticleCell]),
    "CuPbdSphExternalForces": (None, [vp, vp, rxParticleCell, flt]),
    "CuPbdSphScalingFactor": (None, [vp, vp, vp, flt, rxParticleCell]),
    "CuPbdSphPositionCorrection": (None, [vp, vp, vp, rxParticleCell]),
    "CuPbdSphCorrectPosition": (None, [vp, vp, u32]),
    "CuPbdSphUpdatePosition": (None, [vp, vp, vp, flt, u32]),
    "CuPbdSphUpdateVelocity": (None, [vp, vp, vp, flt, u32]),
    "CuXSphViscosity": (None, [vp, vp, vp, vp, flt, rxParticleCell]),
    "CuPbdSphCalDensityFluctuation": (flt, [vp, vp, vp, flt, u32]),
    "CuPbdSphIntegrate": (None, [vp, vp, vp, vp, vp, vp, flt, u32]),
    "CuPbdSphCheckDelete": (None, [vp, vp, vp, fp, fp, fp, u32]),
    "CuPbdSphCheckDeleteX": (None, [vp, vp, vp, flt, fp, u32]),
    "CuPbdSphNormal": (None, [vp, vp, rxParticleCell]),
    "CuGetLastErrorCode": (i32, []),
    "CuGetLastErrorString": (C.c_char_p, []),
    "CuClearLastError": (None, []),
    "CuGetStream": (C.c_void_p, []),
    # rx_cu_mesh.h (anisotropic kernels; marching cubes, whose C++ references are passed as pointers)
    "CuSphCalUpdatedPosition": (None, [vp, vp, flt, flt, rxParticleCell]),
    "CuSphCalCovarianceMatrix": (None, [vp, vp, flt, rxParticleCell]),
    "CuSphSVDecomposition": (None, [vp, vp, vp, vp, u32]),
    "CuSphCalTransformMatrix": (None, [vp, vp, vp, u32]),
    "CuSphGridDensityAniso": (None, [vp, vp, flt, rxParticleCell, i32, i32, i32, flt, flt, flt, flt, flt, flt]),
    "CuMCInit": (C.c_bool, []),
    "CuInitMCTable": (None, []),
    "CuCleanMCTable": (None, []),
    "CuMCCalTriNum": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp, uint3, u32, float3, flt, C.POINTER(u32), C.POINTER(u32),
                             C.POINTER(u32)]),
    "CuMCCalEdgeVrts": (None, [vp, vp, vp, vp, vp, C.POINTER(uint3), C.POINTER(u32), uint3, u32, float3, float3, flt,
                               C.POINTER(u32)]),
    "CuMCCalTri": (None, [vp, vp, vp, vp, vp, C.POINTER(uint3), C.POINTER(u32), uint3, u32, float3, flt, u32, u32, u32]),
    "CuMCCalNrm": (None, [vp, vp, vp, u32, u32]),
    "rxParticlesWrite": (i32, [C.c_char_p, vp, vp, i32]),
    "rxParticlesRead": (i32, [C.c_char_p, vp, vp, i32]),
    # entry points of rx_cu_funcs.cuh:138-253 exported as error-reporting stubs (include/rx_cu_unsupported.h)
    "CuSphES": (None, []),
    "CuAddWTurb3D": (None, []),
    "CuSetWaveletTile3D": (None, []),
    "CuSetWaveletTile3DB": (None, []),
    "CuSetWaveletTile3DBT": (None, []),
    "CuSubInit": (None, []),
    "CuSubCalEt": (None, []),
    "CuSubCalPos": (None, []),
    "CuSubUpdateChildAxis": (None, []),
    "CuSubInit2": (None, []),
    "CuSubAdd": (None, []),
    "CuSubUpdate": (None, []),
    "CuSubSetUnsortArray": (None, []),
    "CuSubCalcHash": (None, []),
    "CuSubReorderDataAndFindCellStart": (None, []),
    "CuSubSphGridDensity": (None, []),
    "CuSubCheckRatio": (None, []),
    "CuCheckNumValidHashData": (u32, []),
    "CuMCCreateMesh": (None, []),
    "CuInitFloatArray": (None, []),
    "CuInitRadArray": (None, []),
    "CuInitDepthMap": (None, []),
    "CuCreateDepthMap": (None, []),
    "CuDepthSmoothing": (None, []),
    "CuDetectSilhouetteEdgeVertex": (None, []),
    "CuCalNodeVertex": (None, []),
    "CuCalMesh": (None, []),
    "CuSilhouetteSmoothing": (None, []),
    "CuTransformBack": (None, []),
    "CuCalVertexNormal": (None, []),
    "CuSetSSMParameters": (None, []),
    "CuInitVPackf": (None, []),
    "CuCleanVPackf": (None, []),
    "CuInitVPacke": (None, []),
    "CuCleanVPacke": (None, []),
    "CuInitTable": (None, []),
    "CuCleanTable": (None, []),
    "CuPbdSphDensityErrorStats": (None, [vp, flt, u32, fp, fp]),
    # rx_sph_host.h
    "rxSphEnvDefault": (None, [C.POINTER(rxSphEnv)]),
    "rxSetupSimParams": (i32, [C.POINTER(rxSphEnv), C.POINTER(rxSimParams)]),
    "rxAbiLayout": (C.c_long, [i32]),
    "rxSceneBox": (i32, [vp, vp, i32, fp, i32, i32, i32, flt, flt, flt, u32]),
    "rxPBDSPH_GPU_Create": (vp, [C.POINTER(rxSphEnv)]),
    "rxPBDSPH_GPU_Destroy": (None, [vp]),
    "rxPBDSPH_GPU_SetParticles": (i32, [vp, vp, vp, i32]),
    "rxPBDSPH_GPU_Update": (i32, [vp, flt, i32]),
    "rxPBDSPH_GPU_GetArray": (i32, [vp, i32, vp, i32]),
    "rxPBDSPH_GPU_GetStats": (i32, [vp, vp, i32]),
    "rxPBDSPH_GPU_Sync": (i32, [vp]),
    "rxPBDSPH_GPU_SetOption": (i32, [vp, i32, flt]),
    "rxPBDSPH_GPU_GetProfile": (i32, [vp, vp, vp, i32]),
    "rxPBDSPH_GPU_SetNetwork": (i32, [vp, vp]),
    "rxPBDSPH_GPU_FeatureScales": (i32, [vp, fp, fp]),
    "rxPBDSPH_GPU_SetThermal": (i32, [vp, vp, vp, vp]),
    "rxCoupled_Create": (vp, [vp, vp, vp, vp, i32, i32, vp, vp]),
    "rxCoupled_Destroy": (None, [vp]),
    "rxCoupled_Update": (i32, [vp, flt, i32]),
    "rxCoupled_Sync": (i32, [vp]),
    "rxCoupled_GetArray": (i32, [vp, i32, vp, i32]),
    "rxCoupled_GetStats": (i32, [vp, vp, i32]),
    "rxCoupled_SetOption": (i32, [vp, i32, flt]),
    "rxCoupled_SetIceTiles": (i32, [vp, i32, vp]),
    "rxCoupled_GetProfile": (i32, [vp, vp, vp, i32]),
    "rxCoupled_SetNetwork": (i32, [vp, vp]),
    "rxSPH_GPU_Create": (vp, [C.POINTER(rxSphEnv)]),
    "rxSPH_GPU_Destroy": (None, [vp]),
    "rxSPH_GPU_SetParticles": (i32, [vp, vp, vp, i32]),
    "rxSPH_GPU_Update": (i32, [vp, flt, i32]),
    "rxSPH_GPU_GetArray": (i32, [vp, i32, vp, i32]),
    "rxSPH_GPU_Sync": (i32, [vp]),
    "rxSPH_GPU_UseGraph": (i32, [vp, i32]),
    # rx_ice_funcs.h
    "rxIceEnvDefault": (None, [C.POINTER(rxIceEnv)]),
    "rxIceSM_Create": (vp, [C.POINTER(rxIceEnv), vp, i32]),
    "rxIceSM_Destroy": (None, [vp]),
    "rxIceSM_SetState": (i32, [vp, vp, vp]),
    "rxIceSM_GetState": (i32, [vp, vp, vp]),
    "rxIceSM_Update": (i32, [vp, i32]),
    "rxIceSM_Sync": (i32, [vp]),
    "rxIceSM_ClusterInfo": (i32, [vp, C.POINTER(i32), C.POINTER(C.c_long), vp, i32]),
    "rxIceSM_GetClusters": (i32, [vp, vp, vp]),
    "rxIceSM_Autotune": (i32, [vp, vp]),
    "rxIceSM_SetTiles": (i32, [vp, vp]),
    "rxIceSM_UseGraph": (i32, [vp, i32]),
    "rxIceSM_KernelTimes": (i32, [vp, fp, fp]),
    "rxIceSM_SetActive": (i32, [vp, vp]),
    # rx_mlp.h
    "rxMLP_NumParams": (i32, []),
    "rxMLP_Create": (vp, [u32]),
    "rxMLP_Destroy": (None, [vp]),
    "rxMLP_SetParams": (i32, [vp, vp]),
    "rxMLP_GetParams": (i32, [vp, vp]),
    "rxMLP_SetNormalization": (i32, [vp, vp, vp, flt, flt]),
    "rxMLP_Forward": (i32, [vp, vp, i32, vp]),
    "rxMLP_Gradients": (i32, [vp, vp, vp, i32, vp, fp]),
    "rxMLP_Train": (i32, [vp, vp, vp, i32, i32, i32, flt, u32, vp]),
}

# C++-linkage entry points (Ice_SM.h / IceObject.h signatures), resolved by their mangled names.
CXX_SYMBOLS = {
    "LaunchShapeMatchingGPU": ("_Z22LaunchShapeMatchingGPUiPfS_S_S_S_S_S_PiS0_f",
                               (None, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, flt])),
    "LaunchShapeMatchingIterationGPU": ("_Z31LaunchShapeMatchingIterationGPUiPfS_S_S_S_S_S_S_S_PiS0_f",
                                        (None, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, flt])),
    "LaunchCalcAverageGPU": ("_Z20LaunchCalcAverageGPUiPfS_S_S_S_S_PiS0_S0_iii",
                             (None, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32])),
    "LaunchInterPolationGPU": ("_Z22LaunchInterPolationGPUiPfS_S_S_", (None, [i32, vp, vp, vp, vp])),
    "LaunchHeatTransferGPU": ("_Z21LaunchHeatTransferGPUPfS_S_PKiS1_PKfPiS4_S1_fffffi",
                              (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp, flt, flt, flt, flt, flt, i32])),
    # the prefix-sum ("path") shape matching: linked with the reference's signatures, records error 5
    "LaunchShapeMatchingUsePathGPU": ("_Z29LaunchShapeMatchingUsePathGPUiPfS_S_S_S_S_S_S_PiS0_f",
                                      (None, [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, flt])),
    "LaunchUpdatePrefixSumGPU": ("_Z24LaunchUpdatePrefixSumGPUiiiiiPiPsPfS1_S0_S1_S1_PjS2_PKfS4_",
                                 (None, [i32] * 5 + [vp] * 11)),
    "LauchUpdateSMFromPath": ("_Z21LauchUpdateSMFromPathiPfS_S_S_S_S_S_S_PiS0_PsS1_S_S_S0_S0_iif",
                              (None, [i32] + [vp] * 17 + [i32, i32, flt])),
}

(RX_ARRAY_POSITION, RX_ARRAY_VELOCITY, RX_ARRAY_DENSITY, RX_ARRAY_LAMBDA_SUM, RX_ARRAY_PRESSURE, RX_ARRAY_FEATURES,
 RX_ARRAY_TEMPERATURE, RX_ARRAY_HEAT, RX_ARRAY_PHASE) = range(9)
RX_PBF_MODE_PLAIN, RX_PBF_MODE_SCALED, RX_PBF_MODE_LEARNED = 0, 1, 2
(RX_PBF_OPT_SYNC_EACH_STEP, RX_PBF_OPT_RECORD_LAMBDA, RX_PBF_OPT_TOL, RX_PBF_OPT_ITERATIONS, RX_PBF_OPT_SKIN, RX_PBF_OPT_PROFILE,
 RX_PBF_OPT_RECORD_FEATURES, RX_PBF_OPT_MODE, RX_PBF_OPT_ADAPTIVE, RX_PBF_OPT_WARM_SCALE,
 RX_PBF_OPT_JUDGE, RX_PBF_OPT_RESET_RUN, RX_PBF_OPT_PROFILE_MASK, RX_PBF_OPT_LIST_TILE, RX_PBF_OPT_HEAT_TD,
 RX_PBF_OPT_HEAT_DT, RX_PBF_OPT_MAX_NEIGHBORS, RX_PBF_OPT_GRAPH, RX_PBF_OPT_ADAPT_SKIN, RX_PBF_OPT_GRAD_MASS,
 RX_PBF_OPT_WARM_CLAMP, RX_PBF_OPT_FIRST_DMAX, RX_PBF_OPT_DEFER) = range(23)
PROF_CLASSES = ["predict", "grid", "list_lambda", "lambda", "dp", "velocity", "xsph", "features", "mlp", "heat", "deferred"]

_lib = None


def load() -> C.CDLL:
    """Loads librxcu.so from the package directory (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for name, (mangled, (res, args)) in CXX_SYMBOLS.items():
        fn = getattr(lib, mangled)
        fn.restype = res
        fn.argtypes = args
        setattr(lib, name, fn)
    _lib = lib
    return lib


def check(lib: C.CDLL | None = None) -> None:
    """Raises RuntimeError when the library recorded an error since the last check."""
    lib = lib or load()
    code = lib.CuGetLastErrorCode()
    if code:
        msg = lib.CuGetLastErrorString().decode(errors="replace")
        lib.CuClearLastError()
        raise RuntimeError(f"rx_cu error {code}: {msg}")

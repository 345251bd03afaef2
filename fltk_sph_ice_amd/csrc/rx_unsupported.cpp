// rx_unsupported.cpp -- the entry points of rx_cu_funcs.cuh:138-253 that this library does not implement
// (wavelet turbulence, sub-particles, the one-call marching-cubes mesh builder, screen-space meshing).  With C
// linkage the symbol carries no parameter types, so each stub is defined without parameters: a caller that
// passes the reference's arguments links and runs, the stub records error 5 and returns 0 / false.
// The three prefix-sum ("path") shape-matching launches of the ice code (Surf_SM.h:34-52, Ice_SM.h:55-69,
// IceObject.h:399-425) have C++ linkage: their stubs keep the reference's parameter lists, so the mangled names are
// the ones its callers bind.
#include "rx_common.h"
#include "rx_ice_funcs.h"

namespace {
void unsupported(const char* fn) {
    rx_set_error(5, "%s: not part of this library (turbulence, sub-particles and screen-space meshes are out of scope)", fn);
}
}  // namespace

extern "C" {
void CuSphES() { unsupported("CuSphES"); }
void CuAddWTurb3D() { unsupported("CuAddWTurb3D"); }
void CuSetWaveletTile3D() { unsupported("CuSetWaveletTile3D"); }
void CuSetWaveletTile3DB() { unsupported("CuSetWaveletTile3DB"); }
void CuSetWaveletTile3DBT() { unsupported("CuSetWaveletTile3DBT"); }
void CuSubInit() { unsupported("CuSubInit"); }
void CuSubCalEt() { unsupported("CuSubCalEt"); }
void CuSubCalPos() { unsupported("CuSubCalPos"); }
void CuSubUpdateChildAxis() { unsupported("CuSubUpdateChildAxis"); }
void CuSubInit2() { unsupported("CuSubInit2"); }
void CuSubAdd() { unsupported("CuSubAdd"); }
void CuSubUpdate() { unsupported("CuSubUpdate"); }
void CuSubSetUnsortArray() { unsupported("CuSubSetUnsortArray"); }
void CuSubCalcHash() { unsupported("CuSubCalcHash"); }
void CuSubReorderDataAndFindCellStart() { unsupported("CuSubReorderDataAndFindCellStart"); }
void CuSubSphGridDensity() { unsupported("CuSubSphGridDensity"); }
void CuSubCheckRatio() { unsupported("CuSubCheckRatio"); }
unsigned int CuCheckNumValidHashData() { unsupported("CuCheckNumValidHashData"); return 0u; }
void CuMCCreateMesh() { unsupported("CuMCCreateMesh"); }
void CuInitFloatArray() { unsupported("CuInitFloatArray"); }
void CuInitRadArray() { unsupported("CuInitRadArray"); }
void CuInitDepthMap() { unsupported("CuInitDepthMap"); }
void CuCreateDepthMap() { unsupported("CuCreateDepthMap"); }
void CuDepthSmoothing() { unsupported("CuDepthSmoothing"); }
void CuDetectSilhouetteEdgeVertex() { unsupported("CuDetectSilhouetteEdgeVertex"); }
void CuCalNodeVertex() { unsupported("CuCalNodeVertex"); }
void CuCalMesh() { unsupported("CuCalMesh"); }
void CuSilhouetteSmoothing() { unsupported("CuSilhouetteSmoothing"); }
void CuTransformBack() { unsupported("CuTransformBack"); }
void CuCalVertexNormal() { unsupported("CuCalVertexNormal"); }
void CuSetSSMParameters() { unsupported("CuSetSSMParameters"); }
void CuInitVPackf() { unsupported("CuInitVPackf"); }
void CuCleanVPackf() { unsupported("CuCleanVPackf"); }
void CuInitVPacke() { unsupported("CuInitVPacke"); }
void CuCleanVPacke() { unsupported("CuCleanVPacke"); }
void CuInitTable() { unsupported("CuInitTable"); }
void CuCleanTable() { unsupported("CuCleanTable"); }
}  // extern "C"

// Ice_SM.h:55-69 (LaunchShapeMatchingUsePathGPU, called by Ice_SM::UpdateUsePathGPU, Ice_SM.cpp:1458-1461).
void LaunchShapeMatchingUsePathGPU(int, float*, float*, float*, float*, float*, float*, float*, float*, int*, int*, float) {
    rx_set_error(5, "LaunchShapeMatchingUsePathGPU: the prefix-sum (path) shape matching is not part of this library; "
                    "use LaunchShapeMatchingGPU / LaunchShapeMatchingIterationGPU or the batched rxIceSM object");
}

// IceObject.h:399-425 (LauchUpdateSMFromPath, called by IceObject::StepObjMoveGPUUsePath, IceObject.cpp:955-998).
void LauchUpdateSMFromPath(int, float*, float*, float*, float*, float*, float*, float*, float*, int*, int*, short*, short*, float*,
                           float*, int*, int*, int, int, float) {
    rx_set_error(5, "LauchUpdateSMFromPath: the prefix-sum (path) shape matching is not part of this library; "
                    "use LaunchShapeMatchingGPU / LaunchShapeMatchingIterationGPU or the batched rxIceSM object");
}

// Surf_SM.h:34-52 (LaunchUpdatePrefixSumGPU, called by Surf_SM::UpdatePrefixSumGPU, Surf_SM.cpp:348,410).
void LaunchUpdatePrefixSumGPU(int, int, int, int, int, int*, short*, float*, float*, short*, float*, float*, unsigned*, unsigned*,
                              const float*, const float*) {
    rx_set_error(5, "LaunchUpdatePrefixSumGPU: the prefix-sum (path) shape matching is not part of this library; "
                    "use LaunchShapeMatchingGPU / LaunchShapeMatchingIterationGPU or the batched rxIceSM object");
}

/* rx_cu_unsupported.h -- the entry points of rx_cu_funcs.cuh:138-253 that this library does not implement:
 * wavelet turbulence, sub-particles, the one-call marching-cubes mesh builder and screen-space meshing
 * (the anisotropic kernels and the marching-cubes stages are in rx_cu_mesh.h).
 * They are exported so that host code written against the reference links unchanged; each one records error 5
 * (CuGetLastErrorCode) naming itself, prints it, leaves its outputs untouched and returns 0 / false.
 * The declarations keep the reference's signatures (C++: some take references); the types they name are only
 * declared here.  The marching-cubes mesh builder is declared in its RX_CUMC_USE_GEOMETRY form. */
#ifndef RX_CU_UNSUPPORTED_H_
#define RX_CU_UNSUPPORTED_H_

#include "rx_cu_common.h"

#ifdef __cplusplus
struct rxSubParticleCell;
struct rxVPackf;
struct rxVPacke;
struct rxSSGridG;
struct rxSsmParams;

extern "C" {
/* wavelet turbulence (:147-155) */
void CuSphES(float* dEt, float scale, float coef_et, float max_et, rxParticleCell cell, float3 cell_size);
void CuAddWTurb3D(float* dVturb, float* dFrc, float* dEt, float* dDens, int d,
                  int first, int nbands, float* dPos, float pmin[3], float pdim[3], uint numParticles, float dt);
void CuSetWaveletTile3D(float* tile, int n, int d);
void CuSetWaveletTile3DB(float* tile, int nx, int ny, int nz, int d);
void CuSetWaveletTile3DBT(float* tile, int nx, int ny, int nz, int d);
/* sub-particles (:158-180) */
void CuSubInit(float* dSubPos, float* dPos, float* dSubAxis, float* dSubChild, uint* dSubRand, float radius, uint maxSubLevel, uint nprts);
void CuSubCalEt(float* dSubEt, float* dEt, float et_coef, float scale, float radius, uint maxSubLevel, rxParticleCell cell);
void CuSubCalPos(float* dSubPos, float* dPos, float* dSubChild, float radius, uint maxSubLevel, uint nprts);
void CuSubUpdateChildAxis(float* dSubChild, float* dSubAxis, float* dSubEt, uint* dSubRand, float radius, float ratioAxisAngle,
                          uint maxSubLevel, rxParticleCell cell, float dt);
void CuSubInit2(float* dSubPos, float* dPos, float* dSubAxis, float* dSubChild, uint* dSubRand, float radius, uint maxSubLevel,
                uint nprts, uint uMaxParticles);
void CuSubAdd(float* dSubPos, float* dPos, float* dSubAxis, float* dSubChild, uint* dSubRand,
              float radius, uint maxSubLevel, uint nprts, uint uMaxParticles, uint uStart, uint uCount);
void CuSubUpdate(float* dPos, float* dEt, float* dSubPos, float* dSubChild, float* dSubAxis, float* dSubEt, uint* dSubRand,
                 float et_coef, float radius, float ratioAxisAngle, uint maxSubLevel, uint nprts, uint uMaxParticles, float scale, float dt);
void CuSubSetUnsortArray(float* dEt, float* dSubPos, rxSubParticleCell& subCell, float et_coef);
void CuSubCalcHash(rxSubParticleCell subCell);
void CuSubReorderDataAndFindCellStart(rxSubParticleCell subCell);
void CuSubSphGridDensity(float* dGridDl, rxSubParticleCell subCell,
                         int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy, float dz);
void CuSubCheckRatio(rxSubParticleCell subCell);
uint CuCheckNumValidHashData(rxSubParticleCell subCell);
/* marching cubes (:184-190): the GL-free mesh builder is declared there but defined nowhere in the reference
 * (rx_mc.cu:392-478 defines only the GL form); the table and per-stage entry points are in rx_cu_mesh.h */
void CuMCCreateMesh(float threshold, unsigned int& nvrts, unsigned int& ntris);
/* screen-space meshing (:215-250) */
void CuInitFloatArray(float* dArray, float val, int n);
void CuInitRadArray(float* drad, float* dval, int n);
void CuInitDepthMap(float* dDMap, int nx, int ny);
void CuCreateDepthMap(float* dDMap, float* dPrtPos, float* dPrtRad, int pnum, int pdim, float* tr, float* pmv,
                      int W, int H, float dw, float dh, int nx, int ny);
void CuDepthSmoothing(float* dDMap, int nx, int ny, int n_filter, float* binomials, float zmax);
void CuDetectSilhouetteEdgeVertex(float* dDMap, int nx, int ny, float dw, float dh, float zmax,
                                  float* dPrtPos, float* dPrtRad, int pnum, int pdim, int W, int H,
                                  rxVPacke& dEdge, rxVPackf& dEdgeVrts, rxVPackf& dBackEdgeVrts, rxSSGridG* dMGrid,
                                  int num_nv, int& num_edge, int& num_front_edge, int& num_back_edge);
void CuCalNodeVertex(float* dDMap, int nx, int ny, float dw, float dh, float zmax,
                     rxVPackf& dNodeVrts, rxSSGridG* dMGrid, int& num_node_vertex);
void CuCalMesh(rxSSGridG* dMGrid, int nx, int ny, float dw, float dh, float zmax, uint* dTriNum, uint* dTriNumScan,
               rxVPackf& dBack2Vrts, float* dVrts, int num_vrts, int& num_back2_vrts, uint*& dTriArray, int& num_mesh);
void CuSilhouetteSmoothing(float* dVrts, int num_vrts, int num_node_vrts, uint* dTriArray, int num_mesh, int n_iter);
void CuTransformBack(float* dSSVrts, int num_vrts, float* mvq, float* q, int W, int H, float* dVrts);
void CuCalVertexNormal(float* dVrts, int num_vrts, uint* dTriArray, int num_mesh, float* dNrms);
void CuSetSSMParameters(rxSsmParams* hostParams);
void CuInitVPackf(rxVPackf& pack, int size);
void CuCleanVPackf(rxVPackf& pack);
void CuInitVPacke(rxVPacke& pack, int size);
void CuCleanVPacke(rxVPacke& pack);
void CuInitTable(void);
void CuCleanTable(void);
}
#endif /* __cplusplus */

#endif /* RX_CU_UNSUPPORTED_H_ */

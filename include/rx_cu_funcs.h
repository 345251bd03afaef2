/*
 * rx_cu_funcs.h -- the extern "C" particle ABI, implemented by HIP kernels for gfx950.
 *
 * Drop-in replacement for the particle-path declarations of rx_cu_funcs.cuh:24-136.  Every function keeps the
 * reference's name, argument order and meaning; each comment cites the reference declaration it replaces.
 *
 * Semantics kept from the reference:
 *   - all pointers are device pointers allocated by the caller (CuAllocateArray) unless named "host";
 *   - particle arrays are float* reinterpreted as float4 (x, y, z, w), stride 4 floats;
 *   - every function returns after its work has completed on the device (synchronous on return);
 *   - errors are printed to stderr and execution continues.  In addition the last error is kept in a
 *     thread-local slot that CuGetLastErrorCode/CuGetLastErrorString return (additive entry points).
 *
 * Functions of rx_cu_funcs.cuh that belong to meshing, turbulence, sub-particles, anisotropic kernels and
 * screen-space meshing (rx_cu_funcs.cuh:138-250) are not part of this library.
 */
#ifndef RX_CU_FUNCS_H_
#define RX_CU_FUNCS_H_

#include <stddef.h>
#include "rx_cu_common.h"

/* Opaque handle type of the reference's GL-interop signatures (rx_cu_funcs.cuh:38-43).  This library has no
 * GL interop: the handle is only ever passed as a null pointer. */
typedef struct rxGLResource cudaGraphicsResource;

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- device, memory (rx_cu_funcs.cuh:26-43) */
void CuInit(int argc, char** argv);                          /* rx_cu_funcs.cuh:26 ("-device=N" picks the GPU) */
void CuSetDevice(int id);                                    /* rx_cu_funcs.cuh:27 */
void CuDeviceProp(void);                                     /* rx_cu_funcs.cuh:29 */
void CuAllocateArray(void** devPtr, int size);               /* rx_cu_funcs.cuh:31 (declared int size) */
void CuSetArrayValue(void* devPtr, int val, size_t size);    /* rx_cu_funcs.cuh:32 */
void CuFreeArray(void* devPtr);                              /* rx_cu_funcs.cuh:33 */
void CuCopyArrayD2D(void* dDst, void* dSrc, int size);       /* rx_cu_funcs.cuh:35 */
void CuThreadSync(void);                                     /* rx_cu_funcs.cuh:36 */
void CuCopyArrayFromDevice(void* host, const void* device, cudaGraphicsResource** resource, int size); /* :38 */
void CuCopyArrayToDevice(void* device, const void* host, int offset, int size);                        /* :39 */
void CuRegisterGLBufferObject(unsigned int vbo, cudaGraphicsResource** resource);  /* :40, reports an error */
void CuUnregisterGLBufferObject(cudaGraphicsResource* resource);                   /* :41, reports an error */
void* CuMapGLBufferObject(cudaGraphicsResource** resource);                        /* :42, returns NULL     */
void CuUnmapGLBufferObject(cudaGraphicsResource* resource);                        /* :43, reports an error */

/* ---------------------------------------------------------------- scan and sort (rx_cu_funcs.cuh:45-48) */
/* Exclusive prefix sum of numElements uints (rx_cu_funcs.cuh:45). */
void CuScan(unsigned int* output, unsigned int* input, unsigned int numElements);
/* Ascending key/value sort of (dHash, dIndex), in place, STABLE: equal keys keep their input order
 * (rx_cu_funcs.cuh:46).  Keys must be cell hashes below max(NumCells, NumCellsB) of the last CuSetParameters
 * (fluid or boundary-particle grid), or the 0x7FFFFFFF marker that CuCalcHash writes for attr == -1
 * particles; any other key sets an error. */
void CuSort(unsigned int* dHash, uint* dIndex, uint numParticles);
/* No counterpart in the reference: CuSort for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same sort and result as CuSort, with its scratch in a
 * static device array of about 24 MB per GPU (held until the process ends) instead of library-allocated memory;
 * at most 2^21 (2,097,152) keys per call, larger counts are rejected with error 1 before any work.  Every call on
 * a GPU shares that GPU's array, so calls that overlap on one GPU race.  Keys that are neither cell hashes nor
 * 0x7FFFFFFF sort as 0x7FFFFFFF as in CuSort, but are not reported (reporting would need a host copy). */
void CuSortNoAlloc(unsigned int* dHash, uint* dIndex, uint numParticles);
/* Exclusive prefix sum of numElements floats (rx_cu_funcs.cuh:48). */
void CuScanf(float* output, float* input, unsigned int numElements);

/* ---------------------------------------------------------------- SPH setup (rx_cu_funcs.cuh:54-58) */
void CuSPHInit(int max_particles);                           /* rx_cu_funcs.cuh:54 (reserves sort scratch) */
void CuSPHClean(void);                                       /* rx_cu_funcs.cuh:55 */
void CuSetParameters(rxSimParams* hostParams);               /* rx_cu_funcs.cuh:57 (copied to __constant__) */
void CuClearData(void);                                      /* rx_cu_funcs.cuh:58 (frees library scratch)  */

/* ---------------------------------------------------------------- neighbour grid (rx_cu_funcs.cuh:61-66) */
/* hash[i] = cell of pos[i] (clamped to the grid), index[i] = i; attr[i] == -1 gives hash 0x7FFFFFFF.
 * Cell coordinate = floor((p - WorldOrigin) / CellWidth) in IEEE FP32 (rx_cu_funcs.cuh:61). */
void CuCalcHash(uint* gridParticleHash, uint* gridParticleIndex, float* pos, int* attr, int numParticles);
/* From the sorted hashes: dCellStart/dCellEnd (empty cells hold 0xFFFFFFFF in both) and the gather of
 * oldPos/oldVel into dSortedPos/dSortedVel (rx_cu_funcs.cuh:62). */
void CuReorderDataAndFindCellStart(rxParticleCell cell, float* oldPos, float* oldVel);
/* Boundary-particle grid variants with explicit grid geometry (rx_cu_funcs.cuh:64-66). */
void CuCalcHashB(uint* dGridParticleHash, uint* dSortedIndex, float* dPos,
                 float3 world_origin, float3 cell_width, uint3 grid_size, int nprts);
void CuReorderDataAndFindCellStartB(rxParticleCell cell, float* oldPos);

/* ---------------------------------------------------------------- WCSPH (rx_cu_funcs.cuh:70-77) */
/* rho_i = sum_{r<=h} m W_poly6, p_i = k (rho_i - rho0); outputs in original particle order (:70). */
void CuSphDensity(float* dDens, float* dPres, rxParticleCell cell);
/* Pressure (symmetric p/rho^2, spiky gradient) + viscosity (Laplacian, / rho_j) + gravity, as an
 * acceleration in dFrc (float4, original order) (:71). */
void CuSphForces(float* dDens, float* dPres, float* dFrc, rxParticleCell cell, float dt);
/* Unit surface normal from the poly6 gradient (:72). */
void CuSphNormal(float* dNrms, float* dDens, rxParticleCell cell);
/* v += dt a; x += dt v; walls keep x at least ParticleRadius inside [BoundaryMin, BoundaryMax]; attr == -1
 * particles are skipped (:73). */
void CuSphIntegrate(float* pos, float* vel, float* frc, float* dens, int* attr, float dt, uint numParticles);
/* CuSphIntegrate with polygon obstacles (:74-75): vrts = float3 per vertex, tris = 3 vertex indices per triangle,
 * cell.dPolyCellStart/End/dSortedPolyIdx = the triangles listed per cell of the fluid grid (start 0xFFFFFFFF when
 * empty).  After the Euler step the segment old -> new position is tested against the triangles of the old
 * position's cell, then of the new position's cell when it differs (rx_sph_cpu.cpp:1040-1077); a triangle is hit
 * only from the side its normal (v1 - v0) x (v2 - v0) faces.  On a hit the position returns to the crossing
 * point plus vr dt d / l and v += vr, vr = -(1 + e') n (n.v), e' = Restitution d / (dt |v|), d = penetration,
 * l = step length.  Then the walls as in CuSphIntegrate.  Particles: cell.uNumParticles, original order, in place. */
void CuSphIntegrateWithPolygon(float* pos, float* vel, float* frc, float* dens, int* attr,
                               float* vrts, int* tris, int tri_num, float dt, rxParticleCell cell);
/* Particles strictly inside the deletion region (minp, maxp) -- or beyond x > xmax for the X variant -- get
 * attr = -1 and are parked at far_point with zero velocity; attr == -1 particles are skipped (:77-78). */
void CuSphCheckDelete(float* pos, float* vel, int* attr, float minp[3], float maxp[3], float far_point[3], uint nprts);
void CuSphCheckDeleteX(float* pos, float* vel, int* attr, float xmax, float farpoint[3], uint nprts);

/* Boundary particles (rx_cu_funcs.cuh:81-83): a second sorted grid (bcell) over rxSimParams::WorldOriginB /
 * CellWidthB / GridSizeB, built with CuCalcHashB + CuSort + CuReorderDataAndFindCellStartB.  Each reports an
 * error and does nothing when GridSizeB does not match bcell.uNumCells.
 * V_b = Mass / sum_{b'} Mass W_poly6(r_bb') (self included), written in original boundary order; the mass
 * argument is not read (rxSimParams::Mass is used, as in rx_sph.cu:501-512) (:81). */
void CuSphBoundaryVolume(float* dVolB, float mass, rxParticleCell cell);
/* rho_i += sum_b rho0 V_b W_poly6(r_ib) and p_i = k (rho_i - rho0); dPos, dDens, dPres in original fluid order (:82). */
void CuSphBoundaryDensity(float* dDens, float* dPres, float* dPos, float* dVolB, rxParticleCell bcell, uint pnum);
/* dFrc_i += sum_b -rho0 V_b (p_i / rho_i^2) grad W_spiky(r_ib) for 1e-4 < r <= h (an acceleration, as the GPU
 * kernel computes it) (:83). */
void CuSphBoundaryForces(float* dDens, float* dPres, float* dPos, float* dVolB, float* dFrc, rxParticleCell bcell, uint pnum);

/* Density on a lattice, as CuPbdSphGridDensity (:86).  CuIceMeshMake leaves out particles whose bIceFlag
 * (original order) is negative (:90-91). */
void CuSphGridDensity(float* dGridD, rxParticleCell cell,
                      int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy, float dz);
void CuIceMeshMake(float* dGridD, rxParticleCell cell,
                   int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy, float dz, float* bIceFlag);

/* Neighbour list of the ice/heat path (rx_cu_funcs.cuh:98): neights[i*30 + k] = neighbour k of particle i
 * (original index, r <= h as in the reference; radius is unused), unused slots -1, slot 29 = count. */
void CuSphDetectNeighborParticles(int* neights, rxParticleCell celldata, float radius, int prtNum);
/* Surface flags from the 30-slot list (rx_cu_funcs.cuh:95). */
void CuSphDetectSurfaceParticles(int* neights, int* surface, rxParticleCell celldata, float* pos, float radius);

/* ---------------------------------------------------------------- PBF (rx_cu_funcs.cuh:103-124) */
void CuPbdSphDensity(float* dDens, rxParticleCell cell);                                       /* :103 */
/* Viscosity + gravity acceleration (dFrc, float4, original order); viscosity divided by rho_i (:104). */
void CuPbdSphExternalForces(float* dDens, float* dFrc, rxParticleCell cell, float dt);
/* lambda_i = -C_i / (sum_k |grad_k C_i|^2 + eps), C_i = rho_i/rho0 - 1, self term included; also writes
 * rho_i.  dPos is not read: positions come from cell.dSortedPos (:106). */
void CuPbdSphScalingFactor(float* dPos, float* dDens, float* dScl, float eps, rxParticleCell cell);
/* dp_i = (1/rho0) sum_j (lambda_i + lambda_j + s_corr) grad W_spiky (:107). */
void CuPbdSphPositionCorrection(float* dPos, float* dScl, float* dDp, rxParticleCell cell);
void CuPbdSphCorrectPosition(float* dPos, float* dDp, uint nprts);                             /* :109 */
void CuPbdSphUpdatePosition(float* pos, float* new_pos, float* new_vel, float dt, uint nprts); /* :110 */
void CuPbdSphUpdateVelocity(float* pos, float* new_pos, float* new_vel, float dt, uint nprts); /* :111 */
/* v_i += c sum_j (m/rho_j)(v_j - v_i) W_poly6; dDens in original order, result in dNewVel (:113). */
void CuXSphViscosity(float* dPos, float* dVel, float* dNewVel, float* dDens, float c, rxParticleCell cell);
/* Mean of max(rho - rho0, 0)/rho0 over the particles; dErr receives the per-particle values (:115). */
float CuPbdSphCalDensityFluctuation(float* dErrScan, float* dErr, float* dDens, float rest_dens, uint nprts);
/* new = integrate(pos, vel, acc) with wall handling; pos/new_pos may alias (:117). */
void CuPbdSphIntegrate(float* pos, float* vel, float* acc, int* attr,
                       float* new_pos, float* new_vel, float dt, uint numParticles);
/* CuPbdSphIntegrate with the polygon obstacles of CuSphIntegrateWithPolygon; results in new_pos/new_vel, which
 * may alias pos/vel (:118-120). */
void CuPbdSphIntegrateWithPolygon(float* pos, float* vel, float* acc, int* attr,
                                  float* new_pos, float* new_vel,
                                  float* vrts, int* tris, int tri_num, float dt, rxParticleCell cell);
void CuPbdSphCheckDelete(float* pos, float* vel, int* attr, float minp[3], float maxp[3], float far_point[3], uint nprts); /* :123 */
void CuPbdSphCheckDeleteX(float* pos, float* vel, int* attr, float xmax, float farpoint[3], uint nprts);                   /* :124 */
void CuPbdSphNormal(float* dNrms, float* dDens, rxParticleCell cell);                                                     /* :128 */
/* Density sum_j m W_poly6 at the nodes (x0 + i dx, y0 + j dy, z0 + k dz) of an nx x ny x nz lattice, written
 * to dGridD[i + j nx + k nx ny] (:126; CuSphGridDensity :86 is the same function). */
void CuPbdSphGridDensity(float* dGridD, rxParticleCell cell,
                         int nx, int ny, int nz, float x0, float y0, float z0, float dx, float dy, float dz);
/* rho_i += sum_b rho0 V_b W_poly6(r_ib); original fluid order (:131). */
void CuPbdSphBoundaryDensity(float* dDens, float* dPos, float* dVolB, rxParticleCell bcell, uint pnum);
/* Fluid: lambda_i and rho_i of CuPbdSphScalingFactor with the boundary terms added to rho_i and to the
 * denominator (sum_b |(rho0 V_b / Mass) grad W_spiky / rho0|^2).  Boundary: lambda_b from the fluid and boundary
 * particles within h of b, written to dSclB in original boundary order (:132-133). */
void CuPbdSphScalingFactorWithBoundary(float* dPos, float* dDens, float* dScl, float eps, rxParticleCell cell,
                                       float* dVolB, float* dSclB, rxParticleCell bcell);
/* dp_i of CuPbdSphPositionCorrection plus sum_b (lambda_i + lambda_b + s_corr) grad W_spiky / rho0, s_corr from
 * (rho0 V_b / Mass) W_poly6 (:134-135). */
void CuPbdSphPositionCorrectionWithBoundary(float* dPos, float* dScl, float* dDp, rxParticleCell cell,
                                            float* dVolB, float* dSclB, rxParticleCell bcell);

/* ================================================================ additive entry points (no reference counterpart) */
/* Last error of this thread: 0 = none.  Cleared by CuClearLastError. */
int CuGetLastErrorCode(void);
const char* CuGetLastErrorString(void);
void CuClearLastError(void);
/* The HIP stream (hipStream_t, non-blocking) on which every entry point of this header runs its work, so that
 * a caller can order or time its own work against it (events, stream waits). */
void* CuGetStream(void);
/* Max over particles of max(rho - rho0, 0)/rho0 and mean of |rho - rho0|/rho0, reduced on the device. */
void CuPbdSphDensityErrorStats(float* dDens, float rest_dens, uint nprts, float* host_max_pos, float* host_mean_abs);

#ifdef __cplusplus
}
#endif

#endif /* RX_CU_FUNCS_H_ */

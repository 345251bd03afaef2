/*
 * rx_ice_funcs.h -- ice shape matching on the GPU.
 *
 * 1. The batched shape-matching object (no reference counterpart as a single object): every cluster of an
 *    ice body in one persistent launch per shape-matching iteration, clusters binned by size with tile
 *    shapes chosen by a built-in autotuner, the substep replayed from a hipGraph.  Semantics per step:
 *      prediction and wall response of OrientedPrtObj::UpdateCluster_SM (ElasticObject_OP.cpp:556-609),
 *      rxShapeMatching::shapeMatching (ShapeMatching.cpp:492-587; alpha, beta, volume conservation, unit
 *      masses, polar decomposition by Jacobi on A^T A, ShapeMatching.h:242-387),
 *      per-particle averaging of ConvoNormal::StepConvolution (Ice_Convolution_Normal.cpp:32-67),
 *      repeated `iterations` times as Ice_CalcMethod_Iteration::StepObjMove (Ice_CalcMethod_Iteration.cpp:32-48),
 *      velocity of OrientedParticle::UpdatePosAndVel (OrientedParticle.cpp:377-393).
 *    Clusters: one per particle, {p} plus every particle within `cluster_radius` of p in the rest state
 *    (IceObject::SetClusterMoveInfoFromNeight, IceObject.cpp:608-633).
 *
 * 2. The C++-linkage launch functions of Ice_SM.h:23-69, IceObject.h:370-425, HeatTransfer.cu:19-36 and
 *    Surf_SM.h:34-52, same names and arguments (the three prefix-sum "path" launches record an error and do
 *    nothing).
 *
 * Entry points returning int give 0 on success; CuGetLastErrorString (rx_cu_funcs.h) holds the message.
 */
#ifndef RX_ICE_FUNCS_H_
#define RX_ICE_FUNCS_H_

#include "rx_cu_common.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rxIceEnv {
    float alpha;            /* stiffness, goal blend per iteration                                   */
    float beta;             /* linear-deformation share of the goal transform                        */
    float dt;
    float gravity[3];
    float box_min[3];       /* simulation space of the clusters (rxShapeMatching::SetSimulationSpace) */
    float box_max[3];
    int iterations;         /* shape-matching iterations per step                                    */
    int volume_conservation;
    float restitution;      /* wall response factor (0.9 in the reference)                           */
    float cluster_radius;   /* neighbour radius that forms the clusters                              */
} rxIceEnv;

void rxIceEnvDefault(rxIceEnv* env);

/* Creates the object for n particles at rest positions rest_pos4 (float4) and builds the clusters. */
void* rxIceSM_Create(const rxIceEnv* env, const float* rest_pos4, int n);
void rxIceSM_Destroy(void* sm);
/* Current state (original order, float4). */
int rxIceSM_SetState(void* sm, const float* pos4, const float* vel4);
int rxIceSM_GetState(void* sm, float* pos4, float* vel4);
/* `steps` steps, asynchronous on the object's stream. */
int rxIceSM_Update(void* sm, int steps);
int rxIceSM_Sync(void* sm);
/* Number of clusters, members, and a size histogram (hist[k] = clusters with k members, k < nhist). */
int rxIceSM_ClusterInfo(void* sm, int* nclusters, long* nmembers, int* hist, int nhist);
/* Cluster CSR (start[nclusters+1], members) copied to the host. */
int rxIceSM_GetClusters(void* sm, int* start, int* members);
/* Times the candidate lanes-per-cluster tiles of every size bin and keeps the fastest; out_tiles receives
 * the tile chosen for each of the 5 bins (<=8, <=16, <=32, <=64, >64 members). */
int rxIceSM_Autotune(void* sm, int* out_tiles);
int rxIceSM_SetTiles(void* sm, const int* tiles);
int rxIceSM_UseGraph(void* sm, int on);
/* Melt mask (n ints, 1 = ice, host memory; NULL removes it): only clusters whose own particle is ice take
 * part, over their ice members (rest centre of mass and Aqq of those members; no linear term below 4 of
 * them); particles that are not ice keep their prediction.  No reference counterpart as a mask: it stands in
 * for the cluster reconstruction of IceObject::StepMelting (IceObject.cpp:1177-1300). */
int rxIceSM_SetActive(void* sm, const int* active);
/* Average milliseconds of the cluster kernel and of the averaging kernel over the last Update (HIP events). */
int rxIceSM_KernelTimes(void* sm, float* cluster_ms, float* average_ms);

#ifdef __cplusplus
}

/* Ice_SM.h:23-36 (LaunchShapeMatchingGPU): one rigid shape-matching pass per cluster (alpha = 1, beta = 0)
 * as OrientedPrtObj::UpdateCluster_SM (ElasticObject_OP.cpp:551-720).  Layout of the reference: prtPos/prtVel
 * float4 per particle; member slots i of cluster c are indxSet[2c] .. indxSet[2c+1] (inclusive);
 * pIndxes[i] = particle of slot i; orgPos/curPos/vel 3 floats per slot; orgCm/curCm 3 floats per cluster.
 * Gravity and walls come from the last CuSetParameters (Gravity, BoundaryMin/Max). */
void LaunchShapeMatchingGPU(int prtNum, float* prtPos, float* prtVel, float* orgPos, float* curPos, float* orgCm, float* curCm,
                            float* vel, int* pIndxes, int* d_IndxSet, float dt);
/* Ice_SM.h:38-53 (LaunchShapeMatchingIterationGPU): one iteration pass per cluster as
 * OrientedPrtObj::UpdateCluster_SM_Itr (ElasticObject_OP.cpp:722-887): the members take the averaged solid
 * positions sldPos (3 floats per particle, from LaunchCalcAverageGPU), the cluster is matched rigidly (Apq's third
 * column negated when det Apq < 0 and the cluster has at least 10 members), and the goal positions, clamped to the
 * walls of the last CuSetParameters, are written to curPos; curCm receives the cluster's centre of mass.  vel is
 * not written (the reference's iteration pass leaves it); prtPos, prtVel, sldVel and dt are not read.  Same slot
 * layout as LaunchShapeMatchingGPU; prtNum is the number of clusters. */
void LaunchShapeMatchingIterationGPU(int prtNum, float* prtPos, float* prtVel, float* sldPos, float* sldVel, float* orgPos, float* curPos,
                                     float* orgCm, float* curCm, float* vel, int* pIndxes, int* d_IndxSet, float dt);
/* IceObject.h:370-386 (LaunchCalcAverageGPU): per particle p < prtNum, the mean position and velocity of its
 * cluster members, ConvoNormal::StepConvolution (Ice_Convolution_Normal.cpp:32-67) in the reference's PtoC layout
 * (IceStructure.cu:63-126): entry j < min(PtoCIndx[p], PtoCMax) names cluster PtoC[base + j] and member number
 * PtoC[base + PtoCMax + j], base = PtoCMax * PtoCParamSize * p, -1 = empty; the member's slot is
 * smIndxSet[2 * cluster] + member number.  smPrtPos/smPrtVel hold 3 floats per slot, sldPrtPos/sldPrtVel 3 per
 * particle.  A particle in no cluster keeps its solid values (the reference kernel wrote uninitialised sums);
 * sphPrtPos, sphPrtVel and PNumMax are not read (as in the reference kernel). */
void LaunchCalcAverageGPU(int prtNum, float* sldPrtPos, float* sldPrtVel, float* sphPrtPos, float* sphPrtVel, float* smPrtPos, float* smPrtVel,
                          int* smIndxSet, int* PtoCIndx, int* PtoC, int PNumMax, int PtoCMax, int PtoCParamSize);
/* IceObject.h:388-395 (LaunchInterPolationGPU): copies the solid positions/velocities into the fluid arrays
 * (blend coefficient 1, IceObject.cpp:876-897). sldPrtPos/Vel hold 3 floats per particle, sphPrtPos/Vel 4. */
void LaunchInterPolationGPU(int prtNum, float* sldPrtPos, float* sldPrtVel, float* sphPrtPos, float* sphPrtVel);
/* HeatTransfer.cu:19-36 / HeatTransfar.h:18-34 (LaunchHeatTransferGPU): one heat-exchange step of prtNum
 * particles, device arrays of prtNum entries: air exchange of surface particles (surfParticles[i] > 0 =
 * neighbour count), SPH diffusion over the 30-slot lists of CuSphDetectNeighborParticles (slot 29 = count;
 * the distance is taken as 1 as in the reference), +300 heat for the particles named in
 * meltPrtIndx[1..999] (a 1000-entry array, -1 = unused), then the latent/sensible phase update
 * (phase -2 ice, -1 melting, 1 freezing, 2 water; phaseChangef[i] set to 1 on a completed change). */
void LaunchHeatTransferGPU(float* heats, float* temps, float* dtemps, const int* surfParticles, const int* neightsParticles,
                           const float* densty, int* phase, int* phaseChangef, const int* meltPrtIndx, float radius, float airTemp,
                           float cffCnHt, float cffCnTd, float dt, int prtNum);
/* Surf_SM.h:34-52 (LaunchUpdatePrefixSumGPU), Ice_SM.h:55-69 (LaunchShapeMatchingUsePathGPU) and
 * IceObject.h:399-425 (LauchUpdateSMFromPath, the reference's spelling): the prefix-sum ("path") variant of shape
 * matching.  Not implemented: all three link with the reference's C++ signatures, do no work and record error 5
 * (CuGetLastErrorCode). */
void LaunchUpdatePrefixSumGPU(int prtNum, int PosSizeX, int PosSizeY, int ApqSizeX, int ApqSizeY, int* md_2DiPTHtoPRT,
                              short* md_2DiPRTtoPTH, float* md_2Df3PrfxPos, float* md_2Df9PrfxApq, short* md_3DiPTHandPrfxSet,
                              float* md_f3OrgPos, float* md_f3OrgCm, unsigned* dgroupPos, unsigned* dgroupApq,
                              const float* md_fPos, const float* md_fVel);
void LaunchShapeMatchingUsePathGPU(int prtNum, float* prtPos, float* prtVel, float* orgPos, float* curPos, float* orgCm,
                                   float* curCm, float* curApq, float* vel, int* pIndxes, int* d_IndxSet, float dt);
void LauchUpdateSMFromPath(int prtNum, float* prtPos, float* prtVel, float* orgPos, float* curPos, float* orgCm, float* curCm,
                           float* clstrApq, float* vel, int* pIndxes, int* startEndSet, short* PRTtoPTH, short* PTHandPrfxSet,
                           float* prfxPos, float* prfxApq, int* CtoP, int* CtoPNum, int CtoPSizeY, int CtoPSizeZ, float dt);
#endif

#endif /* RX_ICE_FUNCS_H_ */

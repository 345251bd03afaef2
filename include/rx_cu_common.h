/*
 * rx_cu_common.h -- parameter and grid types shared by the host and the HIP particle kernels.
 *
 * Byte-identical replacement for the structs of rx_cu_common.cuh:122-259 (matrix3x3, rxSimParams,
 * rxParticleCell).  Field order, types and the MAX_*_NUM array sizes are the reference's, so host code
 * written against rx_cu_common.cuh fills these structs unchanged.  float3/uint3 come from HIP's vector
 * types (12 bytes, 4-byte aligned, the same layout as the reference's vector_types.h).
 *
 * Host code that includes this header with a plain C++ compiler needs -D__HIP_PLATFORM_AMD__ and
 * -I/opt/rocm/include (hipcc defines both).
 */
#ifndef RX_CU_COMMON_H_
#define RX_CU_COMMON_H_

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

typedef unsigned int uint;
typedef unsigned char uchar;

/* Threads per block used by the reference launch helpers (rx_cu_common.cuh:62). */
#define THREAD_NUM 256

#define MAX_POLY_NUM 10
#define MAX_BOX_NUM 10
#define MAX_SPHERE_NUM 10

/* rx_cu_common.cuh:123-126 */
struct matrix3x3 {
    float3 e[3];
};

/* rx_cu_common.cuh:142-237.  Only the fields the particle path reads are documented here; the others are
 * kept for layout compatibility. */
struct rxSimParams {
    float3 Gravity;          /* gravity acceleration added to every particle                     */
    float GlobalDamping;
    float ParticleRadius;    /* wall margin: positions are kept ParticleRadius inside the box    */

    float3 BoundaryMax;      /* simulation box (walls)                                           */
    float3 BoundaryMin;

    uint3 GridSize;          /* cells per axis of the neighbour grid                            */
    uint NumCells;           /* GridSize.x*GridSize.y*GridSize.z                                  */
    float3 WorldOrigin;      /* grid origin                                                      */
    float3 WorldMax;
    float3 CellWidth;        /* cell edge per axis (the particle path uses cell = h)             */

    uint NumBodies;          /* number of particles                                              */
    uint MaxParticlesPerCell;

    uint3 GridSizeB;         /* boundary-particle grid                                           */
    uint NumCellsB;
    float3 WorldOriginB;
    float3 WorldMaxB;
    float3 CellWidthB;
    uint NumBodiesB;

    float EffectiveRadius;   /* kernel support h                                                 */
    float Mass;              /* particle mass                                                    */
    float VorticityConfinement;

    float Buoyancy;

    float Density;           /* rest density rho0                                                */
    float Pressure;

    float Tension;
    float Viscosity;         /* viscosity coefficient mu                                         */
    float GasStiffness;      /* WCSPH gas constant k (p = k (rho - rho0))                        */

    float Volume;
    float KernelParticles;
    float Restitution;

    float Threshold;

    float InitDensity;

    float Dt;                /* time step (used by the PBF s_corr term)                          */

    float Wpoly6;            /* 315/(64 pi h^9)                                                  */
    float GWpoly6;           /* -945/(32 pi h^9)                                                 */
    float LWpoly6;           /* -945/(32 pi h^9)                                                 */
    float Wspiky;            /* 15/(pi h^6)                                                      */
    float GWspiky;           /* -45/(pi h^6)                                                     */
    float LWspiky;           /* -90/(pi h^6)                                                     */
    float Wvisc;             /* 15/(2 pi h^3)                                                    */
    float GWvisc;            /* 15/(2 pi h^3)                                                    */
    float LWvisc;            /* 45/(pi h^6)                                                      */

    float Wd2;
    float Wd3;
    float GWd2;
    float GWd3;
    float Wd1;

    float NoiseScl;
    float NoiseEthr;
    float NoiseMag;

    int AP;                  /* PBF artificial pressure (s_corr) on/off                          */
    float AP_K;              /* s_corr k                                                         */
    float AP_N;              /* s_corr exponent n                                                */
    float AP_Q;              /* s_corr delta q / h                                               */
    float AP_WQ;             /* W_poly6(AP_Q * h)                                                */

    uint PolyNum;
    float3 PolyVel[MAX_POLY_NUM];

    uint BoxNum;
    float3 BoxCen[MAX_BOX_NUM];
    float3 BoxExt[MAX_BOX_NUM];
    struct matrix3x3 BoxRot[MAX_BOX_NUM];
    struct matrix3x3 BoxInvRot[MAX_BOX_NUM];
    uint BoxFlg[MAX_BOX_NUM];

    uint SphereNum;
    float3 SphereCen[MAX_SPHERE_NUM];
    float SphereRad[MAX_SPHERE_NUM];
    uint SphereFlg[MAX_SPHERE_NUM];
};

/* rx_cu_common.cuh:241-259.  Passed by value; all pointers are device pointers owned by the caller. */
struct rxParticleCell {
    float4* dSortedPos;       /* positions in sorted (cell) order                                  */
    float4* dSortedVel;       /* velocities in sorted order                                        */

    uint* dSortedIndex;       /* sorted slot -> original particle index                            */
    uint* dGridParticleHash;  /* cell hash per slot (sort key)                                     */
    uint* dCellStart;         /* first slot of each cell, 0xFFFFFFFF when the cell is empty        */
    uint* dCellEnd;           /* one past the last slot of each cell, 0xFFFFFFFF when empty        */
    uint uNumParticles;
    uint uNumCells;
    uint uNumArdGrid;

    uint* dSortedPolyIdx;
    uint* dGridPolyHash;
    uint* dPolyCellStart;
    uint* dPolyCellEnd;
    uint uNumPolyHash;
};

#ifdef __cplusplus
static_assert(sizeof(float3) == 12, "float3 must be 12 bytes");
static_assert(sizeof(struct matrix3x3) == 36, "matrix3x3 must be 36 bytes");
#endif

#endif /* RX_CU_COMMON_H_ */

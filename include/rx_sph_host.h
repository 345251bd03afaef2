/*
 * rx_sph_host.h -- host side of the particle path: scene parameters, the solver classes of rx_sph.h
 * (rxSPH_GPU, rxPBDSPH_GPU) and the batched shape-matching object, exported through a C ABI for callers that
 * cannot use the C++ classes (ctypes, other languages).
 *
 * All entry points return 0 on success and a non-zero error code otherwise (CuGetLastErrorString holds
 * the message).  Host arrays are float4 per particle (x, y, z, w) unless stated otherwise.
 */
#ifndef RX_SPH_HOST_H_
#define RX_SPH_HOST_H_

#include "rx_cu_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Scene parameters.  Mirrors rxSPHEnviroment (rx_sph.h:95-182) for the fields of the particle path, with the
 * kernel support h, rest density and lattice spacing given explicitly (the particle mass is then solved so
 * that a lattice of that spacing has the rest density). */
typedef struct rxSphEnv {
    int max_particles;
    float boundary_cen[3];   /* box centre                                                    */
    float boundary_ext[3];   /* box half extents                                              */
    float h;                 /* kernel support                                                */
    float rest_dens;         /* rho0                                                          */
    float mass;              /* <= 0: solved from rest_dens, h and spacing                    */
    float spacing;           /* lattice spacing of the rest state (= 2 x particle radius)     */
    float particle_radius;   /* wall margin; <= 0: spacing / 2                                */
    float dt;
    float viscosity;
    float gas_k;
    float gravity[3];
    float epsilon;           /* PBF relaxation                                                */
    float eta;               /* PBF mean density-error target of the reference loop           */
    int min_iter, max_iter;  /* PBF Jacobi iterations                                         */
    int use_ap;              /* s_corr on/off                                                 */
    float ap_k, ap_n, ap_q;
    float xsph_c;            /* XSPH coefficient after the PBF solve (0 = off)                */
    float cell_width;        /* neighbour-grid cell; <= 0: h                                  */
    float expansion;         /* grid covers the box expanded by this fraction (0.01)          */
} rxSphEnv;

/* Defaults of rxSPHEnviroment (rx_sph.h:149-181) where they apply. */
void rxSphEnvDefault(rxSphEnv* env);
/* rxSPH_GPU::Initialize / rxPBDSPH_GPU::Initialize parameter setup (rx_sph_gpu.cpp:157-283,
 * rx_sph_pbd_gpu.cpp:168-300) and setupCells (rx_sph_gpu.cpp:2235-2260) with cell = env->cell_width. */
int rxSetupSimParams(const rxSphEnv* env, rxSimParams* out);
/* Size / field-offset probe of the ABI structs for language bindings: 0 sizeof(rxSimParams),
 * 1 sizeof(rxParticleCell), 2 offsetof(rxSimParams, EffectiveRadius), 3 offsetof(rxSimParams, AP),
 * 4 offsetof(rxSimParams, BoxNum), 5 offsetof(rxSimParams, SphereFlg), 6 offsetof(rxParticleCell, uNumParticles),
 * 7 offsetof(rxParticleCell, uNumPolyHash). */
long rxAbiLayout(int which);

/* Seeded jittered lattice of nx*ny*nz particles, the scene of rxParticleSystemBase::AddBox (rx_ps.cpp:195-261)
 * with a fixed MT19937 seed: position = origin + (i + 0.5) * spacing + U(-jitter, jitter) per axis,
 * velocity = N(0, vel_sigma) per axis.  Returns the number of particles written (<= max_n). */
int rxSceneBox(float* pos4, float* vel4, int max_n, const float origin[3], int nx, int ny, int nz, float spacing,
               float jitter, float vel_sigma, unsigned seed);

#ifdef __cplusplus
}
#endif

#endif /* RX_SPH_HOST_H_ */

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

/* Particle files of rxParticleSystemBase::OutputParticles / InputParticles (rx_ps.cpp:581-644): uint32 n, the
 * n positions (3 float32 each), then the n velocities.  Write: vel4 may be NULL (zeros are written); returns 0
 * on success.  Read: returns the count in the file (-1 on error) and fills at most max_n particles (float4, w = 0);
 * a file that ends after the positions gives zero velocities. */
int rxParticlesWrite(const char* path, const float* pos4, const float* vel4, int n);
int rxParticlesRead(const char* path, float* pos4, float* vel4, int max_n);

/* ---------------------------------------------------------------- solver objects */
/* Arrays returned by the *_GetArray entry points (original particle order). */
enum rxArrayType {
    RX_ARRAY_POSITION = 0,   /* float4 per particle */
    RX_ARRAY_VELOCITY = 1,   /* float4 per particle */
    RX_ARRAY_DENSITY = 2,    /* float per particle  */
    RX_ARRAY_LAMBDA_SUM = 3, /* float per particle: sum of the multipliers the last step applied (PBF)  */
    RX_ARRAY_PRESSURE = 4,   /* float per particle (WCSPH) */
    RX_ARRAY_FEATURES = 5,   /* 16 floats per particle: warm-start features of the last step (PBF) */
    RX_ARRAY_TEMPERATURE = 6,/* float per particle (PBF thermal state)                               */
    RX_ARRAY_HEAT = 7,       /* float per particle: stored latent/sensible heat                      */
    RX_ARRAY_PHASE = 8       /* float per particle: -2 ice, -1 melting, 1 freezing, 2 water           */
};

/* PBF solve modes (rxPBDSPH_GPU_SetOption(RX_PBF_OPT_MODE)). */
enum rxPbfMode {
    RX_PBF_MODE_PLAIN = 0,   /* reference Jacobi: lambda from scratch every iteration                    */
    RX_PBF_MODE_SCALED = 1,  /* first correction uses WARM_SCALE x the first lambda (no network: the
                                constant over-relaxation the learned start is measured against)       */
    RX_PBF_MODE_LEARNED = 2  /* first correction uses the network's prediction from the step's features */
};

/* Options of rxPBDSPH_GPU_SetOption. */
enum rxPbfOption {
    RX_PBF_OPT_SYNC_EACH_STEP = 0, /* 1: Update returns after the step completed on the device        */
    RX_PBF_OPT_RECORD_LAMBDA = 1,  /* 1: accumulate sum_k lambda_k (RX_ARRAY_LAMBDA_SUM)              */
    RX_PBF_OPT_TOL = 2,            /* density error that ends an adaptive solve (default 0.01)        */
    RX_PBF_OPT_ITERATIONS = 3,     /* Jacobi iterations per step (the maximum of an adaptive solve)   */
    RX_PBF_OPT_SKIN = 4,           /* neighbour-list skin as a fraction of h (default 0.15)           */
    RX_PBF_OPT_PROFILE = 5,        /* 1: time every kernel class with HIP events (rxPBDSPH_GPU_GetProfile) */
    RX_PBF_OPT_RECORD_FEATURES = 6,/* 1: compute the warm-start features every step (RX_ARRAY_FEATURES)   */
    RX_PBF_OPT_MODE = 7,           /* rxPbfMode                                                           */
    RX_PBF_OPT_ADAPTIVE = 8,       /* 1: skip the remaining corrections once the judged density error at the
                                      start of an iteration is <= TOL (the loop of rx_sph_pbd_gpu.cpp:862-920) */
    RX_PBF_OPT_WARM_SCALE = 9,     /* factor on the first multiplier: the network's (learned mode) or the
                                      first lambda's (scaled mode); default 1                              */
    RX_PBF_OPT_JUDGE = 10,         /* density error judged against TOL: 0 = max (rho - rho0)+ / rho0,
                                      1 = its mean over the particles (the measure of
                                      CuPbdSphCalDensityFluctuation, rx_sph.cu:668-690)                   */
    RX_PBF_OPT_RESET_RUN = 11,     /* clears the run accumulators (stats slots 48..63)                     */
    RX_PBF_OPT_PROFILE_MASK = 12,  /* bit c set: time kernel class c (rxProfileClass) while profiling;
                                      default all                                                          */
    RX_PBF_OPT_LIST_TILE = 13,     /* list pass: 0 = one thread per slot, global reads (default); LDS-staged
                                      row segments: 1 = 16 cells, 48 KB tile; 2 = 16 cells, 32 KB; 3 = 8
                                      cells, 24 KB, 128 threads; 4 = 16-cell segments without the tile     */
    RX_PBF_OPT_HEAT_TD = 14,       /* thermal diffusion coefficient (HeatTransfar mTD; default 0)         */
    RX_PBF_OPT_HEAT_DT = 15,       /* heat time step (0: the step's dt)                                    */
    RX_PBF_OPT_MAX_NEIGHBORS = 16, /* neighbour-list capacity per particle, a multiple of 4 (default 96)    */
    RX_PBF_OPT_GRAPH = 17,         /* 1: after the first step, each step's launches are recorded once into a
                                      hipGraph per buffer configuration and replayed (for launch-bound sizes);
                                      profiled steps (RX_PBF_OPT_PROFILE) run eagerly.  0 (default): eager */
    RX_PBF_OPT_ADAPT_SKIN = 18,    /* 1 (default): size the neighbour lists from the steps already completed:
                                      the skin grows to 2, 3 or 4 times RX_PBF_OPT_SKIN when more than 3% of a
                                      step's neighbour passes walked the grid instead of their lists (to the
                                      level that holds the step's largest in-step displacement, or a lower one
                                      at which the particles past half the skin are few enough to walk alone),
                                      and shrinks back after 2 steps that needed less; the list capacity grows
                                      with the cutoff volume and by 64 (up to 512) when lists overflow.  Step s
                                      uses the figures of steps up to s - 2 (Update waits for step s - 2 if it
                                      has not finished), so the sizes, and the run, do not depend on timing.
                                      Eager steps only (replayed graphs keep their sizes). */
    RX_PBF_OPT_GRAD_MASS = 19,     /* 1: the constraint gradients carry the particle mass, grad C_i =
                                      (m / rho0) sum_j grad W_ij, in the multiplier's denominator and in the
                                      position correction (Macklin & Mueller 2013, section 3, with rho = sum m W).
                                      0 (default): the reference's grad W / rho0 (rx_sph_pbd.cpp:782-786,
                                      rx_sph_pbd.cpp:970-973), which scales each correction by m */
    RX_PBF_OPT_WARM_CLAMP = 20,    /* learned mode: keep the network's first multiplier between 0 and this
                                      many times the plain first multiplier lambda_0 (its sign, at most this
                                      multiple of its size); 0 (default): as predicted                       */
    RX_PBF_OPT_FIRST_DMAX = 21,    /* largest displacement of a step's first correction, as a fraction of h
                                      (the warm start's correction); 0 (default): unbounded                  */
    RX_PBF_OPT_DEFER = 22          /* 1 (default): a particle past half the skin that walks the grid in a
                                      neighbour pass is walked after the pass by a whole wavefront; 0: in its
                                      own lane, holding its wavefront.  The neighbours are the same; the
                                      sums' order differs                                                */
};

/* Kernel classes of rxPBDSPH_GPU_GetProfile. */
enum rxProfileClass {
    RX_PROF_PREDICT = 0,      /* density + viscosity + prediction          */
    RX_PROF_GRID = 1,         /* hash, counting sort, state gather         */
    RX_PROF_LIST_LAMBDA = 2,  /* neighbour-list build + first lambda       */
    RX_PROF_LAMBDA = 3,       /* lambda pass                               */
    RX_PROF_DP = 4,           /* position correction + apply + walls       */
    RX_PROF_VELOCITY = 5,
    RX_PROF_XSPH = 6,
    RX_PROF_FEATURES = 7,     /* learned solve: feature pass               */
    RX_PROF_MLP = 8,          /* learned solve: network inference          */
    RX_PROF_HEAT = 9,         /* thermal state: diffusion + phase          */
    RX_PROF_DEFERRED = 10     /* walks deferred by the neighbour passes and the mover list's ordering */
};

/* Per-step statistics of rxPBDSPH_GPU_GetStats (float slots):
 *   0       max (rho - rho0)+ / rho0 at the step's start positions (= end of the previous step)
 *   1 + k   the same at the start of Jacobi iteration k (k < 16)
 *   17      largest displacement since the neighbour-list build
 *   18      particles whose neighbour list overflowed
 *   19      corrections executed in the step
 *   20      neighbour passes of particles that walked the grid instead of their list
 *   22      1 when a pass of the step needed a fallback walk beyond its 4-cell cap (the pass then missed
 *           pairs).  Update returns 3 once the host has seen such a completed step, and the next step starts
 *           from a fresh grid (the list is rebuilt)
 *   23      mean |rho - rho0| / rho0 at the step's start positions
 *   24      mean (rho - rho0)+ / rho0 at the step's start positions
 *   25      movers: particles that moved more than half the skin since the list build (at most 4096 are
 *           listed; beyond that every particle bounds displacements by the largest one)
 *   27 + l  corrections that carried a particle past half the skin of level l (l < 4: 1 + l times the set
 *           skin), which the adaptive sizing reads
 *   31      16-byte list chunks the step's list build stored (the passes read these plus a 16-byte head per
 *           particle)
 *   32 + k  slot 24 at the start of Jacobi iteration k (k < 16)
 * Run accumulators (since SetParticles or RX_PBF_OPT_RESET_RUN), over the steps judged so far (a step is
 * judged by the prediction pass of the step after it):
 *   48      max of slot 0        49  max of slot 24      50  corrections executed
 *   51      steps judged         52  steps whose judged error exceeded TOL
 *   53      1 once any pass since the reset needed a walk beyond the cap (slot 22 of any step)
 *   54      list chunks stored by the list builds since the reset (sum of slot 31)
 *   55-57   internal: the running pass's deferred walks (zero between calls), the mover list's sorted
 *           length
 * Host state at the call:
 *   60      list skin in use, as a fraction of h      61  list capacity in use                     */

/* Fused PBF solver (rxPBDSPH_GPU, rx_sph.h:1284-1601): Initialize (rx_sph_pbd_gpu.cpp:168-300) and Update
 * (rx_sph_pbd_gpu.cpp:790-957) with min_iter = max_iter Jacobi iterations, CPU-path lambda and collisions,
 * and the XSPH pass when env->xsph_c != 0.  Update runs `steps` steps asynchronously on the solver's stream. */
void* rxPBDSPH_GPU_Create(const rxSphEnv* env);
void rxPBDSPH_GPU_Destroy(void* solver);
int rxPBDSPH_GPU_SetParticles(void* solver, const float* pos4, const float* vel4, int n);
int rxPBDSPH_GPU_Update(void* solver, float dt, int steps);
int rxPBDSPH_GPU_GetArray(void* solver, int type, float* host, int count);
int rxPBDSPH_GPU_GetStats(void* solver, float* out, int count);
int rxPBDSPH_GPU_Sync(void* solver);
int rxPBDSPH_GPU_SetOption(void* solver, int option, float value);
/* Milliseconds and launch counts per rxProfileClass since the last call (n entries each). */
int rxPBDSPH_GPU_GetProfile(void* solver, float* ms, int* count, int n);
/* Network of the learned warm start (an rxMLP object, see rx_mlp.h).  The solver copies the network's inference
 * weights, biases, standardisation and output scaling at this call (after the network's pending work); later
 * changes to the network, or destroying it, do not affect the solver until SetNetwork is called again. */
int rxPBDSPH_GPU_SetNetwork(void* solver, void* mlp);
/* Thermal state (host arrays in original order; heats and phase may be NULL; temps NULL removes it): after
 * every step the temperatures diffuse over the neighbours within h at the final positions
 * (HeatTransfar::heatParticleAndParticle, HeatTransfar.cpp:429-473: dt (T_j - T_i) / max(rho_j, 0.1) times
 * the cubic spline of radius h, scaled by RX_PBF_OPT_HEAT_TD) and the phase state advances
 * (HeatTransfar::calcTempAndHeat, HeatTransfar.cpp:280-350; melting point 250 on the reference's scale). */
int rxPBDSPH_GPU_SetThermal(void* solver, const float* temps, const float* heats, const int* phase);
/* Feature scales of the scene: lambda denominator and neighbour count of the rest lattice. */
int rxPBDSPH_GPU_FeatureScales(void* solver, float* sd_ref, float* nref);

/* Fused WCSPH solver (rxSPH_GPU, rx_sph.h:432-773): Update (rx_sph_gpu.cpp:756-923) with the CPU path's
 * force and wall rules (rx_sph_cpu.cpp:323-380).  After the first step each step replays a captured
 * hipGraph (rxSPH_GPU_UseGraph(solver, 0) launches the kernels directly instead). */
void* rxSPH_GPU_Create(const rxSphEnv* env);
void rxSPH_GPU_Destroy(void* solver);
int rxSPH_GPU_SetParticles(void* solver, const float* pos4, const float* vel4, int n);
int rxSPH_GPU_Update(void* solver, float dt, int steps);
int rxSPH_GPU_GetArray(void* solver, int type, float* host, int count);
int rxSPH_GPU_Sync(void* solver);
int rxSPH_GPU_UseGraph(void* solver, int on);

/* ---------------------------------------------------------------- fluid + melting ice (no single reference
 * counterpart: the ice-mode frame of rx_fltk_glcanvas.cpp:1134-1145 -- fluid step, ice object step, heat
 * transfer, ice structure -- as one object).  Particles 0 .. n_fluid-1 are fluid, the next n_ice the ice body
 * (rest shape = their initial positions, clusters from ice_env->cluster_radius).  temps/phase: thermal state
 * as rxPBDSPH_GPU_SetThermal (temps required; melting point 250 on the reference's scale).  Per step: PBF step
 * over all particles with heat exchange, then one shape-matching substep of the particles still frozen
 * (phase < 0) written back over the fluid result (blend 1).  GetArray/GetStats/SetOption act on the PBF
 * solver; SetIceTiles(autotune = 1) runs the shape-matching autotuner. */
struct rxIceEnv;
void* rxCoupled_Create(const rxSphEnv* fluid_env, const struct rxIceEnv* ice_env, const float* pos4, const float* vel4, int n_fluid,
                       int n_ice, const float* temps, const int* phase);
void rxCoupled_Destroy(void* coupled);
int rxCoupled_Update(void* coupled, float dt, int steps);
int rxCoupled_Sync(void* coupled);
int rxCoupled_GetArray(void* coupled, int type, float* host, int count);
int rxCoupled_GetStats(void* coupled, float* out, int count);
int rxCoupled_SetOption(void* coupled, int option, float value);
int rxCoupled_SetIceTiles(void* coupled, int autotune, int* out_tiles);
int rxCoupled_GetProfile(void* coupled, float* ms, int* count, int n);
/* Network of the fluid's learned warm start (as rxPBDSPH_GPU_SetNetwork; weights copied): with
 * rxCoupled_SetOption(RX_PBF_OPT_MODE, RX_PBF_MODE_LEARNED) every PBF step of the coupled scene starts from the
 * network's multiplier (the ice particles take part in the fluid solve as in the plain mode). */
int rxCoupled_SetNetwork(void* coupled, void* mlp);

#ifdef __cplusplus
}
#endif

#endif /* RX_SPH_HOST_H_ */

/*
 * oracle.h -- CPU restatement of the reference's particle path.  TEST INFRASTRUCTURE ONLY.
 *
 * The HIP path is compared with this library: only the tests, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py load it.  Nothing in the product links it.
 *
 * Each function follows a reference CPU routine (cited in oracle.cpp).  Storage precision is selectable:
 * the *_f entry points store particle state in float (the reference's RXREAL) and evaluate kernels in
 * double (as rx_kernel.h does); the *_d entry points store everything in double (the FP64 drift oracle).
 * Particle arrays are float4/double4 AoS with stride 4 (the reference's DIM = 4).
 */
#ifndef ORACLE_H_
#define ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrcParams {
    /* fluid */
    double h;             /* kernel support                                       */
    double mass;          /* particle mass                                        */
    double rest_dens;     /* rho0                                                 */
    double gas_k;         /* WCSPH gas constant                                   */
    double viscosity;     /* mu                                                   */
    double gravity[3];
    double radius;        /* particle radius = wall margin                        */
    double bmin[3];       /* walls                                                */
    double bmax[3];
    /* neighbour grid (values are exactly representable in float)                */
    double origin[3];
    double cell[3];
    int gsize[3];
    /* PBF */
    double eps;           /* relaxation epsilon                                   */
    int use_ap;           /* s_corr on/off                                        */
    double ap_k, ap_n, ap_q;
    double xsph_c;        /* XSPH coefficient (0 = off)                           */
    int iterations;       /* Jacobi iterations per step                           */
    double dt;
    int grad_mass;        /* PBF: 1 = constraint gradients carry the mass (m / rho0) sum grad W;
                             0 = the reference's grad W / rho0                    */
    double first_scale;   /* PBF: factor on the multipliers of a step's first correction (0 or 1: none;
                             the fused solver's RX_PBF_MODE_SCALED warm start)    */
    double first_dmax;    /* PBF: largest displacement of a step's first correction (0: none;
                             the fused solver's RX_PBF_OPT_FIRST_DMAX, in length units) */
} OrcParams;

/* Per-step PBF statistics. */
typedef struct OrcPbfStats {
    double max_err[16];   /* max_i max(rho_i - rho0, 0)/rho0 at the start of iteration k */
    double mean_abs[16];  /* mean_i |rho_i - rho0|/rho0 at the start of iteration k      */
    int iterations;
} OrcPbfStats;

/* ---- kernels (rx_kernel.h) */
double orc_kernel_coef(int kind, double h, int type);   /* kind 0 poly6, 1 spiky, 2 visc; type 1 W, 2 grad, 3 lap */
double orc_kernel_poly6(double r, double h, double a);
double orc_kernel_spiky_grad(double r, double h, double a);  /* scalar factor a (h-r)^2 / r, 0 outside (0,h] */
double orc_kernel_visc_lap(double r, double h, double a);
double orc_rest_density(double mass, double h, double spacing);

/* ---- grid (rx_nnsearch.h) */
void orc_calc_hash_f(const float* pos4, const int* attr, int n, const OrcParams* P, uint32_t* hash, uint32_t* index);
void orc_stable_sort(uint32_t* hash, uint32_t* index, int n);
void orc_cell_bounds(const uint32_t* sorted_hash, int n, int ncells, uint32_t* start, uint32_t* end);
/* Neighbours of every particle within r <= h (float distance test of the GPU), brute force; returns the
 * total count and fills counts[n] and, when list != NULL, list[i*maxn + k] (ascending original index). */
long orc_neighbors_brute_f(const float* pos4, int n, float h, int maxn, int* counts, int* list);

/* ---- WCSPH (rx_sph_cpu.cpp) */
void orc_sph_density_f(const float* pos4, int n, const OrcParams* P, float* dens, float* pres);
void orc_sph_force_f(const float* pos4, const float* vel4, const float* dens, const float* pres, int n,
                     const OrcParams* P, float* acc4);
void orc_sph_step_f(float* pos4, float* vel4, int n, const OrcParams* P, float dt, float* dens, float* pres, float* acc4);
void orc_sph_step_d(double* pos4, double* vel4, int n, const OrcParams* P, double dt, double* dens);

/* ---- PBF (rx_sph_pbd.cpp) */
void orc_pbf_density_f(const float* pos4, int n, const OrcParams* P, float* dens);
void orc_pbf_ext_force_f(const float* pos4, const float* vel4, const float* dens, int n, const OrcParams* P, float* acc4);
void orc_pbf_scaling_factor_f(const float* pos4, int n, const OrcParams* P, float* dens, float* lambda);
void orc_pbf_position_correction_f(const float* pos4, const float* lambda, int n, const OrcParams* P, float dt, float* dp4);
void orc_pbf_xsph_f(const float* pos4, const float* vel4, const float* dens, int n, const OrcParams* P, float c, float* out_vel4);
/* rxSPH::integrate with polygon obstacles (rx_sph_cpu.cpp:956-1077): vrts 3 floats per vertex, tris 3 ints per
 * triangle, pstart/pend per grid cell (start 0xFFFFFFFF = none), pidx triangle numbers; walls as in the steps. */
void orc_integrate_polygon_f(const float* pos4, const float* vel4, const float* acc4, const int* attr, int n, const OrcParams* P,
                             const float* vrts, const int* tris, int ntri, const uint32_t* pstart, const uint32_t* pend,
                             const uint32_t* pidx, double restitution, float dt, float* out_pos4, float* out_vel4);
/* Density on a lattice (rx_sph_kernel.cu:1039-1126); ice may be null. */
void orc_grid_density_f(const float* pos4, int n, const OrcParams* P, int nx, int ny, int nz, float x0, float y0, float z0, float dx,
                        float dy, float dz, const float* ice, float* out);
/* Boundary particles (RX_USE_BOUNDARY_PARTICLE): b4 positions, vol volumes, nb of them.
 * rx_sph_cpu.cpp:667-693 (volumes), 527-548 (density term), rx_sph_kernel.cu:662-727 (pressure term, GPU form),
 * rx_sph_pbd.cpp:699-914 (multipliers with boundary, boundary multipliers), 925-1026 (position correction). */
void orc_boundary_volume_f(const float* b4, int nb, const OrcParams* P, float* vol);
void orc_boundary_density_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P, float* dens,
                            float* pres);
void orc_boundary_force_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P,
                          const float* dens, const float* pres, float* acc4);
void orc_pbf_scaling_factor_wb_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P,
                                 float* dens, float* lambda, float* lambda_b);
void orc_pbf_position_correction_wb_f(const float* pos4, const float* lambda, int n, const float* b4, const float* vol,
                                      const float* lambda_b, int nb, const OrcParams* P, float dt, float* dp4);
/* Per-particle magnitudes (double) the FP32 results are judged against: the sum over the summed terms of |term|
 * (forces, the position correction), plus |g| where gravity is added; for lambda (rho_i / rho0 + |C_i|) / (sd_i + eps).
 * Boundary arguments may be 0 / NULL (fluid only); orc_boundary_force_scale_f adds the boundary terms to scale. */
void orc_sph_force_scale_f(const float* pos4, const float* vel4, const float* dens, const float* pres, int n, const OrcParams* P,
                           double* scale);
void orc_pbf_ext_force_scale_f(const float* pos4, const float* vel4, const float* dens, int n, const OrcParams* P, double* scale);
void orc_boundary_force_scale_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P,
                                const float* dens, const float* pres, double* scale);
void orc_pbf_lambda_scale_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P, double* scale,
                            double* scale_b);
void orc_pbf_dp_scale_f(const float* pos4, const float* lambda, int n, const float* b4, const float* vol, const float* lambda_b, int nb,
                        const OrcParams* P, float dt, double* scale);
/* One PBF step of the reference CPU path with a fixed number of Jacobi iterations, collisions and the
 * optional XSPH pass; lambda_sum (may be NULL) receives sum_k lambda_k; dens the last lambda-pass density. */
void orc_pbf_step_f(float* pos4, float* vel4, int n, const OrcParams* P, float* dens, float* lambda_sum, OrcPbfStats* st);
void orc_pbf_step_d(double* pos4, double* vel4, int n, const OrcParams* P, double* dens, OrcPbfStats* st);

/* ---- shape matching (ShapeMatching.h / ShapeMatching.cpp / ElasticObject_OP.cpp / Ice_Convolution_Normal.cpp) */
int orc_jacobi_eigen3(double* a9, double* v9, double eps, int iter_max);
void orc_polar_decomposition(const double* A9, double* R9, double* S9);
typedef struct OrcSmParams {
    double alpha, beta, dt;
    double gravity[3];
    double bmin[3], bmax[3];
    int iterations;
    int volume_conservation;
    double restitution;   /* wall response factor (0.9 in the reference) */
} OrcSmParams;
/* One batched shape-matching step over all clusters (CSR: clusters c have members
 * members[cl_start[c] .. cl_start[c+1]-1]); rest positions of the members come from rest_pos4. */
void orc_sm_step_f(float* pos4, float* vel4, int n, const int* cl_start, const int* members, int nclusters,
                   const float* rest_pos4, const OrcSmParams* S);
void orc_sm_step_d(double* pos4, double* vel4, int n, const int* cl_start, const int* members, int nclusters,
                   const double* rest_pos4, const OrcSmParams* S);
/* The same step with a melt mask: only clusters whose own particle is active, over their active members (rest
 * centre of mass and Aqq of the active members; no linear term below 4 of them); inactive particles keep
 * their prediction. */
void orc_sm_step_masked_f(float* pos4, float* vel4, int n, const int* cl_start, const int* members, int nclusters,
                          const float* rest_pos4, const OrcSmParams* S, const int* active);

/* Legacy ice launches in the reference's slot layout (Ice_SM.h:23-53, IceObject.h:370-395).
 * orc_legacy_cluster_*: itr = 0 is OrientedPrtObj::UpdateCluster_SM (LaunchShapeMatchingGPU: prediction, wall
 * response, rigid matching, vel), itr = 1 is UpdateCluster_SM_Itr (LaunchShapeMatchingIterationGPU: members take
 * sldPos (3 per particle), rigid matching with the reflection fix for >= 10 members, walls); S gives gravity,
 * walls and restitution.  orc_legacy_average_f: the per-particle mean over its clusters (LaunchCalcAverageGPU). */
void orc_legacy_cluster_f(int nc, const float* prtPos4, const float* prtVel4, const float* sldPos3, const float* orgPos, float* curPos,
                          const float* orgCm, float* curCm, float* vel, const int* pIdx, const int* indxSet, float dt,
                          const OrcSmParams* S, int itr);
void orc_legacy_cluster_d(int nc, const double* prtPos4, const double* prtVel4, const double* sldPos3, const double* orgPos,
                          double* curPos, const double* orgCm, double* curCm, double* vel, const int* pIdx, const int* indxSet,
                          double dt, const OrcSmParams* S, int itr);
void orc_legacy_average_f(int n, float* sldPos, float* sldVel, const float* smPos, const float* smVel, const int* indxSet,
                          const int* PtoCIndx, const int* PtoC, int PtoCMax, int PtoCParamSize);

/* ---- heat (HeatTransfar.cpp / HeatTransfer.cu, SPH diffusion) */
void orc_heat_diffuse_f(const float* pos4, const float* dens, const float* temp, int n, const OrcParams* P,
                        float diffusivity, float dt, float* temp_out);

/* LaunchHeatTransferGPU (HeatTransfer.cu:44-204): air exchange, diffusion over the 30-slot lists (distance
 * fixed to 1), melt heat for meltPrtIndx[1..min(999, n-1)], phase update; FP32 in the reference's order. */
void orc_heat_transfer_gpu_f(float* heats, float* temps, float* dtemps, const int* surf, const int* neighs, const float* dens,
                             int* phase, int* changed, const int* melt, float radius, float air, float ht, float td,
                             float dt, int n);
/* HeatTransfar::heatParticleAndParticle + calcTempAndHeat (HeatTransfar.cpp:280-473) over the neighbours within
 * h at the given positions; therm4 = (T, heat, phase, changed) per particle, updated in place. */
void orc_heat_cpu_f(const float* pos4, const float* dens, float* therm4, int n, const OrcParams* P, float td, float dt);
/* detectNeighborParticles (rx_sph_kernel.cu:1166-1208): every j with |x_i - x_j| <= h (FP32, no contraction),
 * in the order (cell hash of j, j), at most 30; out[i*30 + k], unused slots -1, slot 29 = count. */
void orc_detect_neighbors_f(const float* pos4, int n, const OrcParams* P, int* out);
/* detectSurfaceParticles (rx_sph_kernel.cu:1214-1290): surf[i] updated in place from the lists. */
void orc_detect_surface_f(const int* neighs, int* surf, const float* pos4, int n, float mass, float radius);

/* ---- MLP (learned lambda warm start), FP64 */
void orc_mlp_forward_d(const double* X, int nrows, int nin, const double* W1, const double* b1, int nh1,
                       const double* W2, const double* b2, int nh2, const double* W3, const double* b3, double* Y);
/* Gradients of L = 0.5/nrows * sum (Y - T)^2 with respect to all weights. */
void orc_mlp_backward_d(const double* X, const double* T, int nrows, int nin,
                        const double* W1, const double* b1, int nh1, const double* W2, const double* b2, int nh2,
                        const double* W3, const double* b3,
                        double* gW1, double* gb1, double* gW2, double* gb2, double* gW3, double* gb3, double* loss);

#ifdef __cplusplus
}
#endif

#endif /* ORACLE_H_ */

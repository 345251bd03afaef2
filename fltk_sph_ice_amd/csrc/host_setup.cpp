// host_setup.cpp -- scene parameters -> rxSimParams, ABI layout probe and the seeded scene generator.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <random>

#include "rx_common.h"
#include "../../include/rx_sph_host.h"

namespace {
const double kPi = 3.14159265358979323846;

// rx_kernel.h:71-80 in double
double w_poly6(double r, double h, double a) {
    if (r >= 0.0 && r <= h) {
        double q = h * h - r * r;
        return a * q * q * q;
    }
    return 0.0;
}

// Lattice sum of rxSPH::calRestDensity (rx_sph_cpu.cpp:706-720) for unit mass.
double lattice_sum(double h, double l) {
    double aw = 315.0 / (64.0 * kPi * pow(h, 9.0));
    int n = (int)ceil(h / l) + 1;
    double s = 0.0;
    for (int x = -n; x <= n; ++x)
        for (int y = -n; y <= n; ++y)
            for (int z = -n; z <= n; ++z) s += w_poly6(sqrt((double)(x * x + y * y + z * z)) * l, h, aw);
    return s;
}

// Uniform in [0, 1) from one MT19937 draw: (u32 >> 8) * 2^-24.
inline double unif(std::mt19937& g) { return (double)(g() >> 8) * (1.0 / 16777216.0); }
}  // namespace

extern "C" {

void rxSphEnvDefault(rxSphEnv* e) {
    memset(e, 0, sizeof(*e));
    e->max_particles = 50000;
    e->boundary_ext[0] = 2.0f;
    e->boundary_ext[1] = 0.8f;
    e->boundary_ext[2] = 0.8f;
    e->h = 0.0f;
    e->rest_dens = 998.29f;
    e->mass = 0.04f;
    e->spacing = 0.0f;
    e->dt = 0.01f;
    e->viscosity = 1.0e-3f;
    e->gas_k = 3.0f;
    e->gravity[1] = -9.8f;
    e->epsilon = 0.001f;
    e->eta = 0.05f;
    e->min_iter = 2;
    e->max_iter = 10;
    e->use_ap = 1;
    e->ap_k = 0.1f;
    e->ap_n = 4.0f;
    e->ap_q = 0.2f;
    e->expansion = 0.01f;
}

int rxSetupSimParams(const rxSphEnv* e, rxSimParams* out) {
    if (!e || !out) {
        rx_set_error(1, "rxSetupSimParams: null argument");
        return 1;
    }
    if (!(e->h > 0.0f) || !(e->rest_dens > 0.0f)) {
        rx_set_error(1, "rxSetupSimParams: h and rest_dens must be positive");
        return 1;
    }
    double h = e->h;
    double spacing = e->spacing > 0.0f ? e->spacing : 0.5 * h;
    double mass = e->mass > 0.0f ? e->mass : e->rest_dens / lattice_sum(h, spacing);
    memset(out, 0, sizeof(*out));
    rxSimParams& p = *out;
    p.Gravity = make_float3(e->gravity[0], e->gravity[1], e->gravity[2]);
    p.ParticleRadius = e->particle_radius > 0.0f ? e->particle_radius : (float)(0.5 * spacing);
    for (int k = 0; k < 3; ++k) {
        (&p.BoundaryMin.x)[k] = e->boundary_cen[k] - e->boundary_ext[k];
        (&p.BoundaryMax.x)[k] = e->boundary_cen[k] + e->boundary_ext[k];
    }
    // rx_sph_gpu.cpp:214-222 world = box expanded by `expansion`; setupCells (rx_sph_gpu.cpp:2255-2257)
    double cw = e->cell_width > 0.0f ? e->cell_width : h;
    double ex = e->expansion;
    uint g[3];
    for (int k = 0; k < 3; ++k) {
        double lo = (double)e->boundary_cen[k] - e->boundary_ext[k];
        double size = 2.0 * e->boundary_ext[k];
        double o = lo - 0.5 * ex * size;
        double ws = size * (1.0 + ex);
        (&p.WorldOrigin.x)[k] = (float)o;
        (&p.WorldMax.x)[k] = (float)(o + ws);
        (&p.CellWidth.x)[k] = (float)cw;
        g[k] = (uint)(ws / cw) + 1;
    }
    p.GridSize = make_uint3(g[0], g[1], g[2]);
    p.NumCells = g[0] * g[1] * g[2];
    p.NumBodies = 0;
    p.EffectiveRadius = (float)h;
    p.Mass = (float)mass;
    p.Density = e->rest_dens;
    p.Viscosity = e->viscosity;
    p.GasStiffness = e->gas_k;
    p.Dt = e->dt;
    p.KernelParticles = 0.0f;
    // kernel constants of rx_sph_pbd_gpu.cpp:191-201
    p.Wpoly6 = (float)(315.0 / (64.0 * kPi * pow(h, 9.0)));
    p.GWpoly6 = (float)(-945.0 / (32.0 * kPi * pow(h, 9.0)));
    p.LWpoly6 = (float)(-945.0 / (32.0 * kPi * pow(h, 9.0)));
    p.Wspiky = (float)(15.0 / (kPi * pow(h, 6.0)));
    p.GWspiky = (float)(-45.0 / (kPi * pow(h, 6.0)));
    p.LWspiky = (float)(-90.0 / (kPi * pow(h, 6.0)));
    p.Wvisc = (float)(15.0 / (2.0 * kPi * pow(h, 3.0)));
    p.GWvisc = (float)(15.0 / (2.0 * kPi * pow(h, 3.0)));
    p.LWvisc = (float)(45.0 / (kPi * pow(h, 6.0)));
    // rx_sph_pbd_gpu.cpp:213-219
    p.AP = e->use_ap;
    p.AP_K = e->ap_k;
    p.AP_N = e->ap_n;
    p.AP_Q = e->ap_q;
    p.AP_WQ = (float)w_poly6(e->ap_q * h, h, 315.0 / (64.0 * kPi * pow(h, 9.0)));
    p.Restitution = 0.0f;
    return 0;
}

long rxAbiLayout(int which) {
    switch (which) {
    case 0: return (long)sizeof(rxSimParams);
    case 1: return (long)sizeof(rxParticleCell);
    case 2: return (long)offsetof(rxSimParams, EffectiveRadius);
    case 3: return (long)offsetof(rxSimParams, AP);
    case 4: return (long)offsetof(rxSimParams, BoxNum);
    case 5: return (long)offsetof(rxSimParams, SphereFlg);
    case 6: return (long)offsetof(rxParticleCell, uNumParticles);
    case 7: return (long)offsetof(rxParticleCell, uNumPolyHash);
    default: return -1;
    }
}

int rxSceneBox(float* pos4, float* vel4, int max_n, const float origin[3], int nx, int ny, int nz, float spacing, float jitter,
               float vel_sigma, unsigned seed) {
    std::mt19937 g(seed);
    int n = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                if (n >= max_n) return n;
                int idx[3] = {x, y, z};
                for (int k = 0; k < 3; ++k) {
                    double u = unif(g);
                    pos4[4 * n + k] = (float)(origin[k] + (idx[k] + 0.5) * spacing + (2.0 * u - 1.0) * jitter);
                }
                pos4[4 * n + 3] = 0.0f;
                for (int k = 0; k < 3; ++k) {
                    double v = 0.0;
                    if (vel_sigma > 0.0f) {  // Box-Muller
                        double u1 = unif(g), u2 = unif(g);
                        v = vel_sigma * sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * kPi * u2);
                    }
                    if (vel4) vel4[4 * n + k] = (float)v;
                }
                if (vel4) vel4[4 * n + 3] = 0.0f;
                ++n;
            }
    return n;
}

// rxParticleSystemBase::OutputParticles (rx_ps.cpp:581-607): uint32 n, then x y z of every particle, then the
// velocities, all float32 in the host's byte order (little-endian on x86-64).
int rxParticlesWrite(const char* path, const float* pos4, const float* vel4, int n) {
    if (!path || !pos4 || n < 0) {
        rx_set_error(1, "rxParticlesWrite: bad arguments");
        return 1;
    }
    FILE* f = fopen(path, "wb");
    if (!f) {
        rx_set_error(1, "rxParticlesWrite: %s couldn't open", path);
        return 1;
    }
    const uint32_t un = (uint32_t)n;
    bool ok = fwrite(&un, sizeof(un), 1, f) == 1;
    for (int pass = 0; pass < 2 && ok; ++pass) {
        const float* a = pass == 0 ? pos4 : vel4;
        for (int i = 0; i < n && ok; ++i) {
            const float zero[3] = {0.0f, 0.0f, 0.0f};
            ok = fwrite(a ? a + 4 * (size_t)i : zero, sizeof(float), 3, f) == 3;
        }
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) rx_set_error(1, "rxParticlesWrite: write to %s failed", path);
    return ok ? 0 : 1;
}

// rxParticleSystemBase::InputParticles (rx_ps.cpp:615-644): returns the particle count stored in the file
// (-1 on error).  Reads at most max_n particles into pos4/vel4 (float4, w = 0); a file that ends after the
// positions leaves the velocities zero.  max_n = 0 (or null arrays) only returns the count.
int rxParticlesRead(const char* path, float* pos4, float* vel4, int max_n) {
    FILE* f = path ? fopen(path, "rb") : nullptr;
    if (!f) {
        rx_set_error(1, "rxParticlesRead: %s couldn't find", path ? path : "(null)");
        return -1;
    }
    uint32_t n = 0;
    if (fread(&n, sizeof(n), 1, f) != 1 || n > 0x7FFFFFFFu) {
        fclose(f);
        rx_set_error(1, "rxParticlesRead: %s has no particle count", path);
        return -1;
    }
    const int m = (pos4 && max_n > 0) ? (int)std::min<uint32_t>(n, (uint32_t)max_n) : 0;
    bool ok = true;
    for (int pass = 0; pass < 2 && ok; ++pass) {
        float* a = pass == 0 ? pos4 : vel4;
        for (uint32_t i = 0; i < n; ++i) {
            float v[3];
            if (fread(v, sizeof(float), 3, f) != 3) {
                ok = pass == 1 && i == 0;  // no velocities at all is accepted
                if (pass == 1 && ok && vel4)
                    for (int j = 0; j < m; ++j) vel4[4 * j] = vel4[4 * j + 1] = vel4[4 * j + 2] = vel4[4 * j + 3] = 0.0f;
                break;
            }
            if (a && (int)i < m) {
                a[4 * i] = v[0], a[4 * i + 1] = v[1], a[4 * i + 2] = v[2], a[4 * i + 3] = 0.0f;
            }
        }
    }
    fclose(f);
    if (!ok) {
        rx_set_error(1, "rxParticlesRead: %s ends inside the particle data", path);
        return -1;
    }
    return (int)n;
}

}  // extern "C"

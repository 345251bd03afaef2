// oracle.cpp -- CPU restatement of the reference's particle path.  TEST INFRASTRUCTURE ONLY: the oracle the HIP
// path is compared with (loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg).
//
// Structure mirrors the reference CPU classes: a grid (rx_nnsearch.h rxNNGrid), per-particle neighbour
// vectors rebuilt by SetParticlesToCell (rx_sph_cpu.cpp:1365-1380), then the solver loops.  Kernel values
// are evaluated in double (rx_kernel.h); particle state is stored in T (float = the reference's RXREAL, or
// double for the FP64 drift oracle).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;  // RX_PI
const double kFeqEps = 1e-10;               // RX_FEQ_EPS (fixed value; rx_utility.h is not available)

// ------------------------------------------------------------------------------------------ kernels
// rx_kernel.h:29-61 (poly6), 133-165 (spiky), 236-268 (visc) -- 3D coefficients.
double coef(int kind, double h, int type) {
    switch (kind) {
    case 0:
        if (type == 1) return 315.0 / (64.0 * kPi * std::pow(h, 9.0));
        return -945.0 / (32.0 * kPi * std::pow(h, 9.0));
    case 1:
        if (type == 1) return 15.0 / (kPi * std::pow(h, 6.0));
        if (type == 2) return -45.0 / (kPi * std::pow(h, 6.0));
        return -90.0 / (kPi * std::pow(h, 6.0));
    default:
        if (type == 1) return 15.0 / (2.0 * kPi * std::pow(h, 3.0));
        if (type == 2) return 15.0 / (2.0 * kPi * std::pow(h, 5.0));
        return 45.0 / (kPi * std::pow(h, 6.0));
    }
}
// rx_kernel.h:71-80
inline double W_poly6(double r, double h, double a) {
    if (r >= 0.0 && r <= h) { double q = h * h - r * r; return a * q * q * q; }
    return 0.0;
}
// rx_kernel.h:194-203 (returned as the scalar multiplying rij)
inline double GW_spiky_s(double r, double h, double a) {
    if (r > 0.0 && r <= h) { double q = h - r; return a * q * q / r; }
    return 0.0;
}
// rx_kernel.h:316-324
inline double LW_visc(double r, double h, double a) {
    if (r > 0.0 && r <= h) return a * (h - r);
    return 0.0;
}

struct V3 {
    double x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

template <class T> inline V3 ld(const T* p4, int i) { return {(double)p4[4 * i], (double)p4[4 * i + 1], (double)p4[4 * i + 2]}; }
template <class T> inline void st(T* p4, int i, V3 v) { p4[4 * i] = (T)v.x; p4[4 * i + 1] = (T)v.y; p4[4 * i + 2] = (T)v.z; }

// Storage-precision helpers: in float mode scalars that the reference keeps as RXREAL are rounded to float.
template <class T> inline double R(double v) { return (double)(T)v; }

// ------------------------------------------------------------------------------------------ grid
// Cell hash: floor((p - origin)/cell) in IEEE FP32, clamped to the grid (the build's fixed formula; the
// reference GPU form of rx_cu_common.cu:120-144 without fast-math).
inline uint32_t hash_of(float px, float py, float pz, const OrcParams& P) {
    int g[3];
    float p[3] = {px, py, pz};
    for (int k = 0; k < 3; ++k) {
        float o = (float)P.origin[k], c = (float)P.cell[k];
        int v = (int)std::floor((p[k] - o) / c);
        v = std::min(std::max(v, 0), P.gsize[k] - 1);
        g[k] = v;
    }
    return (uint32_t)((g[2] * P.gsize[1] + g[1]) * P.gsize[0] + g[0]);
}

struct Neigh {
    int idx;
    float dist2;
};

// rxNNGrid (rx_nnsearch.h:70-599) with the build's grid rules: cell = h, grid from the caller, stable sort,
// empty cells 0xFFFFFFFF in both start and end.
template <class T> struct Grid {
    std::vector<uint32_t> hash, index, start, end;
    int ncells = 0;

    void build(const T* p4, int n, const OrcParams& P) {
        ncells = P.gsize[0] * P.gsize[1] * P.gsize[2];
        hash.resize(n);
        index.resize(n);
        start.assign(ncells, 0xFFFFFFFFu);
        end.assign(ncells, 0xFFFFFFFFu);
        for (int i = 0; i < n; ++i) {  // rx_nnsearch.h:237-248
            hash[i] = hash_of((float)p4[4 * i], (float)p4[4 * i + 1], (float)p4[4 * i + 2], P);
            index[i] = (uint32_t)i;
        }
        sort_stable(hash.data(), index.data(), n);  // rx_nnsearch.h:251-261 (stable here)
        for (int i = 0; i < n; ++i) {                // rx_nnsearch.h:265-286
            uint32_t h = hash[i];
            if (i == 0 || h != hash[i - 1]) {
                start[h] = (uint32_t)i;
                if (i > 0) end[hash[i - 1]] = (uint32_t)i;
            }
            if (i == n - 1) end[h] = (uint32_t)(i + 1);
        }
    }

    static void sort_stable(uint32_t* h, uint32_t* idx, int n) {
        std::vector<std::pair<uint32_t, uint32_t>> kv(n);
        for (int i = 0; i < n; ++i) kv[i] = {h[i], idx[i]};
        std::stable_sort(kv.begin(), kv.end(), [](const std::pair<uint32_t, uint32_t>& a,
                                                  const std::pair<uint32_t, uint32_t>& b) { return a.first < b.first; });
        for (int i = 0; i < n; ++i) { h[i] = kv[i].first; idx[i] = kv[i].second; }
    }

    // rxNNGrid::GetNN + getNeighborsInCell (rx_nnsearch.h:454-515): cells within (int)(h/cw)+1 of the
    // particle's cell, test Dist2 <= h^2 with Dist2 stored as RXREAL.
    void query(const T* p4, int i, const OrcParams& P, std::vector<Neigh>& out) const {
        out.clear();
        V3 pos = ld(p4, i);
        double h = R<T>(P.h);
        double h2 = R<T>(h * h);
        int c[3];
        double pp[3] = {pos.x, pos.y, pos.z};
        for (int k = 0; k < 3; ++k) c[k] = (int)((pp[k] - P.origin[k]) / P.cell[k]);
        int nard = (int)(h / P.cell[0]) + 1;
        for (int dz = -nard; dz <= nard; ++dz)
            for (int dy = -nard; dy <= nard; ++dy)
                for (int dx = -nard; dx <= nard; ++dx) {
                    int x = c[0] + dx, y = c[1] + dy, z = c[2] + dz;
                    if (x < 0 || x >= P.gsize[0] || y < 0 || y >= P.gsize[1] || z < 0 || z >= P.gsize[2]) continue;
                    uint32_t gh = (uint32_t)((z * P.gsize[1] + y) * P.gsize[0] + x);
                    uint32_t s = start[gh];
                    if (s == 0xFFFFFFFFu) continue;
                    for (uint32_t j = s; j < end[gh]; ++j) {
                        int idx = (int)index[j];
                        V3 d = pos - ld(p4, idx);
                        double d2 = R<T>(dot(d, d));
                        if (d2 <= h2) out.push_back({idx, (float)d2});
                    }
                }
    }
};

template <class T> void neighbour_lists(const T* p4, int n, const OrcParams& P, std::vector<std::vector<Neigh>>& nb) {
    Grid<T> g;
    g.build(p4, n, P);
    nb.resize(n);
    for (int i = 0; i < n; ++i) g.query(p4, i, P, nb[i]);
}

// ------------------------------------------------------------------------------------------ walls
// rxSolidBox::GetDistanceR + AABB_point_dist (rx_sph_solid.cpp:44-108, 244-262) with sgn = 1 (a box that
// contains the fluid): the contact point keeps the particle centre at least r inside the box.
V3 wall_contact(V3 pos, const OrcParams& P, bool* hit) {
    const V3 normals[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    double cen[3], ext[3], sp[3];
    double pp[3] = {pos.x, pos.y, pos.z};
    for (int k = 0; k < 3; ++k) {
        cen[k] = 0.5 * (P.bmin[k] + P.bmax[k]);
        ext[k] = 0.5 * (P.bmax[k] - P.bmin[k]);
        sp[k] = pp[k] - cen[k];
    }
    double r = P.radius;
    double d0[6];
    bool bout[6] = {false, false, false, false, false, false};
    int idx0 = -1, cnt = 0;
    for (int k = 0; k < 3; ++k) {
        if ((d0[2 * k] = (sp[k] - r) - (-ext[k])) < 0.0) { bout[2 * k] = true; idx0 = 2 * k; }
        if ((d0[2 * k + 1] = ext[k] - (sp[k] + r)) < 0.0) { bout[2 * k + 1] = true; idx0 = 2 * k + 1; }
    }
    for (int k = 0; k < 6; ++k) cnt += bout[k] ? 1 : 0;
    double d;
    V3 nrm;
    if (cnt == 0) {
        double min_d = 1e10;
        int idx1 = -1;
        for (int k = 0; k < 6; ++k)
            if (d0[k] <= min_d) { min_d = d0[k]; idx1 = k; }
        d = min_d;
        nrm = (idx1 != -1) ? normals[idx1] : V3{0, 0, 0};
    } else if (cnt == 1) {
        d = d0[idx0];
        nrm = normals[idx0];
    } else {
        double x[3] = {0, 0, 0};
        for (int k = 0; k < 3; ++k) {
            if (bout[2 * k]) x[k] = d0[2 * k];
            else if (bout[2 * k + 1]) x[k] = -d0[2 * k + 1];
        }
        V3 xv{x[0], x[1], x[2]};
        double l = norm(xv);
        d = -l;
        nrm = (-1.0 / l) * xv;
    }
    *hit = (d <= 0.0);
    return pos + std::fabs(d) * nrm;
}

// ------------------------------------------------------------------------------------------ polygons
// Polygon obstacles: vertices (3 floats each), triangles (3 vertex indices each), and the triangles listed per
// cell of the particle grid (start/end per cell, start 0xFFFFFFFF = none; idx = triangle numbers);
// e = the restitution coefficient.
struct Polys {
    const float* vrts;
    const int* tris;
    int ntri;
    const uint32_t* start;
    const uint32_t* end;
    const uint32_t* idx;
    double e;
};

inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// Segment P0 -> P1 against a triangle, in double: the published test of intersectSegmentTriangle
// (rx_cu_common.cu:891-960; the CPU build's copy sits in a header the reference does not ship).  A hit (return
// value 1) needs the segment to enter from the side the normal (V1-V0) x (V2-V0) faces and to end past the plane.
bool segment_triangle(V3 P0, V3 P1, V3 V0, V3 V1, V3 V2, V3& I, V3& n) {
    V3 u = V1 - V0, v = V2 - V0;
    V3 c = cross(u, v);
    double cl = norm(c);
    if (!(cl > 0.0)) return false;
    n = (1.0 / cl) * c;
    V3 dir = P1 - P0;
    double a = dot(n, P0 - V0), b = dot(n, dir);
    if (std::fabs(b) < 1e-10) return false;
    double r = -a / b;
    if (r < 0.0 || std::fabs(a) > std::fabs(b) || b > 0.0) return false;
    I = P0 + r * dir;
    double uu = dot(u, u), uv = dot(u, v), vv = dot(v, v);
    V3 w = I - V0;
    double wu = dot(w, u), wv = dot(w, v), D = uv * uv - uu * vv;
    double s = (uv * wv - vv * wu) / D;
    if (s < 0.0 || s > 1.0) return false;
    double t = (uv * wu - uu * wv) / D;
    if (t < 0.0 || s + t > 1.0) return false;
    return true;
}

// rxSPH::calCollisionPolygon (rx_sph_cpu.cpp:956-993): the triangles of one cell, in list order.
void collide_cell(uint32_t h, V3 x0, V3& x1, V3& v, const Polys& Q, double dt) {
    uint32_t s = Q.start[h];
    if (s == 0xFFFFFFFFu) return;
    for (uint32_t j = s; j < Q.end[h]; ++j) {
        uint32_t t = Q.idx[j];
        if (t >= (uint32_t)Q.ntri) continue;
        V3 vt[3];
        for (int q = 0; q < 3; ++q) {
            int vi = Q.tris[3 * t + q];
            vt[q] = {(double)Q.vrts[3 * vi], (double)Q.vrts[3 * vi + 1], (double)Q.vrts[3 * vi + 2]};
        }
        V3 cp, n;
        if (!segment_triangle(x0, x1, vt[0], vt[1], vt[2], cp, n)) continue;
        double d = norm(x1 - cp);
        double res = (Q.e > 0) ? (Q.e * std::fabs(d) / (dt * norm(v))) : 0.0;
        V3 vr = (-(1.0 + res) * dot(n, v)) * n;
        double l = norm(x1 - x0);
        x1 = cp + (dt * d / l) * vr;
        v = v + vr;
    }
}

// rxSPH::integrate / rxPBDSPH::integrate (rx_sph_cpu.cpp:1040-1077, rx_sph_pbd.cpp:1409-1446) with
// calCollisionSolid (rx_sph_cpu.cpp:1005-1029): the velocity is left untouched by the wall.  With polygons the
// triangles of the old position's cell are tested, then those of the new position's cell when it differs.
template <class T> void integrate(const T* pos, const T* vel, const T* acc, T* pos_new, T* vel_new, int n,
                                  const OrcParams& P, double dt, const Polys* Q = nullptr) {
    for (int i = 0; i < n; ++i) {
        V3 x = ld(pos, i), v = ld(vel, i), a = ld(acc, i);
        V3 x_old = x;
        v = v + dt * a;
        x = x + dt * v;
        if (Q && Q->ntri > 0) {
            uint32_t h0 = hash_of((float)x_old.x, (float)x_old.y, (float)x_old.z, P);
            collide_cell(h0, x_old, x, v, *Q, dt);
            uint32_t h1 = hash_of((float)x.x, (float)x.y, (float)x.z, P);
            if (h1 != h0) collide_cell(h1, x_old, x, v, *Q, dt);
        }
        bool hit;
        V3 c = wall_contact(x, P, &hit);
        if (hit) x = c;
        st(pos_new, i, x);
        st(vel_new, i, v);
    }
}

// ------------------------------------------------------------------------------------------ WCSPH
// rxSPH::calDensity (rx_sph_cpu.cpp:495-552)
template <class T> void sph_density(const T* p4, int n, const std::vector<std::vector<Neigh>>& nb, const OrcParams& P, T* dens) {
    double h = R<T>(P.h), m = R<T>(P.mass), aw = R<T>(coef(0, h, 1));
    for (int i = 0; i < n; ++i) {
        T d = 0;
        for (const Neigh& e : nb[i]) {
            // RXREAL r = sqrt(Dist2) (rx_sph_cpu.cpp:515): a float square root in float mode
            double r = (sizeof(T) == 8) ? norm(ld(p4, i) - ld(p4, e.idx)) : (double)std::sqrt((float)e.dist2);
            d = (T)((double)d + m * W_poly6(r, h, aw));
        }
        dens[i] = d;
    }
}

// rxSPH::calForce (rx_sph_cpu.cpp:558-656); result is an acceleration.
// scale (may be null): per particle sum_j |pressure term_ij| + |viscosity term_ij| + |g|, the magnitude the
// summed acceleration is judged against (its terms nearly cancel in a uniform lattice).
template <class T> void sph_force(const T* p4, const T* v4, const T* dens, T* pres, int n,
                                  const std::vector<std::vector<Neigh>>& nb, const OrcParams& P, T* acc, double* scale = nullptr) {
    double h = R<T>(P.h), m = R<T>(P.mass), ag = R<T>(coef(1, h, 2)), al = R<T>(coef(2, h, 3));
    double r0 = R<T>(P.rest_dens), mu = R<T>(P.viscosity), k = R<T>(P.gas_k);
    for (int i = 0; i < n; ++i) pres[i] = (T)(k * ((double)dens[i] - r0));
    for (int i = 0; i < n; ++i) {
        V3 pos0 = ld(p4, i), vel0 = ld(v4, i);
        double prsi = R<T>((double)pres[i] / ((double)dens[i] * (double)dens[i]));
        V3 Fp{0, 0, 0}, Fev{0, 0, 0};
        double sc = 0.0;
        for (const Neigh& e : nb[i]) {
            int j = e.idx;
            if (j == i) continue;
            V3 rij = pos0 - ld(p4, j);
            V3 vji = ld(v4, j) - vel0;
            double r = norm(rij);
            double prsj = R<T>((double)pres[j] / ((double)dens[j] * (double)dens[j]));
            V3 tp = (m * (prsi + prsj) * GW_spiky_s(r, h, ag)) * rij;
            V3 tv = (m / (double)dens[j] * LW_visc(r, h, al)) * vji;
            Fp = Fp + tp;
            Fev = Fev + tv;
            sc += norm(tp) + std::fabs(mu) * norm(tv);
        }
        V3 f = (-1.0) * Fp + mu * Fev;
        V3 g{P.gravity[0], P.gravity[1], P.gravity[2]};
        f = f + g;
        if (scale) scale[i] = sc + norm(g);
        st(acc, i, f);
        acc[4 * i + 3] = 0;
    }
}

template <class T> void sph_step(T* pos, T* vel, int n, const OrcParams& P, double dt, T* dens, T* pres, T* acc) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos, n, P, nb);  // SetParticlesToCell (rx_sph_cpu.cpp:1365-1380)
    sph_density(pos, n, nb, P, dens);
    sph_force(pos, vel, dens, pres, n, nb, P, acc);
    integrate(pos, vel, acc, pos, vel, n, P, dt);
}

// ------------------------------------------------------------------------------------------ PBF
// rxPBDSPH::calForceExtAndVisc (rx_sph_pbd.cpp:651-689): viscosity divided by rho_i, plus gravity.
// scale (may be null): per particle mu sum_j |term_ij| + |g|.
template <class T> void pbf_ext_force(const T* p4, const T* v4, const T* dens, int n,
                                      const std::vector<std::vector<Neigh>>& nb, const OrcParams& P, T* acc,
                                      double* scale = nullptr) {
    double h = R<T>(P.h), m = R<T>(P.mass), al = R<T>(coef(2, h, 3)), mu = R<T>(P.viscosity);
    for (int i = 0; i < n; ++i) {
        V3 pos0 = ld(p4, i), vel0 = ld(v4, i);
        V3 Fev{0, 0, 0};
        double sc = 0.0;
        for (const Neigh& e : nb[i]) {
            int j = e.idx;
            if (j == i) continue;
            V3 rij = pos0 - ld(p4, j);
            V3 vji = ld(v4, j) - vel0;
            double r = norm(rij);
            V3 t = (m / (double)dens[i] * LW_visc(r, h, al)) * vji;
            Fev = Fev + t;
            sc += norm(t);
        }
        V3 g{P.gravity[0], P.gravity[1], P.gravity[2]};
        V3 f = mu * Fev + g;
        if (scale) scale[i] = std::fabs(mu) * sc + norm(g);
        st(acc, i, f);
        acc[4 * i + 3] = 0;
    }
}

// ------------------------------------------------------------------------------------------ boundary particles
// Boundary particles of the RX_USE_BOUNDARY_PARTICLE configuration: positions b4 (nb of them) and volumes vol.
template <class T> struct Boundary {
    const T* b4 = nullptr;
    const T* vol = nullptr;
    int nb = 0;
};

// GetNearestNeighborsB (rx_sph_cpu.cpp:1352-1362): the boundary particles with Dist2 <= h^2 (Dist2 held as
// RXREAL), found by a scan of all of them in index order.
template <class T> void boundary_neighbours(V3 pos, const T* b4, int nb, const OrcParams& P, std::vector<int>& out) {
    out.clear();
    double h = R<T>(P.h), h2 = R<T>(h * h);
    for (int j = 0; j < nb; ++j) {
        V3 d = pos - ld(b4, j);
        if (R<T>(dot(d, d)) <= h2) out.push_back(j);
    }
}

// rxSPH::calBoundaryVolumes (rx_sph_cpu.cpp:667-693): V_b = m / sum_{b'} m W(r), itself included.
template <class T> void boundary_volumes(const T* b4, int nb, const OrcParams& P, T* vol) {
    double h = R<T>(P.h), m = R<T>(P.mass), aw = R<T>(coef(0, h, 1));
    std::vector<int> nbr;
    for (int i = 0; i < nb; ++i) {
        V3 pos0 = ld(b4, i);
        boundary_neighbours(pos0, b4, nb, P, nbr);
        T mw = 0;
        for (int j : nbr) mw = (T)((double)mw + m * W_poly6(norm(ld(b4, j) - pos0), h, aw));
        vol[i] = (T)(m / (double)mw);
    }
}

// Boundary density term of rxSPH::calDensity / rxPBDSPH::calScalingFactor (rx_sph_cpu.cpp:527-548,
// rx_sph_pbd.cpp:729-748): sum_b rho0 V_b W(r).
template <class T> T boundary_density(V3 pos0, const Boundary<T>& B, const OrcParams& P, std::vector<int>& nbr) {
    double h = R<T>(P.h), aw = R<T>(coef(0, h, 1)), r0 = R<T>(P.rest_dens);
    boundary_neighbours(pos0, B.b4, B.nb, P, nbr);
    T brho = 0;
    for (int j : nbr) brho = (T)((double)brho + r0 * (double)B.vol[j] * W_poly6(norm(ld(B.b4, j) - pos0), h, aw));
    return brho;
}

// Boundary pressure term as the GPU kernel computes it (rx_sph_kernel.cu:662-727): an acceleration,
// sum_b -rho0 V_b (p_i / rho_i^2) grad W_spiky for 1e-4 < r <= h.  (The CPU loop, rx_sph_cpu.cpp:617-651,
// divides this sum by the mass.)
// scale (may be null): sum_b |term_ib| is added to scale[i].
template <class T> void boundary_force(const T* p4, int n, const Boundary<T>& B, const OrcParams& P, const T* dens, const T* pres,
                                       T* acc, double* scale = nullptr) {
    double h = R<T>(P.h), ag = R<T>(coef(1, h, 2)), r0 = R<T>(P.rest_dens);
    std::vector<int> nbr;
    for (int i = 0; i < n; ++i) {
        V3 pos0 = ld(p4, i);
        double prsi = R<T>((double)pres[i] / ((double)dens[i] * (double)dens[i]));
        boundary_neighbours(pos0, B.b4, B.nb, P, nbr);
        V3 bp{0, 0, 0};
        for (int j : nbr) {
            V3 rij = pos0 - ld(B.b4, j);
            double r = norm(rij);
            if (r <= 0.0001) continue;
            V3 t = (-r0 * (double)B.vol[j] * prsi * GW_spiky_s(r, h, ag)) * rij;
            bp = bp + t;
            if (scale) scale[i] += norm(t);
        }
        for (int k = 0; k < 3; ++k) acc[4 * i + k] = (T)((double)acc[4 * i + k] + (k == 0 ? bp.x : k == 1 ? bp.y : bp.z));
    }
}

// rxPBDSPH::calScalingFactor (rx_sph_pbd.cpp:699-822), fluid-only branch.
// With boundary particles (RX_USE_BOUNDARY_PARTICLE) the density and the denominator gain the boundary terms
// (rx_sph_pbd.cpp:729-748, 790-812).
// scale (may be null): per particle (rho_i / rho0 + |C_i|) / (sd_i + eps), the magnitude that bounds the rounding
// of lambda (C = rho/rho0 - 1 cancels).
// P.grad_mass != 0: every constraint gradient carries the particle mass, grad C_i = (m / rho0) sum grad W
// (Macklin & Mueller 2013, section 3, with rho_i = sum m W); the reference's gradient has no mass factor.
template <class T> void pbf_lambda(const T* p4, int n, const std::vector<std::vector<Neigh>>& nb, const OrcParams& P,
                                   T* dens, T* lambda, const Boundary<T>* B = nullptr, double* scale = nullptr) {
    double h = R<T>(P.h), m = R<T>(P.mass), aw = R<T>(coef(0, h, 1)), ag = R<T>(coef(1, h, 2));
    double r0 = R<T>(P.rest_dens), eps = R<T>(P.eps);
    const double gm = P.grad_mass ? m : 1.0;
    std::vector<int> bn;
    for (int i = 0; i < n; ++i) {
        V3 pos0 = ld(p4, i);
        T d = 0;
        for (const Neigh& e : nb[i]) {
            double r = norm(pos0 - ld(p4, e.idx));
            d = (T)((double)d + m * W_poly6(r, h, aw));
        }
        if (B) d = (T)((double)d + (double)boundary_density(pos0, *B, P, bn));
        dens[i] = d;
        T C = (T)((double)d / r0 - 1.0);
        T sd = 0;
        for (const Neigh& e : nb[i]) {
            int k = e.idx;
            V3 dp{0, 0, 0};
            if (k == i) {
                for (const Neigh& f : nb[i]) {
                    V3 rij = pos0 - ld(p4, f.idx);
                    double ri = norm(rij);
                    dp = dp + (gm * GW_spiky_s(ri, h, ag) / r0) * rij;
                }
            } else {
                V3 rik = pos0 - ld(p4, k);
                double r = norm(rik);
                dp = (-gm * GW_spiky_s(r, h, ag) / r0) * rik;
            }
            sd = (T)((double)sd + dot(dp, dp));
        }
        if (B) {
            for (int j : bn) {
                V3 rij = pos0 - ld(B->b4, j);
                double r = norm(rij);
                V3 dpb = (gm * (r0 * (double)B->vol[j] / m) * GW_spiky_s(r, h, ag) / r0) * rij;
                sd = (T)((double)sd + dot(dpb, dpb));
            }
        }
        lambda[i] = (T)(-(double)C / ((double)sd + eps));
        if (scale) scale[i] = ((double)d / r0 + std::fabs((double)C)) / ((double)sd + eps);
    }
}

// rxPBDSPH::calScalingFactor, boundary-particle loop (rx_sph_pbd.cpp:825-906): lambda_b from the fluid and
// boundary particles within h of b.
template <class T> void pbf_boundary_lambda(const T* p4, int n, const Boundary<T>& B, const OrcParams& P, T* lambda_b,
                                            double* scale = nullptr) {
    double h = R<T>(P.h), m = R<T>(P.mass), aw = R<T>(coef(0, h, 1)), ag = R<T>(coef(1, h, 2));
    double r0 = R<T>(P.rest_dens), eps = R<T>(P.eps), h2 = R<T>(h * h);
    const double gm = P.grad_mass ? m : 1.0;
    std::vector<int> fn, bn;
    for (int i = 0; i < B.nb; ++i) {
        V3 pos0 = ld(B.b4, i);
        fn.clear();
        for (int j = 0; j < n; ++j) {  // GetNearestNeighbors: Dist2 <= h^2
            V3 d = pos0 - ld(p4, j);
            if (R<T>(dot(d, d)) <= h2) fn.push_back(j);
        }
        boundary_neighbours(pos0, B.b4, B.nb, P, bn);
        T brho = 0;
        for (int j : fn) brho = (T)((double)brho + m * W_poly6(norm(ld(p4, j) - pos0), h, aw));
        for (int j : bn) brho = (T)((double)brho + r0 * (double)B.vol[j] * W_poly6(norm(ld(B.b4, j) - pos0), h, aw));
        T C = (T)((double)brho / r0 - 1.0);
        T sd = 0;
        for (int k : fn) {
            V3 rik = pos0 - ld(p4, k);
            V3 dp = (gm * GW_spiky_s(norm(rik), h, ag) / r0) * rik;
            sd = (T)((double)sd + dot(dp, dp));
        }
        for (int j : bn) {
            V3 rij = pos0 - ld(B.b4, j);
            V3 dpb = (gm * (r0 * (double)B.vol[j] / m) * GW_spiky_s(norm(rij), h, ag) / r0) * rij;
            sd = (T)((double)sd + dot(dpb, dpb));
        }
        lambda_b[i] = (T)(-(double)C / ((double)sd + eps));
        if (scale) scale[i] = ((double)brho / r0 + std::fabs((double)C)) / ((double)sd + eps);
    }
}

// rxPBDSPH::calPositionCorrection (rx_sph_pbd.cpp:925-1026), fluid-only branch; s_corr carries dt^2 as in
// the reference (rx_sph_pbd.cpp:970).
// With boundary particles: plus sum_b (lambda_i + lambda_b + s_corr) grad W / rho0, s_corr's W scaled by
// rho0 V_b / m (rx_sph_pbd.cpp:978-1018).
// scale (may be null): per particle sum_j |term_ij| (boundary terms included).
template <class T> void pbf_dp(const T* p4, const T* lambda, int n, const std::vector<std::vector<Neigh>>& nb,
                               const OrcParams& P, double dt, T* dp4, const Boundary<T>* B = nullptr, const T* lambda_b = nullptr,
                               double* scale = nullptr) {
    double h = R<T>(P.h), aw = R<T>(coef(0, h, 1)), ag = R<T>(coef(1, h, 2)), r0 = R<T>(P.rest_dens);
    const double gm = P.grad_mass ? R<T>(P.mass) : 1.0;
    double k = R<T>(P.ap_k), nn = R<T>(P.ap_n), dq = R<T>(P.ap_q * h);
    double wq = R<T>(W_poly6(dq, h, aw));
    std::vector<int> bn;
    for (int i = 0; i < n; ++i) {
        V3 pos0 = ld(p4, i);
        V3 dpij{0, 0, 0};
        double sc = 0.0;
        for (const Neigh& e : nb[i]) {
            int j = e.idx;
            V3 rij = pos0 - ld(p4, j);
            double r = norm(rij);
            if (r > h) continue;
            double scorr = 0.0;
            if (P.use_ap) {
                double ww = R<T>(W_poly6(r, h, aw) / wq);
                scorr = R<T>(-k * std::pow(ww, nn) * dt * dt);
            }
            V3 t = (gm * ((double)lambda[i] + (double)lambda[j] + scorr) * GW_spiky_s(r, h, ag) / r0) * rij;
            dpij = dpij + t;
            sc += norm(t);
        }
        if (B) {
            double m = R<T>(P.mass);
            boundary_neighbours(pos0, B->b4, B->nb, P, bn);
            V3 dpb{0, 0, 0};
            for (int j : bn) {
                V3 rij = pos0 - ld(B->b4, j);
                double r = norm(rij);
                if (r > h) continue;
                double scorr = 0.0;
                if (P.use_ap) {
                    double ww = R<T>((r0 * (double)B->vol[j] / m) * W_poly6(r, h, aw) / wq);
                    scorr = R<T>(-k * std::pow(ww, nn) * dt * dt);
                }
                V3 t = (gm * ((double)lambda[i] + (double)lambda_b[j] + scorr) * GW_spiky_s(r, h, ag) / r0) * rij;
                dpb = dpb + t;
                sc += norm(t);
            }
            dpij = dpij + dpb;
        }
        if (scale) scale[i] = sc;
        st(dp4, i, dpij);
        dp4[4 * i + 3] = 0;
    }
}

// XSPH (the form of rx_pbdsph_kernel.cu:1090-1162): v_i + c sum_j (m/rho_j)(v_j - v_i) W_poly6.
template <class T> void pbf_xsph(const T* p4, const T* v4, const T* dens, int n, const std::vector<std::vector<Neigh>>& nb,
                                 const OrcParams& P, double c, T* out) {
    double h = R<T>(P.h), m = R<T>(P.mass), aw = R<T>(coef(0, h, 1));
    for (int i = 0; i < n; ++i) {
        V3 pos0 = ld(p4, i), vel0 = ld(v4, i);
        V3 acc{0, 0, 0};
        for (const Neigh& e : nb[i]) {
            int j = e.idx;
            double r = norm(pos0 - ld(p4, j));
            acc = acc + (m / (double)dens[j] * W_poly6(r, h, aw)) * (ld(v4, j) - vel0);
        }
        st(out, i, vel0 + c * acc);
        out[4 * i + 3] = 0;
    }
}

template <class T> void density_stats(const T* dens, int n, double r0, double* max_pos, double* mean_abs) {
    double mx = 0.0, sum = 0.0;
    for (int i = 0; i < n; ++i) {
        double e = ((double)dens[i] - r0) / r0;
        mx = std::max(mx, std::max(e, 0.0));
        sum += std::fabs(e);
    }
    *max_pos = mx;
    *mean_abs = n ? sum / n : 0.0;
}

// rxPBDSPH::Update (rx_sph_pbd.cpp:377-484) with min = max = P.iterations (a fixed iteration count),
// followed by the XSPH pass of the PBF configuration when P.xsph_c != 0.
template <class T> void pbf_step(T* pos, T* vel, int n, const OrcParams& P, T* dens_out, T* lambda_sum, OrcPbfStats* st) {
    double dt = P.dt;
    std::vector<std::vector<Neigh>> nb;
    std::vector<T> dens(n), acc(4 * n), ppos(4 * n), pvel(4 * n), lam(n), dp(4 * n), zero(4 * n, (T)0);
    neighbour_lists(pos, n, P, nb);
    sph_density(pos, n, nb, P, dens.data());  // rxPBDSPH::calDensity (rx_sph_pbd.cpp:599-646)
    pbf_ext_force(pos, vel, dens.data(), n, nb, P, acc.data());
    integrate(pos, vel, acc.data(), ppos.data(), pvel.data(), n, P, dt);
    if (lambda_sum) std::fill(lambda_sum, lambda_sum + n, (T)0);
    int iters = P.iterations;
    for (int it = 0; it < iters; ++it) {
        neighbour_lists(ppos.data(), n, P, nb);  // the reference re-searches every iteration (rx_sph_pbd.cpp:427)
        pbf_lambda(ppos.data(), n, nb, P, dens.data(), lam.data());
        if (st && it < 16) density_stats(dens.data(), n, R<T>(P.rest_dens), &st->max_err[it], &st->mean_abs[it]);
        if (it == 0 && P.first_scale != 0.0 && P.first_scale != 1.0)  // warm start by a constant factor
            for (int i = 0; i < n; ++i) lam[i] = R<T>((double)lam[i] * R<T>(P.first_scale));
        if (lambda_sum) for (int i = 0; i < n; ++i) lambda_sum[i] = (T)((double)lambda_sum[i] + (double)lam[i]);
        pbf_dp(ppos.data(), lam.data(), n, nb, P, dt, dp.data());
        if (it == 0 && P.first_dmax > 0.0)  // the first correction held to first_dmax
            for (int i = 0; i < n; ++i) {
                double l2 = 0.0;
                for (int k = 0; k < 3; ++k) l2 += (double)dp[4 * i + k] * (double)dp[4 * i + k];
                const double dmax = R<T>(P.first_dmax);
                if (l2 > dmax * dmax)
                    for (int k = 0; k < 3; ++k) dp[4 * i + k] = R<T>((double)dp[4 * i + k] * (dmax / std::sqrt(l2)));
            }
        for (int i = 0; i < 4 * n; ++i) ppos[i] = (T)((double)ppos[i] + (double)dp[i]);
        integrate(ppos.data(), zero.data(), zero.data(), ppos.data(), pvel.data(), n, P, dt);  // collision only
    }
    if (st) st->iterations = iters;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            vel[4 * i + k] = (T)(((double)ppos[4 * i + k] - (double)pos[4 * i + k]) / dt);
            pos[4 * i + k] = ppos[4 * i + k];
        }
    if (P.xsph_c != 0.0) {
        neighbour_lists(pos, n, P, nb);
        std::vector<T> nv(4 * n);
        pbf_xsph(pos, vel, dens.data(), n, nb, P, P.xsph_c, nv.data());
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) vel[4 * i + k] = nv[4 * i + k];
    }
    if (dens_out) std::copy(dens.begin(), dens.end(), dens_out);
}

// ------------------------------------------------------------------------------------------ shape matching
// EigenJacobiMethod (ShapeMatching.h:242-343) for n = 3; v must hold the starting basis (identity).
int jacobi3(double* a, double* v, double eps, int iter_max) {
    const int n = 3;
    double bim[3], bjm[3];
    int cnt = 0;
    for (;;) {
        int i = -1, j = -1;
        double x = 0.0;
        for (int ia = 0; ia < n; ++ia)
            for (int ja = 0; ja < n; ++ja) {
                int idx = ia * n + ja;
                if (ia != ja && std::fabs(a[idx]) > x) { i = ia; j = ja; x = std::fabs(a[idx]); }
            }
        if (i == -1 || j == -1) return 0;
        double aii = a[i * n + i], ajj = a[j * n + j], aij = a[i * n + j];
        double alpha = (aii - ajj) / 2.0;
        double beta = std::sqrt(alpha * alpha + aij * aij);
        double ct = std::sqrt((1.0 + std::fabs(alpha) / beta) / 2.0);
        double st_ = (((aii - ajj) >= 0.0) ? 1.0 : -1.0) * aij / (2.0 * beta * ct);
        for (int m = 0; m < n; ++m) {
            if (m == i || m == j) continue;
            double aim = a[i * n + m], ajm = a[j * n + m];
            bim[m] = aim * ct + ajm * st_;
            bjm[m] = -aim * st_ + ajm * ct;
        }
        double bii = aii * ct * ct + 2.0 * aij * ct * st_ + ajj * st_ * st_;
        double bjj = aii * st_ * st_ - 2.0 * aij * ct * st_ + ajj * ct * ct;
        for (int m = 0; m < n; ++m) {
            if (m == i || m == j) continue;
            a[i * n + m] = a[m * n + i] = bim[m];
            a[j * n + m] = a[m * n + j] = bjm[m];
        }
        a[i * n + i] = bii;
        a[i * n + j] = 0.0;
        a[j * n + j] = bjj;
        a[j * n + i] = 0.0;
        for (int m = 0; m < n; ++m) {
            double vmi = v[m * n + i], vmj = v[m * n + j];
            bim[m] = vmi * ct + vmj * st_;
            bjm[m] = -vmi * st_ + vmj * ct;
        }
        for (int m = 0; m < n; ++m) { v[m * n + i] = bim[m]; v[m * n + j] = bjm[m]; }
        double e = 0.0;
        for (int ja = 0; ja < n; ++ja)
            for (int ia = 0; ia < n; ++ia)
                if (ia != ja) e += std::fabs(a[ja * n + ia]);
        if (e < eps) break;
        cnt++;
        if (cnt > iter_max) break;
    }
    return cnt;
}

// PolarDecomposition (ShapeMatching.h:352-387): R = A (A^T A)^(-1/2), S = R^T A.
void polar(const double* A, double* Rm, double* Sm) {
    double ata[9], U[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += A[k * 3 + r] * A[k * 3 + c];
            ata[r * 3 + c] = s;
        }
    jacobi3(ata, U, 1e-8, 100);
    double l[3];
    for (int k = 0; k < 3; ++k) l[k] = (ata[k * 4] <= 0.0) ? 0.0 : 1.0 / std::sqrt(ata[k * 4]);
    double S1[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) S1[r * 3 + c] = l[0] * U[r * 3 + 0] * U[c * 3 + 0] + l[1] * U[r * 3 + 1] * U[c * 3 + 1] + l[2] * U[r * 3 + 2] * U[c * 3 + 2];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += A[r * 3 + k] * S1[k * 3 + c];
            Rm[r * 3 + c] = s;
        }
    if (Sm)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int k = 0; k < 3; ++k) s += Rm[k * 3 + r] * A[k * 3 + c];
                Sm[r * 3 + c] = s;
            }
}

double det3(const double* a) {
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}
bool inv3(const double* a, double* o) {
    double d = det3(a);
    if (d == 0.0) return false;
    double id = 1.0 / d;
    o[0] = (a[4] * a[8] - a[5] * a[7]) * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = (a[5] * a[6] - a[3] * a[8]) * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = (a[3] * a[7] - a[4] * a[6]) * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return true;
}

// One batched shape-matching step:
//   prediction and wall response of OrientedPrtObj::UpdateCluster_SM (ElasticObject_OP.cpp:556-609),
//   per-cluster matching of rxShapeMatching::shapeMatching (ShapeMatching.cpp:492-587: alpha, beta, volume
//   conservation; unit masses), per-particle averaging of ConvoNormal::StepConvolution
//   (Ice_Convolution_Normal.cpp:32-67), repeated S.iterations times as Ice_CalcMethod_Iteration::StepObjMove
//   (Ice_CalcMethod_Iteration.cpp:32-48), and the velocity update of OrientedParticle::UpdatePosAndVel
//   (OrientedParticle.cpp:377-393).
// With `active` (may be NULL) a cluster takes part only when its own particle is active, and only its active
// members count: the rest centre of mass and Aqq are those of the active members; with fewer than 4 active
// members (Aqq singular in 3-D) the linear term is dropped.  Inactive particles keep their prediction.
template <class T> void sm_step(T* pos, T* vel, int n, const int* cs, const int* mem, int nc, const T* rest, const OrcSmParams& S,
                                const int* active = nullptr) {
    double dt = S.dt;
    std::vector<V3> x(n), xs(n);
    for (int i = 0; i < n; ++i) {
        V3 p = ld(pos, i), v = ld(vel, i);
        double g[3] = {S.gravity[0], S.gravity[1], S.gravity[2]};
        double pv[3] = {p.x, p.y, p.z}, vv[3] = {v.x, v.y, v.z}, c[3];
        for (int k = 0; k < 3; ++k) c[k] = pv[k] + (vv[k] + g[k] * dt) * dt;
        for (int k = 0; k < 3; ++k) {
            if (c[k] < S.bmin[k] || c[k] > S.bmax[k]) {
                c[k] = pv[k] - (vv[k] + g[k] * dt) * dt * S.restitution;
                for (int o = 0; o < 3; ++o)
                    if (o != k) c[o] = pv[o];
            }
        }
        for (int k = 0; k < 3; ++k) c[k] = std::min(std::max(c[k], S.bmin[k]), S.bmax[k]);
        x[i] = p;
        xs[i] = {R<T>(c[0]), R<T>(c[1]), R<T>(c[2])};
    }
    // rest centres of mass and Aqq^-1 per cluster (Ice_SM.cpp:1319-1338 CalcOrgCm, unit masses)
    std::vector<V3> cm0(nc);
    std::vector<double> aqqi(9 * (size_t)nc);
    for (int c = 0; c < nc; ++c) {
        V3 s{0, 0, 0};
        int k = cs[c + 1] - cs[c];
        for (int t = cs[c]; t < cs[c + 1]; ++t) s = s + ld(rest, mem[t]);
        cm0[c] = (1.0 / k) * s;
        double aqq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int t = cs[c]; t < cs[c + 1]; ++t) {
            V3 q = ld(rest, mem[t]) - cm0[c];
            double qq[3] = {q.x, q.y, q.z};
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) aqq[r * 3 + cc] += qq[r] * qq[cc];
        }
        if (!inv3(aqq, &aqqi[9 * (size_t)c])) std::fill(&aqqi[9 * (size_t)c], &aqqi[9 * (size_t)c] + 9, 0.0);
    }
    std::vector<V3> acc(n);
    std::vector<int> cnt(n);
    std::vector<int> am;  // active members of the current cluster
    for (int it = 0; it < S.iterations; ++it) {
        std::fill(acc.begin(), acc.end(), V3{0, 0, 0});
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int c = 0; c < nc; ++c) {
            if (active && !active[c]) continue;
            am.clear();
            for (int t = cs[c]; t < cs[c + 1]; ++t)
                if (!active || active[mem[t]]) am.push_back(mem[t]);
            int k = (int)am.size();
            if (k <= 1) {  // shapeMatching returns early for a single vertex (ShapeMatching.cpp:494)
                for (int p : am) { acc[p] = acc[p] + xs[p]; cnt[p]++; }
                continue;
            }
            const bool partial = k < cs[c + 1] - cs[c];
            V3 c0 = cm0[c];
            double aqqa[9];
            const double* ai = &aqqi[9 * (size_t)c];
            bool linear = S.beta != 0.0;
            if (partial) {  // rest quantities of the active members
                V3 s0{0, 0, 0};
                for (int p : am) s0 = s0 + ld(rest, p);
                c0 = (1.0 / k) * s0;
                double aqq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int p : am) {
                    V3 q = ld(rest, p) - c0;
                    double qq[3] = {q.x, q.y, q.z};
                    for (int r = 0; r < 3; ++r)
                        for (int cc = 0; cc < 3; ++cc) aqq[r * 3 + cc] += qq[r] * qq[cc];
                }
                if (!inv3(aqq, aqqa)) std::fill(aqqa, aqqa + 9, 0.0);
                ai = aqqa;
                if (k < 4) linear = false;
            }
            V3 cm{0, 0, 0};
            for (int p : am) cm = cm + xs[p];
            cm = (1.0 / k) * cm;
            double apq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int pi : am) {
                V3 p = xs[pi] - cm, q = ld(rest, pi) - c0;
                double pp[3] = {p.x, p.y, p.z}, qq[3] = {q.x, q.y, q.z};
                for (int r = 0; r < 3; ++r)
                    for (int cc = 0; cc < 3; ++cc) apq[r * 3 + cc] += pp[r] * qq[cc];
            }
            double Rm[9], RL[9];
            polar(apq, Rm, nullptr);
            if (linear) {
                double A[9];
                for (int r = 0; r < 3; ++r)
                    for (int cc = 0; cc < 3; ++cc) {
                        double s = 0;
                        for (int q = 0; q < 3; ++q) s += apq[r * 3 + q] * ai[q * 3 + cc];
                        A[r * 3 + cc] = s;
                    }
                if (S.volume_conservation) {
                    double d = std::fabs(det3(A));
                    if (d > kFeqEps) {
                        d = 1.0 / std::sqrt(d);
                        if (d > 2.0) d = 2.0;
                        for (int q = 0; q < 9; ++q) A[q] *= d;
                    }
                }
                for (int q = 0; q < 9; ++q) RL[q] = S.beta * A[q] + (1.0 - S.beta) * Rm[q];
            } else {
                for (int q = 0; q < 9; ++q) RL[q] = Rm[q];
            }
            for (int p : am) {
                V3 q = ld(rest, p) - c0;
                V3 gp{RL[0] * q.x + RL[1] * q.y + RL[2] * q.z + cm.x, RL[3] * q.x + RL[4] * q.y + RL[5] * q.z + cm.y,
                      RL[6] * q.x + RL[7] * q.y + RL[8] * q.z + cm.z};
                V3 np = xs[p] + S.alpha * (gp - xs[p]);
                acc[p] = acc[p] + np;
                cnt[p]++;
            }
        }
        for (int i = 0; i < n; ++i)
            if (cnt[i] > 0) {
                V3 a = (1.0 / cnt[i]) * acc[i];
                xs[i] = {R<T>(a.x), R<T>(a.y), R<T>(a.z)};
            }
    }
    for (int i = 0; i < n; ++i) {
        V3 v = (1.0 / dt) * (xs[i] - x[i]);
        st(vel, i, v);
        st(pos, i, xs[i]);
    }
}

// ------------------------------------------------------------------------------------------ legacy ice launches
// The reference's per-cluster slot layout (Ice_SM.h:23-53, IceObject.h:370-395): member slots of cluster c are
// indxSet[2c] .. indxSet[2c+1] (inclusive), pIdx[slot] = particle, orgPos/curPos/vel 3 values per slot,
// orgCm/curCm 3 per cluster, particle arrays prtPos/prtVel 4 per particle, sldPos 3 per particle.  Unit masses.
//
// itr = false: OrientedPrtObj::UpdateCluster_SM (ElasticObject_OP.cpp:551-720): prediction x + (v + g dt) dt
//   with the wall response, rigid matching (alpha = 1), goal written to curPos, vel = (goal - x) / dt.
// itr = true: OrientedPrtObj::UpdateCluster_SM_Itr (ElasticObject_OP.cpp:722-887): curPos = sldPos of the
//   member, rigid matching with Apq's third column negated when det Apq < 0 and the cluster has at least 10
//   members, goal written to curPos and clamped to the walls; vel untouched.
// A cluster of one member keeps its first positions (both return before the matching).
template <class T> void legacy_cluster(int nc, const T* prtPos, const T* prtVel, const T* sldPos, const T* orgPos, T* curPos,
                                       const T* orgCm, T* curCm, T* vel, const int* pIdx, const int* indxSet, double dt,
                                       const double* g, const double* lo, const double* hi, double res, bool itr) {
    for (int c = 0; c < nc; ++c) {
        const int s = indxSet[2 * c], e = indxSet[2 * c + 1];
        for (int i = s; i <= e; ++i) {
            const int p = pIdx[i];
            if (itr) {
                for (int a = 0; a < 3; ++a) curPos[3 * i + a] = sldPos[3 * p + a];
                continue;
            }
            double x[3], v[3], cp[3];
            for (int a = 0; a < 3; ++a) {
                x[a] = prtPos[4 * p + a];
                v[a] = prtVel[4 * p + a];
                cp[a] = R<T>(x[a] + (v[a] + g[a] * dt) * dt);
            }
            for (int a = 0; a < 3; ++a)
                if (cp[a] < lo[a] || cp[a] > hi[a]) {
                    cp[a] = R<T>(x[a] - (v[a] + g[a] * dt) * dt * res);
                    for (int o = 0; o < 3; ++o)
                        if (o != a) cp[o] = x[o];
                }
            for (int a = 0; a < 3; ++a) curPos[3 * i + a] = (T)std::min(std::max(cp[a], lo[a]), hi[a]);
        }
        const int k = e - s + 1;
        if (k <= 1) continue;
        double cm[3] = {0, 0, 0};
        for (int i = s; i <= e; ++i)
            for (int a = 0; a < 3; ++a) cm[a] += (double)curPos[3 * i + a];
        for (int a = 0; a < 3; ++a) cm[a] /= k;
        double apq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = s; i <= e; ++i) {
            double pp[3], qq[3];
            for (int a = 0; a < 3; ++a) {
                pp[a] = (double)curPos[3 * i + a] - cm[a];
                qq[a] = (double)orgPos[3 * i + a] - (double)orgCm[3 * c + a];
            }
            for (int r = 0; r < 3; ++r)
                for (int cc = 0; cc < 3; ++cc) apq[r * 3 + cc] += pp[r] * qq[cc];
        }
        if (itr && k >= 10 && det3(apq) < 0.0)
            for (int r = 0; r < 3; ++r) apq[r * 3 + 2] = -apq[r * 3 + 2];
        double Rm[9];
        polar(apq, Rm, nullptr);
        for (int i = s; i <= e; ++i) {
            double qq[3];
            for (int a = 0; a < 3; ++a) qq[a] = (double)orgPos[3 * i + a] - (double)orgCm[3 * c + a];
            for (int a = 0; a < 3; ++a) {
                double gp = Rm[a * 3] * qq[0] + Rm[a * 3 + 1] * qq[1] + Rm[a * 3 + 2] * qq[2] + cm[a];
                if (itr) gp = std::min(std::max(R<T>(gp), lo[a]), hi[a]);
                curPos[3 * i + a] = (T)gp;
            }
        }
        for (int a = 0; a < 3; ++a) curCm[3 * c + a] = (T)cm[a];
        if (!itr)
            for (int i = s; i <= e; ++i) {
                const int p = pIdx[i];
                for (int a = 0; a < 3; ++a) vel[3 * i + a] = (T)(((double)curPos[3 * i + a] - (double)prtPos[4 * p + a]) / dt);
            }
    }
}

// ConvoNormal::StepConvolution (Ice_Convolution_Normal.cpp:32-67) in the PtoC layout of LaunchCalcAverageGPU
// (IceStructure.cu:63-126): particle p belongs to PtoCIndx[p] entries; entry j names cluster
// PtoC[base + j] and member number PtoC[base + PtoCMax + j] (base = PtoCMax * PtoCParamSize * p, -1 = none).
// The mean member position and velocity over the clusters holding p become sldPos / sldVel of p; a particle in
// no cluster keeps its values.
template <class T> void legacy_average(int n, T* sldPos, T* sldVel, const T* smPos, const T* smVel, const int* indxSet,
                                       const int* PtoCIndx, const int* PtoC, int PtoCMax, int PtoCParamSize) {
    for (int p = 0; p < n; ++p) {
        const long base = (long)PtoCMax * PtoCParamSize * p;
        const int cnt = std::min(PtoCIndx[p], PtoCMax);
        double sp[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
        int k = 0;
        for (int j = 0; j < cnt; ++j) {
            const int c = PtoC[base + j], o = PtoC[base + PtoCMax + j];
            if (c == -1 || o == -1) continue;
            const long slot = (long)indxSet[2 * c] + o;
            for (int a = 0; a < 3; ++a) {
                sp[a] += (double)smPos[3 * slot + a];
                sv[a] += (double)smVel[3 * slot + a];
            }
            ++k;
        }
        if (k == 0) continue;
        for (int a = 0; a < 3; ++a) {
            sldPos[3 * p + a] = (T)(sp[a] / k);
            sldVel[3 * p + a] = (T)(sv[a] / k);
        }
    }
}

// ------------------------------------------------------------------------------------------ MLP
inline double silu(double z) { return z / (1.0 + std::exp(-z)); }
inline double dsilu(double z) { double s = 1.0 / (1.0 + std::exp(-z)); return s * (1.0 + z * (1.0 - s)); }

}  // namespace


// ============================================================================================ C API
extern "C" {

double orc_kernel_coef(int kind, double h, int type) { return coef(kind, h, type); }
double orc_kernel_poly6(double r, double h, double a) { return W_poly6(r, h, a); }
double orc_kernel_spiky_grad(double r, double h, double a) { return GW_spiky_s(r, h, a); }
double orc_kernel_visc_lap(double r, double h, double a) { return LW_visc(r, h, a); }

// rxSPH::calRestDensity (rx_sph_cpu.cpp:706-720) with lattice spacing l = 2 r = spacing.
double orc_rest_density(double mass, double h, double spacing) {
    double r0 = 0.0, aw = coef(0, h, 1);
    int nn = (int)std::ceil(h / spacing) + 1;
    for (int x = -nn; x <= nn; ++x)
        for (int y = -nn; y <= nn; ++y)
            for (int z = -nn; z <= nn; ++z) {
                V3 rij{x * spacing, y * spacing, z * spacing};
                r0 += mass * W_poly6(norm(rij), h, aw);
            }
    return r0;
}

void orc_calc_hash_f(const float* pos4, const int* attr, int n, const OrcParams* P, uint32_t* hash, uint32_t* index) {
    for (int i = 0; i < n; ++i) {
        hash[i] = (attr && attr[i] == -1) ? 2147483647u : hash_of(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], *P);
        index[i] = (uint32_t)i;
    }
}
void orc_stable_sort(uint32_t* hash, uint32_t* index, int n) { Grid<float>::sort_stable(hash, index, n); }
void orc_cell_bounds(const uint32_t* h, int n, int ncells, uint32_t* start, uint32_t* end) {
    std::fill(start, start + ncells, 0xFFFFFFFFu);
    std::fill(end, end + ncells, 0xFFFFFFFFu);
    for (int i = 0; i < n; ++i) {
        if (h[i] >= (uint32_t)ncells) continue;
        if (i == 0 || h[i] != h[i - 1]) {
            start[h[i]] = (uint32_t)i;
            if (i > 0 && h[i - 1] < (uint32_t)ncells) end[h[i - 1]] = (uint32_t)i;
        }
        if (i == n - 1 || h[i + 1] != h[i]) end[h[i]] = (uint32_t)(i + 1);
    }
}

long orc_neighbors_brute_f(const float* p, int n, float h, int maxn, int* counts, int* list) {
    long tot = 0;
    for (int i = 0; i < n; ++i) {
        int c = 0;
        for (int j = 0; j < n; ++j) {
            float dx = p[4 * i] - p[4 * j], dy = p[4 * i + 1] - p[4 * j + 1], dz = p[4 * i + 2] - p[4 * j + 2];
            float r = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (r <= h) {
                if (list && c < maxn) list[(long)i * maxn + c] = j;
                c++;
            }
        }
        counts[i] = c;
        tot += c;
    }
    return tot;
}

void orc_sph_density_f(const float* pos4, int n, const OrcParams* P, float* dens, float* pres) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    sph_density(pos4, n, nb, *P, dens);
    double k = R<float>(P->gas_k), r0 = R<float>(P->rest_dens);
    if (pres)
        for (int i = 0; i < n; ++i) pres[i] = (float)(k * ((double)dens[i] - r0));
}
void orc_sph_force_f(const float* pos4, const float* vel4, const float* dens, const float* pres, int n, const OrcParams* P, float* acc4) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    std::vector<float> pr(pres, pres + n);
    sph_force(pos4, vel4, dens, pr.data(), n, nb, *P, acc4);
}
void orc_sph_step_f(float* pos4, float* vel4, int n, const OrcParams* P, float dt, float* dens, float* pres, float* acc4) {
    sph_step(pos4, vel4, n, *P, (double)dt, dens, pres, acc4);
}
void orc_sph_step_d(double* pos4, double* vel4, int n, const OrcParams* P, double dt, double* dens) {
    std::vector<double> pres(n), acc(4 * (size_t)n);
    sph_step(pos4, vel4, n, *P, dt, dens, pres.data(), acc.data());
}

void orc_pbf_density_f(const float* pos4, int n, const OrcParams* P, float* dens) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    sph_density(pos4, n, nb, *P, dens);  // rxPBDSPH::calDensity (rx_sph_pbd.cpp:599-646) is the WCSPH sum
}
void orc_pbf_ext_force_f(const float* pos4, const float* vel4, const float* dens, int n, const OrcParams* P, float* acc4) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    pbf_ext_force(pos4, vel4, dens, n, nb, *P, acc4);
}
void orc_pbf_scaling_factor_f(const float* pos4, int n, const OrcParams* P, float* dens, float* lambda) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    pbf_lambda(pos4, n, nb, *P, dens, lambda);
}
void orc_pbf_position_correction_f(const float* pos4, const float* lambda, int n, const OrcParams* P, float dt, float* dp4) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    pbf_dp(pos4, lambda, n, nb, *P, (double)dt, dp4);
}
void orc_pbf_xsph_f(const float* pos4, const float* vel4, const float* dens, int n, const OrcParams* P, float c, float* out_vel4) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    pbf_xsph(pos4, vel4, dens, n, nb, *P, (double)c, out_vel4);
}
void orc_boundary_volume_f(const float* b4, int nb, const OrcParams* P, float* vol) { boundary_volumes(b4, nb, *P, vol); }
void orc_boundary_density_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P, float* dens,
                            float* pres) {
    Boundary<float> B{b4, vol, nb};
    std::vector<int> bn;
    double k = R<float>(P->gas_k), r0 = R<float>(P->rest_dens);
    for (int i = 0; i < n; ++i) {
        dens[i] = (float)((double)dens[i] + (double)boundary_density(ld(pos4, i), B, *P, bn));
        if (pres) pres[i] = (float)(k * ((double)dens[i] - r0));
    }
}
void orc_boundary_force_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P,
                          const float* dens, const float* pres, float* acc4) {
    boundary_force(pos4, n, Boundary<float>{b4, vol, nb}, *P, dens, pres, acc4);
}
void orc_pbf_scaling_factor_wb_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P,
                                 float* dens, float* lambda, float* lambda_b) {
    std::vector<std::vector<Neigh>> nbl;
    neighbour_lists(pos4, n, *P, nbl);
    Boundary<float> B{b4, vol, nb};
    pbf_lambda(pos4, n, nbl, *P, dens, lambda, &B);
    if (lambda_b) pbf_boundary_lambda(pos4, n, B, *P, lambda_b);
}
void orc_pbf_position_correction_wb_f(const float* pos4, const float* lambda, int n, const float* b4, const float* vol,
                                      const float* lambda_b, int nb, const OrcParams* P, float dt, float* dp4) {
    std::vector<std::vector<Neigh>> nbl;
    neighbour_lists(pos4, n, *P, nbl);
    Boundary<float> B{b4, vol, nb};
    pbf_dp(pos4, lambda, n, nbl, *P, (double)dt, dp4, &B, lambda_b);
}
void orc_integrate_polygon_f(const float* pos4, const float* vel4, const float* acc4, const int* attr, int n, const OrcParams* P,
                             const float* vrts, const int* tris, int ntri, const uint32_t* pstart, const uint32_t* pend,
                             const uint32_t* pidx, double restitution, float dt, float* out_pos4, float* out_vel4) {
    Polys Q{vrts, tris, ntri, pstart, pend, pidx, restitution};
    for (int i = 0; i < n; ++i) {
        if (attr && attr[i] == -1) {
            for (int k = 0; k < 4; ++k) { out_pos4[4 * i + k] = pos4[4 * i + k]; out_vel4[4 * i + k] = vel4[4 * i + k]; }
            continue;
        }
        integrate(pos4 + 4 * i, vel4 + 4 * i, acc4 + 4 * i, out_pos4 + 4 * i, out_vel4 + 4 * i, 1, *P, (double)dt, &Q);
    }
}

// rxSPH::CalGridDensity's GPU kernels (rx_sph_kernel.cu:1039-1126): sum_j m W_poly6 at the lattice nodes
// (x0 + i dx, y0 + j dy, z0 + k dz), node (i, j, k) at i + j nx + k nx ny; particles with ice[j] < 0 are left
// out when `ice` is given.  A scan of all particles (Dist2 <= h^2 in float).
void orc_grid_density_f(const float* pos4, int n, const OrcParams* P, int nx, int ny, int nz, float x0, float y0, float z0, float dx,
                        float dy, float dz, const float* ice, float* out) {
    double h = R<float>(P->h), m = R<float>(P->mass), aw = R<float>(coef(0, h, 1)), h2 = R<float>(h * h);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                V3 g{(double)(float)((double)x0 + (double)i * dx), (double)(float)((double)y0 + (double)j * dy),
                     (double)(float)((double)z0 + (double)k * dz)};
                float d = 0;
                for (int q = 0; q < n; ++q) {
                    if (ice && ice[q] < 0.0f) continue;
                    V3 r = g - ld(pos4, q);
                    double r2 = R<float>(dot(r, r));
                    if (r2 <= h2) d = (float)((double)d + m * W_poly6(std::sqrt(r2), h, aw));
                }
                out[(size_t)i + (size_t)j * nx + (size_t)k * nx * ny] = d;
            }
}

// Per-particle magnitudes the FP32 results are judged against (sum over the terms of |term|; see the templates).
void orc_sph_force_scale_f(const float* pos4, const float* vel4, const float* dens, const float* pres, int n, const OrcParams* P,
                           double* scale) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    std::vector<float> pr(pres, pres + n), acc(4 * (size_t)n);
    sph_force(pos4, vel4, dens, pr.data(), n, nb, *P, acc.data(), scale);
}
void orc_pbf_ext_force_scale_f(const float* pos4, const float* vel4, const float* dens, int n, const OrcParams* P, double* scale) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    std::vector<float> acc(4 * (size_t)n);
    pbf_ext_force(pos4, vel4, dens, n, nb, *P, acc.data(), scale);
}
void orc_boundary_force_scale_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P,
                                const float* dens, const float* pres, double* scale) {
    std::vector<float> acc(4 * (size_t)n);
    boundary_force(pos4, n, Boundary<float>{b4, vol, nb}, *P, dens, pres, acc.data(), scale);
}
void orc_pbf_lambda_scale_f(const float* pos4, int n, const float* b4, const float* vol, int nb, const OrcParams* P, double* scale,
                            double* scale_b) {
    std::vector<std::vector<Neigh>> nbl;
    neighbour_lists(pos4, n, *P, nbl);
    std::vector<float> d(n), l(n), lb(nb > 0 ? nb : 1);
    Boundary<float> B{b4, vol, nb};
    pbf_lambda(pos4, n, nbl, *P, d.data(), l.data(), nb > 0 ? &B : nullptr, scale);
    if (nb > 0 && scale_b) pbf_boundary_lambda(pos4, n, B, *P, lb.data(), scale_b);
}
void orc_pbf_dp_scale_f(const float* pos4, const float* lambda, int n, const float* b4, const float* vol, const float* lambda_b, int nb,
                        const OrcParams* P, float dt, double* scale) {
    std::vector<std::vector<Neigh>> nbl;
    neighbour_lists(pos4, n, *P, nbl);
    std::vector<float> dp(4 * (size_t)n);
    Boundary<float> B{b4, vol, nb};
    pbf_dp(pos4, lambda, n, nbl, *P, (double)dt, dp.data(), nb > 0 ? &B : nullptr, lambda_b, scale);
}

void orc_pbf_step_f(float* pos4, float* vel4, int n, const OrcParams* P, float* dens, float* lambda_sum, OrcPbfStats* st) {
    pbf_step(pos4, vel4, n, *P, dens, lambda_sum, st);
}
void orc_pbf_step_d(double* pos4, double* vel4, int n, const OrcParams* P, double* dens, OrcPbfStats* st) {
    pbf_step(pos4, vel4, n, *P, dens, (double*)nullptr, st);
}

int orc_jacobi_eigen3(double* a9, double* v9, double eps, int iter_max) { return jacobi3(a9, v9, eps, iter_max); }
void orc_polar_decomposition(const double* A9, double* R9, double* S9) { polar(A9, R9, S9); }

void orc_sm_step_f(float* pos4, float* vel4, int n, const int* cs, const int* mem, int nc, const float* rest, const OrcSmParams* S) {
    sm_step(pos4, vel4, n, cs, mem, nc, rest, *S);
}
void orc_sm_step_masked_f(float* pos4, float* vel4, int n, const int* cs, const int* mem, int nc, const float* rest,
                          const OrcSmParams* S, const int* active) {
    sm_step(pos4, vel4, n, cs, mem, nc, rest, *S, active);
}
void orc_sm_step_d(double* pos4, double* vel4, int n, const int* cs, const int* mem, int nc, const double* rest, const OrcSmParams* S) {
    sm_step(pos4, vel4, n, cs, mem, nc, rest, *S);
}
void orc_legacy_cluster_f(int nc, const float* prtPos4, const float* prtVel4, const float* sldPos3, const float* orgPos, float* curPos,
                          const float* orgCm, float* curCm, float* vel, const int* pIdx, const int* indxSet, float dt,
                          const OrcSmParams* S, int itr) {
    legacy_cluster(nc, prtPos4, prtVel4, sldPos3, orgPos, curPos, orgCm, curCm, vel, pIdx, indxSet, (double)dt, S->gravity, S->bmin,
                   S->bmax, S->restitution, itr != 0);
}
void orc_legacy_cluster_d(int nc, const double* prtPos4, const double* prtVel4, const double* sldPos3, const double* orgPos,
                          double* curPos, const double* orgCm, double* curCm, double* vel, const int* pIdx, const int* indxSet,
                          double dt, const OrcSmParams* S, int itr) {
    legacy_cluster(nc, prtPos4, prtVel4, sldPos3, orgPos, curPos, orgCm, curCm, vel, pIdx, indxSet, dt, S->gravity, S->bmin, S->bmax,
                   S->restitution, itr != 0);
}
void orc_legacy_average_f(int n, float* sldPos, float* sldVel, const float* smPos, const float* smVel, const int* indxSet,
                          const int* PtoCIndx, const int* PtoC, int PtoCMax, int PtoCParamSize) {
    legacy_average(n, sldPos, sldVel, smPos, smVel, indxSet, PtoCIndx, PtoC, PtoCMax, PtoCParamSize);
}

// SPH heat diffusion over the r <= h neighbours: T_i' = T_i + dt * kappa * sum_j (m/rho_j)(T_j - T_i) LW_visc.
// (HeatTransfer.cu:100-123 diffuses over a neighbour list with the distance fixed to 1; this restatement
// uses the true distance.)
void orc_heat_diffuse_f(const float* pos4, const float* dens, const float* temp, int n, const OrcParams* P,
                        float diffusivity, float dt, float* temp_out) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    double h = R<float>(P->h), m = R<float>(P->mass), al = R<float>(coef(2, h, 3));
    for (int i = 0; i < n; ++i) {
        V3 p0 = ld(pos4, i);
        double s = 0.0;
        for (const Neigh& e : nb[i]) {
            int j = e.idx;
            if (j == i) continue;
            double r = norm(p0 - ld(pos4, j));
            s += m / (double)dens[j] * ((double)temp[j] - (double)temp[i]) * LW_visc(r, h, al);
        }
        temp_out[i] = (float)((double)temp[i] + (double)dt * (double)diffusivity * s);
    }
}

// HeatTransfar::calcTempAndHeat (HeatTransfar.cpp:280-350) for one particle; th = (T, heat, phase, changed).
static void heat_phase(float* th, float dT) {
    const float latent = 165.0f;
    float t = th[0], q = th[1], chg = th[3];
    int ph = (int)th[2];
    if (ph == -2 && t > 250.0f) { ph = -1; t = 250.0f; q = 0.0f; }
    else if (ph == 2 && t < 250.0f) { ph = 1; t = 250.0f; q = latent; }
    if (ph == -1 || ph == 1) {
        float c = 2.1f + (2.1f * q / latent);
        q += dT * c * 1.0f;
        if (ph == -1 && q < -latent / 300.0f) { ph = -2; t = 249.0f; q = 0.0f; }
        else if (ph == 1 && q > latent) { ph = 2; t = 251.0f; q = 0.0f; }
        if (ph == -1 && q > latent) { ph = 2; t = 251.0f; q = 0.0f; chg = 1.0f; }
        else if (ph == 1 && q < 0.0f) { ph = -2; t = 249.0f; q = 0.0f; chg = 1.0f; }
    } else {
        float c = (t > 250.0f) ? 4.2f : 2.1f;
        t = t + q / (c * 1.0f);
        q = 0.0f;
        t += dT;
        t = std::min(std::max(t, 0.0f), 500.0f);
    }
    th[0] = t; th[1] = q; th[2] = (float)ph; th[3] = chg;
}

// HeatTransfar::heatParticleAndParticle (HeatTransfar.cpp:429-473) over the neighbours within h (true
// distances), then calcTempAndHeat for every particle.  therm4 = (T, heat, phase, changed) per particle.
void orc_heat_cpu_f(const float* pos4, const float* dens, float* therm4, int n, const OrcParams* P, float td, float dt) {
    std::vector<std::vector<Neigh>> nb;
    neighbour_lists(pos4, n, *P, nb);
    double h = R<float>(P->h), ws = 10.0 / (7.0 * kPi * h * h);
    std::vector<float> dT(n);
    for (int i = 0; i < n; ++i) {
        double tmp = 0.0;
        V3 p0 = ld(pos4, i);
        for (const Neigh& e : nb[i]) {
            int j = e.idx;
            if (j == i) continue;
            double r = norm(p0 - ld(pos4, j)), q = r / h;
            double w = q < 1.0 ? ws * (1.0 - 1.5 * q * q + 0.75 * q * q * q) : (q < 2.0 ? ws * 0.25 * (2.0 - q) * (2.0 - q) * (2.0 - q) : 0.0);
            float d = dens[j] < 0.1f ? 0.1f : dens[j];
            tmp += (double)dt * ((double)therm4[4 * j] - (double)therm4[4 * i]) / d * w;
        }
        dT[i] = (float)(tmp * td);
    }
    for (int i = 0; i < n; ++i) heat_phase(&therm4[4 * i], dT[i]);
}

// HeatTransfer.cu:62-75
static float heat_spline(float d, float r, float w) {
    float q = d / r;
    if (q >= 0.0f && q < 1.0f) return w * (1.0f - 1.5f * q * q + 0.75f * q * q * q);
    if (q >= 1.0f && q < 2.0f) return w * 0.25f * (2.0f - q) * (2.0f - q) * (2.0f - q);
    return 0.0f;
}

// LaunchHeatTransferGPU (HeatTransfer.cu:44-57) and its four kernels (:79-204), one particle at a time.
void orc_heat_transfer_gpu_f(float* heats, float* temps, float* dtemps, const int* surf, const int* neighs, const float* dens,
                             int* phase, int* changed, const int* melt, float radius, float air, float ht, float td,
                             float dt, int n) {
    const float pi = 3.14159265359f, latent = 165.0f;
    float wspline = 10.0f / (7.0f * pi * (float)std::pow((double)radius, 2.0));
    for (int i = 0; i < n; ++i) {  // UpdateAirAndParticle (:79-96)
        if (surf[i] <= 0) continue;
        float an = 20.0f - (float)surf[i];
        if (an < 0) an = 0.0f;
        float area = an / 20.0f;
        area = std::min(std::max(area, 0.0f), 1.0f);
        heats[i] += ht * (air - temps[i]) * area;
    }
    for (int i = 0; i < n; ++i) {  // UpdateParticleAndParticle (:100-123)
        float tmp = 0.0f, now = temps[i];
        int nn = neighs[i * 30 + 29];
        for (int j = 0; j < nn && j < 30; ++j) {
            int id = neighs[i * 30 + j];
            if (id == i) continue;
            float d = dens[id];
            if (d < 0.05f) d = 0.05f;
            tmp += dt * 1.0f * (temps[id] - now) / d * heat_spline(1.0f, radius, wspline);
        }
        dtemps[i] = tmp * td;
    }
    if (melt)  // UpdateMeltParticle (:127-137): threads 1..999 of the launch
        for (int k = 1; k < 1000 && k < n; ++k)
            if (melt[k] >= 0 && melt[k] < n) heats[melt[k]] += 300.0f;
    for (int i = 0; i < n; ++i) {  // UpdatePhase (:140-204)
        if (phase[i] == -2 && temps[i] > 250.0f) {
            phase[i] = -1; temps[i] = 250.0f; heats[i] = 0;
        } else if (phase[i] == 2 && temps[i] < 250.0f) {
            phase[i] = 1; temps[i] = 250.0f; heats[i] = latent;
        }
        if (phase[i] == -1 || phase[i] == 1) {
            float c = 2.1f + (2.1f * heats[i] / latent);
            heats[i] += dtemps[i] * c * 1.0f;
            if (phase[i] == -1 && heats[i] < -latent / 300.0f) {
                phase[i] = -2; temps[i] = 249.0f; heats[i] = 0;
            } else if (phase[i] == 1 && heats[i] > latent) {
                phase[i] = 2; temps[i] = 251.0f; heats[i] = 0;
            }
            if (phase[i] == -1 && heats[i] > latent) {
                phase[i] = 2; temps[i] = 251.0f; heats[i] = 0; changed[i] = 1;
            } else if (phase[i] == 1 && heats[i] < 0.0f) {
                phase[i] = -2; temps[i] = 249.0f; heats[i] = 0; changed[i] = 1;
            }
        } else {
            float c = (temps[i] > 250.0f) ? 4.2f : 2.1f;
            temps[i] = temps[i] + heats[i] / (c * 1.0f);
            heats[i] = 0.0f;
            temps[i] += dtemps[i];
            if (temps[i] > 500.0f) temps[i] = 500.0f;
            if (temps[i] < 0.0f) temps[i] = 0.0f;
        }
    }
}

// detectNeighborParticles (rx_sph_kernel.cu:1134-1208) by brute force: the GPU visits the cells of the box
// [x - h, x + h] in z, y, x order and each cell in sorted (stable) order, i.e. candidates by (hash, index).
void orc_detect_neighbors_f(const float* p, int n, const OrcParams* P, int* out) {
    std::vector<uint32_t> hash(n);
    for (int i = 0; i < n; ++i) hash[i] = hash_of(p[4 * i], p[4 * i + 1], p[4 * i + 2], *P);
    std::vector<std::pair<uint32_t, int>> c;
    float h = (float)P->h;
    for (int i = 0; i < n; ++i) {
        c.clear();
        for (int j = 0; j < n; ++j) {
            float dx = p[4 * i] - p[4 * j], dy = p[4 * i + 1] - p[4 * j + 1], dz = p[4 * i + 2] - p[4 * j + 2];
            float r = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (r <= h) c.push_back({hash[j], j});
        }
        std::sort(c.begin(), c.end());
        int cnt = std::min<int>((int)c.size(), 30);
        for (int k = 0; k < 29; ++k) out[i * 30 + k] = k < cnt ? c[k].second : -1;
        out[i * 30 + 29] = cnt;  // with 30 hits the 30th neighbour is overwritten by the count
    }
}

// CalDistToNormalizedMassCenter + detectSurfaceParticles (rx_sph_kernel.cu:1214-1290)
void orc_detect_surface_f(const int* nb, int* surf, const float* p, int n, float mass, float radius) {
    for (int i = 0; i < n; ++i) {
        int nn = nb[i * 30 + 29];
        float sx = 0, sy = 0, sz = 0, sm = 0;
        for (int t = 0; t < nn - 1; ++t) {
            int j = nb[i * 30 + t];
            sx += mass * (p[4 * i] - p[4 * j]);
            sy += mass * (p[4 * i + 1] - p[4 * j + 1]);
            sz += mass * (p[4 * i + 2] - p[4 * j + 2]);
            sm += mass;
        }
        float d = 1.17549435e-38f;
        if (sm > 0.0f) {
            sx /= sm; sy /= sm; sz /= sm;
            d = std::sqrt(sx * sx + sy * sy + sz * sz);
        }
        if (nn <= 3) surf[i] = nn;
        else if (surf[i] > 0) surf[i] = (d < 0.25f * radius) ? 0 : nn;
        else surf[i] = (d > 0.35f * radius || d < 0.0f) ? nn : 0;
    }
}

void orc_mlp_forward_d(const double* X, int nrows, int nin, const double* W1, const double* b1, int nh1,
                       const double* W2, const double* b2, int nh2, const double* W3, const double* b3, double* Y) {
    std::vector<double> h1(nh1), h2(nh2);
    for (int r = 0; r < nrows; ++r) {
        const double* x = X + (size_t)r * nin;
        for (int o = 0; o < nh1; ++o) {
            double s = b1[o];
            for (int i = 0; i < nin; ++i) s += W1[o * nin + i] * x[i];
            h1[o] = silu(s);
        }
        for (int o = 0; o < nh2; ++o) {
            double s = b2[o];
            for (int i = 0; i < nh1; ++i) s += W2[o * nh1 + i] * h1[i];
            h2[o] = silu(s);
        }
        double s = b3[0];
        for (int i = 0; i < nh2; ++i) s += W3[i] * h2[i];
        Y[r] = s;
    }
}

void orc_mlp_backward_d(const double* X, const double* T, int nrows, int nin,
                        const double* W1, const double* b1, int nh1, const double* W2, const double* b2, int nh2,
                        const double* W3, const double* b3,
                        double* gW1, double* gb1, double* gW2, double* gb2, double* gW3, double* gb3, double* loss) {
    std::fill(gW1, gW1 + nh1 * nin, 0.0);
    std::fill(gb1, gb1 + nh1, 0.0);
    std::fill(gW2, gW2 + nh2 * nh1, 0.0);
    std::fill(gb2, gb2 + nh2, 0.0);
    std::fill(gW3, gW3 + nh2, 0.0);
    gb3[0] = 0.0;
    double L = 0.0;
    std::vector<double> z1(nh1), h1(nh1), z2(nh2), h2(nh2), d2(nh2), d1(nh1);
    for (int r = 0; r < nrows; ++r) {
        const double* x = X + (size_t)r * nin;
        for (int o = 0; o < nh1; ++o) {
            double s = b1[o];
            for (int i = 0; i < nin; ++i) s += W1[o * nin + i] * x[i];
            z1[o] = s;
            h1[o] = silu(s);
        }
        for (int o = 0; o < nh2; ++o) {
            double s = b2[o];
            for (int i = 0; i < nh1; ++i) s += W2[o * nh1 + i] * h1[i];
            z2[o] = s;
            h2[o] = silu(s);
        }
        double y = b3[0];
        for (int i = 0; i < nh2; ++i) y += W3[i] * h2[i];
        double e = y - T[r];
        L += 0.5 * e * e;
        double dy = e / nrows;
        gb3[0] += dy;
        for (int i = 0; i < nh2; ++i) {
            gW3[i] += dy * h2[i];
            d2[i] = dy * W3[i] * dsilu(z2[i]);
            gb2[i] += d2[i];
        }
        for (int o = 0; o < nh2; ++o)
            for (int i = 0; i < nh1; ++i) gW2[o * nh1 + i] += d2[o] * h1[i];
        for (int i = 0; i < nh1; ++i) {
            double s = 0;
            for (int o = 0; o < nh2; ++o) s += d2[o] * W2[o * nh1 + i];
            d1[i] = s * dsilu(z1[i]);
            gb1[i] += d1[i];
        }
        for (int o = 0; o < nh1; ++o)
            for (int i = 0; i < nin; ++i) gW1[o * nin + i] += d1[o] * x[i];
    }
    *loss = L / nrows;
}

}  // extern "C"

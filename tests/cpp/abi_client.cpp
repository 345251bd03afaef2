// A C++ caller of librxcu.so through the public headers only (include/*.h), the way the reference's host code
// calls the library once it is linked against it.  The tests compile it with hipcc and compare what it writes
// with the ctypes path and the oracle.
//
//   abi_client layout          one line per ABI struct size / field offset: "<name> <this compiler> <library>"
//   abi_client scene <out>     rxSphEnvDefault, rxSetupSimParams and rxSceneBox (CPU only): writes the env, the
//                              parameters and a 12 x 10 x 8 seeded block
//   abi_client grid <in> <out> CuSetParameters, CuCalcHash, CuSort, CuReorderDataAndFindCellStart
//                              in:  rxSimParams bytes, int32 n, n float4 positions, n float4 velocities
//                              out: n raw hashes, n sorted hashes, n sorted indices, NumCells starts, NumCells ends
//   abi_client pbf <in> <out>  rxPBDSPH_GPU_Create / SetParticles / Update / GetArray
//                              in:  rxSphEnv bytes, int32 n, int32 steps, float dt, n float4 pos, n float4 vel
//                              out: n float4 positions, n float4 velocities (original particle order)
// Exit status 0 on success; a failed library call prints CuGetLastErrorString and exits 1.
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rx_cu_funcs.h"
#include "rx_sph_host.h"

namespace {

int fail(const char* what) {
    std::fprintf(stderr, "abi_client: %s failed: %s\n", what, CuGetLastErrorString());
    return 1;
}

bool read_all(const char* path, std::vector<unsigned char>& buf) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize((size_t)len);
    bool ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
    std::fclose(f);
    return ok;
}

struct Writer {
    FILE* f;
    explicit Writer(const char* path) : f(std::fopen(path, "wb")) {}
    ~Writer() {
        if (f) std::fclose(f);
    }
    void put(const void* p, size_t n) { std::fwrite(p, 1, n, f); }
};

int layout() {
    struct Row {
        const char* name;
        long ours;
        int which;
    } rows[] = {
        {"sizeof_rxSimParams", (long)sizeof(rxSimParams), 0},
        {"sizeof_rxParticleCell", (long)sizeof(rxParticleCell), 1},
        {"offsetof_EffectiveRadius", (long)offsetof(rxSimParams, EffectiveRadius), 2},
        {"offsetof_AP", (long)offsetof(rxSimParams, AP), 3},
        {"offsetof_BoxNum", (long)offsetof(rxSimParams, BoxNum), 4},
        {"offsetof_SphereFlg", (long)offsetof(rxSimParams, SphereFlg), 5},
        {"offsetof_uNumParticles", (long)offsetof(rxParticleCell, uNumParticles), 6},
        {"offsetof_uNumPolyHash", (long)offsetof(rxParticleCell, uNumPolyHash), 7},
    };
    for (const Row& r : rows) std::printf("%s %ld %ld\n", r.name, r.ours, rxAbiLayout(r.which));
    std::printf("sizeof_rxSphEnv %ld -1\n", (long)sizeof(rxSphEnv));
    return 0;
}

int scene(const char* out) {
    rxSphEnv e;
    rxSphEnvDefault(&e);
    const int nx = 12, ny = 10, nz = 8, n = nx * ny * nz;
    e.max_particles = n;
    e.h = 0.1f;
    e.spacing = 0.05f;
    for (int k = 0; k < 3; ++k) {
        e.boundary_cen[k] = 0.5f;
        e.boundary_ext[k] = 0.5f;
    }
    rxSimParams p;
    if (rxSetupSimParams(&e, &p)) return fail("rxSetupSimParams");
    std::vector<float> pos(4 * (size_t)n), vel(4 * (size_t)n);
    const float origin[3] = {0.1f, 0.05f, 0.2f};
    const int got = rxSceneBox(pos.data(), vel.data(), n, origin, nx, ny, nz, 0.05f, 0.0005f, 0.05f, 7u);
    if (got != n) return fail("rxSceneBox");
    Writer w(out);
    w.put(&e, sizeof(e));
    w.put(&p, sizeof(p));
    w.put(&n, sizeof(n));
    w.put(pos.data(), pos.size() * sizeof(float));
    w.put(vel.data(), vel.size() * sizeof(float));
    return 0;
}

int grid(const char* in, const char* out) {
    std::vector<unsigned char> buf;
    if (!read_all(in, buf) || buf.size() < sizeof(rxSimParams) + sizeof(int)) return fail("reading the input");
    rxSimParams p;
    std::memcpy(&p, buf.data(), sizeof(p));
    int n;
    std::memcpy(&n, buf.data() + sizeof(p), sizeof(n));
    if (buf.size() != sizeof(p) + sizeof(n) + 32 * (size_t)n) return fail("input size");
    const float* hpos = (const float*)(buf.data() + sizeof(p) + sizeof(n));
    const float* hvel = hpos + 4 * (size_t)n;
    const uint nc = p.NumCells;

    CuSetParameters(&p);
    float *dpos = nullptr, *dvel = nullptr;
    rxParticleCell c;
    std::memset(&c, 0, sizeof(c));
    CuAllocateArray((void**)&dpos, 16 * n);
    CuAllocateArray((void**)&dvel, 16 * n);
    CuAllocateArray((void**)&c.dSortedPos, 16 * n);
    CuAllocateArray((void**)&c.dSortedVel, 16 * n);
    CuAllocateArray((void**)&c.dSortedIndex, 4 * n);
    CuAllocateArray((void**)&c.dGridParticleHash, 4 * n);
    CuAllocateArray((void**)&c.dCellStart, 4 * (int)nc);
    CuAllocateArray((void**)&c.dCellEnd, 4 * (int)nc);
    if (CuGetLastErrorCode()) return fail("CuAllocateArray");
    c.uNumParticles = (uint)n;
    c.uNumCells = nc;
    CuCopyArrayToDevice(dpos, hpos, 0, 16 * n);
    CuCopyArrayToDevice(dvel, hvel, 0, 16 * n);

    std::vector<uint> raw(n), hs(n), ids(n), cs(nc), ce(nc);
    CuCalcHash(c.dGridParticleHash, c.dSortedIndex, dpos, nullptr, n);
    CuCopyArrayFromDevice(raw.data(), c.dGridParticleHash, nullptr, 4 * n);
    CuSort(c.dGridParticleHash, c.dSortedIndex, (uint)n);
    CuReorderDataAndFindCellStart(c, dpos, dvel);
    CuCopyArrayFromDevice(hs.data(), c.dGridParticleHash, nullptr, 4 * n);
    CuCopyArrayFromDevice(ids.data(), c.dSortedIndex, nullptr, 4 * n);
    CuCopyArrayFromDevice(cs.data(), c.dCellStart, nullptr, 4 * (int)nc);
    CuCopyArrayFromDevice(ce.data(), c.dCellEnd, nullptr, 4 * (int)nc);
    if (CuGetLastErrorCode()) return fail("the grid build");
    for (void* d : {(void*)dpos, (void*)dvel, (void*)c.dSortedPos, (void*)c.dSortedVel, (void*)c.dSortedIndex,
                    (void*)c.dGridParticleHash, (void*)c.dCellStart, (void*)c.dCellEnd})
        CuFreeArray(d);
    Writer w(out);
    w.put(raw.data(), 4 * (size_t)n);
    w.put(hs.data(), 4 * (size_t)n);
    w.put(ids.data(), 4 * (size_t)n);
    w.put(cs.data(), 4 * (size_t)nc);
    w.put(ce.data(), 4 * (size_t)nc);
    return 0;
}

int pbf(const char* in, const char* out) {
    std::vector<unsigned char> buf;
    const size_t head = sizeof(rxSphEnv) + 2 * sizeof(int) + sizeof(float);
    if (!read_all(in, buf) || buf.size() < head) return fail("reading the input");
    rxSphEnv e;
    std::memcpy(&e, buf.data(), sizeof(e));
    int n, steps;
    float dt;
    std::memcpy(&n, buf.data() + sizeof(e), sizeof(n));
    std::memcpy(&steps, buf.data() + sizeof(e) + sizeof(n), sizeof(steps));
    std::memcpy(&dt, buf.data() + sizeof(e) + 2 * sizeof(int), sizeof(dt));
    if (buf.size() != head + 32 * (size_t)n) return fail("input size");
    const float* hpos = (const float*)(buf.data() + head);
    const float* hvel = hpos + 4 * (size_t)n;

    void* s = rxPBDSPH_GPU_Create(&e);
    if (!s) return fail("rxPBDSPH_GPU_Create");
    if (rxPBDSPH_GPU_SetParticles(s, hpos, hvel, n)) return fail("rxPBDSPH_GPU_SetParticles");
    if (rxPBDSPH_GPU_Update(s, dt, steps)) return fail("rxPBDSPH_GPU_Update");
    std::vector<float> pos(4 * (size_t)n), vel(4 * (size_t)n);
    if (rxPBDSPH_GPU_GetArray(s, RX_ARRAY_POSITION, pos.data(), n)) return fail("rxPBDSPH_GPU_GetArray");
    if (rxPBDSPH_GPU_GetArray(s, RX_ARRAY_VELOCITY, vel.data(), n)) return fail("rxPBDSPH_GPU_GetArray");
    rxPBDSPH_GPU_Destroy(s);
    Writer w(out);
    w.put(pos.data(), pos.size() * sizeof(float));
    w.put(vel.data(), vel.size() * sizeof(float));
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && !std::strcmp(argv[1], "layout")) return layout();
    if (argc >= 3 && !std::strcmp(argv[1], "scene")) return scene(argv[2]);
    if (argc >= 4 && !std::strcmp(argv[1], "grid")) return grid(argv[2], argv[3]);
    if (argc >= 4 && !std::strcmp(argv[1], "pbf")) return pbf(argv[2], argv[3]);
    std::fprintf(stderr, "usage: abi_client layout | scene <out> | grid <in> <out> | pbf <in> <out>\n");
    return 2;
}

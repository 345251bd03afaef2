// pbf_solver.h -- the fused PBF solver class (the rxPBDSPH_GPU interface of rx_sph.h:1284-1601).
#pragma once

#include <vector>

#include "fused_grid.h"
#include "mlp.h"
#include "../../include/rx_sph_host.h"

struct PbfK;

class rxPBDSPH_GPU {
public:
    rxPBDSPH_GPU();
    ~rxPBDSPH_GPU();
    int Initialize(const rxSphEnv& env);     // rx_sph_pbd_gpu.cpp:168-300
    void Finalize();
    int SetParticles(const float* pos4, const float* vel4, int n);
    int Update(float dt, int step);          // rx_sph_pbd_gpu.cpp:790-957
    int GetArray(int type, float* host, int count);
    int GetStats(float* out, int count);
    int SetOption(int opt, float value);
    int Sync();
    int SetNetwork(void* mlp_object);  // inference weights of an rxMLP object (copied; the object may go)
    int SetThermal(const float* temps, const float* heats, const int* phase);
    void FeatureScales(float* sd_ref, float* nref) const {
        if (sd_ref) *sd_ref = m_sd_ref;
        if (nref) *nref = m_nref;
    }
    // Kernel timing with HIP events on the solver's stream (enabled by RX_PBF_OPT_PROFILE):
    // total milliseconds and launch count per kernel class since the last reset.
    int GetProfile(float* ms, int* count, int n);
    hipStream_t stream() const { return m_stream; }
    // Device view of the cell-ordered state after Update (valid until the next Update): positions, velocities,
    // slot -> particle ids, list-build positions, thermal state (NULL when not thermal), statistics array and
    // the slot of the largest-displacement statistic.
    struct DeviceView {
        float4* x;
        float4* v;
        const uint* id;
        const float4* xb;
        const float4* therm;
        float* stats;
        int disp2_slot;
    };
    DeviceView device_view() const;
    // The state was moved outside the solver (more than the skin): the next step starts from a fresh grid.
    void invalidate_list() { m_have_list = false; }
    uint size() const { return m_n; }

private:
    bool alloc(uint n);
    PbfK consts(float dt) const;
    void reorder(float4*& a, hipStream_t s);
    void lattice_refs();

    rxSphEnv m_env{};
    rxSimParams m_params{};
    FusedGrid m_grid;
    hipStream_t m_stream = nullptr;
    uint m_n = 0;
    float m_skin = 0.0f;
    int m_maxn = 96;
    float m_tol = 0.01f;
    bool m_have_list = false;
    bool m_sync_each_step = false;
    bool m_record_features = false;
    int m_mode = RX_PBF_MODE_PLAIN;
    bool m_adaptive = false;
    float m_warm_scale = 1.0f;
    float m_warm_kappa = 0.0f;   // RX_PBF_OPT_WARM_CLAMP
    float m_first_dmax = 0.0f;   // RX_PBF_OPT_FIRST_DMAX, as a fraction of h
    int m_defer_on = 1;          // RX_PBF_OPT_DEFER
    int m_judge = 0;
    bool m_grad_mass = false;  // RX_PBF_OPT_GRAD_MASS
    int m_list_tile = 0;  // LDS tiling variant of the list pass (RX_PBF_OPT_LIST_TILE)
    float m_sd_ref = 1.0f, m_nref = 1.0f;
    MlpDev m_mlp{};            // points into m_mlp_mem: the solver's own copy of the network
    void* m_mlp_mem = nullptr;
    bool m_mlp_ready = false;
    int m_step = 0;
    // device state, all in cell order
    float4* m_x = nullptr;   // positions at the step start
    float4* m_xs = nullptr;  // predicted / corrected positions
    float4* m_xl = nullptr;  // (position, lambda) written by lambda passes
    float4* m_xb = nullptr;  // positions at the list build
    float4* m_v = nullptr;   // velocities
    float4* m_t = nullptr;   // gather / output scratch
    float* m_dens = nullptr;
    float* m_lamacc = nullptr;  // sum of lambda over the iterations of the last step
    float* m_ftmp = nullptr;
    float4* m_therm = nullptr;   // (temperature, heat, phase, changed) per particle, when thermal
    float4* m_therm_t = nullptr;
    float m_heat_td = 0.0f, m_heat_dt = 0.0f;
    float* m_feat = nullptr;     // 16 features per particle (learned warm start)
    float4* m_ftmp16 = nullptr;
    uint* m_id = nullptr;       // slot -> original particle index
    uint* m_idt = nullptr;
    uint* m_list = nullptr;     // skin neighbour list: uint4 chunks of 8 16-bit entries, [chunk][slot]
    uint4* m_cnt = nullptr;     // list head per slot: 3 plane bases + count (bit 31: use the grid instead)
    float* m_stats = nullptr;
    float4* m_movers = nullptr;  // particles that moved more than skin/2 since the list build (bounded list)
    uint* m_defer = nullptr;     // slots a neighbour pass left to its deferred kernel
    // profiling
    bool m_profile = false;
    unsigned m_profile_mask = ~0u;  // rxProfileClass bits timed while profiling
    struct Mark {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<Mark> m_marks;
    std::vector<hipEvent_t> m_event_pool;
    size_t m_event_next = 0;
    double m_prof_ms[16] = {0};
    long m_prof_cnt[16] = {0};
    int begin_mark(int cls);
    void end_mark(int id);
    // hipGraph replay of the step (RX_PBF_OPT_GRAPH): the step rotates some buffers, so one instantiated graph
    // is kept per configuration of those buffers, keyed by the pointers and every value the launches capture
    // (a step rotates them through six configurations).
    struct StepPtrs {
        float4 *x, *xs, *xl, *xb, *v, *t, *therm, *therm_t;
        float *dens, *lamacc, *ftmp;
        uint *id, *idt, *perm, *slot;
    };
    struct StepGraph {
        std::vector<unsigned char> key;
        StepPtrs after;  // the buffer configuration the step leaves
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> m_graphs;
    bool m_use_graph = false;
    StepPtrs save_ptrs() const;
    void load_ptrs(const StepPtrs& p);
    std::vector<unsigned char> graph_key(float dt) const;
    void drop_graphs();
    int enqueue_step(float dt);
    void collect_marks();
    // Adaptive list sizes (RX_PBF_OPT_ADAPT_SKIN).  Behind each eager step its largest displacement, overflow
    // and fallback counts (4 stats slots) are copied into a pinned ring with an event; before a step the host
    // reads the records whose events have completed, without waiting, and sets the skin level and list
    // capacity the next list build uses.
    static constexpr int kRecs = 8;
    static constexpr int kRecLen = 14;  // stats slots copied per record (ST_DISP2 .. ST_CROSS + 3)
    float* m_hrec = nullptr;
    long m_bad_steps = 0, m_bad_reported = 0;  // completed steps whose fallback reach was capped, and those reported
    hipEvent_t m_rec_ev[kRecs] = {};
    long m_rec_head = 0, m_rec_tail = 0;
    bool m_adapt = true;
    float m_skin_base = 0.0f;  // level 0; level l uses (1 + l) times it
    int m_level = 0, m_level_target = 0, m_calm = 0, m_calm_want = 0;
    int m_maxn_target = 96;
    int m_maxn_cap = 512;     // largest capacity the adaptive sizing asks for
    void poll_records();
    void push_record();
    bool apply_targets(hipStream_t s);
};

// Solver handle internals shared by the plain, batched and best-response
// code paths.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "hu_kernels.h"
#include "hu_tree.h"

struct mtt_solver;  // opaque in the C ABI

namespace mtt {

struct BatchedState;  // hu_batched.hip

// Per-kernel timing with HIP events recorded on the solver's stream.
struct KernelTimer {
    bool enabled = false;
    struct Mark { const char* name; hipEvent_t a, b; };
    std::vector<Mark> pending;
    std::vector<std::string> names;
    std::vector<double> ms;
    std::vector<long> count;
    void begin(const char* name, hipStream_t s, Mark& m) {
        m.name = name;
        (void)hipEventCreate(&m.a);
        (void)hipEventCreate(&m.b);
        (void)hipEventRecord(m.a, s);
    }
    void end(Mark& m, hipStream_t s) {
        (void)hipEventRecord(m.b, s);
        pending.push_back(m);
    }
    void collect() {
        for (auto& m : pending) {
            float t = 0.0f;
            (void)hipEventSynchronize(m.b);
            (void)hipEventElapsedTime(&t, m.a, m.b);
            size_t k = 0;
            while (k < names.size() && names[k] != m.name) ++k;
            if (k == names.size()) { names.push_back(m.name); ms.push_back(0.0); count.push_back(0); }
            ms[k] += t;
            count[k]++;
            (void)hipEventDestroy(m.a);
            (void)hipEventDestroy(m.b);
        }
        pending.clear();
    }
};

// Launches a kernel and, when timing is on, brackets it with events.
#define MTT_TIMED(S, NAME, ...)                                   \
    do {                                                          \
        ::mtt::KernelTimer::Mark _m;                              \
        if ((S).timer.enabled) (S).timer.begin(NAME, (S).stream, _m); \
        __VA_ARGS__;                                              \
        if ((S).timer.enabled) (S).timer.end(_m, (S).stream);     \
    } while (0)

struct SolverImpl {
    mtt_hu_game game;
    mtt_solver_cfg cfg;
    HuTree tree;
    hipStream_t stream = nullptr;
    // root-alive combos
    int n_alive = 0;
    uint8_t c1[1326], c2[1326];
    int gidx[1326];
    std::vector<uint8_t> deck;            // cards not on the root board
    std::vector<float> range[2];          // root-alive layout (n_alive)
    double pair_mass = 0;                 // sum_ij r0 r1 nb
    // tree topology
    std::vector<std::vector<int>> depth_lists;  // decision + chance nodes
    std::vector<int> terminals;
    int iters_done = 0;
    KernelTimer timer;
    // device (plain + shared)
    DeviceArena arena;
    mtt_flat_node* d_nodes = nullptr;
    uint8_t* d_deck = nullptr;
    mtt_alive_combo* d_combos = nullptr;
    std::vector<int*> d_depth;            // per depth
    float* d_reach = nullptr;             // [nodes][2][1216]
    float* d_cfv = nullptr;
    float* d_R = nullptr;                 // [infosets][A][1216]
    float* d_S = nullptr;
    float* d_sigma = nullptr;
    float* d_mats = nullptr;              // [n_mat][1216][1216]
    long long* d_sums = nullptr;          // [terms][2][64] card sums (2^48 fixed point)
    PlainTerminalPlan plan;
    // batched path
    BatchedState* batched = nullptr;       // owned; freed by batched_destroy
    ~SolverImpl();
};

int solver_gen_matrices(SolverImpl& s, float* d_out);
int solver_alloc_plain(SolverImpl& s);
int solver_run_plain(SolverImpl& s, int iters);
void iteration_weights(const mtt_solver_cfg& cfg, int t, float& wp, float& wn, float& ws);

// batched path (hu_batched.hip)
int batched_create(SolverImpl& s);
int batched_run(SolverImpl& s, int iters);
int batched_jit_info(const SolverImpl& s, int* active, double* compile_s);
int batched_river_showdown(SolverImpl& s, int board, float* reach, float* values, uint8_t* cards, uint16_t* grank,
                           int* nlive, int* ncol);
int batched_strategy_sums(SolverImpl& s, float* d_S_ref);  // scatter into the reference layout
int batched_export(SolverImpl& s, float* d_out, int which);  // 0 strategy sums, 1 regrets
void batched_destroy(BatchedState* b);

// FP64 best response (br.hip): exploitability of the average strategy held
// (normalised, reference layout) in d_strat.
int br_exploitability(SolverImpl& s, const float* d_strat, double* out6);

// The same for a strategy held in host memory (reference layout, normalised),
// on a plain-mode handle created for the strategy's game (solver.hip).
int solver_exploitability_of(const mtt_hu_game& game, const float* r0_1326, const float* r1_1326,
                             const float* host_strat, size_t n_floats, int n_infosets, double* out6);

}  // namespace mtt

// Per-template sweeps of the batched heads-up path, generated at run time.
//
// The ahead-of-time sweep kernels interpret a template's DFS event records:
// every event is decoded on the scalar unit and the regrets are read through
// indexed registers.  For the turn and river levels (few templates, each
// instantiated over many boards) the walk is instead emitted as straight-line
// HIP per template -- slots, columns, terminal offsets and pots become
// literals, regrets become named registers -- and compiled once with hipRTC
// for the device.  The generated kernels take the same arguments, grids and
// LDS as the interpreted ones and compute the same recurrence.
#pragma once
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "bt_dev.h"

namespace mtt {

// One street level's templates (level 0: turn, 1: river or the single street).
struct JitLevel {
    int level = 0;
    int np = 0;                              // padded combos per board
    int ntmpl = 0;
    const std::vector<bt::DNode>* nodes = nullptr;  // DFS preorder per template, concatenated
    const std::vector<int>* tnode0 = nullptr;       // [ntmpl+1] first node of each template
    const std::vector<int>* tev0 = nullptr;         // [ntmpl+1] first event of each template
    const std::vector<int>* events = nullptr;       // template-local node << 2 | EV_*
    std::vector<int> slot0, term0, col0;            // [ntmpl+1] level-global firsts
};

// Launch shapes of the generated sweeps.  The forward always runs two
// workgroups of 576 per board (as the interpreted kernels).  The turn
// backward runs ceil(1152 / bwd_nt) workgroups of bwd_nt threads (a
// multiple of 64 in 192..576) per (board, template); the river backward does
// too when bwd_passes = 1, and with bwd_passes = 2 (the default) runs one
// workgroup of bwd_nt threads per (board, template) that walks the combos in
// ceil(1088 / bwd_nt) slices in turn, so the card sums of the template's
// terminals are staged and totalled once.  256 threads (the default): the
// walk fits its 187 registers without spills and two independent 4-wave
// workgroups share a CU, so one's loads overlap the other's walk; 576
// threads (one 9-wave workgroup per CU, 168 registers with spills) took 29%
// longer.  *_waves > 0 asks the compiler to fit that many waves per SIMD
// (amdgpu_waves_per_eu).
struct JitConfig {
    int bwd_nt = 256;
    int bwd_waves = 0;
    int fwd_waves = 0;
    int bwd_passes = 2;
    // river showdown columns: 0 = BF16 hi + lo (Bh, Bl) for the GEMM, 1 = FP32 (Bf) for the strength-order sweep
    int sd_fp32 = 0;
    // copies of the forward's card-sum bins in LDS (1 or 2, by lane parity)
    int fwd_bins = 2;
    // river forward workgroup: 576 = two workgroups per (board, template),
    // one per half of the combos; less = one workgroup walking the combos in
    // slices (its card sums go to the first half of the partials only, which
    // the backward then reads alone).  384 (three slices, the default): river
    // forward 3.48 -> 3.30 ms and backward 2.31 -> 2.26 ms against 576.  The
    // turn level keeps the two half workgroups (49 boards give too few
    // workgroups to slice).
    int fwd_nt = 384;
    // river forward strategy sums: 0 = read with the regrets in the prologue; k > 0 = read, k decisions
    // at a time once their reaches are known, by bounded buffer loads that skip lanes whose acting
    // player has zero reach.  64 (the default) puts all of a template's decisions in one group, updated
    // after the walk: river forward 3.12 -> 2.87 ms (iterations 41-60; 3.22 -> 3.09 on another GPU)
    int fwd_smask = 64;
    // river backward regrets: 0 = all loaded in the prologue; k > 0 = each decision's loaded k decisions
    // ahead of its entry in walk order
    int bwd_rpre = 0;
    // river backward regrets of the next slice: 1 (the default) = loaded into a second register set while
    // the current slice walks (with all regrets in the prologue, bwd_rpre = 0; 229 VGPRs, still two
    // workgroups per CU): river backward 2.32 -> 2.26 ms; 0 = each slice loads its own in its prologue
    int bwd_xpre = 1;
    // river forward regrets of the next slice, the same way: 1 (the default) = in flight during the current
    // slice's walk (158 VGPRs, still two 384-thread workgroups per CU): river forward 3.13 -> 3.07 ms
    int fwd_xpre = 1;
};

// Passes of the backward of a level (0: turn, 1: river) under cfg.
inline int bwd_passes(int level, const JitConfig& cfg) {
    return level == 1 && cfg.bwd_passes > 1 ? (1088 + cfg.bwd_nt - 1) / cfg.bwd_nt : 1;
}

// Workgroups per (board, template) of the backward of a level, of bwd_nt
// threads each.
inline int bwd_groups(int level, const JitConfig& cfg) {
    return bwd_passes(level, cfg) > 1 ? 1 : (1152 + cfg.bwd_nt - 1) / cfg.bwd_nt;
}

struct JitSweeps {
    bool active = false;
    double compile_s = 0.0;                  // hipRTC compile + load (0 when cached)
    JitConfig cfg;
    hipFunction_t fwd[2] = {nullptr, nullptr};
    hipFunction_t bwd[2] = {nullptr, nullptr};
};

// Whether the generated forward of a level walks the combos in slices (one
// workgroup per (board, template)) or runs two half workgroups, and its
// workgroup shape.
inline bool fwd_sliced(int level, const JitConfig& cfg) { return level == 1 && cfg.fwd_nt != 576; }

// Threads per workgroup of the generated forward at a level.
inline int fwd_threads(int level, const JitConfig& cfg) { return fwd_sliced(level, cfg) ? cfg.fwd_nt : 576; }
inline int fwd_groups(int level, const JitConfig& cfg) { return fwd_sliced(level, cfg) ? 1 : 2; }

// LDS bytes of a generated forward over templates of at most maxterm terminals.
inline size_t fwd_lds(int maxterm, const JitConfig& cfg) { return (size_t)maxterm * 104 * 8 * cfg.fwd_bins; }

// Generates, compiles and loads the sweeps of the given levels for the
// current device.  Compiled modules are cached per (device, source) for the
// life of the process.  Returns MTT_OK or an error code with
// mtt_last_error() set.
int bt_jit_build(const JitLevel* levels, int n, const JitConfig& cfg, JitSweeps* out);

// The generated source for the given levels (for inspection and tests).
std::string bt_jit_source(const JitLevel* levels, int n, const JitConfig& cfg);

// The default launch shapes, with MTT_JIT_BWD_NT / MTT_JIT_BWD_WAVES /
// MTT_JIT_FWD_WAVES / MTT_JIT_BWD_PASSES overrides from the environment
// (invalid values ignored).
JitConfig bt_jit_config();

// hipRTC compile of a generated source for gfx950 (no device needed).
int bt_jit_compile(const std::string& src, std::vector<char>* code);

// Whether generated sweeps are enabled: MTT_JIT unset or not "0".
bool bt_jit_enabled();

}  // namespace mtt

// FP64 best response and exploitability on the GPU (new: the reference has
// no best-response code).  One level-synchronous pass over the reference
// FlatNode tree computes both players' best-response values at once:
// forward, both players' reaches under the average strategy; terminals, the
// same fold/showdown values as the CFR path in FP64; backward, a player's
// own decision takes the max over actions of its values and the opponent's
// values sum over actions (the actor's reach is already in the children).
// With mode = 0 the actor's values are sigma-weighted instead, giving the
// profile's own values.
#include <algorithm>

#include "solver.h"

namespace mtt {

constexpr int NP = kPadFlop;
constexpr int AS = MTT_MAX_ACTIONS;

__device__ __forceinline__ uint64_t br_dealt_mask(int code, const uint8_t* __restrict__ deck) {
    uint64_t m = 0;
    int a = code & 63, b = code >> 6;
    if (a) m |= 1ull << deck[a - 1];
    if (b) m |= 1ull << deck[b - 1];
    return m;
}

__global__ __launch_bounds__(256) void k_br_forward(const int* __restrict__ depth_nodes, int num,
                                                    const mtt_flat_node* __restrict__ nodes,
                                                    const uint8_t* __restrict__ deck,
                                                    const mtt_alive_combo* __restrict__ combos,
                                                    const float* __restrict__ strat, double* __restrict__ reach,
                                                    int n_alive) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= NP) return;
    uint64_t mask = tid < n_alive ? combos[tid].mask : ~0ull;
    for (int ni = blockIdx.y; ni < num; ni += gridDim.y) {
        int id = depth_nodes[ni];
        mtt_flat_node n = nodes[id];
        double r0 = reach[((size_t)id * 2 + 0) * NP + tid];
        double r1 = reach[((size_t)id * 2 + 1) * NP + tid];
        if (n.type == MTT_NODE_CHANCE) {
            for (int k = 0; k < n.num_actions; ++k) {
                int di = (n.turn_idx >= 0 && k >= n.turn_idx) ? k + 1 : k;
                bool dead = (mask >> deck[di]) & 1ull;
                size_t ch = (size_t)(n.children_offset + k);
                reach[(ch * 2 + 0) * NP + tid] = dead ? 0.0 : r0;
                reach[(ch * 2 + 1) * NP + tid] = dead ? 0.0 : r1;
            }
        } else if (n.type == MTT_NODE_DECISION) {
            size_t base = (size_t)n.infoset_offset * AS * NP + tid;
            for (int a = 0; a < n.num_actions; ++a) {
                double s = strat[base + (size_t)a * NP];
                size_t ch = (size_t)(n.children_offset + a);
                reach[(ch * 2 + 0) * NP + tid] = n.player == 0 ? r0 * s : r0;
                reach[(ch * 2 + 1) * NP + tid] = n.player == 1 ? r1 * s : r1;
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_br_term_sums(const int* __restrict__ terms, int num,
                                                     const double* __restrict__ reach,
                                                     const int16_t* __restrict__ card_combos,
                                                     double* __restrict__ sums) {
    __shared__ double s[2][64];
    int t = blockIdx.x;
    if (t >= num) return;
    int id = terms[t];
    int c = threadIdx.x;
    for (int p = 0; p < 2; ++p) {
        double acc = 0.0;
        if (c < 52) {
            const double* r = reach + ((size_t)id * 2 + p) * NP;
            for (int k = 0; k < 51; ++k) {
                int j = card_combos[c * 51 + k];
                if (j >= 0) acc += r[j];
            }
        }
        s[p][c] = c < 52 ? acc : 0.0;
    }
    __syncthreads();
    if (c < 2) {
        double tot = 0.0;
        for (int k = 0; k < 52; ++k) tot += s[c][k];
        s[c][52] = 0.5 * tot;
    }
    __syncthreads();
    sums[((size_t)t * 2 + 0) * 64 + c] = s[0][c];
    sums[((size_t)t * 2 + 1) * 64 + c] = s[1][c];
}

__device__ __forceinline__ double br_mass(const double* __restrict__ s, const mtt_alive_combo& c, double r_self) {
    return s[52] - s[c.c1] - s[c.c2] + r_self;
}

__global__ __launch_bounds__(256) void k_br_fold(const int* __restrict__ terms, const int* __restrict__ fold_list,
                                                 int n_fold, const mtt_flat_node* __restrict__ nodes,
                                                 const uint8_t* __restrict__ deck,
                                                 const mtt_alive_combo* __restrict__ combos,
                                                 const double* __restrict__ reach, const double* __restrict__ sums,
                                                 double* __restrict__ val, int n_alive) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= NP) return;
    for (int f = blockIdx.y; f < n_fold; f += gridDim.y) {
        int t = fold_list[f];
        int id = terms[t];
        mtt_flat_node n = nodes[id];
        int w = n.active_mask == 1 ? 0 : 1, l = 1 - w;
        double v[2] = {0.0, 0.0};
        if (tid < n_alive) {
            mtt_alive_combo c = combos[tid];
            if (!(c.mask & br_dealt_mask(n.river_idx, deck))) {
                const double* sw = sums + ((size_t)t * 2 + w) * 64;
                const double* sl = sums + ((size_t)t * 2 + l) * 64;
                double rw = reach[((size_t)id * 2 + w) * NP + tid];
                double rl = reach[((size_t)id * 2 + l) * NP + tid];
                v[w] = ((double)n.pot_size - (double)n.committed[w]) * br_mass(sl, c, rl);
                v[l] = -(double)n.committed[l] * br_mass(sw, c, rw);
            }
        }
        val[((size_t)id * 2 + 0) * NP + tid] = v[0];
        val[((size_t)id * 2 + 1) * NP + tid] = v[1];
    }
}

struct BrTile { int mat, col0, ncol; };

// FP64 grouped showdown: 64 heroes x 32 columns per tile, 4x2 per thread.
__global__ __launch_bounds__(256) void k_br_showdown(const BrTile* __restrict__ tiles, const int* __restrict__ col_term,
                                                     const uint8_t* __restrict__ col_player,
                                                     const int* __restrict__ terms,
                                                     const mtt_flat_node* __restrict__ nodes,
                                                     const uint8_t* __restrict__ deck,
                                                     const mtt_alive_combo* __restrict__ combos,
                                                     const float* __restrict__ mats, const double* __restrict__ reach,
                                                     const double* __restrict__ sums, double* __restrict__ val,
                                                     int n_alive) {
    __shared__ double As[16][64];
    __shared__ double Bs[16][33];
    __shared__ const double* bptr[32];
    BrTile tile = tiles[blockIdx.y];
    const int h0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    if (threadIdx.x < 32) {
        int c = threadIdx.x;
        const double* p = nullptr;
        if (c < tile.ncol) {
            int gc = tile.col0 + c;
            int id = terms[col_term[gc]];
            p = reach + ((size_t)id * 2 + (1 - col_player[gc])) * NP;
        }
        bptr[c] = p;
    }
    __syncthreads();
    const float* E = mats + (size_t)tile.mat * NP * NP;
    double acc[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int j0 = 0; j0 < n_alive; j0 += 16) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int e = threadIdx.x + 256 * k;
            int jj = e >> 6, hh = e & 63;
            int j = j0 + jj;
            As[jj][hh] = j < n_alive ? (double)E[(size_t)j * NP + h0 + hh] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            int e = threadIdx.x + 256 * k;
            int cc = e >> 4, jj = e & 15;
            int j = j0 + jj;
            const double* p = bptr[cc];
            Bs[jj][cc] = (p && j < n_alive) ? p[j] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < 16; ++jj) {
            double a[4], b[2];
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = As[jj][tx * 4 + q];
            b[0] = Bs[jj][ty * 2];
            b[1] = Bs[jj][ty * 2 + 1];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                acc[x][0] = fma(a[x], b[0], acc[x][0]);
                acc[x][1] = fma(a[x], b[1], acc[x][1]);
            }
        }
        __syncthreads();
    }
    for (int y = 0; y < 2; ++y) {
        int c = ty * 2 + y;
        if (c >= tile.ncol) continue;
        int gc = tile.col0 + c;
        int t = col_term[gc];
        int id = terms[t];
        int pl = col_player[gc];
        mtt_flat_node n = nodes[id];
        uint64_t dm = br_dealt_mask(n.river_idx, deck);
        const double* so = sums + ((size_t)t * 2 + (1 - pl)) * 64;
        const double* ro = reach + ((size_t)id * 2 + (1 - pl)) * NP;
        double* out = val + ((size_t)id * 2 + pl) * NP;
        for (int x = 0; x < 4; ++x) {
            int h = h0 + tx * 4 + x;
            double v = 0.0;
            if (h < n_alive) {
                mtt_alive_combo cb = combos[h];
                if (!(cb.mask & dm))
                    v = (double)n.pot_size * acc[x][y] - (double)n.committed[pl] * br_mass(so, cb, ro[h]);
            }
            out[h] = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_br_backward(const int* __restrict__ depth_nodes, int num,
                                                     const mtt_flat_node* __restrict__ nodes,
                                                     const uint8_t* __restrict__ deck,
                                                     const mtt_alive_combo* __restrict__ combos,
                                                     const float* __restrict__ strat, double* __restrict__ val,
                                                     int n_alive, int br_mode) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= NP) return;
    uint64_t mask = tid < n_alive ? combos[tid].mask : ~0ull;
    for (int ni = blockIdx.y; ni < num; ni += gridDim.y) {
        int id = depth_nodes[ni];
        mtt_flat_node n = nodes[id];
        if (n.type == MTT_NODE_CHANCE) {
            double s0 = 0.0, s1 = 0.0;
            for (int k = 0; k < n.num_actions; ++k) {
                int di = (n.turn_idx >= 0 && k >= n.turn_idx) ? k + 1 : k;
                if ((mask >> deck[di]) & 1ull) continue;
                size_t ch = (size_t)(n.children_offset + k);
                s0 += val[(ch * 2 + 0) * NP + tid];
                s1 += val[(ch * 2 + 1) * NP + tid];
            }
            double inv = 1.0 / (double)(n.num_actions - 4);
            bool live = tid < n_alive;
            val[((size_t)id * 2 + 0) * NP + tid] = live ? s0 * inv : 0.0;
            val[((size_t)id * 2 + 1) * NP + tid] = live ? s1 * inv : 0.0;
        } else if (n.type == MTT_NODE_DECISION) {
            int p = n.player, o = 1 - p;
            size_t base = (size_t)n.infoset_offset * AS * NP + tid;
            double best = br_mode ? -1e300 : 0.0, opp = 0.0;
            for (int a = 0; a < n.num_actions; ++a) {
                size_t ch = (size_t)(n.children_offset + a);
                double vp = val[(ch * 2 + p) * NP + tid];
                if (br_mode) best = vp > best ? vp : best;
                else best += (double)strat[base + (size_t)a * NP] * vp;
                opp += val[(ch * 2 + o) * NP + tid];
            }
            val[((size_t)id * 2 + p) * NP + tid] = best;
            val[((size_t)id * 2 + o) * NP + tid] = opp;
        }
    }
}

static inline dim3 ngrid(int num) { return dim3((NP + 255) / 256, std::min(std::max(num, 1), 65535)); }

int br_exploitability(SolverImpl& s, const float* d_strat, double* out) {
    const auto& F = s.tree.flat;
    size_t nn = F.size();
    DeviceArena tmp;
    double *d_reach = nullptr, *d_val = nullptr, *d_sums = nullptr;
    float* d_mats = s.d_mats;
    int rc = tmp.alloc(nn * 2 * NP, &d_reach);
    rc = rc ? rc : tmp.alloc(nn * 2 * NP, &d_val);
    rc = rc ? rc : tmp.alloc(s.terminals.size() * 2 * 64, &d_sums);
    if (!rc && !d_mats) {
        rc = tmp.alloc((size_t)s.tree.n_matrices * NP * NP, &d_mats);
        rc = rc ? rc : solver_gen_matrices(s, d_mats);
    }
    PlainTerminalPlan plan = s.plan;
    if (!rc && !plan.d_terms) rc = build_terminal_plan(F, s.terminals, s.n_alive, s.c1, s.c2, plan, tmp);
    if (rc) return rc;
    // BR tiles: 32 columns
    std::vector<int> col_term_h;
    std::vector<uint8_t> col_player_h;
    std::vector<BrTile> tiles;
    {
        std::vector<std::vector<int>> by_mat;
        for (int t = 0; t < (int)s.terminals.size(); ++t) {
            const mtt_flat_node& n = F[s.terminals[t]];
            if (n.type != MTT_NODE_SHOWDOWN) continue;
            if ((int)by_mat.size() <= n.turn_idx) by_mat.resize(n.turn_idx + 1);
            by_mat[n.turn_idx].push_back(t);
        }
        for (int m = 0; m < (int)by_mat.size(); ++m) {
            int start = (int)col_term_h.size();
            for (int t : by_mat[m]) {
                col_term_h.push_back(t); col_player_h.push_back(0);
                col_term_h.push_back(t); col_player_h.push_back(1);
            }
            int n = (int)col_term_h.size() - start;
            for (int c = 0; c < n; c += 32) tiles.push_back(BrTile{m, start + c, std::min(32, n - c)});
        }
    }
    int* d_ct = nullptr;
    uint8_t* d_cp = nullptr;
    BrTile* d_tiles = nullptr;
    rc = tmp.upload(col_term_h, &d_ct);
    rc = rc ? rc : tmp.upload(col_player_h, &d_cp);
    rc = rc ? rc : tmp.upload(tiles, &d_tiles);
    if (rc) return rc;
    std::vector<double> root(2 * NP, 0.0);
    for (int p = 0; p < 2; ++p)
        for (int i = 0; i < s.n_alive; ++i) root[p * NP + i] = s.range[p][i];
    int D = (int)s.depth_lists.size();
    double res[2][2];
    for (int mode = 1; mode >= 0; --mode) {
        MTT_HIP(hipMemsetAsync(d_reach, 0, nn * 2 * NP * sizeof(double), s.stream));
        MTT_HIP(hipMemcpyAsync(d_reach, root.data(), root.size() * sizeof(double), hipMemcpyHostToDevice, s.stream));
        for (int d = 0; d < D; ++d)
            hipLaunchKernelGGL(k_br_forward, ngrid((int)s.depth_lists[d].size()), dim3(256), 0, s.stream, s.d_depth[d],
                               (int)s.depth_lists[d].size(), s.d_nodes, s.d_deck, s.d_combos, d_strat, d_reach,
                               s.n_alive);
        int nt = (int)s.terminals.size();
        if (nt) {
            hipLaunchKernelGGL(k_br_term_sums, dim3(nt), dim3(64), 0, s.stream, plan.d_terms, nt, d_reach,
                               plan.d_card_combos, d_sums);
            if (plan.n_fold)
                hipLaunchKernelGGL(k_br_fold, ngrid(plan.n_fold), dim3(256), 0, s.stream, plan.d_terms,
                                   plan.d_fold_list, plan.n_fold, s.d_nodes, s.d_deck, s.d_combos, d_reach, d_sums,
                                   d_val, s.n_alive);
            for (int t0 = 0; t0 < (int)tiles.size(); t0 += 65535) {
                int cnt = std::min(65535, (int)tiles.size() - t0);
                hipLaunchKernelGGL(k_br_showdown, dim3(NP / 64, cnt), dim3(256), 0, s.stream, d_tiles + t0, d_ct, d_cp,
                                   plan.d_terms, s.d_nodes, s.d_deck, s.d_combos, d_mats, d_reach, d_sums, d_val,
                                   s.n_alive);
            }
        }
        for (int d = D - 1; d >= 0; --d)
            hipLaunchKernelGGL(k_br_backward, ngrid((int)s.depth_lists[d].size()), dim3(256), 0, s.stream,
                               s.d_depth[d], (int)s.depth_lists[d].size(), s.d_nodes, s.d_deck, s.d_combos, d_strat,
                               d_val, s.n_alive, mode);
        MTT_HIP(hipGetLastError());
        std::vector<double> v(2 * NP);
        MTT_HIP(hipMemcpyAsync(v.data(), d_val, v.size() * sizeof(double), hipMemcpyDeviceToHost, s.stream));
        MTT_HIP(hipStreamSynchronize(s.stream));
        for (int p = 0; p < 2; ++p) {
            double acc = 0.0;
            for (int i = 0; i < s.n_alive; ++i) acc += (double)s.range[p][i] * v[p * NP + i];
            res[mode][p] = acc;
        }
    }
    double M = s.pair_mass, pot = s.game.pot;
    out[0] = res[1][0];
    out[1] = res[1][1];
    out[2] = res[0][0];
    out[3] = res[0][1];
    out[4] = M;
    out[5] = 100.0 * (out[0] + out[1] - pot * M) / (2.0 * M * pot);
    return MTT_OK;
}

}  // namespace mtt

// Plain FP32 heads-up CFR path: the reference's three-kernel structure
// (level-synchronous forward sweep, terminal evaluation, backward sweep;
// honest_multistreet_solver.cu:88-249) written for wave64 CDNA4, with the
// corrected payoffs (full-range fold sums with card removal, zero-sum
// showdown, exact chance normalisation, dead combos valued 0).
//
// Layouts are the reference's: per-node state[(node*2+p)*1216 + combo],
// per-action state[(infoset*MTT_MAX_ACTIONS + a)*1216 + combo].  The grouped
// showdown kernel evaluates all showdown terminals that share an equity
// matrix as one FP32 VALU product (no matrix cores), so that each 1216^2
// matrix is streamed once per iteration instead of once per terminal.
#include "hu_kernels.h"

namespace mtt {

constexpr int NP = kPadFlop;   // 1216
constexpr int AS = MTT_MAX_ACTIONS;

__device__ __forceinline__ uint64_t dealt_mask(int code, const uint8_t* __restrict__ deck) {
    uint64_t m = 0;
    int a = code & 63, b = code >> 6;
    if (a) m |= 1ull << deck[a - 1];
    if (b) m |= 1ull << deck[b - 1];
    return m;
}

__device__ __forceinline__ int chance_card(const mtt_flat_node& n, int k, const uint8_t* __restrict__ deck) {
    int di = (n.turn_idx >= 0 && k >= n.turn_idx) ? k + 1 : k;
    return deck[di];
}

// ------------------------------------------------------------- forward --
// kernel_forward_sweep (honest_multistreet_solver.cu:88-126).  grid.x
// covers the 1216 combos, grid.y strides over the nodes of one depth.
__global__ __launch_bounds__(256) void k_hu_forward(const int* __restrict__ depth_nodes, int num,
                                                    const mtt_flat_node* __restrict__ nodes,
                                                    const uint8_t* __restrict__ deck,
                                                    const mtt_alive_combo* __restrict__ combos,
                                                    const float* __restrict__ R, float* __restrict__ sigma,
                                                    float* __restrict__ reach, int n_alive, int as) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= NP) return;
    uint64_t mask = tid < n_alive ? combos[tid].mask : ~0ull;
    for (int ni = blockIdx.y; ni < num; ni += gridDim.y) {
        int id = depth_nodes[ni];
        mtt_flat_node n = nodes[id];
        float r0 = reach[((size_t)id * 2 + 0) * NP + tid];
        float r1 = reach[((size_t)id * 2 + 1) * NP + tid];
        if (n.type == MTT_NODE_CHANCE) {
            for (int k = 0; k < n.num_actions; ++k) {
                int c = chance_card(n, k, deck);
                bool dead = (mask >> c) & 1ull;
                size_t ch = (size_t)(n.children_offset + k);
                reach[(ch * 2 + 0) * NP + tid] = dead ? 0.0f : r0;
                reach[(ch * 2 + 1) * NP + tid] = dead ? 0.0f : r1;
            }
        } else if (n.type == MTT_NODE_DECISION && n.num_actions <= as) {
            int na = n.num_actions;
            const float* Rn = R + (size_t)n.infoset_offset * as * NP + tid;
            float* Sg = sigma + (size_t)n.infoset_offset * as * NP + tid;
            float pos[AS], sum = 0.0f;
#pragma unroll
            for (int a = 0; a < AS; ++a) {
                pos[a] = (a < na && tid < n_alive) ? fmaxf(Rn[(size_t)a * NP], 0.0f) : 0.0f;
                sum += pos[a];
            }
            float inv = sum > 1e-7f ? 1.0f / sum : 0.0f;
#pragma unroll
            for (int a = 0; a < AS; ++a) {
                if (a >= na) break;
                float s = tid < n_alive ? (sum > 1e-7f ? pos[a] * inv : 1.0f / (float)na) : 0.0f;
                Sg[(size_t)a * NP] = s;
                size_t ch = (size_t)(n.children_offset + a);
                reach[(ch * 2 + 0) * NP + tid] = n.player == 0 ? r0 * s : r0;
                reach[(ch * 2 + 1) * NP + tid] = n.player == 1 ? r1 * s : r1;
            }
        }
    }
}

// ------------------------------------------------------ terminal sums --
// Per terminal and player: sum of the player's reach over the combos that
// hold each card (52) and the total (slot 52), in 64-bit fixed point with 48
// fraction bits: exact integers, so the blocker-adjusted mass
// tot - R[c1] - R[c2] + r_i, which cancels heavily when most of the
// opponent's reach shares a card with the hero, carries no rounding until
// its one conversion to FP32.  Integer arithmetic only (reach < 2^15):
// r * 2^48 truncated, from the integer and fraction parts of r * 2^16 (both
// exact in FP32).
__device__ __forceinline__ long long fx48_of(float r) {
    const float t = r * 65536.0f;
    const uint32_t hi = (uint32_t)t;
    const uint32_t lo = (uint32_t)((t - (float)hi) * 4294967296.0f);
    return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float fx48_to_float(long long v) { return ldexpf((float)v, -48); }

__global__ __launch_bounds__(64) void k_hu_term_sums(const int* __restrict__ terms, int num,
                                                     const float* __restrict__ reach,
                                                     const int16_t* __restrict__ card_combos,  // [52][51]
                                                     long long* __restrict__ sums) {           // [num][2][64]
    __shared__ long long s[2][64];
    int t = blockIdx.x;
    if (t >= num) return;
    int id = terms[t];
    int c = threadIdx.x;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        long long acc = 0;
        if (c < 52) {
            const float* r = reach + ((size_t)id * 2 + p) * NP;
            for (int k = 0; k < 51; ++k) {
                int j = card_combos[c * 51 + k];
                if (j >= 0) acc += fx48_of(r[j]);
            }
        }
        s[p][c] = c < 52 ? acc : 0;
    }
    __syncthreads();
    if (c < 2) {
        long long tot = 0;
        for (int k = 0; k < 52; ++k) tot += s[c][k];
        s[c][52] = tot >> 1;  // every combo holds two cards: the sum is even
    }
    __syncthreads();
    sums[((size_t)t * 2 + 0) * 64 + c] = s[0][c];
    sums[((size_t)t * 2 + 1) * 64 + c] = s[1][c];
}

// mass_p(i) = sum_j nb(i,j) r_opp(j) = tot - R[c1] - R[c2] + r_opp(i)
__device__ __forceinline__ float blocker_mass(const long long* __restrict__ s, const mtt_alive_combo& c,
                                              float r_self) {
    return fx48_to_float(s[52] - s[c.c1] - s[c.c2] + fx48_of(r_self));
}

// ---------------------------------------------------------------- folds --
__global__ __launch_bounds__(256) void k_hu_fold_values(const int* __restrict__ terms, const int* __restrict__ fold_list,
                                                        int n_fold, const mtt_flat_node* __restrict__ nodes,
                                                        const uint8_t* __restrict__ deck,
                                                        const mtt_alive_combo* __restrict__ combos,
                                                        const float* __restrict__ reach,
                                                        const long long* __restrict__ sums, float* __restrict__ cfv,
                                                        int n_alive) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= NP) return;
    for (int f = blockIdx.y; f < n_fold; f += gridDim.y) {
        int t = fold_list[f];
        int id = terms[t];
        mtt_flat_node n = nodes[id];
        int w = n.active_mask == 1 ? 0 : 1, l = 1 - w;
        float v[2] = {0.0f, 0.0f};
        if (tid < n_alive) {
            mtt_alive_combo c = combos[tid];
            if (!(c.mask & dealt_mask(n.river_idx, deck))) {
                const long long* sw = sums + ((size_t)t * 2 + w) * 64;
                const long long* sl = sums + ((size_t)t * 2 + l) * 64;
                float rw = reach[((size_t)id * 2 + w) * NP + tid];
                float rl = reach[((size_t)id * 2 + l) * NP + tid];
                v[w] = (n.pot_size - n.committed[w]) * blocker_mass(sl, c, rl);
                v[l] = -n.committed[l] * blocker_mass(sw, c, rw);
            }
        }
        cfv[((size_t)id * 2 + 0) * NP + tid] = v[0];
        cfv[((size_t)id * 2 + 1) * NP + tid] = v[1];
    }
}

// ------------------------------------------------- grouped showdowns ----
// For one equity matrix E (villain-major, E[j][h]) and the columns of its
// showdown terminals: out[col][h] = sum_j E[j][h] * r_opp(col)[j].
// Tile 64 heroes x 64 columns, 256 threads with 4x4 each, K chunk 32.
// cfv = pot * (E r_opp) - committed * mass.
struct SdTile {
    int mat;        // matrix index
    int col0;       // first column of this tile in the global column list
    int ncol;       // columns in this tile (<= 64)
};

__global__ __launch_bounds__(256) void k_hu_showdown_grouped(
    const SdTile* __restrict__ tiles, const int* __restrict__ col_term, const uint8_t* __restrict__ col_player,
    const int* __restrict__ terms, const mtt_flat_node* __restrict__ nodes, const uint8_t* __restrict__ deck,
    const mtt_alive_combo* __restrict__ combos, const float* __restrict__ mats, const float* __restrict__ reach,
    const long long* __restrict__ sums, float* __restrict__ cfv, int n_alive) {
    __shared__ float As[32][64];
    __shared__ float Bs[32][65];
    __shared__ const float* bptr[64];
    SdTile tile = tiles[blockIdx.y];
    const int h0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    if (threadIdx.x < 64) {
        int c = threadIdx.x;
        const float* p = nullptr;
        if (c < tile.ncol) {
            int gc = tile.col0 + c;
            int id = terms[col_term[gc]];
            int pl = col_player[gc];
            p = reach + ((size_t)id * 2 + (1 - pl)) * NP;
        }
        bptr[c] = p;
    }
    __syncthreads();
    const float* E = mats + (size_t)tile.mat * NP * NP;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0f;
    const int kmax = n_alive;  // rows >= n_alive are zero
    for (int j0 = 0; j0 < kmax; j0 += 32) {
        // A: 32 rows j x 64 heroes -> 2048 floats, 8 per thread
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            int e = threadIdx.x + 256 * k;
            int jj = e >> 6, hh = e & 63;
            int j = j0 + jj, h = h0 + hh;
            As[jj][hh] = (j < kmax && h < NP) ? E[(size_t)j * NP + h] : 0.0f;
        }
        // B: 32 rows j x 64 columns
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            int e = threadIdx.x + 256 * k;
            int cc = e >> 5, jj = e & 31;
            int j = j0 + jj;
            const float* p = bptr[cc];
            Bs[jj][cc] = (p && j < kmax) ? p[j] : 0.0f;
        }
        __syncthreads();
#pragma unroll 8
        for (int jj = 0; jj < 32; ++jj) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) { a[q] = As[jj][tx * 4 + q]; b[q] = Bs[jj][ty * 4 + q]; }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] = fmaf(a[x], b[y], acc[x][y]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int y = 0; y < 4; ++y) {
        int c = ty * 4 + y;
        if (c >= tile.ncol) continue;
        int gc = tile.col0 + c;
        int t = col_term[gc];
        int id = terms[t];
        int pl = col_player[gc];
        mtt_flat_node n = nodes[id];
        uint64_t dm = dealt_mask(n.river_idx, deck);
        const long long* so = sums + ((size_t)t * 2 + (1 - pl)) * 64;
        const float* ro = reach + ((size_t)id * 2 + (1 - pl)) * NP;
        float* out = cfv + ((size_t)id * 2 + pl) * NP;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            int h = h0 + tx * 4 + x;
            if (h >= NP) continue;
            float v = 0.0f;
            if (h < n_alive) {
                mtt_alive_combo cb = combos[h];
                if (!(cb.mask & dm)) v = n.pot_size * acc[x][y] - n.committed[pl] * blocker_mass(so, cb, ro[h]);
            }
            out[h] = v;
        }
    }
}

// Reference-structured terminal kernel (honest_multistreet_solver.cu:128-196):
// one terminal per grid.y, one hero per thread, showdown as a per-terminal
// FP32 mat-vec.  Card sums are recomputed per block from the reach vectors.
__global__ __launch_bounds__(256) void k_hu_evaluate_terminals(const int* __restrict__ terms, int num,
                                                               const mtt_flat_node* __restrict__ nodes,
                                                               const uint8_t* __restrict__ deck,
                                                               const mtt_alive_combo* __restrict__ combos,
                                                               const float* __restrict__ reach,
                                                               const float* __restrict__ mats,
                                                               float* __restrict__ cfv, int n_alive) {
    __shared__ long long card[2][64];
    __shared__ float rr[2][NP];
    int t = blockIdx.y;
    if (t >= num) return;
    int id = terms[t];
    mtt_flat_node n = nodes[id];
    for (int j = threadIdx.x; j < NP; j += blockDim.x) {
        rr[0][j] = j < n_alive ? reach[((size_t)id * 2 + 0) * NP + j] : 0.0f;
        rr[1][j] = j < n_alive ? reach[((size_t)id * 2 + 1) * NP + j] : 0.0f;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        int c = threadIdx.x;
        long long a0 = 0, a1 = 0;
        if (c < 52)
            for (int j = 0; j < n_alive; ++j) {
                mtt_alive_combo cb = combos[j];
                if (cb.c1 == c || cb.c2 == c) { a0 += fx48_of(rr[0][j]); a1 += fx48_of(rr[1][j]); }
            }
        card[0][c] = a0;
        card[1][c] = a1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        long long tot = 0;
        for (int k = 0; k < 52; ++k) tot += card[threadIdx.x][k];
        card[threadIdx.x][52] = tot >> 1;
    }
    __syncthreads();
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= NP) return;
    float v0 = 0.0f, v1 = 0.0f;
    if (h < n_alive) {
        mtt_alive_combo cb = combos[h];
        if (!(cb.mask & dealt_mask(n.river_idx, deck))) {
            float m0 = blocker_mass(card[1], cb, rr[1][h]);  // player 0 faces player 1's reach
            float m1 = blocker_mass(card[0], cb, rr[0][h]);
            if (n.type == MTT_NODE_FOLD) {
                int w = n.active_mask == 1 ? 0 : 1;
                float mw = w == 0 ? m0 : m1, ml = w == 0 ? m1 : m0;
                float vw = (n.pot_size - n.committed[w]) * mw, vl = -n.committed[1 - w] * ml;
                v0 = w == 0 ? vw : vl;
                v1 = w == 0 ? vl : vw;
            } else {
                const float* E = mats + (size_t)n.turn_idx * NP * NP;
                float e0 = 0.0f, e1 = 0.0f;
                for (int j = 0; j < n_alive; ++j) {
                    float e = E[(size_t)j * NP + h];
                    e0 = fmaf(e, rr[1][j], e0);
                    e1 = fmaf(e, rr[0][j], e1);
                }
                v0 = n.pot_size * e0 - n.committed[0] * m0;
                v1 = n.pot_size * e1 - n.committed[1] * m1;
            }
        }
    }
    cfv[((size_t)id * 2 + 0) * NP + h] = v0;
    cfv[((size_t)id * 2 + 1) * NP + h] = v1;
}

// ------------------------------------------------------------ backward --
// kernel_backward_sweep (honest_multistreet_solver.cu:198-249) with the
// exact chance normalisation and the DCFR-generalised regret update.
__global__ __launch_bounds__(256) void k_hu_backward(const int* __restrict__ depth_nodes, int num,
                                                     const mtt_flat_node* __restrict__ nodes,
                                                     const uint8_t* __restrict__ deck,
                                                     const mtt_alive_combo* __restrict__ combos,
                                                     const float* __restrict__ sigma,
                                                     const float* __restrict__ reach, float* __restrict__ cfv,
                                                     float* __restrict__ R, float* __restrict__ S, float w_pos,
                                                     float w_neg, float w_strat, int n_alive, int as) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= NP) return;
    uint64_t mask = tid < n_alive ? combos[tid].mask : ~0ull;
    for (int ni = blockIdx.y; ni < num; ni += gridDim.y) {
        int id = depth_nodes[ni];
        mtt_flat_node n = nodes[id];
        if (n.type == MTT_NODE_CHANCE) {
            float s0 = 0.0f, s1 = 0.0f;
            for (int k = 0; k < n.num_actions; ++k) {
                int c = chance_card(n, k, deck);
                if ((mask >> c) & 1ull) continue;
                size_t ch = (size_t)(n.children_offset + k);
                s0 += cfv[(ch * 2 + 0) * NP + tid];
                s1 += cfv[(ch * 2 + 1) * NP + tid];
            }
            float inv = 1.0f / (float)(n.num_actions - 4);
            bool live = tid < n_alive;
            cfv[((size_t)id * 2 + 0) * NP + tid] = live ? s0 * inv : 0.0f;
            cfv[((size_t)id * 2 + 1) * NP + tid] = live ? s1 * inv : 0.0f;
        } else if (n.type == MTT_NODE_DECISION && n.num_actions <= as) {
            int na = n.num_actions, p = n.player, o = 1 - p;
            size_t base = (size_t)n.infoset_offset * as * NP + tid;
            float av[AS], node_v = 0.0f, opp = 0.0f;
#pragma unroll
            for (int a = 0; a < AS; ++a) {
                if (a >= na) { av[a] = 0.0f; continue; }
                size_t ch = (size_t)(n.children_offset + a);
                av[a] = cfv[(ch * 2 + p) * NP + tid];
                node_v = fmaf(sigma[base + (size_t)a * NP], av[a], node_v);
                opp += cfv[(ch * 2 + o) * NP + tid];
            }
            cfv[((size_t)id * 2 + p) * NP + tid] = node_v;
            cfv[((size_t)id * 2 + o) * NP + tid] = opp;
            if (tid < n_alive) {
                float rp = reach[((size_t)id * 2 + p) * NP + tid];
#pragma unroll
                for (int a = 0; a < AS; ++a) {
                    if (a >= na) break;
                    size_t k = base + (size_t)a * NP;
                    float q = R[k] + (av[a] - node_v);
                    R[k] = q * (q > 0.0f ? w_pos : w_neg);
                    S[k] = fmaf(w_strat * rp, sigma[k], S[k]);
                }
            }
        }
    }
}

// ---------------------------------------------------------- launchers --
static inline dim3 node_grid(int num) { return dim3((NP + 255) / 256, num < 65535 ? (num > 0 ? num : 1) : 65535); }

int launch_hu_forward(const int* d_nodes_at_depth, int num, const mtt_flat_node* d_nodes, const uint8_t* d_deck,
                      const mtt_alive_combo* d_combos, const float* d_R, float* d_sigma, float* d_reach, int n_alive,
                      hipStream_t s, int as) {
    if (num <= 0) return MTT_OK;
    hipLaunchKernelGGL(k_hu_forward, node_grid(num), dim3(256), 0, s, d_nodes_at_depth, num, d_nodes, d_deck,
                       d_combos, d_R, d_sigma, d_reach, n_alive, as);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

int launch_hu_backward(const int* d_nodes_at_depth, int num, const mtt_flat_node* d_nodes, const uint8_t* d_deck,
                       const mtt_alive_combo* d_combos, const float* d_sigma, const float* d_reach, float* d_cfv,
                       float* d_R, float* d_S, float w_pos, float w_neg, float w_strat, int n_alive, hipStream_t s,
                       int as) {
    if (num <= 0) return MTT_OK;
    hipLaunchKernelGGL(k_hu_backward, node_grid(num), dim3(256), 0, s, d_nodes_at_depth, num, d_nodes, d_deck,
                       d_combos, d_sigma, d_reach, d_cfv, d_R, d_S, w_pos, w_neg, w_strat, n_alive, as);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

int launch_hu_terminals_ref(const int* d_terms, int num, const mtt_flat_node* d_nodes, const uint8_t* d_deck,
                            const mtt_alive_combo* d_combos, const float* d_reach, const float* d_mats, float* d_cfv,
                            int n_alive, hipStream_t s) {
    if (num <= 0) return MTT_OK;
    if (num > 65535) { set_error("terminal count %d exceeds one launch", num); return MTT_ERR_ARG; }
    hipLaunchKernelGGL(k_hu_evaluate_terminals, dim3((NP + 255) / 256, num), dim3(256), 0, s, d_terms, num, d_nodes,
                       d_deck, d_combos, d_reach, d_mats, d_cfv, n_alive);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

int launch_hu_terminals_grouped(const PlainTerminalPlan& plan, const mtt_flat_node* d_nodes, const uint8_t* d_deck,
                                const mtt_alive_combo* d_combos, const float* d_reach, const float* d_mats,
                                long long* d_sums, float* d_cfv, int n_alive, hipStream_t s) {
    if (plan.n_terms <= 0) return MTT_OK;
    hipLaunchKernelGGL(k_hu_term_sums, dim3(plan.n_terms), dim3(64), 0, s, plan.d_terms, plan.n_terms, d_reach,
                       plan.d_card_combos, d_sums);
    if (plan.n_fold > 0)
        hipLaunchKernelGGL(k_hu_fold_values, node_grid(plan.n_fold), dim3(256), 0, s, plan.d_terms, plan.d_fold_list,
                           plan.n_fold, d_nodes, d_deck, d_combos, d_reach, d_sums, d_cfv, n_alive);
    for (int t0 = 0; t0 < plan.n_tiles; t0 += 65535) {
        int nt = plan.n_tiles - t0 < 65535 ? plan.n_tiles - t0 : 65535;
        hipLaunchKernelGGL(k_hu_showdown_grouped, dim3(NP / 64, nt), dim3(256), 0, s,
                           reinterpret_cast<const SdTile*>(plan.d_tiles) + t0, plan.d_col_term, plan.d_col_player,
                           plan.d_terms, d_nodes, d_deck, d_combos, d_mats, d_reach, d_sums, d_cfv, n_alive);
    }
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

int build_terminal_plan(const std::vector<mtt_flat_node>& nodes, const std::vector<int>& terms, int n_alive,
                        const uint8_t* c1, const uint8_t* c2, PlainTerminalPlan& plan, DeviceArena& arena) {
    plan = PlainTerminalPlan{};
    plan.n_terms = (int)terms.size();
    std::vector<int> fold_list;
    std::vector<std::vector<int>> by_mat;
    for (int t = 0; t < (int)terms.size(); ++t) {
        const mtt_flat_node& n = nodes[terms[t]];
        if (n.type == MTT_NODE_FOLD) fold_list.push_back(t);
        else {
            if ((int)by_mat.size() <= n.turn_idx) by_mat.resize(n.turn_idx + 1);
            by_mat[n.turn_idx].push_back(t);
        }
    }
    std::vector<int> col_term;
    std::vector<uint8_t> col_player;
    std::vector<SdTile> tiles;
    for (int m = 0; m < (int)by_mat.size(); ++m) {
        int start = (int)col_term.size();
        for (int t : by_mat[m]) {
            col_term.push_back(t); col_player.push_back(0);
            col_term.push_back(t); col_player.push_back(1);
        }
        int n = (int)col_term.size() - start;
        for (int c = 0; c < n; c += 64) tiles.push_back(SdTile{m, start + c, n - c < 64 ? n - c : 64});
    }
    std::vector<int16_t> cc(52 * 51, -1);
    std::vector<int> fill(52, 0);
    for (int i = 0; i < n_alive; ++i) {
        cc[c1[i] * 51 + fill[c1[i]]++] = (int16_t)i;
        cc[c2[i] * 51 + fill[c2[i]]++] = (int16_t)i;
    }
    plan.n_fold = (int)fold_list.size();
    plan.n_tiles = (int)tiles.size();
    int rc = MTT_OK;
    rc = rc ? rc : arena.upload(terms, &plan.d_terms);
    rc = rc ? rc : arena.upload(fold_list, &plan.d_fold_list);
    rc = rc ? rc : arena.upload(col_term, &plan.d_col_term);
    rc = rc ? rc : arena.upload(col_player, &plan.d_col_player);
    rc = rc ? rc : arena.upload(tiles, reinterpret_cast<SdTile**>(&plan.d_tiles));
    rc = rc ? rc : arena.upload(cc, &plan.d_card_combos);
    return rc;
}

}  // namespace mtt

using namespace mtt;

// ================================================================ C ABI ===
extern "C" {

int mtt_hu_forward_sweep(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node* d_nodes,
                         const uint8_t* d_deck, const mtt_alive_combo* d_combos, const float* d_cumulative_regrets,
                         float* d_current_strategy, float* d_reach_probs, int n_alive, int action_stride,
                         hipStream_t stream) {
    MTT_ARG(num_nodes_at_depth >= 0, "mtt_hu_forward_sweep: negative node count");
    MTT_ARG(n_alive > 0 && n_alive <= NP, "mtt_hu_forward_sweep: n_alive %d", n_alive);
    MTT_ARG(action_stride >= 2 && action_stride <= AS, "mtt_hu_forward_sweep: action stride %d outside 2..%d",
            action_stride, AS);
    if (!num_nodes_at_depth) return MTT_OK;
    MTT_ARG(d_depth_node_indices && d_nodes && d_deck && d_combos && d_cumulative_regrets && d_current_strategy &&
                d_reach_probs, "mtt_hu_forward_sweep: null pointer");
    return launch_hu_forward(d_depth_node_indices, num_nodes_at_depth, d_nodes, d_deck, d_combos,
                             d_cumulative_regrets, d_current_strategy, d_reach_probs, n_alive, stream, action_stride);
}

int mtt_hu_evaluate_terminals(const int* d_terminal_node_indices, int num_terminals, const mtt_flat_node* d_nodes,
                              const float* d_reach_probs, const float* d_turn_matrices, float* d_cfvs,
                              const uint8_t* d_deck, const mtt_alive_combo* d_combos, int n_alive, hipStream_t stream) {
    MTT_ARG(num_terminals >= 0, "mtt_hu_evaluate_terminals: negative count");
    MTT_ARG(n_alive > 0 && n_alive <= NP, "mtt_hu_evaluate_terminals: n_alive %d", n_alive);
    if (!num_terminals) return MTT_OK;
    MTT_ARG(d_terminal_node_indices && d_nodes && d_deck && d_combos && d_reach_probs && d_turn_matrices && d_cfvs,
            "mtt_hu_evaluate_terminals: null pointer");
    return launch_hu_terminals_ref(d_terminal_node_indices, num_terminals, d_nodes, d_deck, d_combos, d_reach_probs,
                                   d_turn_matrices, d_cfvs, n_alive, stream);
}

int mtt_hu_backward_sweep(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node* d_nodes,
                          const uint8_t* d_deck, const mtt_alive_combo* d_combos, const float* d_current_strategy,
                          const float* d_reach_probs, float* d_cfvs, float* d_cumulative_regrets,
                          float* d_cumulative_strategy_sums, float iter_weight, int n_alive, int action_stride,
                          hipStream_t stream) {
    MTT_ARG(num_nodes_at_depth >= 0, "mtt_hu_backward_sweep: negative node count");
    MTT_ARG(n_alive > 0 && n_alive <= NP, "mtt_hu_backward_sweep: n_alive %d", n_alive);
    MTT_ARG(action_stride >= 2 && action_stride <= AS, "mtt_hu_backward_sweep: action stride %d outside 2..%d",
            action_stride, AS);
    if (!num_nodes_at_depth) return MTT_OK;
    MTT_ARG(d_depth_node_indices && d_nodes && d_deck && d_combos && d_current_strategy && d_reach_probs && d_cfvs &&
                d_cumulative_regrets && d_cumulative_strategy_sums, "mtt_hu_backward_sweep: null pointer");
    return launch_hu_backward(d_depth_node_indices, num_nodes_at_depth, d_nodes, d_deck, d_combos, d_current_strategy,
                              d_reach_probs, d_cfvs, d_cumulative_regrets, d_cumulative_strategy_sums, 1.0f, 0.0f,
                              iter_weight, n_alive, stream, action_stride);
}

int mtt_hu_backward_sweep_dcfr(const int* d_depth_node_indices, int num_nodes_at_depth, const mtt_flat_node* d_nodes,
                               const uint8_t* d_deck, const mtt_alive_combo* d_combos, const float* d_current_strategy,
                               const float* d_reach_probs, float* d_cfvs, float* d_cumulative_regrets,
                               float* d_cumulative_strategy_sums, float w_pos, float w_neg, float w_strat, int n_alive,
                               int action_stride, hipStream_t stream) {
    MTT_ARG(num_nodes_at_depth >= 0, "mtt_hu_backward_sweep_dcfr: negative node count");
    MTT_ARG(n_alive > 0 && n_alive <= NP, "mtt_hu_backward_sweep_dcfr: n_alive %d", n_alive);
    MTT_ARG(action_stride >= 2 && action_stride <= AS, "mtt_hu_backward_sweep_dcfr: action stride %d outside 2..%d",
            action_stride, AS);
    if (!num_nodes_at_depth) return MTT_OK;
    MTT_ARG(d_depth_node_indices && d_nodes && d_deck && d_combos && d_current_strategy && d_reach_probs && d_cfvs &&
                d_cumulative_regrets && d_cumulative_strategy_sums, "mtt_hu_backward_sweep_dcfr: null pointer");
    return launch_hu_backward(d_depth_node_indices, num_nodes_at_depth, d_nodes, d_deck, d_combos, d_current_strategy,
                              d_reach_probs, d_cfvs, d_cumulative_regrets, d_cumulative_strategy_sums, w_pos, w_neg,
                              w_strat, n_alive, stream, action_stride);
}

}  // extern "C"

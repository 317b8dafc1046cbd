// Evaluator and equity builders on CDNA4 (gfx950):
//   - device batch evaluation
//   - exhaustive flop win/tie counts (config-2 builder)
//   - the subsampled EV dump (mtt_ev_dumper.cu)
//   - turn / river equity matrices of the reference layout
//
// All of it is integer VALU work.  The count builder precomputes one score
// per (runout, hand) and then streams LDS tiles of those scores through a
// register-blocked compare kernel, so each pair-runout comparison costs one
// compare and one add; the evaluator itself is never called per pair.
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <vector>

#include "mtt_common.h"

namespace mtt {

// ------------------------------------------------------------ errors -----
static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    return e == hipErrorOutOfMemory ? MTT_ERR_OOM : MTT_ERR_HIP;
}

// ------------------------------------------------- device batch evaluator --
__global__ void k_eval7_batch(const uint8_t* __restrict__ cards, int64_t n, uint32_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint8_t c[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) c[k] = cards[7 * i + k];
    out[i] = eval7(c);
}

// ------------------------------------------------ exhaustive flop counts --
// Per flop f: deck[49] (cards not on the flop, ascending), the 1176 alive
// combos (lexicographic) and their output indices.  Runout b = (x<y) pair of
// deck indices in lexicographic order: 1176 runouts, padded to kRunPad.
constexpr int kRunouts = 1176;
constexpr int kRunPad = 1184;     // multiple of the LDS chunk
constexpr int kHandPad = 1280;    // 10 tiles of 128
constexpr int kTile = 128;
constexpr int kChunk = 16;

struct FlopTables {
    uint8_t deck[49];
    uint8_t pad[3];
    uint8_t c1[kHandPad];
    uint8_t c2[kHandPad];
    int16_t out_idx[kHandPad];     // output row/col index of alive combo
    int16_t alive_of_out[1326];    // inverse: output index -> alive index or -1
};

__constant__ uint8_t c_runout_x[kRunouts];
__constant__ uint8_t c_runout_y[kRunouts];

// scores[f][b][h]: score of alive hand h on runout b, 0 if the hand holds a
// runout card, and 0 for padded hands and padded runouts.
__global__ void k_fc_scores(const FlopTables* __restrict__ tabs, const uint8_t* __restrict__ flops,
                            uint32_t* __restrict__ scores) {
    int b = blockIdx.x, f = blockIdx.y;
    const FlopTables& t = tabs[f];
    uint32_t* row = scores + ((size_t)f * kRunPad + b) * kHandPad;
    if (b >= kRunouts) {
        for (int h = threadIdx.x; h < kHandPad; h += blockDim.x) row[h] = 0;
        return;
    }
    int x = t.deck[c_runout_x[b]], y = t.deck[c_runout_y[b]];
    uint64_t board = card_bits16(flops[3 * f]) | card_bits16(flops[3 * f + 1]) |
                     card_bits16(flops[3 * f + 2]) | card_bits16(x) | card_bits16(y);
    uint64_t dead = (1ull << x) | (1ull << y);
    for (int h = threadIdx.x; h < kHandPad; h += blockDim.x) {
        uint32_t s = 0;
        if (h < kAliveFlop) {
            int a = t.c1[h], c = t.c2[h];
            if (!(dead & ((1ull << a) | (1ull << c)))) s = score_packed(board | card_bits16(a) | card_bits16(c));
        }
        row[h] = s;
    }
}

// W tile: 128 heroes x 128 villains per workgroup, 8x8 per thread.
// Hero score 0 (conflict/pad) never wins; villain conflict becomes
// 0xFFFFFFFF so it is never beaten.
__global__ __launch_bounds__(256) void k_fc_compare(const FlopTables* __restrict__ tabs,
                                                    const uint32_t* __restrict__ scores,
                                                    int32_t* __restrict__ W, int S) {
    __shared__ uint4 lds_h[kChunk][kTile / 4];
    __shared__ uint4 lds_v[kChunk][kTile / 4];
    const int f = blockIdx.z;
    const int I0 = blockIdx.y * kTile, J0 = blockIdx.x * kTile;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const uint32_t* sf = scores + (size_t)f * kRunPad * kHandPad;
    uint32_t acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0;

    for (int b0 = 0; b0 < kRunPad; b0 += kChunk) {
        // 16 runouts x (128 hero + 128 villain) scores -> 4096 u32, 16 per thread
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            int e = tid + 256 * k;         // 0..1023 uint4 slots
            int r = e >> 6, q = e & 63;    // runout in chunk, uint4 within 256 scores
            const uint4* src = reinterpret_cast<const uint4*>(sf + (size_t)(b0 + r) * kHandPad);
            if (q < 32) {
                lds_h[r][q] = src[(I0 >> 2) + q];
            } else {
                uint4 v = src[(J0 >> 2) + (q - 32)];
                v.x = v.x ? v.x : 0xFFFFFFFFu;
                v.y = v.y ? v.y : 0xFFFFFFFFu;
                v.z = v.z ? v.z : 0xFFFFFFFFu;
                v.w = v.w ? v.w : 0xFFFFFFFFu;
                lds_v[r][q - 32] = v;
            }
        }
        __syncthreads();
#pragma unroll 2
        for (int r = 0; r < kChunk; ++r) {
            uint4 h0 = lds_h[r][ty * 2], h1 = lds_h[r][ty * 2 + 1];
            uint4 v0 = lds_v[r][tx * 2], v1 = lds_v[r][tx * 2 + 1];
            uint32_t hv[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
            uint32_t vv[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int c = 0; c < 8; ++c) acc[a][c] += (hv[a] > vv[c]) ? 1u : 0u;
        }
        __syncthreads();
    }
    const FlopTables& t = tabs[f];
    int32_t* Wf = W + (size_t)f * S * S;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        int i = I0 + ty * 8 + a;
        if (i >= kAliveFlop) continue;
        uint64_t mi = (1ull << t.c1[i]) | (1ull << t.c2[i]);
        int oi = t.out_idx[i];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            int j = J0 + tx * 8 + c;
            if (j >= kAliveFlop) continue;
            uint64_t mj = (1ull << t.c1[j]) | (1ull << t.c2[j]);
            Wf[(size_t)oi * S + t.out_idx[j]] = (mi & mj) ? 0 : (int32_t)acc[a][c];
        }
    }
}

// T = 990 - W - W^T on unblocked pairs; zero every entry outside the alive
// set (W included) and on blocked pairs.
__global__ void k_fc_finish(const FlopTables* __restrict__ tabs, int32_t* __restrict__ W,
                            int32_t* __restrict__ T, int S) {
    int f = blockIdx.z;
    int oi = blockIdx.y;
    const FlopTables& t = tabs[f];
    int32_t* Wf = W + (size_t)f * S * S;
    int32_t* Tf = T + (size_t)f * S * S;
    int ai = oi < 1326 ? t.alive_of_out[oi] : -1;
    for (int oj = blockIdx.x * blockDim.x + threadIdx.x; oj < S; oj += gridDim.x * blockDim.x) {
        int aj = oj < 1326 ? t.alive_of_out[oj] : -1;
        size_t ij = (size_t)oi * S + oj;
        if (ai < 0 || aj < 0) {
            Wf[ij] = 0;
            Tf[ij] = 0;
            continue;
        }
        uint64_t mi = (1ull << t.c1[ai]) | (1ull << t.c2[ai]);
        uint64_t mj = (1ull << t.c1[aj]) | (1ull << t.c2[aj]);
        if (mi & mj) { Tf[ij] = 0; continue; }
        Tf[ij] = 990 - Wf[ij] - Wf[(size_t)oj * S + oi];
    }
}

static void build_flop_tables(const uint8_t* flop, int layout, FlopTables& t) {
    memset(&t, 0, sizeof(t));
    uint64_t fm = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
    int nd = 0;
    for (int c = 0; c < 52; ++c)
        if (!(fm & (1ull << c))) t.deck[nd++] = (uint8_t)c;
    int n = alive_combos(fm, t.c1, t.c2);
    for (int k = 0; k < 1326; ++k) t.alive_of_out[k] = -1;
    for (int i = 0; i < n; ++i) {
        int o = layout == MTT_LAYOUT_GLOBAL1326 ? combo_index(t.c1[i], t.c2[i]) : i;
        t.out_idx[i] = (int16_t)o;
        t.alive_of_out[o] = (int16_t)i;
    }
}

static std::once_flag g_runout_once;
static int upload_runouts() {
    static int status = MTT_OK;
    std::call_once(g_runout_once, [] {
        uint8_t xs[kRunouts], ys[kRunouts];
        int b = 0;
        for (int x = 0; x < 49; ++x)
            for (int y = x + 1; y < 49; ++y) { xs[b] = (uint8_t)x; ys[b] = (uint8_t)y; b++; }
        hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_runout_x), xs, sizeof(xs));
        if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_runout_y), ys, sizeof(ys));
        if (e != hipSuccess) status = hip_fail(e, "upload runout table", __FILE__, __LINE__);
    });
    return status;
}

static bool valid_flop(const uint8_t* f) {
    return f[0] < 52 && f[1] < 52 && f[2] < 52 && f[0] != f[1] && f[0] != f[2] && f[1] != f[2];
}

// ------------------------------------------------------------- EV dump --
// Runout grid visited by kernel_evaluate_flop_ev (mtt_ev_dumper.cu:98-101):
// turn deck index 0,2,..,46; river r1+1, r1+4, .. < 47.
struct EvGrid {
    int n = 0;
    uint8_t r1[400], r2[400];
    EvGrid() {
        for (int a = 0; a < 47; a += 2)
            for (int b = a + 1; b < 47; b += 3) { r1[n] = (uint8_t)a; r2[n] = (uint8_t)b; n++; }
    }
};

__global__ void k_ev_scores(const mtt_alive_combo* __restrict__ combos, const uint8_t* __restrict__ deck,
                            const uint8_t* __restrict__ flop, const uint8_t* __restrict__ g1,
                            const uint8_t* __restrict__ g2, uint32_t* __restrict__ sc) {
    int p = blockIdx.x;
    int x = deck[g1[p]], y = deck[g2[p]];
    uint64_t board = card_bits16(flop[0]) | card_bits16(flop[1]) | card_bits16(flop[2]) |
                     card_bits16(x) | card_bits16(y);
    for (int h = threadIdx.x; h < kAliveFlop; h += blockDim.x) {
        mtt_alive_combo c = combos[h];
        sc[(size_t)p * kPadFlop + h] = score_packed(board | card_bits16(c.c1) | card_bits16(c.c2));
    }
}

// Per hero: villains j = 0,4,..; counts over the grid skipping dead cards;
// eq = (2W+T)/(2n) and fused multiply-adds as in mtt_ev_dumper.cu:104-113.
__global__ void k_ev_flop(const mtt_alive_combo* __restrict__ combos, const uint8_t* __restrict__ deck,
                          const uint8_t* __restrict__ g1, const uint8_t* __restrict__ g2, int npts,
                          const uint32_t* __restrict__ sc, const float* __restrict__ r0,
                          const float* __restrict__ r1, float* __restrict__ ev0, float* __restrict__ ev1) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kAliveFlop) return;
    mtt_alive_combo hero = combos[i];
    float t0 = 0.0f, t1 = 0.0f;
    for (int j = 0; j < kAliveFlop; j += 4) {
        mtt_alive_combo v = combos[j];
        if (hero.mask & v.mask) continue;
        uint64_t dead = hero.mask | v.mask;
        int wins = 0, ties = 0, total = 0;
        for (int p = 0; p < npts; ++p) {
            if (dead & ((1ull << deck[g1[p]]) | (1ull << deck[g2[p]]))) continue;
            uint32_t sh = sc[(size_t)p * kPadFlop + i], sv = sc[(size_t)p * kPadFlop + j];
            if (sh > sv) wins += 2;
            else if (sh == sv) ties += 1;
            total += 2;
        }
        if (total > 0) {
            float eq = (float)(wins + ties) / (float)total;
            t0 = __fmaf_rn(r1[j], __fmaf_rn(eq, 2.0f, -1.0f), t0);
            t1 = __fmaf_rn(r0[j], __fmaf_rn(1.0f - eq, 2.0f, -1.0f), t1);
        }
    }
    ev0[i] = t0;
    ev1[i] = t1;
}

// -------------------------------------------------- turn / river matrices --
// scores[t][r][h] of alive hand h on board flop+deck[t]+deck[r] (t != r),
// 0 if h holds one of those cards.
__global__ void k_board_scores49(const mtt_alive_combo* __restrict__ combos, const uint8_t* __restrict__ deck,
                                 const uint8_t* __restrict__ flop, uint32_t* __restrict__ sc) {
    int t = blockIdx.y, r = blockIdx.x;
    uint32_t* row = sc + ((size_t)t * 49 + r) * kPadFlop;
    int tc = deck[t], rc = deck[r];
    uint64_t board = card_bits16(flop[0]) | card_bits16(flop[1]) | card_bits16(flop[2]) |
                     card_bits16(tc) | card_bits16(rc);
    uint64_t dead = (1ull << tc) | (1ull << rc);
    for (int h = threadIdx.x; h < kPadFlop; h += blockDim.x) {
        uint32_t s = 0;
        if (h < kAliveFlop && t != r) {
            mtt_alive_combo c = combos[h];
            if (!(c.mask & dead)) s = score_packed(board | card_bits16(c.c1) | card_bits16(c.c2));
        }
        row[h] = s;
    }
}

// E[t][j][h] villain-major; one workgroup per (turn, villain row).
__global__ void k_turn_matrix(const mtt_alive_combo* __restrict__ combos, const uint8_t* __restrict__ deck,
                              const uint32_t* __restrict__ sc, float* __restrict__ out) {
    int t = blockIdx.y, j = blockIdx.x;
    float* row = out + ((size_t)t * kPadFlop + j) * kPadFlop;
    uint64_t tbit = 1ull << deck[t];
    bool jdead = j >= kAliveFlop;
    uint64_t vm = jdead ? 0 : combos[j].mask;
    if (vm & tbit) jdead = true;
    for (int h = threadIdx.x; h < kPadFlop; h += blockDim.x) {
        float e = 0.0f;
        if (!jdead && h < kAliveFlop) {
            uint64_t hm = combos[h].mask;
            if (!(hm & vm) && !(hm & tbit)) {
                uint64_t dead = hm | vm | tbit;
                int w = 0, ti = 0, n = 0;
                for (int r = 0; r < 49; ++r) {
                    if (dead & (1ull << deck[r])) continue;
                    uint32_t a = sc[((size_t)t * 49 + r) * kPadFlop + h];
                    uint32_t v = sc[((size_t)t * 49 + r) * kPadFlop + j];
                    w += a > v;
                    ti += a == v;
                    n++;
                }
                e = n > 0 ? ((float)w + 0.5f * (float)ti) / (float)n : 0.0f;
            }
        }
        row[h] = e;
    }
}

__global__ void k_river_matrix(const mtt_alive_combo* __restrict__ combos, const uint8_t* __restrict__ deck,
                               const int16_t* __restrict__ boards, const uint32_t* __restrict__ sc,
                               float* __restrict__ out) {
    int bi = blockIdx.y, j = blockIdx.x;
    int t = boards[2 * bi], r = boards[2 * bi + 1];
    float* row = out + ((size_t)bi * kPadFlop + j) * kPadFlop;
    uint64_t bm = (1ull << deck[t]) | (1ull << deck[r]);
    const uint32_t* s = sc + ((size_t)t * 49 + r) * kPadFlop;
    bool jdead = j >= kAliveFlop || (combos[j].mask & bm);
    uint64_t vm = j < kAliveFlop ? combos[j].mask : 0;
    uint32_t sv = j < kAliveFlop ? s[j] : 0;
    for (int h = threadIdx.x; h < kPadFlop; h += blockDim.x) {
        float e = 0.0f;
        if (!jdead && h < kAliveFlop) {
            uint64_t hm = combos[h].mask;
            if (!(hm & vm) && !(hm & bm)) {
                uint32_t a = s[h];
                e = a > sv ? 1.0f : (a == sv ? 0.5f : 0.0f);
            }
        }
        row[h] = e;
    }
}

}  // namespace mtt

using namespace mtt;

// ================================================================ C ABI ===

// ---------------------------------------------- preflop flop matrices ----
// kernel_gen_100_flop_matrices (preflop_grid_generator.cu:130-183): one
// workgroup per (hero, flop).  The hero's 47-card deck and its scores on
// the sampled runouts (turn deck index 0,2,..,46; river r1+1, r1+4, .. < 47)
// are staged in LDS once; threads stride over villains and evaluate only
// the villain's hand.  eq = (2W + T/2) / (2n) as in the reference (a tie
// counts a quarter).  Heroes that hold a flop card write nothing; blocked
// villains get 0; output transposed E^T[f][j][h].
constexpr int FM_PTS = 200;
__global__ __launch_bounds__(256) void k_flop_matrices(const mtt_subset_flop* __restrict__ flops,
                                                       const mtt_alive_combo* __restrict__ combos,
                                                       float* __restrict__ out) {
    __shared__ uint32_t hs[FM_PTS];
    __shared__ uint64_t pm[FM_PTS];
    __shared__ uint64_t pb[FM_PTS];
    __shared__ uint8_t deck[47];
    __shared__ int npts;
    const int h = blockIdx.x, f = blockIdx.y;
    const mtt_subset_flop fl = flops[f];
    const uint64_t fmask = (1ull << fl.c1) | (1ull << fl.c2) | (1ull << fl.c3);
    const mtt_alive_combo hero = combos[h];
    if (hero.mask & fmask) return;
    const uint64_t fb = card_bits16(fl.c1) | card_bits16(fl.c2) | card_bits16(fl.c3);
    if (threadIdx.x == 0) {
        int d = 0;
        for (int c = 0; c < 52; ++c)
            if (!((hero.mask | fmask) & (1ull << c))) deck[d++] = (uint8_t)c;
        int n = 0;
        for (int r1 = 0; r1 < 47; r1 += 2)
            for (int r2 = r1 + 1; r2 < 47; r2 += 3) {
                pm[n] = (1ull << deck[r1]) | (1ull << deck[r2]);
                pb[n] = card_bits16(deck[r1]) | card_bits16(deck[r2]);
                n++;
            }
        npts = n;
    }
    __syncthreads();
    const int np = npts;
    for (int p = threadIdx.x; p < np; p += blockDim.x)
        hs[p] = score_packed(fb | card_bits16(hero.c1) | card_bits16(hero.c2) | pb[p]);
    __syncthreads();
    for (int j = threadIdx.x; j < kAliveFlop; j += blockDim.x) {
        const mtt_alive_combo v = combos[j];
        float eq = 0.0f;
        if (!((hero.mask & v.mask) || (v.mask & fmask))) {
            const uint64_t vb = fb | card_bits16(v.c1) | card_bits16(v.c2);
            int wins = 0, ties = 0, total = 0;
            for (int p = 0; p < np; ++p) {
                if (v.mask & pm[p]) continue;
                const uint32_t sv = score_packed(vb | pb[p]), sh = hs[p];
                if (sh > sv) wins += 2;
                else if (sh == sv) ties += 1;
                total += 2;
            }
            eq = total > 0 ? ((float)wins + 0.5f * (float)ties) / (float)total : 0.0f;
        }
        out[((size_t)f * kPadFlop + j) * kPadFlop + h] = eq;
    }
}

extern "C" {

const char* mtt_last_error(void) { return g_err.c_str(); }
const char* mtt_version(void) { return "mtt-mi355x 0.1 (gfx950)"; }

uint32_t mtt_eval7(const uint8_t cards[7]) { return eval7(cards); }

int mtt_eval7_host_batch(const uint8_t* cards, int64_t n, uint32_t* out) {
    MTT_ARG(cards && out && n >= 0, "mtt_eval7_host_batch: null pointer or negative n");
    for (int64_t i = 0; i < n; ++i) out[i] = eval7(cards + 7 * i);
    return MTT_OK;
}

int mtt_eval7_batch(const uint8_t* d_cards, int64_t n, uint32_t* d_out, hipStream_t stream) {
    MTT_ARG(n >= 0, "mtt_eval7_batch: negative n");
    if (n == 0) return MTT_OK;
    MTT_ARG(d_cards && d_out, "mtt_eval7_batch: null pointer");
    int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_eval7_batch, dim3((unsigned)blocks), dim3(256), 0, stream, d_cards, n, d_out);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

int mtt_canonical_class(int c1, int c2) {
    int r1 = c1 >> 2, s1 = c1 & 3, r2 = c2 >> 2, s2 = c2 & 3;
    int hi = r1 > r2 ? r1 : r2, lo = r1 < r2 ? r1 : r2;
    if (r1 == r2) return 12 - hi;
    int base = (s1 == s2) ? 13 : 91;
    for (int r = 12; r > hi; --r) base += r;
    return base + (hi - 1 - lo);
}

int mtt_canonical_flops(uint8_t* out) {
    if (!out) { set_error("mtt_canonical_flops: null pointer"); return -1; }
    int n = 0;
    auto put = [&](int a, int b, int c, int w) {
        out[4 * n] = (uint8_t)a; out[4 * n + 1] = (uint8_t)b; out[4 * n + 2] = (uint8_t)c; out[4 * n + 3] = (uint8_t)w;
        n++;
    };
    auto mc = [](int r, int s) { return r * 4 + s; };
    for (int r = 12; r >= 0; --r) put(mc(r, 0), mc(r, 1), mc(r, 2), 4);
    for (int rp = 12; rp >= 0; --rp)
        for (int rk = 12; rk >= 0; --rk) {
            if (rp == rk) continue;
            put(mc(rp, 0), mc(rp, 1), mc(rk, 0), 12);
            put(mc(rp, 0), mc(rp, 1), mc(rk, 2), 12);
        }
    for (int r1 = 12; r1 >= 2; --r1)
        for (int r2 = r1 - 1; r2 >= 1; --r2)
            for (int r3 = r2 - 1; r3 >= 0; --r3) {
                put(mc(r1, 0), mc(r2, 0), mc(r3, 0), 4);
                put(mc(r1, 0), mc(r2, 0), mc(r3, 1), 12);
                put(mc(r1, 0), mc(r2, 1), mc(r3, 0), 12);
                put(mc(r1, 1), mc(r2, 0), mc(r3, 0), 12);
                put(mc(r1, 0), mc(r2, 1), mc(r3, 2), 24);
            }
    return n;
}

int mtt_combo_index(int a, int b) {
    if (a > b) { int t = a; a = b; b = t; }
    return combo_index(a, b);
}

int mtt_flop_counts(const uint8_t* flops, int n_flops, int32_t* d_W, int32_t* d_T, int layout,
                    hipStream_t stream) {
    MTT_ARG(n_flops >= 0, "mtt_flop_counts: negative n_flops");
    if (n_flops == 0) return MTT_OK;
    MTT_ARG(flops && d_W && d_T, "mtt_flop_counts: null pointer");
    MTT_ARG(layout == MTT_LAYOUT_GLOBAL1326 || layout == MTT_LAYOUT_ALIVE1216,
            "mtt_flop_counts: unknown layout %d", layout);
    for (int f = 0; f < n_flops; ++f)
        MTT_ARG(valid_flop(flops + 3 * f), "mtt_flop_counts: flop %d has invalid or repeated cards", f);
    int rc = upload_runouts();
    if (rc) return rc;
    const int S = layout == MTT_LAYOUT_GLOBAL1326 ? 1326 : kPadFlop;
    // Work in batches so the score workspace stays bounded (6 MB per flop).
    const int kBatch = 64;
    int nb = n_flops < kBatch ? n_flops : kBatch;
    std::vector<FlopTables> tabs(nb);
    FlopTables* d_tabs = nullptr;
    uint8_t* d_flops = nullptr;
    uint32_t* d_sc = nullptr;
    auto cleanup = [&]() {
        hipFree(d_tabs); hipFree(d_flops); hipFree(d_sc);
    };
    hipError_t e = hipMalloc(&d_tabs, sizeof(FlopTables) * nb);
    if (e == hipSuccess) e = hipMalloc(&d_flops, 3 * nb);
    if (e == hipSuccess) e = hipMalloc(&d_sc, sizeof(uint32_t) * (size_t)nb * kRunPad * kHandPad);
    if (e != hipSuccess) { cleanup(); return hip_fail(e, "mtt_flop_counts workspace", __FILE__, __LINE__); }
    for (int f0 = 0; f0 < n_flops; f0 += nb) {
        int cnt = n_flops - f0 < nb ? n_flops - f0 : nb;
        for (int k = 0; k < cnt; ++k) build_flop_tables(flops + 3 * (f0 + k), layout, tabs[k]);
        e = hipMemcpyAsync(d_tabs, tabs.data(), sizeof(FlopTables) * cnt, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_flops, flops + 3 * f0, 3 * cnt, hipMemcpyHostToDevice, stream);
        if (e != hipSuccess) { cleanup(); return hip_fail(e, "mtt_flop_counts upload", __FILE__, __LINE__); }
        int32_t* W = d_W + (size_t)f0 * S * S;
        int32_t* T = d_T + (size_t)f0 * S * S;
        hipLaunchKernelGGL(k_fc_scores, dim3(kRunPad, cnt), dim3(256), 0, stream, d_tabs, d_flops, d_sc);
        hipLaunchKernelGGL(k_fc_compare, dim3(kHandPad / kTile, kHandPad / kTile, cnt), dim3(256), 0, stream,
                           d_tabs, d_sc, W, S);
        hipLaunchKernelGGL(k_fc_finish, dim3((S + 255) / 256, S, cnt), dim3(256), 0, stream, d_tabs, W, T, S);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(stream);  // tabs reused next batch
        if (e != hipSuccess) { cleanup(); return hip_fail(e, "mtt_flop_counts kernels", __FILE__, __LINE__); }
    }
    cleanup();
    return MTT_OK;
}

int mtt_evaluate_flop_ev(const mtt_alive_combo* d_combos, const uint8_t* d_deck, const uint8_t* d_flop,
                         const float* d_r0, const float* d_r1, float* d_ev0, float* d_ev1, hipStream_t stream) {
    MTT_ARG(d_combos && d_deck && d_flop && d_r0 && d_r1 && d_ev0 && d_ev1, "mtt_evaluate_flop_ev: null pointer");
    static const EvGrid grid;
    uint8_t* d_g = nullptr;
    uint32_t* d_sc = nullptr;
    hipError_t e = hipMalloc(&d_g, 2 * 400);
    if (e == hipSuccess) e = hipMalloc(&d_sc, sizeof(uint32_t) * (size_t)grid.n * kPadFlop);
    if (e == hipSuccess) e = hipMemcpyAsync(d_g, grid.r1, 400, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_g + 400, grid.r2, 400, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ev_scores, dim3(grid.n), dim3(256), 0, stream, d_combos, d_deck, d_flop, d_g,
                           d_g + 400, d_sc);
        hipLaunchKernelGGL(k_ev_flop, dim3((kAliveFlop + 63) / 64), dim3(64), 0, stream, d_combos, d_deck, d_g,
                           d_g + 400, grid.n, d_sc, d_r0, d_r1, d_ev0, d_ev1);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    hipFree(d_g);
    hipFree(d_sc);
    if (e != hipSuccess) return hip_fail(e, "mtt_evaluate_flop_ev", __FILE__, __LINE__);
    return MTT_OK;
}

#pragma clang fp contract(off)
int mtt_ev_dump_tensor(int first, int count, float* out) {
    MTT_ARG(out && first >= 0 && count >= 0 && first + count <= 1755, "mtt_ev_dump_tensor: bad range");
    uint8_t flops[1755 * 4];
    mtt_canonical_flops(flops);
    memset(out, 0, sizeof(float) * (size_t)count * 1352);
    mtt_alive_combo* d_combos = nullptr;
    uint8_t *d_deck = nullptr, *d_flop = nullptr;
    float *d_r = nullptr, *d_ev = nullptr;
    hipError_t e = hipMalloc(&d_combos, sizeof(mtt_alive_combo) * kPadFlop);
    if (e == hipSuccess) e = hipMalloc(&d_deck, 52);
    if (e == hipSuccess) e = hipMalloc(&d_flop, 4);
    if (e == hipSuccess) e = hipMalloc(&d_r, sizeof(float) * kPadFlop);
    if (e == hipSuccess) e = hipMalloc(&d_ev, sizeof(float) * 2 * kPadFlop);
    std::vector<float> reach(kPadFlop, 1.0f / (float)kAliveFlop);
    if (e == hipSuccess) e = hipMemcpy(d_r, reach.data(), sizeof(float) * kPadFlop, hipMemcpyHostToDevice);
    int rc = MTT_OK;
    if (e != hipSuccess) rc = hip_fail(e, "mtt_ev_dump_tensor setup", __FILE__, __LINE__);
    std::vector<mtt_alive_combo> alive(kPadFlop);
    std::vector<float> ev(2 * kPadFlop);
    for (int k = 0; k < count && rc == MTT_OK; ++k) {
        const uint8_t* fl = flops + 4 * (first + k);
        uint64_t fm = (1ull << fl[0]) | (1ull << fl[1]) | (1ull << fl[2]);
        uint8_t deck[52];
        int nd = 0;
        for (int c = 0; c < 52; ++c)
            if (!(fm & (1ull << c))) deck[nd++] = (uint8_t)c;
        uint8_t c1s[1326], c2s[1326];
        int na = alive_combos(fm, c1s, c2s);
        memset(alive.data(), 0, sizeof(mtt_alive_combo) * kPadFlop);
        for (int i = 0; i < na; ++i) alive[i] = {c1s[i], c2s[i], (1ull << c1s[i]) | (1ull << c2s[i])};
        e = hipMemcpy(d_flop, fl, 3, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_deck, deck, nd, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_combos, alive.data(), sizeof(mtt_alive_combo) * kPadFlop, hipMemcpyHostToDevice);
        if (e != hipSuccess) { rc = hip_fail(e, "mtt_ev_dump_tensor upload", __FILE__, __LINE__); break; }
        rc = mtt_evaluate_flop_ev(d_combos, d_deck, d_flop, d_r, d_r, d_ev, d_ev + kPadFlop, 0);
        if (rc) break;
        e = hipMemcpy(ev.data(), d_ev, sizeof(float) * 2 * kPadFlop, hipMemcpyDeviceToHost);
        if (e != hipSuccess) { rc = hip_fail(e, "mtt_ev_dump_tensor download", __FILE__, __LINE__); break; }
        // Host accumulation in combo order, multiply then add (mtt_ev_dumper.cu:160-168).
        for (int s = 0; s < 4; ++s) {
            float* base = out + ((size_t)k * 4 + s) * 2 * 169;
            float mult = 1.0f + 0.1f * (float)s;
            for (int c = 0; c < na; ++c) {
                int cls = mtt_canonical_class(c1s[c], c2s[c]);
                float a = ev[c] * mult;
                float b = ev[kPadFlop + c] * mult;
                base[cls] = base[cls] + a;
                base[169 + cls] = base[169 + cls] + b;
            }
        }
    }
    hipFree(d_combos); hipFree(d_deck); hipFree(d_flop); hipFree(d_r); hipFree(d_ev);
    return rc;
}
#pragma clang fp contract(on)

int mtt_ev_dump(int limit, const char* path) {
    MTT_ARG(limit >= 0 && limit <= 1755, "mtt_ev_dump: limit %d outside 0..1755", limit);
    std::vector<float> t((size_t)limit * 1352);
    int rc = mtt_ev_dump_tensor(0, limit, t.data());
    if (rc) return rc;
    const char* p = path ? path : "mtt_ev_dump_1755.bin";
    FILE* fp = fopen(p, "wb");
    if (!fp) { set_error("mtt_ev_dump: cannot open %s", p); return MTT_ERR_IO; }
    size_t w = fwrite(t.data(), sizeof(float), t.size(), fp);
    int cl = fclose(fp);
    if (w != t.size() || cl != 0) { set_error("mtt_ev_dump: short write to %s", p); return MTT_ERR_IO; }
    return MTT_OK;
}

int mtt_gen_turn_matrices(const mtt_alive_combo* d_combos, const uint8_t* d_deck, const uint8_t* d_flop,
                          float* d_out, hipStream_t stream) {
    MTT_ARG(d_combos && d_deck && d_flop && d_out, "mtt_gen_turn_matrices: null pointer");
    uint32_t* d_sc = nullptr;
    MTT_HIP(hipMalloc(&d_sc, sizeof(uint32_t) * 49 * 49 * kPadFlop));
    hipLaunchKernelGGL(k_board_scores49, dim3(49, 49), dim3(256), 0, stream, d_combos, d_deck, d_flop, d_sc);
    hipLaunchKernelGGL(k_turn_matrix, dim3(kPadFlop, 49), dim3(256), 0, stream, d_combos, d_deck, d_sc, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    hipFree(d_sc);
    if (e != hipSuccess) return hip_fail(e, "mtt_gen_turn_matrices", __FILE__, __LINE__);
    return MTT_OK;
}

int mtt_gen_river_matrices(const mtt_alive_combo* d_combos, const uint8_t* d_deck, const uint8_t* d_flop,
                           const int16_t* boards, int n_boards, float* d_out, hipStream_t stream) {
    MTT_ARG(n_boards >= 0, "mtt_gen_river_matrices: negative n_boards");
    if (n_boards == 0) return MTT_OK;
    MTT_ARG(d_combos && d_deck && d_flop && d_out && boards, "mtt_gen_river_matrices: null pointer");
    for (int i = 0; i < n_boards; ++i)
        MTT_ARG(boards[2 * i] >= 0 && boards[2 * i] < 49 && boards[2 * i + 1] >= 0 && boards[2 * i + 1] < 49 &&
                    boards[2 * i] != boards[2 * i + 1],
                "mtt_gen_river_matrices: board %d has bad deck indices", i);
    uint32_t* d_sc = nullptr;
    int16_t* d_b = nullptr;
    MTT_HIP(hipMalloc(&d_sc, sizeof(uint32_t) * 49 * 49 * kPadFlop));
    hipError_t e = hipMalloc(&d_b, sizeof(int16_t) * 2 * n_boards);
    if (e == hipSuccess) e = hipMemcpyAsync(d_b, boards, sizeof(int16_t) * 2 * n_boards, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_board_scores49, dim3(49, 49), dim3(256), 0, stream, d_combos, d_deck, d_flop, d_sc);
        hipLaunchKernelGGL(k_river_matrix, dim3(kPadFlop, n_boards), dim3(256), 0, stream, d_combos, d_deck, d_b,
                           d_sc, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    hipFree(d_sc);
    hipFree(d_b);
    if (e != hipSuccess) return hip_fail(e, "mtt_gen_river_matrices", __FILE__, __LINE__);
    return MTT_OK;
}

int mtt_gen_flop_matrices(const mtt_subset_flop* d_flops, int n_flops, const mtt_alive_combo* d_combos,
                          float* d_flop_matrices, hipStream_t stream) {
    MTT_ARG(n_flops >= 0 && n_flops <= 65535, "mtt_gen_flop_matrices: n_flops %d out of 0..65535", n_flops);
    if (!n_flops) return MTT_OK;
    MTT_ARG(d_flops && d_combos && d_flop_matrices, "mtt_gen_flop_matrices: null pointer");
    hipLaunchKernelGGL(k_flop_matrices, dim3(kAliveFlop, n_flops), dim3(256), 0, stream, d_flops, d_combos,
                       d_flop_matrices);
    MTT_HIP(hipGetLastError());
    return MTT_OK;
}

}  // extern "C"

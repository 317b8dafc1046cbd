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
#include "solver.h"

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
    // the runtime keeps a failed call's error as the thread's last error; it is reported here, so clear
    // it (else the next hipGetLastError after a good launch would return it)
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? MTT_ERR_OOM : MTT_ERR_HIP;
}

// Owned device buffer, freed when it goes out of scope (after the blocking
// copies that end every use).
template <typename T>
struct DeviceBuf {
    T* p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc(&p, sizeof(T) * (n ? n : 1)); }
    ~DeviceBuf() {
        if (p) (void)hipFree(p);
    }
};

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
// Per flop: deck[49] (cards not on the flop, ascending), the 1176 alive
// combos (lexicographic) and their output indices.  Runout b = (x<y) pair of
// deck indices in lexicographic order: 1176 runouts, taken in 588 pairs.
//
// Counting (all integer VALU):
//   1. k_fc_codes: per (runout pair, hand) the two scores as order-preserving
//      16-bit codes, packed in one dword: v = code + 1, or 0xFFFF when the
//      hand holds a runout card (or is padding).
//   2. k_fc_pairs: per unordered pair {i, j} (upper-triangle tiles only) and
//      two runouts at a time in 16-bit lanes:
//          x = min(sat_sub(v_i + 1, v_j), 2)  =  2 (i wins), 1 (tie), 0 (i loses
//                                                or either hand is dead)
//          s1 += x,  s2 += x*x                    (v_pk_sub_u16 clamp,
//                                                  v_pk_min_u16, v_pk_add_u16,
//                                                  v_pk_mad_u16)
//      so s1 = 2W + T and s2 = 4W + T give W[i][j] and T[i][j] from one pass;
//      W[j][i] = 990 - W[i][j] - T[i][j] (every unblocked pair has exactly
//      C(45,2) = 990 live runouts).  Two VALU ops per pair-runout.
//   0. k_fc_zero clears the outputs first, so the entries of output indices
//      that are not alive combos of the flop stay 0.
constexpr int kRunouts = 1176;
constexpr int kRunPairs = 588;
constexpr int kHandPad = 1280;    // 10 tiles of 128
constexpr int kTile = 128;
constexpr int kTiles = 10;
constexpr int kTilePairs = kTiles * (kTiles + 1) / 2;
constexpr int kChunk = 12;        // runout pairs per LDS stage, 588 = 49 x 12
constexpr int kFcBatch = 64;      // flops per launch of the allocating entry point
constexpr int kFcLive = 990;      // runouts avoiding both hands of an unblocked pair

struct FlopTables {
    uint8_t deck[52];
    uint8_t pad[12];
    uint8_t c1[kHandPad];
    uint8_t c2[kHandPad];
    int16_t out_idx[kHandPad];     // output row/col index of alive combo
    int16_t alive_of_out[1328];    // inverse: output index -> alive index or -1
};
static_assert(sizeof(FlopTables) % 16 == 0, "FlopTables keeps 16-byte alignment in arrays");

struct FlopBatch {                  // flop cards by value (no host buffer to upload)
    uint8_t c[3 * kFcBatch];
};

struct RunoutTable {
    uint8_t x[kRunouts], y[kRunouts];
};
constexpr RunoutTable make_runouts() {
    RunoutTable t{};
    int b = 0;
    for (int x = 0; x < 49; ++x)
        for (int y = x + 1; y < 49; ++y) { t.x[b] = (uint8_t)x; t.y[b] = (uint8_t)y; b++; }
    return t;
}
__constant__ RunoutTable c_runouts = make_runouts();

struct TilePairTable {
    uint8_t i[kTilePairs], j[kTilePairs];
};
constexpr TilePairTable make_tile_pairs() {
    TilePairTable t{};
    int k = 0;
    for (int i = 0; i < kTiles; ++i)
        for (int j = i; j < kTiles; ++j) { t.i[k] = (uint8_t)i; t.j[k] = (uint8_t)j; k++; }
    return t;
}
__constant__ TilePairTable c_tile_pairs = make_tile_pairs();

// One wave per flop: lane a (< 52) owns the combos {a, b > a}; exclusive
// wave scans of the per-lane counts place the deck and the alive combos (in
// lexicographic order).
__device__ __forceinline__ int fc_excl_scan(int v) {
    const int lane = threadIdx.x;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    return x - v;
}

__global__ __launch_bounds__(64) void k_fc_tables(FlopBatch fb, int layout, FlopTables* __restrict__ tabs) {
    const int f = blockIdx.x, a = threadIdx.x;
    FlopTables& t = tabs[f];
    const uint64_t fm = (1ull << fb.c[3 * f]) | (1ull << fb.c[3 * f + 1]) | (1ull << fb.c[3 * f + 2]);
    const bool global = layout == MTT_LAYOUT_GLOBAL1326;
    for (int o = a; o < 1328; o += 64) t.alive_of_out[o] = (!global && o < kAliveFlop) ? (int16_t)o : (int16_t)-1;
    for (int h = kAliveFlop + a; h < kHandPad; h += 64) { t.c1[h] = 0; t.c2[h] = 0; t.out_idx[h] = -1; }
    const bool card_live = a < 52 && !(fm & (1ull << a));
    // deck: cards not on the flop, ascending
    const int dpos = fc_excl_scan(card_live ? 1 : 0);
    if (card_live) t.deck[dpos] = (uint8_t)a;
    if (a >= 49 && a < 52) t.deck[a] = 0;
    // combos {a, b > a}: alive if neither card is on the flop
    const uint64_t above = a < 52 ? (~0ull << (a + 1)) & ((1ull << 52) - 1) : 0ull;
    const int n_alive_a = card_live ? __builtin_popcountll(above & ~fm) : 0;
    int n = fc_excl_scan(n_alive_a);
    int g = a < 52 ? a * 51 - a * (a - 1) / 2 : 0;          // global index of {a, a+1}
    for (int b = a + 1; b < 52; ++b, ++g) {
        if (card_live && !(fm & (1ull << b))) {
            t.c1[n] = (uint8_t)a;
            t.c2[n] = (uint8_t)b;
            t.out_idx[n] = (int16_t)(global ? g : n);
            n++;
        }
    }
    __syncthreads();
    // inverse map (global layout) after every lane has placed its combos
    if (global)
        for (int h = a; h < kAliveFlop; h += 64) t.alive_of_out[t.out_idx[h]] = (int16_t)h;
}

// codes[f][rp][h] = v(runout 2rp) | v(runout 2rp+1) << 16.
__global__ __launch_bounds__(256) void k_fc_codes(const FlopTables* __restrict__ tabs, FlopBatch fb,
                                                  uint32_t* __restrict__ codes) {
    const int rp = blockIdx.x, f = blockIdx.y;
    const FlopTables& t = tabs[f];
    const uint64_t flop = card_bits16(fb.c[3 * f]) | card_bits16(fb.c[3 * f + 1]) | card_bits16(fb.c[3 * f + 2]);
    const int x0 = t.deck[c_runouts.x[2 * rp]], y0 = t.deck[c_runouts.y[2 * rp]];
    const int x1 = t.deck[c_runouts.x[2 * rp + 1]], y1 = t.deck[c_runouts.y[2 * rp + 1]];
    const uint64_t b0 = flop | card_bits16(x0) | card_bits16(y0);
    const uint64_t b1 = flop | card_bits16(x1) | card_bits16(y1);
    const uint64_t d0 = (1ull << x0) | (1ull << y0), d1 = (1ull << x1) | (1ull << y1);
    uint32_t* row = codes + ((size_t)f * kRunPairs + rp) * kHandPad;
    for (int h = threadIdx.x; h < kHandPad; h += blockDim.x) {
        uint32_t v0 = 0xFFFFu, v1 = 0xFFFFu;
        if (h < kAliveFlop) {
            const int a = t.c1[h], c = t.c2[h];
            const uint64_t hm = (1ull << a) | (1ull << c);
            const uint64_t hb = card_bits16(a) | card_bits16(c);
            if (!(d0 & hm)) v0 = code16(score_packed(b0 | hb)) + 1u;
            if (!(d1 & hm)) v1 = code16(score_packed(b1 | hb)) + 1u;
        }
        row[h] = v0 | (v1 << 16);
    }
}

typedef unsigned short fc_u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ fc_u16x2 fc_bits(uint32_t v) { return __builtin_bit_cast(fc_u16x2, v); }
__device__ __forceinline__ uint32_t fc_word(fc_u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// One workgroup per (tile pair I <= J, flop): 128 hero x 128 villain alive
// indices as four 64 x 64 wave tiles, 8 x 8 pairs per lane.  Wave tiles that
// lie wholly below the diagonal or wholly in the padding past 1176 only help
// stage the LDS chunks.
__global__ __launch_bounds__(256) void k_fc_pairs(const FlopTables* __restrict__ tabs,
                                                  const uint32_t* __restrict__ codes,
                                                  int32_t* __restrict__ W, int32_t* __restrict__ T, int S) {
    __shared__ uint4 lds_h[kChunk][kTile / 4];
    __shared__ uint4 lds_v[kChunk][kTile / 4];
    const int f = blockIdx.y;
    const int I = c_tile_pairs.i[blockIdx.x], J = c_tile_pairs.j[blockIdx.x];
    const int I0 = I * kTile, J0 = J * kTile;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int wr = w >> 1, wc = w & 1, lr = lane & 7, lc = lane >> 3;
    const bool active = !(I == J && wr > wc) && (I0 + 64 * wr < kAliveFlop) && (J0 + 64 * wc < kAliveFlop);
    const uint32_t* cf = codes + (size_t)f * kRunPairs * kHandPad;
    const fc_u16x2 two = {2, 2};
    const fc_u16x2 one = {1, 1};
    fc_u16x2 s1[8][8], s2[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) { s1[a][b] = fc_u16x2{0, 0}; s2[a][b] = fc_u16x2{0, 0}; }

    for (int c0 = 0; c0 < kRunPairs; c0 += kChunk) {
        // kChunk runout pairs x (128 hero + 128 villain) dwords: 3 uint4 per thread
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int e = tid + 256 * k;        // 0..767
            const int r = e >> 6, q = e & 63;
            const uint4* src = reinterpret_cast<const uint4*>(cf + (size_t)(c0 + r) * kHandPad);
            if (q < 32) {
                // hero form v + 1 in each 16-bit half: a dead hand (0xFFFF) becomes 0
                uint4 v = src[(I0 >> 2) + q];
                v.x = fc_word(fc_bits(v.x) + one);
                v.y = fc_word(fc_bits(v.y) + one);
                v.z = fc_word(fc_bits(v.z) + one);
                v.w = fc_word(fc_bits(v.w) + one);
                lds_h[r][q] = v;
            } else {
                lds_v[r][q - 32] = src[(J0 >> 2) + (q - 32)];
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll 2
            for (int r = 0; r < kChunk; ++r) {
                const uint4 h0 = lds_h[r][wr * 16 + lr * 2], h1 = lds_h[r][wr * 16 + lr * 2 + 1];
                const uint4 v0 = lds_v[r][wc * 16 + lc * 2], v1 = lds_v[r][wc * 16 + lc * 2 + 1];
                const uint32_t hv[8] = {h0.x, h0.y, h0.z, h0.w, h1.x, h1.y, h1.z, h1.w};
                const uint32_t vv[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
                for (int a = 0; a < 8; ++a)
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const fc_u16x2 x = __builtin_elementwise_min(
                            __builtin_elementwise_sub_sat(fc_bits(hv[a]), fc_bits(vv[b])), two);
                        s1[a][b] += x;
                        s2[a][b] += x * x;
                    }
            }
        }
        __syncthreads();
    }
    if (!active) return;
    const FlopTables& t = tabs[f];
    int32_t* Wf = W + (size_t)f * S * S;
    int32_t* Tf = T + (size_t)f * S * S;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const int i = I0 + 64 * wr + 8 * lr + a;
        if (i >= kAliveFlop) continue;
        const uint64_t mi = (1ull << t.c1[i]) | (1ull << t.c2[i]);
        const size_t oi = (size_t)t.out_idx[i];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int j = J0 + 64 * wc + 8 * lc + b;
            if (j >= kAliveFlop || j < i) continue;
            const size_t oj = (size_t)t.out_idx[j];
            const uint64_t mj = (1ull << t.c1[j]) | (1ull << t.c2[j]);
            int32_t wij = 0, tij = 0, wji = 0;
            if (!(mi & mj)) {
                const int x1 = (int)s1[a][b].x + (int)s1[a][b].y;     // 2W + T
                const int x2 = (int)s2[a][b].x + (int)s2[a][b].y;     // 4W + T
                wij = (x2 - x1) >> 1;
                tij = 2 * x1 - x2;
                wji = kFcLive - wij - tij;
            }
            Wf[oi * S + oj] = wij;
            Tf[oi * S + oj] = tij;
            Wf[oj * S + oi] = wji;
            Tf[oj * S + oi] = tij;
        }
    }
}

// The outputs of a batch start at zero (16-byte stores, grid-stride): the
// pair kernel then writes every alive pair, and the entries of output
// indices that are not alive combos stay 0.  A fill kernel rather than a
// memset node, so that the launch sequence can be graph-captured.
template <typename V>
__global__ __launch_bounds__(256) void k_fc_zero(V* __restrict__ W, V* __restrict__ T, size_t n) {
    const V z{};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        W[i] = z;
        T[i] = z;
    }
}

// Workspace of the allocation-free entry point: kNoAllocFlops flops' tables
// and codes in one file-scope device array (each GPU the library runs on
// holds its own copy until the process ends).  Host code never uses the
// array's name as an address: the per-device address comes from
// hipGetSymbolAddress with that device current.
constexpr int kNoAllocFlops = 16;
constexpr size_t kFcFlopBytes = sizeof(FlopTables) + sizeof(uint32_t) * (size_t)kRunPairs * kHandPad;
constexpr size_t kNoAllocBytes = kNoAllocFlops * kFcFlopBytes;   // 48,294,400 bytes
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static int noalloc_scratch(void** out) {
    int dev = -1;
    MTT_HIP(hipGetDevice(&dev));
    MTT_ARG(dev >= 0 && dev < kMaxDevices, "mtt_flop_counts_noalloc: device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "mtt_flop_counts_noalloc: scratch address", __FILE__, __LINE__);
        if (!p) { set_error("mtt_flop_counts_noalloc: no scratch address on device %d", dev); return MTT_ERR_HIP; }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MTT_OK;
}

static bool valid_flop(const uint8_t* f) {
    return f[0] < 52 && f[1] < 52 && f[2] < 52 && f[0] != f[1] && f[0] != f[2] && f[1] != f[2];
}

static size_t fc_workspace_bytes(int n_flops) { return (size_t)n_flops * kFcFlopBytes; }

static int fc_check(const char* fn, const uint8_t* flops, int n_flops, const int32_t* d_W, const int32_t* d_T,
                    int layout) {
    MTT_ARG(n_flops >= 0, "%s: negative n_flops", fn);
    if (n_flops == 0) return MTT_OK;
    MTT_ARG(flops && d_W && d_T, "%s: null pointer", fn);
    MTT_ARG(layout == MTT_LAYOUT_GLOBAL1326 || layout == MTT_LAYOUT_ALIVE1216, "%s: unknown layout %d", fn, layout);
    for (int f = 0; f < n_flops; ++f)
        MTT_ARG(valid_flop(flops + 3 * f), "%s: flop %d has invalid or repeated cards", fn, f);
    return MTT_OK;
}

// Per-kernel HIP-event timing of the count builder (off unless enabled).
static std::mutex g_fc_timer_mu;
static KernelTimer g_fc_timer;

// Launches the four kernels for `cnt` flops (cnt <= kFcBatch) on `stream`,
// with the tables and codes in `ws` (fc_workspace_bytes(cnt) bytes).
static hipError_t fc_launch(const uint8_t* flops, int cnt, int32_t* W, int32_t* T, int layout, void* ws,
                            hipStream_t stream) {
    const int S = layout == MTT_LAYOUT_GLOBAL1326 ? 1326 : kPadFlop;
    FlopBatch fb{};
    memcpy(fb.c, flops, 3 * (size_t)cnt);
    FlopTables* tabs = reinterpret_cast<FlopTables*>(ws);
    uint32_t* codes = reinterpret_cast<uint32_t*>(tabs + cnt);
    std::lock_guard<std::mutex> lock(g_fc_timer_mu);
    const bool timed = g_fc_timer.enabled;
    KernelTimer::Mark m;
    if (timed) g_fc_timer.begin("fc_zero", stream, m);
    // cnt * S * S int32 per output, a multiple of 4 (S is even): 16-byte stores when both outputs are
    // 16-byte aligned, else 4-byte stores
    const size_t n32 = (size_t)cnt * S * S;
    if ((((uintptr_t)W | (uintptr_t)T) & 15u) == 0)
        hipLaunchKernelGGL(k_fc_zero<uint4>, dim3(2048), dim3(256), 0, stream, reinterpret_cast<uint4*>(W),
                           reinterpret_cast<uint4*>(T), n32 / 4);
    else
        hipLaunchKernelGGL(k_fc_zero<int32_t>, dim3(2048), dim3(256), 0, stream, W, T, n32);
    if (timed) { g_fc_timer.end(m, stream); g_fc_timer.begin("fc_tables", stream, m); }
    hipLaunchKernelGGL(k_fc_tables, dim3(cnt), dim3(64), 0, stream, fb, layout, tabs);
    if (timed) { g_fc_timer.end(m, stream); g_fc_timer.begin("fc_codes", stream, m); }
    hipLaunchKernelGGL(k_fc_codes, dim3(kRunPairs, cnt), dim3(256), 0, stream, tabs, fb, codes);
    if (timed) { g_fc_timer.end(m, stream); g_fc_timer.begin("fc_pairs", stream, m); }
    hipLaunchKernelGGL(k_fc_pairs, dim3(kTilePairs, cnt), dim3(256), 0, stream, tabs, codes, W, T, S);
    if (timed) g_fc_timer.end(m, stream);
    return hipGetLastError();
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

__device__ __forceinline__ void k_ev_scores_body(const mtt_alive_combo* __restrict__ combos,
                                                 const uint8_t* __restrict__ deck, const uint8_t* __restrict__ flop,
                                                 const uint8_t* __restrict__ g1, const uint8_t* __restrict__ g2,
                                                 uint32_t* __restrict__ sc, int p) {
    int x = deck[g1[p]], y = deck[g2[p]];
    uint64_t board = card_bits16(flop[0]) | card_bits16(flop[1]) | card_bits16(flop[2]) |
                     card_bits16(x) | card_bits16(y);
    for (int h = threadIdx.x; h < kAliveFlop; h += blockDim.x) {
        mtt_alive_combo c = combos[h];
        sc[(size_t)p * kPadFlop + h] = score_packed(board | card_bits16(c.c1) | card_bits16(c.c2));
    }
}

__global__ void k_ev_scores(const mtt_alive_combo* __restrict__ combos, const uint8_t* __restrict__ deck,
                            const uint8_t* __restrict__ flop, const uint8_t* __restrict__ g1,
                            const uint8_t* __restrict__ g2, uint32_t* __restrict__ sc) {
    k_ev_scores_body(combos, deck, flop, g1, g2, sc, blockIdx.x);
}

// Per hero: villains j = 0,4,..; counts over the grid skipping dead cards;
// eq = (2W+T)/(2n) and fused multiply-adds as in mtt_ev_dumper.cu:104-113.
// Reaches are per-villain arrays, or the scalars u0 / u1 when those are null.
__device__ __forceinline__ void k_ev_flop_body(const mtt_alive_combo* __restrict__ combos,
                                               const uint8_t* __restrict__ deck, const uint8_t* __restrict__ g1,
                                               const uint8_t* __restrict__ g2, int npts,
                                               const uint32_t* __restrict__ sc, float u0, float u1,
                                               const float* __restrict__ r0, const float* __restrict__ r1, int i,
                                               float& out0, float& out1) {
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
            const float w1 = r1 ? r1[j] : u1, w0 = r0 ? r0[j] : u0;
            t0 = __fmaf_rn(w1, __fmaf_rn(eq, 2.0f, -1.0f), t0);
            t1 = __fmaf_rn(w0, __fmaf_rn(1.0f - eq, 2.0f, -1.0f), t1);
        }
    }
    out0 = t0;
    out1 = t1;
}

__global__ void k_ev_flop(const mtt_alive_combo* __restrict__ combos, const uint8_t* __restrict__ deck,
                          const uint8_t* __restrict__ g1, const uint8_t* __restrict__ g2, int npts,
                          const uint32_t* __restrict__ sc, const float* __restrict__ r0,
                          const float* __restrict__ r1, float* __restrict__ ev0, float* __restrict__ ev1) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kAliveFlop) return;
    float t0, t1;
    k_ev_flop_body(combos, deck, g1, g2, npts, sc, 0.0f, 0.0f, r0, r1, i, t0, t1);
    ev0[i] = t0;
    ev1[i] = t1;
}

// Batched forms for the EV dump: blockIdx.y = flop of the batch, each with
// its own combos [f][1216], deck [f][52], flop [f][4] and score rows
// [f][npts][1216]; the arithmetic per flop is that of k_ev_scores / k_ev_flop.
__global__ void k_ev_scores_b(const mtt_alive_combo* __restrict__ combos, const uint8_t* __restrict__ deck,
                              const uint8_t* __restrict__ flop, const uint8_t* __restrict__ g1,
                              const uint8_t* __restrict__ g2, int npts, uint32_t* __restrict__ sc) {
    const int f = blockIdx.y;
    k_ev_scores_body(combos + (size_t)f * kPadFlop, deck + (size_t)f * 52, flop + (size_t)f * 4, g1, g2,
                     sc + (size_t)f * npts * kPadFlop, blockIdx.x);
}

__global__ __launch_bounds__(64) void k_ev_flop_b(const mtt_alive_combo* __restrict__ combos,
                                                  const uint8_t* __restrict__ deck, const uint8_t* __restrict__ g1,
                                                  const uint8_t* __restrict__ g2, int npts,
                                                  const uint32_t* __restrict__ sc, float reach,
                                                  float* __restrict__ ev) {
    const int f = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kAliveFlop) return;
    float t0, t1;
    k_ev_flop_body(combos + (size_t)f * kPadFlop, deck + (size_t)f * 52, g1, g2, npts,
                   sc + (size_t)f * npts * kPadFlop, reach, reach, nullptr, nullptr, i, t0, t1);
    ev[((size_t)f * 2) * kPadFlop + i] = t0;
    ev[((size_t)f * 2 + 1) * kPadFlop + i] = t1;
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
    if (c1 < 0 || c1 >= 52 || c2 < 0 || c2 >= 52 || c1 == c2) {
        set_error("mtt_canonical_class: cards %d, %d are not two distinct cards below 52", c1, c2);
        return -1;
    }
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
    int rc = fc_check("mtt_flop_counts", flops, n_flops, d_W, d_T, layout);
    if (rc || n_flops == 0) return rc;
    const size_t S = layout == MTT_LAYOUT_GLOBAL1326 ? 1326 : kPadFlop;
    // Batches of up to 64 flops share one workspace (193 MB), reused in
    // stream order; it is freed once the stream has drained.
    const int nb = n_flops < kFcBatch ? n_flops : kFcBatch;
    void* ws = nullptr;
    hipError_t e = hipMalloc(&ws, fc_workspace_bytes(nb));
    if (e != hipSuccess) return hip_fail(e, "mtt_flop_counts workspace", __FILE__, __LINE__);
    for (int f0 = 0; f0 < n_flops && e == hipSuccess; f0 += nb) {
        const int cnt = n_flops - f0 < nb ? n_flops - f0 : nb;
        e = fc_launch(flops + 3 * (size_t)f0, cnt, d_W + (size_t)f0 * S * S, d_T + (size_t)f0 * S * S, layout, ws,
                      stream);
    }
    hipError_t e2 = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = e2;
    hipFree(ws);
    if (e != hipSuccess) return hip_fail(e, "mtt_flop_counts kernels", __FILE__, __LINE__);
    return MTT_OK;
}

int mtt_flop_counts_noalloc(const uint8_t* flops, int n_flops, int32_t* d_W, int32_t* d_T, int layout,
                            hipStream_t stream) {
    int rc = fc_check("mtt_flop_counts_noalloc", flops, n_flops, d_W, d_T, layout);
    if (rc || n_flops == 0) return rc;
    MTT_ARG(n_flops <= kNoAllocFlops,
            "mtt_flop_counts_noalloc: %d flops exceed the %d-flop workspace (%zu bytes per GPU); "
            "split the call or use mtt_flop_counts", n_flops, kNoAllocFlops, kNoAllocBytes);
    void* ws = nullptr;
    rc = noalloc_scratch(&ws);
    if (rc) return rc;
    // Every workspace byte a kernel reads is written earlier in the same
    // launch sequence, so nothing is cleared.  No synchronisation: the call
    // is asynchronous on `stream`, and calls that overlap on one GPU race
    // on the shared workspace.
    hipError_t e = fc_launch(flops, n_flops, d_W, d_T, layout, ws, stream);
    if (e != hipSuccess) return hip_fail(e, "mtt_flop_counts_noalloc kernels", __FILE__, __LINE__);
    return MTT_OK;
}

int mtt_flop_counts_timing(int enable) {
    std::lock_guard<std::mutex> lock(g_fc_timer_mu);
    g_fc_timer.collect();
    g_fc_timer.enabled = enable != 0;
    g_fc_timer.names.clear();
    g_fc_timer.ms.clear();
    g_fc_timer.count.clear();
    return MTT_OK;
}

int mtt_flop_counts_kernel_times(int idx, char* name, int name_cap, double* total_ms, int64_t* launches) {
    MTT_ARG(name && total_ms && launches && name_cap > 0, "mtt_flop_counts_kernel_times: bad arguments");
    std::lock_guard<std::mutex> lock(g_fc_timer_mu);
    g_fc_timer.collect();
    if (idx < 0 || idx >= (int)g_fc_timer.names.size()) return -1;
    snprintf(name, name_cap, "%s", g_fc_timer.names[idx].c_str());
    *total_ms = g_fc_timer.ms[idx];
    *launches = g_fc_timer.count[idx];
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
    if (count == 0) return MTT_OK;
    static const EvGrid grid;
    uint8_t flops[1755 * 4];
    mtt_canonical_flops(flops);
    memset(out, 0, sizeof(float) * (size_t)count * 1352);
    // Flops go in batches of up to 256 per launch pair (scores 0.9 MB per flop).
    const int B = count < 256 ? count : 256;
    DeviceBuf<mtt_alive_combo> d_combos;
    DeviceBuf<uint8_t> d_deck, d_flop, d_g;
    DeviceBuf<uint32_t> d_sc;
    DeviceBuf<float> d_ev;
    hipError_t e = d_combos.alloc((size_t)B * kPadFlop);
    if (e == hipSuccess) e = d_deck.alloc((size_t)B * 52);
    if (e == hipSuccess) e = d_flop.alloc((size_t)B * 4);
    if (e == hipSuccess) e = d_g.alloc(800);
    if (e == hipSuccess) e = d_sc.alloc((size_t)B * grid.n * kPadFlop);
    if (e == hipSuccess) e = d_ev.alloc((size_t)B * 2 * kPadFlop);
    if (e == hipSuccess) e = hipMemcpy(d_g.p, grid.r1, 400, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_g.p + 400, grid.r2, 400, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "mtt_ev_dump_tensor setup", __FILE__, __LINE__);
    std::vector<mtt_alive_combo> alive((size_t)B * kPadFlop);
    std::vector<uint8_t> decks((size_t)B * 52), fls((size_t)B * 4);
    std::vector<float> ev((size_t)B * 2 * kPadFlop);
    std::vector<uint8_t> c1s((size_t)B * 1326), c2s((size_t)B * 1326);
    std::vector<int> nas(B);
    const float reach = 1.0f / (float)kAliveFlop;
    for (int k0 = 0; k0 < count; k0 += B) {
        const int n = count - k0 < B ? count - k0 : B;
        memset(alive.data(), 0, sizeof(mtt_alive_combo) * alive.size());
        memset(decks.data(), 0, decks.size());
        for (int k = 0; k < n; ++k) {
            const uint8_t* fl = flops + 4 * (first + k0 + k);
            const uint64_t fm = (1ull << fl[0]) | (1ull << fl[1]) | (1ull << fl[2]);
            int nd = 0;
            for (int c = 0; c < 52; ++c)
                if (!(fm & (1ull << c))) decks[(size_t)k * 52 + nd++] = (uint8_t)c;
            for (int q = 0; q < 4; ++q) fls[(size_t)k * 4 + q] = q < 3 ? fl[q] : 0;
            uint8_t* a1 = &c1s[(size_t)k * 1326];
            uint8_t* a2 = &c2s[(size_t)k * 1326];
            nas[k] = alive_combos(fm, a1, a2);
            for (int i = 0; i < nas[k]; ++i)
                alive[(size_t)k * kPadFlop + i] = {a1[i], a2[i], (1ull << a1[i]) | (1ull << a2[i])};
        }
        e = hipMemcpy(d_combos.p, alive.data(), sizeof(mtt_alive_combo) * (size_t)n * kPadFlop, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_deck.p, decks.data(), (size_t)n * 52, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(d_flop.p, fls.data(), (size_t)n * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "mtt_ev_dump_tensor upload", __FILE__, __LINE__);
        hipLaunchKernelGGL(k_ev_scores_b, dim3(grid.n, n), dim3(256), 0, 0, d_combos.p, d_deck.p, d_flop.p, d_g.p,
                           d_g.p + 400, grid.n, d_sc.p);
        hipLaunchKernelGGL(k_ev_flop_b, dim3((kAliveFlop + 63) / 64, n), dim3(64), 0, 0, d_combos.p, d_deck.p, d_g.p,
                           d_g.p + 400, grid.n, d_sc.p, reach, d_ev.p);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(ev.data(), d_ev.p, sizeof(float) * (size_t)n * 2 * kPadFlop,
                                           hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(e, "mtt_ev_dump_tensor kernels", __FILE__, __LINE__);
        // Host accumulation in combo order, multiply then add (mtt_ev_dumper.cu:160-168).
        for (int k = 0; k < n; ++k) {
            const float* evk = &ev[(size_t)k * 2 * kPadFlop];
            const uint8_t* a1 = &c1s[(size_t)k * 1326];
            const uint8_t* a2 = &c2s[(size_t)k * 1326];
            for (int s = 0; s < 4; ++s) {
                float* base = out + ((size_t)(k0 + k) * 4 + s) * 2 * 169;
                float mult = 1.0f + 0.1f * (float)s;
                for (int c = 0; c < nas[k]; ++c) {
                    int cls = mtt_canonical_class(a1[c], a2[c]);
                    float x = evk[c] * mult;
                    float y = evk[kPadFlop + c] * mult;
                    base[cls] = base[cls] + x;
                    base[169 + cls] = base[169 + cls] + y;
                }
            }
        }
    }
    return MTT_OK;
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

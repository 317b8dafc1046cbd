// River showdown values by an exact sweep in strength order: the per-board
// program builder (host) and the sweep kernel.
#include "bt_showdown.h"

#include <algorithm>

namespace mtt {
namespace bt {
namespace {

// A program is a sequence of visits.  A visit makes one 32-row block of the
// columns the ring block and runs its records; its first record is a header:
// word 0 = OP_VISIT, word 1 = the block to request for two visits ahead
// (NO_BLOCK for none), word 2 = the visit's record count (header included).
// Other records: word 0 = opcode (bits 0-3), member count k (bits 4-7: 5, or
// 15 for HALF and READ), FLUSH flag (bit 9: afterwards write rows 0..n-1 of the
// 16-row output tile to C, n - 1 in bits 10-13); words 1..k = members as
// row | card a << 11 | card b << 17 | valid << 31.  Records are padded to
// 5 or 15 members with copies of their first member, marked invalid and
// placed first: an invalid member adds no reach, and whatever it writes is
// written again by the member it copies.
enum : uint32_t {
    OP_NOP = 0,
    OP_SINGLE = 1,   // k singleton groups (consecutive rows, distinct cards): read, value, add 2R
    OP_HALF = 2,     // k members of a tie group: add R (LDS atomics: cards may repeat)
    OP_READ = 3,     // k members of a tie group: value at the half state
    OP_END = 4,
    OP_VISIT = 5,
};
constexpr uint32_t F_FLUSH = 1u << 9;
constexpr uint32_t NO_BLOCK = 0xFFu;
constexpr int REC = SD_REC;
constexpr int VISIT = SD_VISIT;

constexpr int RB = 32;              // rows per ring block
constexpr int OB = 16;              // rows per output block
constexpr int GRP = 5;              // members of a small record (and of every SINGLE record)
constexpr int MAXM = REC - 1;       // members of a large HALF / READ record
constexpr int MAXS = 4;             // real members per SINGLE record (one group)
constexpr uint32_t M_VALID = 1u << 31;  // member flag: clear on padding members (their reach counts as 0)
constexpr int OSTR = 68;            // output tile row stride (floats): two lanes per bank on write and flush
constexpr int NRB = NPR / RB;       // 34 ring blocks
static_assert(NPR % RB == 0 && RB % OB == 0, "row blocks");
static_assert(VISIT * REC / 4 == 64, "one visit is one uint4 per lane");

__device__ __forceinline__ unsigned long long fx_of(float r) { return fx48_of(r); }
// C = W - L, summed in 2^48 units (two's complement)
__device__ __forceinline__ float val_of(unsigned long long v) { return fx48_to_float((long long)v); }
__device__ __forceinline__ uint32_t rfl(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
// Word i of a record held (uniform) in four vector registers, as a scalar.
__device__ __forceinline__ uint32_t qword(const uint4& q0, const uint4& q1, const uint4& q2, const uint4& q3, int i) {
    const uint4& q = i < 4 ? q0 : (i < 8 ? q1 : (i < 12 ? q2 : q3));
    const int j = i & 3;
    return rfl(j == 0 ? q.x : (j == 1 ? q.y : (j == 2 ? q.z : q.w)));
}

// One wave per (64 columns, board).  LDS: card sums H[card][lane] (2^48
// fixed point, two's complement, offset by -csT), the ring block of the
// columns ring[row][lane] (FP32, taken to fixed point where used), the output
// tile otile[row][lane]
// and the current visit's records (read by every lane at one address, then
// made uniform).  At each visit boundary the next visit's records and the
// column block two visits ahead are requested; the register set a block
// arrives in alternates with the visit's parity in a loop unrolled by two,
// so every wait on them is for loads issued a whole visit earlier.
__global__ __launch_bounds__(64) void k_sd_sweep(BatchedDev D, const uint32_t* __restrict__ prog,
                                                 const uint4* __restrict__ hdrs, const int* __restrict__ colinfo,
                                                 int ncol) {
    __shared__ unsigned long long H[52 * 64];
    __shared__ float ring[RB * 64];
    __shared__ float otile[OB * OSTR];
    __shared__ uint4 pbuf[VISIT * REC / 4];
    const LevelDev& L = D.L[1];
    const int b = blockIdx.y, lane = threadIdx.x, c0 = blockIdx.x * 64, col = c0 + lane;
    const int ncolp = L.ncolp;
    const bool cl = col < ncol;
    const uint4* gp = reinterpret_cast<const uint4*>(prog);   // record r = gp[4r .. 4r+3]
    const uint4 hd = hdrs[b];

    // visit 0 into LDS
    pbuf[lane] = gp[(size_t)hd.x * (REC / 4) + lane];

    // H = -csT (the column's card totals, both forward halves), HP = -Ptot; the totals are
    // requested 26 cards at a time (from a valid address for every lane) before any is used
    unsigned long long tot = 0;
    {
        const int ci = cl ? colinfo[col] : colinfo[0];
        const int p = (int)((unsigned)ci >> 31), t = ci & 0x7FFFFFFF;
        const unsigned long long* P0 = L.part + ((size_t)b * L.nterm + t) * 208 + (1 - p) * 52;
        const unsigned long long km = cl ? ~0ull : 0ull;
#pragma unroll
        for (int cb = 0; cb < 52; cb += 26) {
            unsigned long long v0[26], v1[26];
#pragma unroll
            for (int x = 0; x < 26; ++x) {
                v0[x] = P0[cb + x];
                v1[x] = P0[104 + cb + x];
            }
#pragma unroll
            for (int x = 0; x < 26; ++x) {
                const unsigned long long v = (v0[x] + v1[x]) & km;
                H[(cb + x) * 64 + lane] = 0ull - v;
                tot += v;
            }
        }
    }
    unsigned long long HP = 0ull - (tot >> 1);

    // a block of this lane's column, loaded unconditionally (from a valid address) so that the
    // number of loads between two visit boundaries never varies; invalid lanes or blocks read as 0
    const float* rcv = D.Bf + ((size_t)b * ncolp + (cl ? col : 0)) * NPR;
#define SD_LOAD(blk, set_, ok_)                                                                \
    do {                                                                                       \
        ok_ = cl && (int)(blk) < NRB;                                                          \
        const float4* q_ = reinterpret_cast<const float4*>(rcv + (size_t)min((int)(blk), NRB - 1) * RB); \
        _Pragma("unroll") for (int i_ = 0; i_ < RB / 4; ++i_) set_[i_] = q_[i_];               \
    } while (0)
#define SD_RING(set_, ok_)                                                                     \
    do {                                                                                       \
        _Pragma("unroll") for (int i_ = 0; i_ < RB / 4; ++i_) {                                \
            ring[(4 * i_ + 0) * 64 + lane] = ok_ ? set_[i_].x : 0.0f;                          \
            ring[(4 * i_ + 1) * 64 + lane] = ok_ ? set_[i_].y : 0.0f;                          \
            ring[(4 * i_ + 2) * 64 + lane] = ok_ ? set_[i_].z : 0.0f;                          \
            ring[(4 * i_ + 3) * 64 + lane] = ok_ ? set_[i_].w : 0.0f;                          \
        }                                                                                      \
    } while (0)
#define SD_FLUSH(h)                                                                            \
    do {                                                                                       \
        const int n_ = (int)(((h) >> 10) & 15u) + 1, hh_ = lane & 15, q_ = lane >> 4;          \
        float* Cb_ = D.C + (size_t)b * ncolp * NPR + (size_t)oblk * OB + hh_;                  \
        _Pragma("unroll") for (int t_ = 0; t_ < 64 / 4; ++t_) {                                \
            const int cc_ = 4 * t_ + q_;                                                       \
            const float v_ = otile[hh_ * OSTR + cc_];                                          \
            if (hh_ < n_ && c0 + cc_ < ncol) Cb_[(size_t)(c0 + cc_) * NPR] = v_;               \
        }                                                                                      \
        ++oblk;                                                                                \
    } while (0)
#define QW(i) qword(cq0, cq1, cq2, cq3, (i))
    // a HALF / READ record of exactly NM members (5 or 15): every read issued before any update
#define SD_HALF(NM)                                                                            \
    do {                                                                                       \
        unsigned long long rr[NM];                                                             \
        _Pragma("unroll") for (int m = 0; m < NM; ++m) {                                       \
            const uint32_t w_ = QW(1 + m);                                                     \
            rr[m] = fx_of(ring[(int)(w_ & (RB - 1)) * 64 + lane]) & ((w_ & M_VALID) ? ~0ull : 0ull); \
        }                                                                                      \
        _Pragma("unroll") for (int m = 0; m < NM; ++m) {                                       \
            const uint32_t w_ = QW(1 + m);                                                     \
            HP += rr[m];                                                                       \
            atomicAdd(&H[(int)((w_ >> 11) & 63u) * 64 + lane], rr[m]);                         \
            atomicAdd(&H[(int)((w_ >> 17) & 63u) * 64 + lane], rr[m]);                         \
        }                                                                                      \
    } while (0)
#define SD_READ(NM)                                                                            \
    do {                                                                                       \
        unsigned long long va[NM], vb[NM];                                                     \
        _Pragma("unroll") for (int m = 0; m < NM; ++m) {                                       \
            const uint32_t w_ = QW(1 + m);                                                     \
            va[m] = H[(int)((w_ >> 11) & 63u) * 64 + lane];                                    \
            vb[m] = H[(int)((w_ >> 17) & 63u) * 64 + lane];                                    \
        }                                                                                      \
        _Pragma("unroll") for (int m = 0; m < NM; ++m)                                         \
            otile[(int)(QW(1 + m) & (OB - 1)) * OSTR + lane] = val_of(HP - va[m] - vb[m]);     \
    } while (0)
    // the records 1..nrec-1 of the visit in pbuf; done = an END was met.  A record holds 5 or
    // 15 members (static word indices, no per-member branches).
#define SD_RUN()                                                                               \
    do {                                                                                       \
        uint4 n0 = pbuf[4], n1 = pbuf[5], n2 = pbuf[6], n3 = pbuf[7];                          \
        for (int r_ = 1; r_ < nrec; ++r_) {                                                    \
            const uint4 cq0 = n0, cq1 = n1, cq2 = n2, cq3 = n3;                                \
            const int nx_ = min(r_ + 1, VISIT - 1) * (REC / 4);                                \
            n0 = pbuf[nx_]; n1 = pbuf[nx_ + 1]; n2 = pbuf[nx_ + 2]; n3 = pbuf[nx_ + 3];        \
            const uint32_t h = QW(0);                                                          \
            const uint32_t op = h & 15u;                                                       \
            const int nm = (int)((h >> 4) & 15u);                                              \
            if (op == OP_END) {                                                                \
                done = true;                                                                   \
                break;                                                                         \
            }                                                                                  \
            if (op == OP_HALF) {                                                               \
                if (nm == MAXM) SD_HALF(MAXM); else SD_HALF(GRP);                              \
            } else if (op == OP_READ) {                                                        \
                if (nm == MAXM) SD_READ(MAXM); else SD_READ(GRP);                              \
            } else if (op == OP_SINGLE) {                                                      \
                unsigned long long rr[GRP], va[GRP], vb[GRP];                                  \
                _Pragma("unroll") for (int m = 0; m < GRP; ++m) {                              \
                    const uint32_t w_ = QW(1 + m);                                             \
                    rr[m] = fx_of(ring[(int)(w_ & (RB - 1)) * 64 + lane]) & ((w_ & M_VALID) ? ~0ull : 0ull); \
                    va[m] = H[(int)((w_ >> 11) & 63u) * 64 + lane];                            \
                    vb[m] = H[(int)((w_ >> 17) & 63u) * 64 + lane];                            \
                }                                                                              \
                _Pragma("unroll") for (int m = 0; m < GRP; ++m) {                              \
                    const uint32_t w_ = QW(1 + m);                                             \
                    otile[(int)(w_ & (OB - 1)) * OSTR + lane] = val_of(HP - va[m] - vb[m] - rr[m]); \
                    HP += 2 * rr[m];                                                           \
                    H[(int)((w_ >> 11) & 63u) * 64 + lane] = va[m] + 2 * rr[m];                \
                    H[(int)((w_ >> 17) & 63u) * 64 + lane] = vb[m] + 2 * rr[m];                \
                }                                                                              \
            }                                                                                  \
            if (h & F_FLUSH) SD_FLUSH(h);                                                      \
        }                                                                                      \
    } while (0)
    // next visit: its records from the prefetch register into LDS, its header read, the
    // visit after it requested; the ring block from set_, which then requests the header's block
#define SD_CROSS(set_, ok_)                                                                    \
    do {                                                                                       \
        vo += nrec;                                                                            \
        pbuf[lane] = pch;                                                                      \
        const uint4 vh_ = pbuf[0];                                                             \
        nrec = (int)rfl(vh_.z);                                                                \
        const uint32_t pf_ = rfl(vh_.y);                                                       \
        pch = gp[(size_t)(vo + nrec) * (REC / 4) + lane];                                      \
        SD_RING(set_, ok_);                                                                    \
        SD_LOAD(pf_, set_, ok_);                                                               \
    } while (0)

    uint32_t vo = hd.x;                                  // first record of the current visit
    int nrec = (int)rfl(pbuf[0].z);
    uint4 pch = gp[(size_t)(vo + nrec) * (REC / 4) + lane];  // visit 1
    // block of visit 0 into the ring; visit 1's block arrives in set a, visit 2's in set b
    // (issued in this order: the waits of the first visit boundaries count on set a's loads
    // being older than set b's)
    float4 sa[RB / 4], sb[RB / 4];
    bool oka, okb;
    SD_LOAD(hd.y, sa, oka);
    SD_RING(sa, oka);
    __builtin_amdgcn_sched_barrier(0);
    SD_LOAD(hd.z, sa, oka);
    __builtin_amdgcn_sched_barrier(0);
    SD_LOAD(hd.w, sb, okb);
    __builtin_amdgcn_sched_barrier(0);
    int oblk = 0;
    bool done = false;
    SD_RUN();
    while (!done) {
        SD_CROSS(sa, oka);
        SD_RUN();
        if (done) break;
        SD_CROSS(sb, okb);
        SD_RUN();
    }
#undef SD_LOAD
#undef SD_RING
#undef SD_FLUSH
#undef QW
#undef SD_RUN
#undef SD_CROSS
#undef SD_HALF
#undef SD_READ
}


// ------------------------------------------------------ prefix scan --
constexpr int SC_PER = NPR / 64;            // 17 combo positions per lane
constexpr int SC_CPW = 16;                  // columns per wave (one wave per workgroup; 16-32 measure alike,
                                            // 2-8 pay the per-wave table and list set-up more often)
constexpr int SC_AZ = NPR + 1;              // As slot that stays 0 (padding positions read it)
constexpr int SC_D = SC_LISTS * (SC_LIST + 1);
constexpr int SC_DZ = SC_D;                 // Ds slot that stays 0
static_assert(NPR % 64 == 0 && NPR % 4 == 0, "combo padding");

__device__ __forceinline__ int sc_excl_scan(int v) {
    const int lane = threadIdx.x & 63;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    return x - v;
}

__device__ __forceinline__ float sc_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// LDS word at a byte offset (the tables hold byte offsets, so each lookup is
// one mask or shift and the array's base is the instruction's immediate).
__device__ __forceinline__ int sc_at(const int* base, uint32_t byte_off) {
    return *reinterpret_cast<const int*>(reinterpret_cast<const char*>(base) + byte_off);
}

// One wave per (board, SC_CPW columns), the columns in turn:
//   1. the column (FP32, 16-byte loads, requested while the previous column
//      was finishing) and its total; R = floor(r 2^sc + 1/2), sc = 27 -
//      exponent of the total, written to LDS;
//   2. card lists in two rounds of 32 lists, two lanes per list (23 entries
//      each, LDS byte offsets held in registers): gathers, prefix, the
//      pair's total by a lane swap, doubled list prefixes D_c to LDS;
//   3. each lane reads 17 consecutive positions: local scan, wave scan of
//      the lane totals; the doubled strength prefixes DP overwrite the
//      column in LDS;
//   4. each lane takes positions lane, lane + 64, .. (their table words held
//      in registers for all the wave's columns): six LDS lookups, one
//      conversion, coalesced stores (padding positions look up zero slots).
template <int CPW>
__global__ __launch_bounds__(64) void k_sd_scan(BatchedDev D, const uint32_t* __restrict__ htab,
                                                const uint32_t* __restrict__ ltab, int ncol) {
    __shared__ int4 As4[(NPR + 4) / 4];
    __shared__ int Ds[SC_D + 4];
    int* As = reinterpret_cast<int*>(As4);
    const LevelDev& L = D.L[1];
    const int b = blockIdx.y, lane = threadIdx.x;
    const int ncolp = L.ncolp;
    const int cw0 = blockIdx.x * CPW;
    const int ncw = min(CPW, ncol - cw0);
    const uint32_t* ht = htab + (size_t)b * 3 * NPR;
    uint32_t lp[2][SC_LHALF_WORDS];
    {
        const uint32_t* lt = ltab + (size_t)b * SC_LTAB_WORDS;
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int k = 0; k < SC_LHALF_WORDS; ++k) lp[r][k] = lt[(r * SC_LHALF_WORDS + k) * 64 + lane];
    }
    if (lane == 0) {
        As[SC_AZ] = 0;
        Ds[SC_DZ] = 0;
    }
    const int half = lane & 1;
    // the column's 272 float4, five per lane (the fifth read on every lane, kept on lanes 0..15)
    const float* Bcol = D.Bf + ((size_t)b * ncolp + cw0) * NPR;
    float4 v[5];
    auto fetch = [&](int q) {
        const float4* rc = reinterpret_cast<const float4*>(Bcol + (size_t)q * NPR);
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = rc[lane + 64 * k];
        v[4] = rc[256 + (lane & 15)];
    };
    fetch(0);
    for (int q = 0; q < ncw; ++q) {
        const int col = cw0 + q;
        // 1. total and fixed point (positions 1081..1087 are padding)
        if (lane >= 16) v[4] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (lane == 14) v[4].y = v[4].z = v[4].w = 0.0f;
        if (lane == 15) v[4] = make_float4(0.f, 0.f, 0.f, 0.f);
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < 5; ++k) s += (v[k].x + v[k].y) + (v[k].z + v[k].w);
        s = sc_sum(s);
        int ex = 0;
        (void)frexpf(s, &ex);
        const int sc = s > 0.0f ? 27 - ex : 0;       // every R_j and their total below 2^27 (+ rounding)
        const float up = ldexpf(1.0f, sc);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int4 r = make_int4((int)fmaf(v[k].x, up, 0.5f), (int)fmaf(v[k].y, up, 0.5f),
                                     (int)fmaf(v[k].z, up, 0.5f), (int)fmaf(v[k].w, up, 0.5f));
            if (k < 4 || lane < 16) As4[lane + 64 * k] = r;
        }
        if (q + 1 < ncw) fetch(q + 1);                // the next column arrives under this one's work
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // 2. card lists: round r holds lists 32 r + lane / 2, half lane & 1 (entries 23 half ..)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int li = 32 * r + (lane >> 1);
            int g[SC_LHALF];
#pragma unroll
            for (int m = 0; m < SC_LHALF; ++m)
                g[m] = sc_at(As, (m & 1) ? lp[r][m >> 1] >> 16 : lp[r][m >> 1] & 0xFFFFu);
            int hs = 0;
#pragma unroll
            for (int m = 0; m < SC_LHALF; ++m) hs += g[m];
            const int other = __shfl_xor(hs, 1, 64);
            const int tc = hs + other;
            int e = half ? other : 0;
            if (li < SC_LISTS) {
                int* Dl = Ds + li * (SC_LIST + 1) + SC_LHALF * half;
#pragma unroll
                for (int m = 0; m < SC_LHALF; ++m) {
                    Dl[m] = 2 * e - tc;
                    e += g[m];
                }
                if (half) Dl[SC_LHALF] = tc;          // entry 46: the list total
            }
        }
        // 3. prefix over 17 consecutive positions per lane, then across lanes; DP over the column
        {
            int x[SC_PER];
#pragma unroll
            for (int t = 0; t < SC_PER; ++t) x[t] = As[SC_PER * lane + t];
            int run = 0;
#pragma unroll
            for (int t = 0; t < SC_PER; ++t) {
                const int r = x[t];
                x[t] = run;                           // exclusive within the lane
                run += r;
            }
            const int base = sc_excl_scan(run);
            const int tot = __shfl(base + run, 63, 64);
            const int off = 2 * base - tot;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int t = 0; t < SC_PER; ++t) As[SC_PER * lane + t] = 2 * x[t] + off;
            if (lane == 63) As[NPR] = tot;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // 4. values
        float* Cc = D.C + ((size_t)b * ncolp + col) * NPR;
        const float unit = ldexpf(1.0f, -(sc + 1));
#pragma unroll
        for (int t = 0; t < SC_PER; ++t) {
            const int j = 64 * t + lane;
            const uint32_t w0 = ht[j], w1 = ht[NPR + j], w2 = ht[2 * NPR + j];
            const int two_c = sc_at(As, w0 & 0xFFFFu) + sc_at(As, w0 >> 16) - sc_at(Ds, w1 & 0xFFFFu) -
                              sc_at(Ds, w1 >> 16) - sc_at(Ds, w2 & 0xFFFFu) - sc_at(Ds, w2 >> 16);
            Cc[j] = (float)two_c * unit;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace

bool sd_program(const uint8_t* cards, const uint16_t* grank, int n, std::vector<uint32_t>* out, uint32_t hdr[4]) {
    if (n < 0 || n > NPR) return false;
    for (int j = 0; j < n; ++j)
        if (cards[2 * j] >= 52 || cards[2 * j + 1] >= 52 || cards[2 * j] == cards[2 * j + 1]) return false;
    for (int j = 1; j < n; ++j)
        if (grank[j] < grank[j - 1]) return false;
    struct Rec {
        uint32_t w[REC];
    };
    struct Visit {
        int block;
        std::vector<Rec> recs;
    };
    std::vector<Visit> visits;
    auto mword = [&](int j) {
        return (uint32_t)j | (uint32_t)cards[2 * j] << 11 | (uint32_t)cards[2 * j + 1] << 17 | M_VALID;
    };
    auto cmask = [&](int j) { return (1ull << cards[2 * j]) | (1ull << cards[2 * j + 1]); };
    // a record needing ring block bl (-1: any), in a new visit if the current one is another block or full
    auto add = [&](const Rec& r, int bl) {
        if (visits.empty()) visits.push_back({bl < 0 ? 0 : bl, {}});
        if ((bl >= 0 && visits.back().block != bl) || (int)visits.back().recs.size() == VISIT - 1) {
            const int nb = bl >= 0 ? bl : visits.back().block;
            visits.push_back({nb, {}});
        }
        visits.back().recs.push_back(r);
    };
    int oblk = 0;
    // after the value of `row` went to the output tile: flush a finished output block
    auto flush_after = [&](int row) {
        if ((row + 1) % OB == 0 || row == n - 1) {
            visits.back().recs.back().w[0] |= F_FLUSH | (uint32_t)(row - oblk * OB) << 10;
            ++oblk;
        }
    };
    // padded to 5 or 15 members: invalid copies of the first member ahead of the members
    auto batch = [&](uint32_t op, const std::vector<int>& mem, int bl) {
        Rec r{};
        const int k = (int)mem.size(), pad = (k <= GRP ? GRP : MAXM) - k;
        r.w[0] = op | (uint32_t)(k + pad) << 4;
        for (int x = 0; x < pad; ++x) r.w[1 + x] = mword(mem[0]) & ~M_VALID;
        for (int x = 0; x < k; ++x) r.w[1 + pad + x] = mword(mem[x]);
        add(r, bl);
    };
    auto half_pass = [&](int i, int e) {  // half of every member's reach, ring block by ring block
        for (int bs = i; bs < e;) {
            const int be = std::min(e, (bs / RB + 1) * RB);
            for (int s0 = bs; s0 < be; s0 += MAXM) {
                std::vector<int> mem;
                for (int j = s0; j < std::min(be, s0 + MAXM); ++j) mem.push_back(j);
                batch(OP_HALF, mem, bs / RB);
            }
            bs = be;
        }
    };
    std::vector<int> single;
    uint64_t smask = 0;
    auto flush_single = [&]() {
        if (single.empty()) return;
        batch(OP_SINGLE, single, single[0] / RB);
        flush_after(single.back());
        single.clear();
        smask = 0;
    };
    int i = 0;
    while (i < n) {
        int e = i + 1;
        while (e < n && grank[e] == grank[i]) ++e;
        if (e - i == 1) {  // consecutive singletons: one output block, distinct cards
            if (!single.empty() &&
                ((int)single.size() == MAXS || (smask & cmask(i)) || single[0] / OB != i / OB))
                flush_single();
            single.push_back(i);
            smask |= cmask(i);
            i = e;
            continue;
        }
        flush_single();
        half_pass(i, e);
        for (int bs = i; bs < e;) {  // every member's value at the half state, output block by output block
            const int be = std::min(e, (bs / OB + 1) * OB);
            for (int s0 = bs; s0 < be; s0 += MAXM) {
                std::vector<int> mem;
                for (int j = s0; j < std::min(be, s0 + MAXM); ++j) mem.push_back(j);
                batch(OP_READ, mem, -1);
            }
            flush_after(be - 1);
            bs = be;
        }
        half_pass(i, e);  // the other half (re-visits the group's blocks)
        i = e;
    }
    flush_single();
    Rec end{};
    end.w[0] = OP_END;
    add(end, -1);
    // serialise: a header record per visit (the block two visits ahead, the record count)
    std::vector<uint32_t>& P = *out;
    hdr[0] = (uint32_t)(P.size() / REC);
    auto blk = [&](size_t v) { return v < visits.size() ? (uint32_t)visits[v].block : NO_BLOCK; };
    hdr[1] = blk(0);
    hdr[2] = blk(1);
    hdr[3] = blk(2);
    for (size_t v = 0; v < visits.size(); ++v) {
        const size_t at = P.size();
        P.resize(at + (size_t)REC * (1 + visits[v].recs.size()), 0u);
        P[at] = OP_VISIT;
        P[at + 1] = blk(v + 2);
        P[at + 2] = (uint32_t)(1 + visits[v].recs.size());
        for (size_t r = 0; r < visits[v].recs.size(); ++r)
            for (int x = 0; x < REC; ++x) P[at + (size_t)REC * (1 + r) + x] = visits[v].recs[r].w[x];
    }
    return true;
}


bool sc_tables(const uint8_t* cards, const uint16_t* grank, int n, uint32_t* htab, uint32_t* ltab) {
    if (n != 1081) return false;
    for (int j = 0; j < n; ++j)
        if (cards[2 * j] >= 52 || cards[2 * j + 1] >= 52 || cards[2 * j] == cards[2 * j + 1]) return false;
    for (int j = 1; j < n; ++j)
        if (grank[j] < grank[j - 1]) return false;
    // live cards in ascending order -> list index; each list: its combos' positions, ascending
    int lid[52];
    for (int c = 0; c < 52; ++c) lid[c] = -1;
    std::vector<int> list[52];
    for (int j = 0; j < n; ++j) {
        list[cards[2 * j]].push_back(j);
        list[cards[2 * j + 1]].push_back(j);
    }
    int nl = 0;
    for (int c = 0; c < 52; ++c)
        if (!list[c].empty()) {
            if ((int)list[c].size() != SC_LIST) return false;
            lid[c] = nl++;
        }
    if (nl != SC_LISTS) return false;
    // padding positions: every lookup reads a zero slot (As[NPR + 1], Ds[47 * 47])
    const uint32_t az = 4u * (NPR + 1), dz = 4u * (SC_LISTS * (SC_LIST + 1));
    for (int j = 0; j < NPR; ++j) {
        htab[j] = az | az << 16;
        htab[NPR + j] = dz | dz << 16;
        htab[2 * NPR + j] = dz | dz << 16;
    }
    for (size_t i = 0; i < (size_t)SC_LTAB_WORDS; ++i) ltab[i] = 0;
    for (int c = 0; c < 52; ++c) {
        if (lid[c] < 0) continue;
        const int r = lid[c] / 32, l2 = 2 * (lid[c] % 32);
        for (int k = 0; k < SC_LIST; ++k) {
            const int h = k / SC_LHALF, m = k % SC_LHALF;
            ltab[(r * SC_LHALF_WORDS + (m >> 1)) * 64 + l2 + h] |= 4u * (uint32_t)list[c][k] << (16 * (m & 1));
        }
    }
    // entries of card c's list before position x
    auto below = [&](int c, int x) {
        return (int)(std::lower_bound(list[c].begin(), list[c].end(), x) - list[c].begin());
    };
    int gs = 0;
    for (int j = 0; j < n; ++j) {
        if (j > 0 && grank[j] != grank[j - 1]) gs = j;
        int ge = j + 1;
        while (ge < n && grank[ge] == grank[j]) ++ge;
        htab[j] = 4u * (uint32_t)gs | 4u * (uint32_t)ge << 16;
        for (int s = 0; s < 2; ++s) {
            const int c = cards[2 * j + s];
            const uint32_t lo = (uint32_t)(lid[c] * (SC_LIST + 1) + below(c, gs));
            const uint32_t hi = (uint32_t)(lid[c] * (SC_LIST + 1) + below(c, ge));
            htab[(size_t)(1 + s) * NPR + j] = 4u * lo | 4u * hi << 16;
        }
    }
    return true;
}

void sc_launch(const BatchedDev& D, const uint32_t* htab, const uint32_t* ltab, int nboards, int ncol,
               hipStream_t st) {
    const dim3 g((ncol + SC_CPW - 1) / SC_CPW, nboards);
    hipLaunchKernelGGL(k_sd_scan<SC_CPW>, g, dim3(64), 0, st, D, htab, ltab, ncol);
}

void sd_launch(const BatchedDev& D, const uint32_t* prog, const uint4* hdrs, const int* colinfo, int nboards, int ncol,
               hipStream_t st) {
    const dim3 g((ncol + 63) / 64, nboards);
    hipLaunchKernelGGL(k_sd_sweep, g, dim3(64), 0, st, D, prog, hdrs, colinfo, ncol);
}

}  // namespace bt
}  // namespace mtt

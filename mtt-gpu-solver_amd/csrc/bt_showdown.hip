// River showdown values by an exact sweep in strength order: the per-board
// program builder (host) and the sweep kernel.
#include "bt_showdown.h"

#include <algorithm>

namespace mtt {
namespace bt {
namespace {

// Program records: 16 words each.  Word 0: opcode (bits 0-3), member count k
// (bits 4-7), ADV flag (bit 8: first make the next block of the program's
// block sequence the ring block, and request the block in bits 16-23, 0xFF
// for none, two steps ahead), FLUSH flag (bit 9: afterwards write rows 0..n-1
// of the 16-row output tile to C, n - 1 in bits 10-13).  Words 1..k: members
// as row | card a << 11 | card b << 17.  Record 0 holds the first three
// blocks of the sequence in words 1-3.
enum : uint32_t {
    OP_NOP = 0,      // flags only
    OP_SINGLE = 1,   // k singleton groups (consecutive rows, distinct cards): read, value, add 2R
    OP_HALF = 2,     // k members of a tie group: add R (LDS atomics: cards may repeat)
    OP_READ = 3,     // k members of a tie group: value at the half state
    OP_END = 4,
};
constexpr uint32_t F_ADV = 1u << 8, F_FLUSH = 1u << 9;
constexpr uint32_t NO_BLOCK = 0xFFu;
constexpr int REC = SD_REC;
constexpr int CHUNK = SD_CHUNK;     // records per program chunk (one uint4 per lane)

constexpr int RB = 32;              // rows per ring block
constexpr int OB = 16;              // rows per output block
constexpr int MAXM = REC - 1;       // members per HALF / READ record
constexpr int MAXS = 4;             // members per SINGLE record
constexpr int OSTR = 68;            // output tile row stride (floats): two lanes per bank on write and flush
constexpr int NRB = NPR / RB;       // 34 ring blocks
static_assert(NPR % RB == 0 && RB % OB == 0, "row blocks");

__device__ __forceinline__ unsigned long long fx_of(float r) { return __double2ull_rn((double)r * FX); }
// C = W - L, summed in 2^48 units (two's complement)
__device__ __forceinline__ float val_of(unsigned long long v) { return (float)((double)(long long)v * 0x1p-48); }
__device__ __forceinline__ uint32_t rfl(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
// Word i of a record held (uniform) in four vector registers, as a scalar.
__device__ __forceinline__ uint32_t qword(const uint4& q0, const uint4& q1, const uint4& q2, const uint4& q3, int i) {
    const uint4& q = i < 4 ? q0 : (i < 8 ? q1 : (i < 12 ? q2 : q3));
    const int j = i & 3;
    return rfl(j == 0 ? q.x : (j == 1 ? q.y : (j == 2 ? q.z : q.w)));
}

// One wave per (64 columns, board).  LDS: card sums H[card][lane] (2^48
// fixed point, two's complement, offset by -csT), the current 32-row block
// of the columns ring[row][lane], the 16-row output tile otile[row][lane]
// and the current 16-record chunk of the program (read by every lane at one
// address, then made uniform).  Column blocks arrive two blocks ahead and
// program chunks one chunk ahead, by vector loads.
__global__ __launch_bounds__(64) void k_sd_sweep(BatchedDev D, const uint32_t* __restrict__ prog,
                                                 const uint32_t* __restrict__ poff, const int* __restrict__ colinfo,
                                                 int ncol) {
    __shared__ unsigned long long H[52 * 64];
    __shared__ float ring[RB * 64];
    __shared__ float otile[OB * OSTR];
    __shared__ uint4 pbuf[CHUNK * REC / 4];
    const LevelDev& L = D.L[1];
    const int b = blockIdx.y, lane = threadIdx.x, c0 = blockIdx.x * 64, col = c0 + lane;
    const int ncolp = L.ncolp;
    const bool cl = col < ncol;
    const float* rc = D.Bf + ((size_t)b * ncolp + col) * NPR;  // read only when cl

    // program: chunk 0 into LDS, chunk 1 in flight
    const uint4* gp = reinterpret_cast<const uint4*>(prog) + (size_t)poff[b] * (REC / 4);
    pbuf[lane] = gp[lane];
    uint4 pch = gp[CHUNK * REC / 4 + lane];

    // H = -csT (the column's card totals, both forward halves), HP = -Ptot
    unsigned long long tot = 0;
    {
        const int ci = cl ? colinfo[col] : 0;
        const int p = (int)((unsigned)ci >> 31), t = ci & 0x7FFFFFFF;
        const unsigned long long* P0 = L.part + ((size_t)b * L.nterm + t) * 208 + (1 - p) * 52;
#pragma unroll 4
        for (int c = 0; c < 52; ++c) {
            const unsigned long long v = cl ? P0[c] + P0[104 + c] : 0ull;
            H[c * 64 + lane] = 0ull - v;
            tot += v;
        }
    }
    unsigned long long HP = 0ull - (tot >> 1);

#define SD_LOAD(blk, set_)                                                                        \
    do {                                                                                       \
        if (cl && (int)(blk) < NRB) {                                                               \
            const float4* q_ = reinterpret_cast<const float4*>(rc + (size_t)(blk) * RB);       \
            _Pragma("unroll") for (int i_ = 0; i_ < RB / 4; ++i_) set_[i_] = q_[i_];              \
        } else {                                                                               \
            _Pragma("unroll") for (int i_ = 0; i_ < RB / 4; ++i_) set_[i_] = make_float4(0.f, 0.f, 0.f, 0.f); \
        }                                                                                      \
    } while (0)
#define SD_RING(set_)                                                                             \
    do {                                                                                       \
        _Pragma("unroll") for (int i_ = 0; i_ < RB / 4; ++i_) {                                \
            ring[(4 * i_ + 0) * 64 + lane] = set_[i_].x;                                          \
            ring[(4 * i_ + 1) * 64 + lane] = set_[i_].y;                                          \
            ring[(4 * i_ + 2) * 64 + lane] = set_[i_].z;                                          \
            ring[(4 * i_ + 3) * 64 + lane] = set_[i_].w;                                          \
        }                                                                                      \
    } while (0)

    // the sequence's first block into the ring, the next two in flight (set v & 1 holds the
    // block of ADV number v + 1)
    const uint32_t b0 = rfl(pbuf[0].y), b1 = rfl(pbuf[0].z), b2 = rfl(pbuf[0].w);
    float4 sa[RB / 4], sb[RB / 4];
    SD_LOAD(b0, sa);
    SD_RING(sa);
    SD_LOAD(b1, sb);
    SD_LOAD(b2, sa);
    int nadv = 0, oblk = 0;

#define SD_ADV(nb)                                                                             \
    do {                                                                                       \
        ++nadv;                                                                                \
        if (nadv & 1) {                                                                        \
            SD_RING(sb);                                                                       \
            SD_LOAD(nb, sb);                                                                   \
        } else {                                                                               \
            SD_RING(sa);                                                                       \
            SD_LOAD(nb, sa);                                                                   \
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
    // record ri + 1 becomes current (cq*); then the read of the next one is issued (at a
    // chunk boundary the next chunk is written to LDS first and the one after requested)
#define SD_NEXT()                                                                              \
    do {                                                                                       \
        cq0 = n0; cq1 = n1; cq2 = n2; cq3 = n3;                                                \
        const int nr_ = ri + 1;                                                                \
        if ((nr_ & (CHUNK - 1)) == 0) {                                                        \
            pbuf[lane] = pch;                                                                  \
            pch = gp[(size_t)(nr_ + CHUNK) * (REC / 4) + lane];                                \
        }                                                                                      \
        const int pb_ = (nr_ & (CHUNK - 1)) * (REC / 4);                                       \
        n0 = pbuf[pb_]; n1 = pbuf[pb_ + 1]; n2 = pbuf[pb_ + 2]; n3 = pbuf[pb_ + 3];            \
    } while (0)
#define QW(i) qword(cq0, cq1, cq2, cq3, (i))

    uint4 cq0, cq1, cq2, cq3;
    uint4 n0 = pbuf[REC / 4], n1 = pbuf[REC / 4 + 1], n2 = pbuf[REC / 4 + 2], n3 = pbuf[REC / 4 + 3];
    int ri = 1;  // record 0 is the header
    for (;;) {
        SD_NEXT();
        const uint32_t h = QW(0);
        const uint32_t op = h & 15u;
        const int k = (int)((h >> 4) & 15u);
        if (op == OP_END) break;
        if (h & F_ADV) SD_ADV(h >> 16);
        if (op == OP_HALF) {
            float r[MAXM];
#pragma unroll
            for (int m = 0; m < MAXM; ++m)
                if (m < k) r[m] = ring[(int)(QW(1 + m) & (RB - 1)) * 64 + lane];
#pragma unroll
            for (int m = 0; m < MAXM; ++m) {
                if (m < k) {
                    const uint32_t w_ = QW(1 + m);
                    const unsigned long long R = fx_of(r[m]);
                    HP += R;
                    atomicAdd(&H[(int)((w_ >> 11) & 63u) * 64 + lane], R);
                    atomicAdd(&H[(int)((w_ >> 17) & 63u) * 64 + lane], R);
                }
            }
        } else if (op == OP_READ) {
            unsigned long long va[MAXM], vb[MAXM];
#pragma unroll
            for (int m = 0; m < MAXM; ++m) {
                if (m < k) {
                    const uint32_t w_ = QW(1 + m);
                    va[m] = H[(int)((w_ >> 11) & 63u) * 64 + lane];
                    vb[m] = H[(int)((w_ >> 17) & 63u) * 64 + lane];
                }
            }
#pragma unroll
            for (int m = 0; m < MAXM; ++m)
                if (m < k) otile[(int)(QW(1 + m) & (OB - 1)) * OSTR + lane] = val_of(HP - va[m] - vb[m]);
        } else if (op == OP_SINGLE) {
            float r[MAXS];
            unsigned long long va[MAXS], vb[MAXS];
#pragma unroll
            for (int m = 0; m < MAXS; ++m) {
                if (m < k) {
                    const uint32_t w_ = QW(1 + m);
                    r[m] = ring[(int)(w_ & (RB - 1)) * 64 + lane];
                    va[m] = H[(int)((w_ >> 11) & 63u) * 64 + lane];
                    vb[m] = H[(int)((w_ >> 17) & 63u) * 64 + lane];
                }
            }
#pragma unroll
            for (int m = 0; m < MAXS; ++m) {
                if (m < k) {
                    const uint32_t w_ = QW(1 + m);
                    const unsigned long long R = fx_of(r[m]);
                    otile[(int)(w_ & (OB - 1)) * OSTR + lane] = val_of(HP - va[m] - vb[m] - R);
                    HP += 2 * R;
                    H[(int)((w_ >> 11) & 63u) * 64 + lane] = va[m] + 2 * R;
                    H[(int)((w_ >> 17) & 63u) * 64 + lane] = vb[m] + 2 * R;
                }
            }
        }
        if (h & F_FLUSH) SD_FLUSH(h);
        ++ri;
    }
#undef SD_LOAD
#undef SD_RING
#undef SD_ADV
#undef SD_FLUSH
#undef SD_NEXT
#undef QW
}

}  // namespace

bool sd_program(const uint8_t* cards, const uint16_t* grank, int n, std::vector<uint32_t>* out) {
    if (n < 0 || n > NPR) return false;
    for (int j = 0; j < n; ++j)
        if (cards[2 * j] >= 52 || cards[2 * j + 1] >= 52 || cards[2 * j] == cards[2 * j + 1]) return false;
    for (int j = 1; j < n; ++j)
        if (grank[j] < grank[j - 1]) return false;
    std::vector<uint32_t>& P = *out;
    const size_t start = P.size();
    uint32_t pend = 0;                   // flags for the next record
    size_t last = 0;                     // first word of the last record
    std::vector<int> seq{0};             // ring blocks in the order the program visits them
    std::vector<size_t> adv_at;          // first word of each ADV record
    auto record = [&](uint32_t hdr, const uint32_t* ws, int nw) {
        last = P.size();
        P.resize(last + REC, 0u);
        P[last] = hdr | pend;
        if (pend & F_ADV) adv_at.push_back(last);
        pend = 0;
        for (int x = 0; x < nw; ++x) P[last + 1 + x] = ws[x];
    };
    record(OP_NOP, nullptr, 0);          // header: first blocks, filled in below
    auto mword = [&](int j) { return (uint32_t)j | (uint32_t)cards[2 * j] << 11 | (uint32_t)cards[2 * j + 1] << 17; };
    auto cmask = [&](int j) { return (1ull << cards[2 * j]) | (1ull << cards[2 * j + 1]); };
    int oblk = 0;
    auto need_block = [&](int bl) {      // ring block bl current
        if (seq.back() == bl) return;
        if (pend & F_ADV) record(OP_NOP, nullptr, 0);
        pend |= F_ADV;
        seq.push_back(bl);
    };
    // after the value of `row` went to the output tile: flush a finished output block
    auto flush_after = [&](int row) {
        if ((row + 1) % OB == 0 || row == n - 1) {
            P[last] |= F_FLUSH | (uint32_t)(row - oblk * OB) << 10;
            ++oblk;
        }
    };
    auto batch = [&](uint32_t op, const std::vector<int>& mem) {
        uint32_t w[MAXM];
        for (size_t x = 0; x < mem.size(); ++x) w[x] = mword(mem[x]);
        record(op | (uint32_t)mem.size() << 4, w, (int)mem.size());
    };
    auto half_pass = [&](int i, int e) {  // half of every member's reach, ring block by ring block
        for (int bs = i; bs < e;) {
            const int be = std::min(e, (bs / RB + 1) * RB);
            need_block(bs / RB);
            for (int s0 = bs; s0 < be; s0 += MAXM) {
                std::vector<int> mem;
                for (int j = s0; j < std::min(be, s0 + MAXM); ++j) mem.push_back(j);
                batch(OP_HALF, mem);
            }
            bs = be;
        }
    };
    std::vector<int> single;
    uint64_t smask = 0;
    auto flush_single = [&]() {
        if (single.empty()) return;
        need_block(single[0] / RB);
        batch(OP_SINGLE, single);
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
                batch(OP_READ, mem);
            }
            flush_after(be - 1);
            bs = be;
        }
        half_pass(i, e);  // the other half (re-visits the group's ring blocks)
        i = e;
    }
    flush_single();
    record(OP_END, nullptr, 0);
    record(OP_NOP, nullptr, 0);  // spare: read ahead by the kernel
    while ((P.size() - start) % (CHUNK * REC)) record(OP_NOP, nullptr, 0);
    // the block each ADV requests (two steps ahead in the sequence), and the first three
    auto at = [&](size_t v) { return v < seq.size() ? (uint32_t)seq[v] : NO_BLOCK; };
    for (size_t v = 0; v < adv_at.size(); ++v) P[adv_at[v]] |= at(v + 3) << 16;
    P[start + 1] = at(0);
    P[start + 2] = at(1);
    P[start + 3] = at(2);
    return true;
}

void sd_launch(const BatchedDev& D, const uint32_t* prog, const uint32_t* poff, const int* colinfo, int nboards,
               int ncol, hipStream_t st) {
    const dim3 g((ncol + 63) / 64, nboards);
    hipLaunchKernelGGL(k_sd_sweep, g, dim3(64), 0, st, D, prog, poff, colinfo, ncol);
}

}  // namespace bt
}  // namespace mtt

// Shared definitions for the MI355X solver core: card encoding, the 7-card
// evaluator (host + device), error plumbing for the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mtt.h"

#define MTT_HD __host__ __device__ __forceinline__

namespace mtt {

// ------------------------------------------------------------ error state --
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define MTT_HIP(call)                                                        \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess) return ::mtt::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define MTT_ARG(cond, ...)                                                   \
    do {                                                                     \
        if (!(cond)) { ::mtt::set_error(__VA_ARGS__); return MTT_ERR_ARG; }  \
    } while (0)

// ------------------------------------------------------------------ cards --
// Card c = rank*4 + suit, rank 0='2'..12='A' (scripts/optimize_subset.py:59-61).
constexpr int kAliveFlop = 1176;   // combos avoiding a flop
constexpr int kPadFlop = 1216;     // reference padding, 19 x 64
constexpr int kAliveTurn = 1128;   // combos avoiding 4 board cards
constexpr int kAliveRiver = 1081;  // combos avoiding 5 board cards
// A player's range may total at most this much: per-card reach sums are
// kept in signed 64-bit fixed point with 48 fraction bits.
constexpr double kMaxRangeTotal = 32768.0;

MTT_HD int hibit(uint32_t x) { return 31 - __builtin_clz(x); }

// Highest 5-run top rank in a 13-bit mask, wheel = 3, none = -1
// (same result as mtt_ev_dumper.cu:57-61; run detection by shifted ANDs).
MTT_HD int straight_top(uint32_t m) {
    uint32_t run = m & (m >> 1) & (m >> 2) & (m >> 3) & (m >> 4);
    if (run) return hibit(run) + 4;
    return ((m & 0x100Fu) == 0x100Fu) ? 3 : -1;
}

// 7-card score from the four 13-bit suit masks of the hand, with the
// reference's total preorder (mtt_ev_dumper.cu:53-83): flush first, quads and
// trips and pairs without kickers, high card = raw mask of all seven ranks.
// Rank multiplicities come from suit-mask intersections instead of a
// per-rank histogram, so the evaluator is branch-light integer VALU work.
MTT_HD uint32_t score_from_suits(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3) {
    uint32_t fm = 0;
    if (__builtin_popcount(s0) >= 5) fm = s0;
    else if (__builtin_popcount(s1) >= 5) fm = s1;
    else if (__builtin_popcount(s2) >= 5) fm = s2;
    else if (__builtin_popcount(s3) >= 5) fm = s3;
    if (fm) {
        int sf = straight_top(fm);
        if (sf >= 0) return (8u << 24) | (uint32_t)sf;
        uint32_t m = fm;
        while (__builtin_popcount(m) > 5) m &= m - 1;  // keep the five highest
        uint32_t sc = 5u << 24;
#pragma unroll
        for (int p = 0; p < 5; ++p) {
            int r = hibit(m);
            sc |= (uint32_t)r << (4 * (4 - p));
            m &= ~(1u << r);
        }
        return sc;
    }
    uint32_t rm = s0 | s1 | s2 | s3;
    uint32_t a01 = s0 & s1, a23 = s2 & s3;
    uint32_t quads = a01 & a23;
    uint32_t ge3 = (a01 & (s2 | s3)) | (a23 & (s0 | s1));
    uint32_t ge2 = a01 | a23 | ((s0 | s1) & (s2 | s3));
    uint32_t trips = ge3 & ~quads, pairs = ge2 & ~ge3;
    if (quads) return (7u << 24) | ((uint32_t)hibit(quads) << 4);
    if (trips) {
        int t1 = hibit(trips);
        uint32_t rest = trips & ~(1u << t1);
        if (rest) return (6u << 24) | ((uint32_t)t1 << 4) | (uint32_t)hibit(rest);
        if (pairs) return (6u << 24) | ((uint32_t)t1 << 4) | (uint32_t)hibit(pairs);
    }
    int st = straight_top(rm);
    if (st >= 0) return (4u << 24) | (uint32_t)st;
    if (trips) return (3u << 24) | ((uint32_t)hibit(trips) << 8);
    if (pairs) {
        int p1 = hibit(pairs);
        uint32_t rest = pairs & ~(1u << p1);
        if (rest) return (2u << 24) | ((uint32_t)p1 << 8) | ((uint32_t)hibit(rest) << 4);
        return (1u << 24) | ((uint32_t)p1 << 12);
    }
    return rm;
}

// Suit masks packed 13 bits per suit into a 64-bit word (suit s at bit 16*s)
// so that adding cards is one OR.
MTT_HD uint64_t card_bits16(int c) { return 1ull << (16 * (c & 3) + (c >> 2)); }
MTT_HD uint32_t score_packed(uint64_t m) {
    return score_from_suits((uint32_t)(m & 0x1FFF), (uint32_t)((m >> 16) & 0x1FFF),
                            (uint32_t)((m >> 32) & 0x1FFF), (uint32_t)((m >> 48) & 0x1FFF));
}

MTT_HD uint32_t eval7(const uint8_t* c) {
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) m |= card_bits16(c[k]);
    return score_packed(m);
}

// Order-preserving 16-bit code of a score: code16(a) < code16(b) exactly when
// a < b, equal exactly when equal.  The preorder has few classes once
// kickers are dropped: a category base plus the payload's position in its
// category, high card and flush by their 13-bit rank masks (numeric order of
// equal-size masks is the order of their descending rank lists).  Range
// [0, 16781); the count builder compares these in 16-bit lanes.
MTT_HD uint32_t code16(uint32_t s) {
    uint32_t cat = s >> 24, p = s & 0xFFFFFFu;
    switch (cat) {
        case 0: return p;                                              // 7-rank mask
        case 1: return 8192u + (p >> 12);                              // pair
        case 2: return 8205u + (p >> 8) * 13u + ((p >> 4) & 15u);      // two pair
        case 3: return 8374u + (p >> 8);                               // trips
        case 4: return 8387u + p - 3u;                                 // straight top 3..12
        case 5: {                                                      // flush: 5 rank nibbles
            uint32_t m = 0;
            for (int k = 0; k < 5; ++k) m |= 1u << ((p >> (4 * k)) & 15u);
            return 8397u + m;
        }
        case 6: return 16589u + (p >> 4) * 13u + (p & 15u);           // full house
        case 7: return 16758u + (p >> 4);                              // quads
        default: return 16771u + p - 3u;                               // straight flush
    }
}
constexpr uint32_t kCode16Range = 16781;

// Lexicographic rank of (a, b), a < b, among the 1,326 two-card combos.
MTT_HD int combo_index(int a, int b) { return a * 51 - a * (a - 1) / 2 + (b - a - 1); }

// Lexicographic (c1<c2) combos avoiding `dead`, as in mtt_ev_dumper.cu:147-152.
inline int alive_combos(uint64_t dead, uint8_t* c1s, uint8_t* c2s) {
    int n = 0;
    for (int a = 0; a < 52; ++a)
        for (int b = a + 1; b < 52; ++b) {
            uint64_t m = (1ull << a) | (1ull << b);
            if (m & dead) continue;
            c1s[n] = (uint8_t)a;
            c2s[n] = (uint8_t)b;
            n++;
        }
    return n;
}

}  // namespace mtt

// River showdown values of the batched heads-up path by an exact sweep in
// strength order (the alternative to the sign x blocker GEMM).
//
// For a river board whose live combos are sorted by strength, a showdown
// column r (the opponent's reach over the combos) gives hero i (cards a, b,
// tie group [gs, ge)) the value
//     C[i] = sum of r_j over non-blocked j in weaker groups
//            - sum of r_j over non-blocked j in stronger groups,
// the same quantity as S r with S = sign x blocker (the GEMM's product).
// With r in 2^48 fixed point (R_j), prefix sums P(x) = sum_{j<x} R_j and
// per-card prefix sums cs_c(x) = sum_{j<x, c in j} R_j:
//     C[i] = P(gs) + P(ge) - Ptot - sum_{c in {a,b}} (cs_c(gs) + cs_c(ge) - csT_c)
// exactly (integers), where Ptot, csT_c are the column's totals -- the card
// sums the forward sweep already accumulates per terminal.
//
// One wave sweeps 64 columns of one board (lane = column), so the card
// indices of each step are wave-uniform and the 52 running card sums of a
// lane sit in LDS at [card][lane].  A per-board program (built on the host
// from the board's strength order) drives the sweep: singleton groups read
// their two card sums and add their reach; a tie group saves the card sums
// it touches, adds half of every member's reach (then every member reads
// exactly cs(gs) + cs(ge)), and restores the sums to old + 2 x half.
// Members are packed into batches whose cards are pairwise distinct, so a
// batch costs one LDS round trip.
#pragma once
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

#include "bt_dev.h"

namespace mtt {
namespace bt {

// Appends the sweep program of one board -- n live combos in strength
// order, cards c[2 * j], c[2 * j + 1] and group ranks g[j] (non-decreasing)
// -- as a sequence of visits: each visit names one 16-row block of the
// columns and holds at most SD_VISIT 16-word records (the first a header).
// hdr receives the board's first record and the first three blocks visited.
// Returns false if n exceeds NPR, a card is invalid or the ranks are not
// sorted.
bool sd_program(const uint8_t* cards, const uint16_t* grank, int n, std::vector<uint32_t>* prog, uint32_t hdr[4]);

// Words per program record; records per visit (the kernel stages one visit in LDS).
constexpr int SD_REC = 16;
constexpr int SD_VISIT = 16;
// Records the kernel may read past the end of the last program.
constexpr int SD_PAD_RECORDS = SD_VISIT;

// ------------------------------------------------------ prefix scan --
// The same values by prefix sums, one wave per column (lanes over combos).
// With the column in integer fixed point R_j (2^sc units, sc chosen so that
// the column total stays below 2^27) and, in doubled form,
//     DP[x]     = 2 P(x) - Ptot                   (P: prefix in strength order)
//     D_c[k]    = 2 ex_c(k) - tot_c               (ex_c: prefix over the 46
//                                                  combos holding card c)
// hero i (cards a, b, tie group [gs, ge), whose group covers entries
// [lo_c, hi_c) of card c's list) has
//     2 C[i] = DP[gs] + DP[ge] - D_a[lo_a] - D_a[hi_a] - D_b[lo_b] - D_b[hi_b]
// exactly.  Per board the tables hold LDS byte offsets (4 x index): per combo
// position j, 4 gs | 4 ge << 16, 4 (a's list * 47 + lo_a) | 4 (a's list * 47
// + hi_a) << 16 and the same for b (htab[3][NPR]; padding positions point at
// zero slots: 4 (NPR + 1) and 4 * 47 * 47), and per live card its 46 combos'
// positions, two per word: ltab[2][12][64], list c in round c / 32 on lanes
// 2 (c % 32) (entries 0..22) and 2 (c % 32) + 1 (entries 23..45).  River
// boards only:
// n = 1081 combos of 47 live cards.  Returns false otherwise, or if a card is
// invalid or the ranks are not sorted.
constexpr int SC_LISTS = 47;
constexpr int SC_LIST = 46;
constexpr int SC_LHALF = SC_LIST / 2;                       // entries per lane
constexpr int SC_LHALF_WORDS = (SC_LHALF + 1) / 2;          // 12 words per lane and round
constexpr int SC_LTAB_WORDS = 2 * SC_LHALF_WORDS * 64;      // per board
bool sc_tables(const uint8_t* cards, const uint16_t* grank, int n, uint32_t* htab, uint32_t* ltab);
// Launches the scan for every board of level 1: columns from D.Bf, values
// to D.C (0 on padding); htab / ltab: the boards' tables, back to back.
void sc_launch(const BatchedDev& D, const uint32_t* htab, const uint32_t* ltab, int nboards, int ncol,
               hipStream_t st);

// Launches the sweep for every board of level 1: columns from D.Bf, card
// totals from the forward's card sums (L.part), values to D.C.
// prog: the concatenated programs followed by SD_PAD_RECORDS zero records;
// hdrs[board] = first record, first three blocks; colinfo[col] =
// terminal | player << 31 of each showdown column.
void sd_launch(const BatchedDev& D, const uint32_t* prog, const uint4* hdrs, const int* colinfo, int nboards, int ncol,
               hipStream_t st);

}  // namespace bt
}  // namespace mtt

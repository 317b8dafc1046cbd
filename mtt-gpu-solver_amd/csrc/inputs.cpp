// Seeded synthetic inputs: boards and ranges.  The reference has no RNG
// library; these generators are the project's own (SplitMix64, partial
// Fisher-Yates, Marsaglia-Tsang gamma with the alpha<1 boost, Box-Muller
// normals) so that every input of the benchmark configurations is a pure
// function of its seed on any host.
#include <cmath>
#include <cstring>

#include "mtt_common.h"

namespace mtt {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
    }
    double gamma(double a) {
        if (a < 1.0) {
            double g = gamma(a + 1.0);
            double u = uniform();
            if (u < 1e-300) u = 1e-300;
            return g * std::pow(u, 1.0 / a);
        }
        double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x = normal(), v = 1.0 + c * x;
            if (v <= 0) continue;
            v = v * v * v;
            double u = uniform();
            if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
            if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
        }
    }
};

}  // namespace mtt

extern "C" {

// n cards without replacement: the first n of a partial Fisher-Yates
// shuffle of 0..51, j = i + floor(u * (52 - i)).
int mtt_seeded_cards(uint64_t seed, int n, uint8_t* out) {
    MTT_ARG(out && n >= 0 && n <= 52, "mtt_seeded_cards: bad arguments");
    mtt::SplitMix64 rng(seed);
    uint8_t deck[52];
    for (int i = 0; i < 52; ++i) deck[i] = (uint8_t)i;
    for (int i = 0; i < n; ++i) {
        int j = i + (int)(rng.uniform() * (52 - i));
        uint8_t t = deck[i];
        deck[i] = deck[j];
        deck[j] = t;
        out[i] = deck[i];
    }
    return MTT_OK;
}

// Board and two ranges from one seed: the board is drawn first (n_board
// cards), then player 0's weights over the 1,326 combos in lexicographic
// order, then player 1's.  dist 0: Uniform[0,1] drawn for unblocked combos
// only; dist 1: Dirichlet(alpha) drawn for all 1,326 combos.  Blocked combos
// get 0 and each range is normalised to sum 1.
int mtt_seeded_inputs(uint64_t seed, int n_board, int dist, float alpha, uint8_t* board_out,
                      float* range0_out, float* range1_out) {
    MTT_ARG(board_out && range0_out && range1_out, "mtt_seeded_inputs: null pointer");
    MTT_ARG(n_board >= 0 && n_board <= 5, "mtt_seeded_inputs: n_board %d", n_board);
    MTT_ARG(dist == 0 || (dist == 1 && alpha > 0), "mtt_seeded_inputs: bad distribution");
    mtt::SplitMix64 rng(seed);
    uint8_t deck[52];
    for (int i = 0; i < 52; ++i) deck[i] = (uint8_t)i;
    uint64_t bm = 0;
    for (int i = 0; i < n_board; ++i) {
        int j = i + (int)(rng.uniform() * (52 - i));
        uint8_t t = deck[i];
        deck[i] = deck[j];
        deck[j] = t;
        board_out[i] = deck[i];
        bm |= 1ull << deck[i];
    }
    for (int p = 0; p < 2; ++p) {
        float* out = p == 0 ? range0_out : range1_out;
        double w[1326], sum = 0;
        int k = 0;
        for (int a = 0; a < 52; ++a)
            for (int b = a + 1; b < 52; ++b, ++k) {
                bool blocked = ((1ull << a) | (1ull << b)) & bm;
                double x;
                if (dist == 0) x = blocked ? 0.0 : rng.uniform();
                else {
                    x = rng.gamma(alpha);
                    if (blocked) x = 0.0;
                }
                w[k] = x;
                sum += x;
            }
        for (k = 0; k < 1326; ++k) out[k] = (float)(sum > 0 ? w[k] / sum : 0.0);
    }
    return MTT_OK;
}

}  // extern "C"

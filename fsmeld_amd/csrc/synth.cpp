// synth.cpp -- seeded synthetic targets and queries (host only).
//
// SplitMix64 seeds a xoshiro256** stream; normals come from Box-Muller.  The
// generator is deterministic for a given seed on a given libm, so the same
// workload is produced on every machine of this image.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/fsmeld.h"

namespace {

uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    bool has_spare = false;
    double spare = 0.0;
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix64(x);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    int64_t below(int64_t n) { return (int64_t)(uniform() * (double)n); }
    double normal() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        const double u1 = 1.0 - uniform();  // (0, 1]
        const double u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double a = 2.0 * M_PI * u2;
        spare = r * std::sin(a);
        has_spare = true;
        return r * std::cos(a);
    }
};

}  // namespace

extern "C" int fsm_synth_target(uint64_t seed, int64_t n, double spike_frac, double spike_sd,
                                int n_const_runs, int const_len, int n_nans, double* out) {
    if (!out || n <= 0 || spike_frac < 0 || n_const_runs < 0 || const_len < 0 || n_nans < 0)
        return FSM_ERR_ARG;
    Rng rng(seed);
    double acc = 0.0;
    for (int64_t i = 0; i < n; i++) {
        acc += rng.normal();
        out[i] = acc;
    }
    if (spike_frac > 0) {
        const int64_t ns = (int64_t)std::llround(spike_frac * (double)n);
        for (int64_t k = 0; k < ns; k++) {
            const int64_t p = rng.below(n);
            out[p] += spike_sd * rng.normal();
        }
    }
    if (const_len > 0 && n > const_len) {
        for (int r = 0; r < n_const_runs; r++) {
            const int64_t a = rng.below(n - const_len);
            for (int64_t k = 1; k < const_len; k++) out[a + k] = out[a];
        }
    }
    for (int k = 0; k < n_nans; k++) out[rng.below(n)] = std::nan("");
    return FSM_OK;
}

extern "C" int fsm_synth_query(uint64_t seed, const double* target, int64_t n, int m,
                               double noise_sd, double* out_q, int64_t* offset) {
    if (!target || !out_q || m < 2 || n < m) return FSM_ERR_ARG;
    Rng rng(seed ^ 0x5bd1e9955bd1e995ull);
    for (int attempt = 0; attempt < 100000; attempt++) {
        const int64_t o = rng.below(n - m + 1);
        bool ok = true;
        for (int i = 0; i < m && ok; i++) {
            if (!std::isfinite(target[o + i])) ok = false;
            if (i > 0 && target[o + i] == target[o + i - 1]) ok = false;
        }
        if (!ok) continue;
        for (int i = 0; i < m; i++) out_q[i] = target[o + i] + noise_sd * rng.normal();
        if (offset) *offset = o;
        return FSM_OK;
    }
    return FSM_ERR_ARG;
}

// gen.cpp - seeded, reproducible workload generators.
//
// Replaces MakeYCSB (LTPG/YCSB/0/Query.h:17-121) and the Random class
// (YCSB/0/Random.h), whose uniform_dist() calls rand() after
// srand(time(0)) from many threads (Random.h:12,52) and so cannot be
// reproduced.  Here every random draw is a pure function of
// (seed, transaction index, slot, attempt): SplitMix64 (Steele, Lea and
// Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014)
// applied to a mixed counter.
//
// Zipf ranks follow Gray et al., "Quickly generating billion-record synthetic
// databases", SIGMOD 1994 (the method of the YCSB ZipfianGenerator).  A rank
// r in [0, rows) maps to row (r * M) mod rows with M = (2^61 - 1) mod rows;
// 2^61 - 1 is prime, so the map is a bijection and the hottest rank is row 0
// (the reference's hot key, YCSB/0/Query.h:34-37).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ltpg.h"

namespace ltpg {
void set_error(const std::string &msg);
}

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Counter-based draw k of (txn, slot) under seed.
inline uint64_t draw(uint64_t seed, uint64_t txn, uint32_t slot, uint32_t k)
{
    uint64_t h = splitmix64(seed ^ 0x4C5450474C545047ull);
    h = splitmix64(h ^ txn);
    h = splitmix64(h ^ (((uint64_t)slot << 32) | k));
    return h;
}

inline double unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

inline uint64_t below(uint64_t x, uint64_t n) { return (uint64_t)(((unsigned __int128)x * n) >> 64); }

// zeta(n, theta) = sum_{i=1..n} i^-theta, summed in fixed chunks and
// combined in order so the result does not depend on the thread count.
double zeta(uint64_t n, double theta)
{
    static std::mutex mu;
    static std::map<std::pair<uint64_t, double>, double> cache;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find({n, theta});
        if (it != cache.end()) return it->second;
    }
    const uint64_t chunk = 1u << 20;
    const uint64_t nchunks = (n + chunk - 1) / chunk;
    std::vector<double> part(nchunks, 0.0);
    unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nt; w++) {
        th.emplace_back([&, w] {
            for (uint64_t c = w; c < nchunks; c += nt) {
                double s = 0;
                const uint64_t lo = c * chunk + 1, hi = std::min(n, (c + 1) * chunk);
                for (uint64_t i = hi; i >= lo; i--) s += pow((double)i, -theta);  // small terms first
                part[c] = s;
            }
        });
    }
    for (auto &t : th) t.join();
    double z = 0;
    for (uint64_t c = nchunks; c-- > 0;) z += part[c];
    std::lock_guard<std::mutex> g(mu);
    cache[{n, theta}] = z;
    return z;
}

struct Zipf {
    uint64_t n;
    double theta, zetan, alpha, eta, half_pow;
    Zipf(uint64_t n_, double theta_) : n(n_), theta(theta_)
    {
        zetan = zeta(n, theta);
        const double zeta2 = 1.0 + pow(0.5, theta);
        alpha = 1.0 / (1.0 - theta);
        eta = (1.0 - pow(2.0 / (double)n, 1.0 - theta)) / (1.0 - zeta2 / zetan);
        half_pow = pow(0.5, theta);
    }
    uint64_t rank(double u) const
    {
        const double uz = u * zetan;
        if (uz < 1.0) return 0;
        if (uz < 1.0 + half_pow) return 1;
        uint64_t r = (uint64_t)((double)n * pow(eta * u - eta + 1.0, alpha));
        return r >= n ? n - 1 : r;
    }
};

}  // namespace

extern "C" {

void ltpg_kv_gen_init(ltpg_kv_gen *g)
{
    if (!g) return;
    memset(g, 0, sizeof(*g));
    g->seed = 0x4C545047ull;  // "LTPG"
    g->rows = 1000000;
    g->fields = 10;
    g->n_txns = 65536;
    g->n_slots = 10;
    g->min_ops = 10;
    g->max_ops = 10;
    g->read_pct = 50;
    g->theta = 0.8;
    g->tid_base = 0;
    g->distinct_keys = 1;
    g->txn_base = 0;
    g->scan_len = 0;
    g->reserved = 0;
}

int ltpg_kv_generate(const ltpg_kv_gen *g, uint64_t *keys, uint32_t *meta, uint32_t *tids,
                     uint32_t threads)
{
    if (!g || !keys || !meta || !tids) {
        ltpg::set_error("null argument");
        return LTPG_ERR_INVALID;
    }
    if (g->rows == 0 || g->fields == 0 || g->fields > 65536 || g->n_slots == 0 ||
        g->n_slots > 64 || g->min_ops > g->max_ops || g->max_ops > g->n_slots ||
        g->read_pct > 100 || g->theta < 0.0 || g->theta >= 1.0 ||
        (g->theta > 0.0 && g->rows < 3) || g->scan_len > LTPG_KV_MAX_SCAN) {
        ltpg::set_error("invalid generator parameters");
        return LTPG_ERR_INVALID;
    }
    if (g->distinct_keys && g->rows < g->max_ops) {
        ltpg::set_error("distinct_keys needs rows >= max_ops");
        return LTPG_ERR_INVALID;
    }
    const uint64_t B = g->n_txns, S = g->n_slots;
    const bool zipf = g->theta > 0.0;
    Zipf z(zipf ? g->rows : 3, zipf ? g->theta : 0.5);
    const uint64_t M = ((1ull << 61) - 1) % g->rows;
    const uint32_t span = g->max_ops - g->min_ops + 1;

    auto work = [&](uint64_t lo, uint64_t hi) {
        uint64_t picked[64];
        for (uint64_t t = lo; t < hi; t++) {
            const uint64_t gt = g->txn_base + t;
            const uint32_t nops = g->min_ops + (uint32_t)below(draw(g->seed, gt, 255, 0), span);
            for (uint32_t s = 0; s < S; s++) {
                const uint64_t i = s * B + t;
                if (s >= nops) {
                    keys[i] = 0;
                    meta[i] = LTPG_KV_META(LTPG_KV_NONE, 0);
                    continue;
                }
                const uint64_t x = draw(g->seed, gt, s, 0);
                const uint32_t kind =
                    below(x, 100) < g->read_pct ? LTPG_KV_READ : LTPG_KV_WRITE;
                const uint32_t field = (uint32_t)below(splitmix64(x), g->fields);
                uint64_t row = 0;
                for (uint32_t k = 1;; k++) {
                    const uint64_t y = draw(g->seed, gt, s, k);
                    if (zipf)
                        row = (uint64_t)(((unsigned __int128)z.rank(unit(y)) * M) % g->rows);
                    else
                        row = below(y, g->rows);
                    if (!g->distinct_keys) break;
                    bool dup = false;
                    for (uint32_t q = 0; q < s; q++) dup |= picked[q] == row;
                    if (!dup) break;
                }
                picked[s] = row;
                keys[i] = row;
                meta[i] = (kind == LTPG_KV_READ && g->scan_len) ? LTPG_KV_SCAN_META(field, g->scan_len)
                                                                 : LTPG_KV_META(kind, field);
            }
            tids[t] = g->tid_base + (uint32_t)t;
        }
    };

    unsigned nt = threads ? threads : std::thread::hardware_concurrency();
    nt = std::max(1u, std::min<unsigned>(nt, 64u));
    if (B < 4096) nt = 1;
    if (nt == 1) {
        work(0, B);
    } else {
        std::vector<std::thread> th;
        const uint64_t per = (B + nt - 1) / nt;
        for (unsigned w = 0; w < nt; w++) {
            const uint64_t lo = w * per, hi = std::min(B, lo + per);
            if (lo < hi) th.emplace_back(work, lo, hi);
        }
        for (auto &t : th) t.join();
    }
    return LTPG_OK;
}

}  // extern "C"

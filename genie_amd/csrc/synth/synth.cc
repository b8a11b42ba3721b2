// Deterministic synthetic workloads (see synth.h).
#include "synth.h"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i)
{
    return mix64(mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ull)) + i);
}

inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

// Bounded Zipf over ranks 1..n: P(r) proportional to r^-s, sampled by
// inverting the cumulative table.
struct Zipf {
    std::vector<double> cdf;
    Zipf(uint32_t n, double s) : cdf(n)
    {
        double acc = 0;
        for (uint32_t r = 1; r <= n; ++r) {
            acc += std::pow((double)r, -s);
            cdf[r - 1] = acc;
        }
        for (double& c : cdf)
            c /= acc;
        cdf.back() = 1.0;
    }
    uint32_t rank(double u) const  // 1-based
    {
        return (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    }
};

template <class F>
void parallel_for(uint64_t n, F f)
{
    unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096)
        T = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t)
        th.emplace_back([=] {
            uint64_t a = n * t / T, b = n * (t + 1) / T;
            f(a, b);
        });
    for (auto& x : th)
        x.join();
}

}  // namespace

extern "C" {

int synth_rows(uint64_t seed, uint32_t n, uint32_t d, int model, double p0, double p1, int32_t lo, int32_t hi,
               uint32_t n_values, int32_t* rows)
{
    if (!rows || n == 0 || d == 0)
        return -1;
    if (model == SYNTH_NORMAL) {
        parallel_for((uint64_t)n * d, [=](uint64_t a, uint64_t b) {
            for (uint64_t i = a; i < b; ++i) {
                // Box-Muller from two uniforms of the same cell
                double u1 = u01(rnd(seed, 1, 2 * i)), u2 = u01(rnd(seed, 1, 2 * i + 1));
                if (u1 < 1e-300)
                    u1 = 1e-300;
                double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                double x = std::nearbyint(p0 + p1 * z);
                rows[i] = (int32_t)std::min<double>(hi, std::max<double>(lo, x));
            }
        });
        return 0;
    }
    if (model == SYNTH_ZIPF_PERM) {
        if (n_values == 0)
            return -1;
        Zipf z(n_values, p0);
        // per-dim permutation of the values (Fisher-Yates from the stream)
        std::vector<int32_t> perm((size_t)d * n_values);
        for (uint32_t j = 0; j < d; ++j) {
            int32_t* p = &perm[(size_t)j * n_values];
            for (uint32_t v = 0; v < n_values; ++v)
                p[v] = (int32_t)v;
            for (uint32_t v = n_values - 1; v > 0; --v) {
                uint32_t r = (uint32_t)(rnd(seed, 2, (uint64_t)j * n_values + v) % (v + 1));
                std::swap(p[v], p[r]);
            }
        }
        parallel_for((uint64_t)n * d, [&](uint64_t a, uint64_t b) {
            for (uint64_t i = a; i < b; ++i) {
                uint32_t j = (uint32_t)(i % d);
                rows[i] = perm[(size_t)j * n_values + z.rank(u01(rnd(seed, 3, i))) - 1];
            }
        });
        return 0;
    }
    if (model == SYNTH_UNIFORM) {
        if (n_values == 0)
            return -1;
        parallel_for((uint64_t)n * d, [=](uint64_t a, uint64_t b) {
            for (uint64_t i = a; i < b; ++i)
                rows[i] = (int32_t)(rnd(seed, 4, i) % n_values);
        });
        return 0;
    }
    return -1;
}

int synth_queries(uint64_t seed, const int32_t* rows, uint32_t n, uint32_t d, uint32_t q, int32_t noise,
                  int32_t lo, int32_t hi, double redraw_p, uint32_t n_values, int32_t* out, uint32_t* out_src)
{
    if (!rows || !out || n == 0)
        return -1;
    for (uint32_t i = 0; i < q; ++i) {
        uint32_t src = (uint32_t)(rnd(seed, 10, i) % n);
        if (out_src)
            out_src[i] = src;
        for (uint32_t j = 0; j < d; ++j) {
            uint64_t cell = (uint64_t)i * d + j;
            int32_t v = rows[(uint64_t)src * d + j];
            if (redraw_p >= 0) {
                if (u01(rnd(seed, 11, cell)) < redraw_p)
                    v = (int32_t)(rnd(seed, 12, cell) % (n_values ? n_values : 1));
            } else if (noise > 0) {
                v += (int32_t)(rnd(seed, 13, cell) % (uint64_t)(2 * noise + 1)) - noise;
                v = std::min(hi, std::max(lo, v));
            }
            out[cell] = v;
        }
    }
    return 0;
}

int synth_sets(uint64_t seed, uint32_t n, uint32_t len_lo, uint32_t len_hi, uint32_t vocab, double s,
               int64_t* row_ptr, int32_t* vals)
{
    if (!row_ptr || len_hi < len_lo || vocab == 0)
        return -1;
    row_ptr[0] = 0;
    for (uint32_t i = 0; i < n; ++i)
        row_ptr[i + 1] = row_ptr[i] + len_lo + (int64_t)(rnd(seed, 20, i) % (uint64_t)(len_hi - len_lo + 1));
    if (!vals)
        return 0;
    Zipf z(vocab, s);
    parallel_for((uint64_t)row_ptr[n], [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i)
            vals[i] = (int32_t)z.rank(u01(rnd(seed, 21, i))) - 1;
    });
    return 0;
}

int synth_uniform_int(uint64_t seed, uint64_t stream, uint64_t n, int32_t lo, int32_t hi, int32_t* out)
{
    if (!out || hi < lo)
        return -1;
    for (uint64_t i = 0; i < n; ++i)
        out[i] = lo + (int32_t)(rnd(seed, 100 + stream, i) % (uint64_t)((int64_t)hi - lo + 1));
    return 0;
}

int synth_exponential(uint64_t seed, uint64_t n, double rate, double* out)
{
    if (!out || rate <= 0)
        return -1;
    for (uint64_t i = 0; i < n; ++i) {
        double u = u01(rnd(seed, 200, i));
        out[i] = -std::log(1.0 - u) / rate;
    }
    return 0;
}

}  // extern "C"

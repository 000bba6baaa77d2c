// filt_fields.cpp -- seeded synthetic property fields (host code).
//
// One implementation serves every path (plain, fused, batched, serving, the
// oracle comparisons), so that all of them see bit-identical inputs.
//   * counter-based PRNG: splitmix64 of (seed, stream, cell index), so a
//     cell's draw does not depend on generation order or thread count;
//   * normal deviates by Box-Muller;
//   * correlated log-normal permeability: white noise convolved with a
//     separable 3D Gaussian of std l/sqrt(2) cells (covariance
//     exp(-r^2 / (2 l^2))), periodic wrap, renormalised to zero mean and unit
//     variance, optionally clipped at +-clip standard deviations, then
//     K = exp(ln(K_mean) + sigma * z).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/filtration_hip.h"
#include "filt_guard.hpp"

namespace {

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Uniform in (0, 1]: 53 random bits, offset by one ulp so log() is finite.
inline double u01(uint64_t seed, uint64_t stream, uint64_t i)
{
    const uint64_t h = splitmix64(seed ^ splitmix64(stream * 0xD1B54A32D192ED03ULL + i));
    return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

inline double normal(uint64_t seed, uint64_t i)
{
    const double a = u01(seed, 1, i), b = u01(seed, 2, i);
    return sqrt(-2.0 * log(a)) * cos(6.283185307179586 * b);
}

template <class F>
void parallel_for(long long n, int threads, F fn)
{
    if (threads <= 1 || n < 2) {
        for (long long i = 0; i < n; i++)
            fn(i);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<bool> failed{false};  // a worker's exception, rethrown below
    const long long chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        const long long lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi)
            break;
        pool.emplace_back([=, &failed] {
            try {
                for (long long i = lo; i < hi; i++)
                    fn(i);
            } catch (...) {
                failed = true;
            }
        });
    }
    for (auto& th : pool)
        th.join();
    if (failed)
        throw std::bad_alloc();
}

int hw_threads(int requested)
{
    if (requested > 0)
        return requested;
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::min(16u, h ? h : 1u);
}

}  // namespace

extern "C" {

int filt_gen_uniform(double* out, size_t n, uint64_t seed, double lo, double hi)
{
    if (!out && n)
        return FILT_ERR_INVALID;
    for (size_t i = 0; i < n; i++)
        out[i] = lo + (hi - lo) * u01(seed, 3, i);
    return FILT_OK;
}

int filt_gen_lognormal(double* out, int nx, int ny, int nz, uint64_t seed, double k_mean,
                       double sigma, double corr_cells, double clip_sigma, int threads)
{
    return filt::guarded([&]() -> int {
        if (!out || nx < 1 || ny < 1 || nz < 1 || !(k_mean > 0) || sigma < 0)
            return FILT_ERR_INVALID;
        const long long n = (long long)nx * ny * nz;
        const int nt = hw_threads(threads);
        std::vector<double> a((size_t)n), b((size_t)n);
        parallel_for(n, nt, [&](long long i) { a[(size_t)i] = normal(seed, (uint64_t)i); });

        if (corr_cells > 0) {
            const double s = corr_cells / sqrt(2.0);
            const int R = (int)ceil(3.0 * s);
            std::vector<double> w(2 * R + 1);
            for (int d = -R; d <= R; d++)
                w[d + R] = exp(-0.5 * d * d / (s * s));
            auto wrap = [](int v, int m) { v %= m; return v < 0 ? v + m : v; };
            // x pass a -> b, y pass b -> a, z pass a -> b
            parallel_for((long long)ny * nz, nt, [&](long long row) {
                const double* src = &a[(size_t)row * nx];
                double* dst = &b[(size_t)row * nx];
                for (int i = 0; i < nx; i++) {
                    double acc = 0;
                    for (int d = -R; d <= R; d++)
                        acc += w[d + R] * src[wrap(i + d, nx)];
                    dst[i] = acc;
                }
            });
            parallel_for((long long)nz, nt, [&](long long k) {
                std::vector<double> col(ny);
                for (int i = 0; i < nx; i++) {
                    for (int j = 0; j < ny; j++)
                        col[j] = b[(size_t)(i + (long long)nx * (j + (long long)ny * k))];
                    for (int j = 0; j < ny; j++) {
                        double acc = 0;
                        for (int d = -R; d <= R; d++)
                            acc += w[d + R] * col[wrap(j + d, ny)];
                        a[(size_t)(i + (long long)nx * (j + (long long)ny * k))] = acc;
                    }
                }
            });
            parallel_for((long long)ny, nt, [&](long long j) {
                std::vector<double> col(nz);
                const long long plane = (long long)nx * ny;
                for (int i = 0; i < nx; i++) {
                    for (int k = 0; k < nz; k++)
                        col[k] = a[(size_t)(i + nx * j + plane * k)];
                    for (int k = 0; k < nz; k++) {
                        double acc = 0;
                        for (int d = -R; d <= R; d++)
                            acc += w[d + R] * col[wrap(k + d, nz)];
                        b[(size_t)(i + nx * j + plane * k)] = acc;
                    }
                }
            });
            a.swap(b);
        }

        // zero mean, unit variance; per-plane sums in a fixed order
        std::vector<double> s1(nz), s2(nz);
        const long long plane = (long long)nx * ny;
        parallel_for((long long)nz, nt, [&](long long k) {
            double p1 = 0, p2 = 0;
            for (long long c = 0; c < plane; c++) {
                const double v = a[(size_t)(k * plane + c)];
                p1 += v;
                p2 += v * v;
            }
            s1[k] = p1;
            s2[k] = p2;
        });
        double m1 = 0, m2 = 0;
        for (int k = 0; k < nz; k++) {
            m1 += s1[k];
            m2 += s2[k];
        }
        const double mean = m1 / n;
        const double var = std::max(m2 / n - mean * mean, 1e-300);
        const double inv_sd = 1.0 / sqrt(var);
        const double lnk = log(k_mean);
        parallel_for(n, nt, [&](long long c) {
            double z = (a[(size_t)c] - mean) * inv_sd;
            if (clip_sigma > 0)
                z = std::min(clip_sigma, std::max(-clip_sigma, z));
            out[c] = exp(lnk + sigma * z);
        });
        return FILT_OK;
    });
}

}  // extern "C"

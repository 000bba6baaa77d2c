// divby_check.hip -- checks DivBy (filt_device.hpp) against the IEEE division
// n / d bit for bit on the GPU: 2^24 numerators (random bit patterns, the
// range edges 2^-600 and 2^601, exact quotients, significands of all ones,
// zeros, subnormals, infinities, NaN) against 64 divisors in [2^-100, 2^100]
// (grid spacings, their doubles, and random ones); and recip_ieee against
// 1 / d on 2^24 divisors (random bit patterns, random magnitudes in and
// around [2^-100, 2^101) of either sign, the range edges, significands of all
// ones, zeros, subnormals, infinities, NaN).  Prints one JSON line and exits
// nonzero on any difference.  Run by tests/test_gpu_plain.py.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "filt_device.hpp"

namespace {

__global__ void k_check(const double* num, long long count, const double* div, int ndiv,
                        unsigned long long* mismatches, double* example)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const double n = num[i];
    for (int k = 0; k < ndiv; k++) {
        const double d = div[k];
        const filt::DivBy by{d, filt::refined_reciprocal(d)};
        const double got = by(n), want = n / d;
        if (__double_as_longlong(got) != __double_as_longlong(want)) {
            const unsigned long long c = atomicAdd(mismatches, 1ull);
            if (c < 4) {
                example[3 * c] = n;
                example[3 * c + 1] = d;
                example[3 * c + 2] = got;
            }
        }
    }
}

__global__ void k_recip(const double* den, long long count, unsigned long long* mismatches,
                        double* example)
{
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const double d = den[i];
    const double got = filt::recip_ieee(d), want = 1.0 / d;
    if (__double_as_longlong(got) != __double_as_longlong(want)) {
        const unsigned long long c = atomicAdd(mismatches, 1ull);
        if (c < 4) {
            example[2 * c] = d;
            example[2 * c + 1] = got;
        }
    }
}

double from_bits(uint64_t b)
{
    double v;
    std::memcpy(&v, &b, sizeof v);
    return v;
}

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::printf("{\"error\": \"%s: %s\"}\n", #x, hipGetErrorString(e_));    \
            return 2;                                                               \
        }                                                                           \
    } while (0)

}  // namespace

int main()
{
    const long long N = 1ll << 24;
    std::vector<double> num(N);
    std::mt19937_64 g(20261203);
    const double specials[] = {0.0, -0.0, INFINITY, -INFINITY, NAN, 5e-324, -5e-324,
                               2.2250738585072014e-308, 1.7976931348623157e308, 1e-300, -1e300};
    for (long long i = 0; i < N; i++) {
        const uint64_t u = g();
        double v;
        switch (i % 8) {
        case 0: v = from_bits(u); break;
        case 1:
        case 2:
            v = std::ldexp((double)(u >> 11) * 0x1p-53 + 0.5, (int)(u % 200) - 100) *
                ((u >> 7) & 1 ? -1 : 1);
            break;
        case 3: v = ((double)(u >> 11) * 0x1p-53 - 0.5) * 1e5; break;
        case 4: v = std::ldexp((double)(u % 100000), (int)((u >> 20) % 40) - 20); break;
        case 5:  // around 2^-600 and 2^601
            v = std::ldexp(1.0 + (double)(u >> 12) * 0x1p-52,
                           (int)(u % 60) - 630 + (int)((u >> 40) & 1) * 1200);
            break;
        case 6: v = specials[u % 11]; break;
        default: {  // significands of all ones, either sign
            uint64_t b = 0x3fffffffffffffffull - (u % 16);
            b += ((u >> 10) % 100) << 52;
            b ^= ((u >> 60) & 1) << 63;
            v = from_bits(b);
        }
        }
        num[i] = v;
    }
    std::vector<double> div = {10.0 / 256, 2 * (10.0 / 256), 1.0 / 3, 2.0 / 3, 0.1, 0.2, 1.0, 3.0,
                               10.0 / 100, 2 * (10.0 / 100), 7.0 / 512, 1.0 - 0x1p-53,
                               1.0 + 0x1p-52, 0x1p-100, 0x1p100, 0x1.fffffffffffffp-1,
                               0x1.fffffffffffffp+0, 1.5, 0.75};
    std::mt19937_64 g2(777);
    while (div.size() < 64) {
        const uint64_t u = g2();
        div.push_back(std::ldexp(1.0 + (double)(u >> 12) * 0x1p-52, (int)(u % 201) - 100));
    }
    double *d_num = nullptr, *d_div = nullptr, *d_ex = nullptr;
    unsigned long long* d_mis = nullptr;
    CHECK(hipMalloc((void**)&d_num, N * sizeof(double)));
    CHECK(hipMalloc((void**)&d_div, div.size() * sizeof(double)));
    CHECK(hipMalloc((void**)&d_ex, 12 * sizeof(double)));
    CHECK(hipMalloc((void**)&d_mis, sizeof(unsigned long long)));
    CHECK(hipMemcpy(d_num, num.data(), N * sizeof(double), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(d_div, div.data(), div.size() * sizeof(double), hipMemcpyHostToDevice));
    CHECK(hipMemset(d_mis, 0, sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_check, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, 0, d_num, N,
                       d_div, (int)div.size(), d_mis, d_ex);
    CHECK(hipGetLastError());
    unsigned long long mis = 0;
    double ex[12] = {};
    CHECK(hipMemcpy(&mis, d_mis, sizeof mis, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(ex, d_ex, sizeof ex, hipMemcpyDeviceToHost));

    // reciprocals: the divisors go into the numerator buffer
    std::mt19937_64 g3(4242);
    for (long long i = 0; i < N; i++) {
        const uint64_t u = g3();
        const double sgn = (u >> 63) ? -1.0 : 1.0;
        double v;
        switch (i % 8) {
        case 0: v = from_bits(u); break;
        case 1:
        case 2:
        case 3:  // in and around the range, log-uniform
            v = sgn * std::ldexp(1.0 + (double)(u >> 12) * 0x1p-52, (int)(u % 211) - 105);
            break;
        case 4:  // the range edges and their neighbours
            v = sgn * std::nextafter(std::ldexp(1.0, (u & 1) ? 101 : -100),
                                     (u & 2) ? INFINITY : 0.0) *
                ((u & 4) ? 1.0 : std::ldexp(1.0, (int)((u >> 8) % 3) - 1));
            break;
        case 5: {  // significands of all ones, near the range
            uint64_t b = 0x000fffffffffffffull - (u % 16);
            b |= (uint64_t)(923 + (int)((u >> 10) % 203) - 1) << 52;
            b ^= ((u >> 62) & 1) << 63;
            v = from_bits(b);
            break;
        }
        case 6: v = specials[u % 11] * ((u >> 40) & 1 ? -1.0 : 1.0); break;
        default: v = sgn * ((double)(u >> 11) * 0x1p-53 + 1e-3) * 1e6; break;  // Newton-like
        }
        num[i] = v;
    }
    CHECK(hipMemcpy(d_num, num.data(), N * sizeof(double), hipMemcpyHostToDevice));
    CHECK(hipMemset(d_mis, 0, sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_recip, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, 0, d_num, N,
                       d_mis, d_ex);
    CHECK(hipGetLastError());
    unsigned long long rmis = 0;
    double rex[8] = {};
    CHECK(hipMemcpy(&rmis, d_mis, sizeof rmis, hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(rex, d_ex, sizeof rex, hipMemcpyDeviceToHost));

    std::printf("{\"numerators\": %lld, \"divisors\": %zu, \"mismatches\": %llu", N, div.size(),
                mis);
    for (unsigned k = 0; k < (mis < 4 ? mis : 4); k++)
        std::printf(", \"example%u\": [\"%a\", \"%a\", \"%a\"]", k, ex[3 * k], ex[3 * k + 1],
                    ex[3 * k + 2]);
    std::printf(", \"reciprocals\": %lld, \"recip_mismatches\": %llu", N, rmis);
    for (unsigned k = 0; k < (rmis < 4 ? rmis : 4); k++)
        std::printf(", \"recip_example%u\": [\"%a\", \"%a\"]", k, rex[2 * k], rex[2 * k + 1]);
    std::printf("}\n");
    (void)hipFree(d_num);
    (void)hipFree(d_div);
    (void)hipFree(d_ex);
    (void)hipFree(d_mis);
    return mis != 0 || rmis != 0;
}

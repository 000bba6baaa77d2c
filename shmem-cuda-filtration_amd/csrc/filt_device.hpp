// filt_device.hpp -- per-cell arithmetic of the filtration step, device side.
//
// Every expression keeps the operation order of the reference's CPU path
// (cpu.cpp:57-74, 77-160, 218-291, 293-320).  The library is compiled with
// -ffp-contract=off, so apart from pow() every FP64 result here rounds
// exactly as the CPU build does (AMDGPU FP64 division is IEEE-correct).
// pow() is ROCm's, except for the exact shortcuts in pow_spec().
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/filtration_hip.h"

namespace filt {

// Flat index i + j*nx + k*nx*ny (defines.h:41-47 layout), 64-bit.
__host__ __device__ inline long long idx3(int i, int j, int k, int nx, int ny)
{
    return (long long)i + (long long)j * nx + (long long)k * nx * (long long)ny;
}

// pow(x, y) with exact shortcuts for the exponents the configurations use.
// y == 2: x*x is the correctly rounded square.  y == 4: x^4 from the exact
// square (fma error term) squared again, rounded once, i.e. within a hair of
// the correctly rounded value, which glibc's pow also returns.  Any other
// case (and non-finite or tiny x) goes to the library pow.
__device__ inline double pow_spec(double x, double y)
{
    if (y == 2.0)
        return x * x;
    if (y == 4.0 && fabs(x) > 1e-60 && fabs(x) < 1e60) {
        const double p = x * x;
        const double e = fma(x, x, -p);          // x*x = p + e exactly
        const double hi = p * p;
        const double lo = fma(p, p, -hi);        // p*p = hi + lo exactly
        return hi + (lo + 2.0 * p * e);          // (p+e)^2 less e^2
    }
    return pow(x, y);
}

// Stage 1 for one cell: cpu.cpp:57-74 (assign_P2_Xi1_Xi2, assign_ro1_ro2).
struct Stage1 {
    double P2, Xi1, Xi2, ro1, ro2;
};

// Stage 1's densities (cpu.cpp:72-73).  ro2 is evaluated at P2 = P1 + 0,
// which has P1's bits but for P1 = -0; the difference P - P_atm, and so the
// density, is the same for -0 and +0.  So both follow from P2 alone.
template <class Params>
__device__ __forceinline__ double ro1_of(double P, const Params& p)
{
    return p.ro01 * (1. + p.beta1 * (P - p.P_atm));
}

template <class Params>
__device__ __forceinline__ double ro2_of(double P, const Params& p)
{
    return p.ro02 * (1. + p.beta2 * (P - p.P_atm));
}

__device__ inline Stage1 stage1_cell(double P1, double S2, double K, double lam, double swr,
                                     const filt_params& p)
{
    Stage1 r;
    const double S_e = (1. - S2 - swr) / (1. - swr);
    const double k1 = pow_spec(S_e, (2. + 3. * lam) / lam);
    const double k2 = (1. - S_e) * (1. - S_e) * (1 - pow_spec(S_e, (2. + lam) / lam));
    const double P_k = 0;  // capillary pressure disabled, cpu.cpp:63
    r.P2 = P1 + P_k;
    r.Xi1 = -1. * K * k1 / p.mu1;
    r.Xi2 = -1. * K * k2 / p.mu2;
    r.ro1 = ro1_of(P1, p);
    r.ro2 = ro2_of(r.P2, p);
    return r;
}

// Upwind flux difference along one axis (cpu.cpp:242-244 and siblings).
// a2, a1: forces on the + and - faces; X*, r*: Xi and ro at centre, -, +.
__device__ inline double upwind(double a2, double a1, double Xc, double rc, double Xm, double rm,
                                double Xp, double rp, double h)
{
    return (((a2 + fabs(a2)) / 2.0 - (a1 - fabs(a1)) / 2.0) * (-1) * Xc * rc -
            (a1 + fabs(a1)) / 2.0 * (-1) * Xm * rm +
            (a2 - fabs(a2)) / 2.0 * (-1) * Xp * rp) / h;
}

// Division by a grid spacing, bit for bit as the IEEE division n / d.
//
// hipcc expands an FP64 division into: v_div_scale of d and of n, y0 =
// v_rcp_f64(d), two Newton steps y <- fma(y, fma(-d, y, 1), y), q = n*y,
// r = fma(-d, q, n), v_div_fmas(r, y, q) = fma(r, y, q) (times 2^+-64 when
// v_div_scale scaled), then v_div_fixup, which handles zeros, infinities,
// NaN, overflow and underflow and otherwise returns its input with the
// quotient's sign.  v_div_scale scales only for a denormal d or 1/d, a
// denormal quotient, exponent(n) - exponent(d) >= 768, or a biased
// exponent(n) <= 53.  For 2^-100 <= d <= 2^101 and 2^-600 <= |n| < 2^601 none
// of that happens, so the expansion is exactly the sequence below with the
// reciprocal computed once per divisor: three FP64 operations per division
// instead of a v_rcp_f64, eight more FP64 operations and the two scale steps.
// The remainder is formed as -(d*q - n): that is n - d*q rounded once, as in
// the expansion, and for a zero numerator it keeps the zero's sign through the
// last step (n = +-0 gives q = +-0, d*q - n = +0, and fma(-0, y, +-0) = +-0, as
// n / d), so zeros need no select; for a nonzero remainder -(d*q - n) and
// n - d*q round to the same value, and a zero remainder leaves q as it is
// either way.  Any other numerator outside the range takes the division
// itself.  tests/hip/divby_check compares the result with n / d on 2^24
// numerators (specials and significands of all ones included) for each of 64
// divisors; an earlier probe of the expansion's own remainder form covered
// 4.3e9 pairs.
__device__ __forceinline__ double refined_reciprocal(double d)
{
    double y = __builtin_amdgcn_rcp(d);
    double e = fma(-d, y, 1.0);
    y = fma(y, e, y);
    e = fma(-d, y, 1.0);
    return fma(y, e, y);
}

// The rare numerators (infinities, NaN, |n| >= 2^601, 0 < |n| < 2^-600):
// out of line, so the registers of the full division sequence are not
// reserved across the kernels that use DivBy.
__device__ __attribute__((noinline)) double div_rare(double n, double d);

struct DivBy {
    double d, y;  // divisor in [2^-100, 2^101], y = refined_reciprocal(d)

    __device__ __forceinline__ double operator()(double n) const
    {
        const double q = n * y;
        const double t = fma(d, q, -n);  // d*q - n, rounded once
        double f = fma(-t, y, q);
        const unsigned e = ((unsigned)__double2hiint(n) >> 20) & 0x7ffu;
        if (!(e - 423u <= 1200u || n == 0.0))  // |n| outside [2^-600, 2^601)
            f = div_rare(n, d);
        return f;
    }
};

__device__ __attribute__((noinline)) inline double div_rare(double n, double d) { return n / d; }

// 1 / d bit for bit as the IEEE division, for a divisor that varies per
// lane.  With numerator 1 the expansion above is, for 2^-100 <= |d| < 2^101,
// y = refined_reciprocal(d) and then the remainder step with q = 1*y = y:
// two FP64 operations after the reciprocal instead of the two scale steps,
// the multiply, the remainder, the scaled fma and the fix-up, and three
// fewer steps on the dependent chain.  The expansion is symmetric in the
// sign of d, so |d| is what counts.  Other divisors (zeros, subnormals,
// infinities, NaN, magnitudes outside the range) take the division out of
// line.  tests/hip/divby_check compares it with 1 / d on 2^24 divisors.
__device__ __forceinline__ double recip_ieee(double d)
{
    const double y = refined_reciprocal(d);
    const double t = fma(d, y, -1.0);  // d*y - 1, rounded once
    double f = fma(-t, y, y);
    const unsigned e = ((unsigned)__double2hiint(d) >> 20) & 0x7ffu;
    if (!(e - 923u <= 200u))  // |d| outside [2^-100, 2^101)
        f = div_rare(1.0, d);
    return f;
}

// The divisors DivBy accepts.
__host__ __device__ inline bool divby_ok(double d)
{
    return d >= 0x1p-100 && d <= 0x1p101;
}

// upwind() with the division by h done by div.
template <class Div>
__device__ inline double upwind_div(double a2, double a1, double Xc, double rc, double Xm,
                                    double rm, double Xp, double rp, const Div& div)
{
    return div(((a2 + fabs(a2)) / 2.0 - (a1 - fabs(a1)) / 2.0) * (-1) * Xc * rc -
               (a1 + fabs(a1)) / 2.0 * (-1) * Xm * rm +
               (a2 - fabs(a2)) / 2.0 * (-1) * Xp * rp);
}

// Bit-pattern equality (also for NaN, and telling -0 from +0).
__device__ __forceinline__ bool same_bits(double a, double b)
{
    return __double_as_longlong(a) == __double_as_longlong(b);
}

// Newton iterations of cpu.cpp:293-320 on one cell, in registers.
// With P_d == 0 the capillary terms P_k = P_d*S_e^(-1/lam) and
// PkS ~ P_d*S_e^(-1/lam-1) are exactly zero whenever those powers are finite,
// which holds for S_e >= 2^-10 and lam >= 0.1 (|exponent| <= 11); the pows
// are then skipped.  Elsewhere they are evaluated as the reference does, so
// inf*0 = NaN propagates identically.
//
// That test needs no division: with d = 1 - S_wr > 0 and n = 1 - S - S_wr,
// RN(n/d) >= 2^-10 exactly when n >= d*2^-10.  (d*2^-10 is exact; the
// values that round to 2^-10 or above are those >= 2^-10*(1 - 2^-54), the
// midpoint to the next double below, and no double n lies strictly between
// pred(d*2^-10) and d*2^-10 while n/d >= 2^-10*(1 - 2^-54).)
//
// An iteration is a function of (P, S) alone.  Once its output repeats its
// input bit for bit (a fixed point) every later iteration returns it too;
// once it repeats the state of two iterations back (a 2-cycle between two
// neighbouring doubles, which about 2% of cells settle into) the states
// alternate.  Stopping there and taking the state the remaining count
// lands on gives the full iteration's result bit for bit.
//
// kLean = true keeps the division and only the fixed-point exit: six fewer
// live registers, for kernels at the edge of an occupancy step.  Params is
// filt_params, or filt_params in the kernel argument segment (KArg).
template <bool kLean = false, class Params>
__device__ inline void newton_cell(double& P, double& S, double roS1, double roS2, double lam,
                                   double swr, double pd, const Params& p);

// One Newton iteration of cpu.cpp:300-317 from (P, S) with the capillary
// terms P_k and PkS.
template <class Params>
__device__ inline void newton_step(double P, double S, double P_k, double PkS, double roS1,
                                   double roS2, const Params& p, double& Pn, double& Sn)
{
    const double ro01 = p.ro01, ro02 = p.ro02, b1 = p.beta1, b2 = p.beta2, Pa = p.P_atm;
    const double F1 = ro01 * (1. + b1 * (P - Pa)) * (1. - S) - roS1;
    const double F2 = ro02 * (1. + b2 * (P + P_k - Pa)) * S - roS2;
    const double F1P = ro01 * b1 * (1 - S);
    const double F2P = ro02 * b2 * S;
    const double F1S = (-1) * ro01 * (1 + b1 * (P - Pa));
    const double F2S = ro02 * (1 + b2 * (P + P_k - Pa + (S * PkS)));
    const double det = F1P * F2S - F1S * F2P;
    const double inv = recip_ieee(det);  // 1 / det, bit for bit
    Pn = P - inv * (F2S * F1 - F1S * F2);
    Sn = S - inv * (F1P * F2 - F2P * F1);
}

// newton_step with P_k = PkS = 0, bit for bit, for P_atm != 0 and S > -inf.
// There (P + 0) - P_atm equals P - P_atm for every P (they differ only for
// P = -0 with P_atm = +-0), and (P - P_atm) + S*0 equals P - P_atm for every
// finite S: S*0 is a zero, and P - P_atm is -0 only when P_atm is zero.  So
// F2S is the product F2 starts from, and one iteration takes six FP64
// operations fewer (P + 0, S*0, the sum, and F2S's own three).
template <class Params>
__device__ inline void newton_step_nocap(double P, double S, double roS1, double roS2,
                                         const Params& p, double& Pn, double& Sn)
{
    const double ro01 = p.ro01, ro02 = p.ro02, b1 = p.beta1, b2 = p.beta2, Pa = p.P_atm;
    const double F1 = ro01 * (1. + b1 * (P - Pa)) * (1. - S) - roS1;
    const double F2S = ro02 * (1 + b2 * (P - Pa));
    const double F2 = F2S * S - roS2;
    const double F1P = ro01 * b1 * (1 - S);
    const double F2P = ro02 * b2 * S;
    const double F1S = (-1) * ro01 * (1 + b1 * (P - Pa));
    const double det = F1P * F2S - F1S * F2P;
    const double inv = recip_ieee(det);  // 1 / det, bit for bit
    Pn = P - inv * (F2S * F1 - F1S * F2);
    Sn = S - inv * (F1P * F2 - F2P * F1);
}

template <bool kLean, class Params>
__device__ inline void newton_cell(double& P, double& S, double roS1, double roS2, double lam,
                                   double swr, double pd, const Params& p)
{
    const bool no_cap = (pd == 0.0) && (lam >= 0.1);
    const double d = 1 - swr;
    const int n = p.newton_iterations;
    int w = 1;
    if (!kLean && no_cap && d > 0.0 && p.P_atm != 0.0) {
        // Capillary-free iterations while S_e >= 2^-10: P_k and PkS are
        // exactly zero (a zero PkS enters F2S only as S*PkS added to a
        // nonzero sum), so no pow and no division for S_e, and with S > -inf
        // (S NaN or +inf fails the S_e test) the iteration is
        // newton_step_nocap.  The first iteration that needs the general
        // form continues below.
        const double d_t = d * 0.0009765625;
        double Pq = P, Sq = S;  // the state two iterations back
        for (; w <= n; w++) {
            if (!((int)(1 - S - swr >= d_t) & (int)(S != -__builtin_inf())))
                break;
            double Pn, Sn;
            newton_step_nocap(P, S, roS1, roS2, p, Pn, Sn);
            // both exit tests evaluated in full (bitwise &, |), so a lane's
            // exit is one mask test rather than a chain of branches
            const int fix = (int)same_bits(Pn, P) & (int)same_bits(Sn, S);
            const int cyc = (int)(w >= 2) & (int)same_bits(Pn, Pq) & (int)same_bits(Sn, Sq);
            if (fix | cyc) {
                // fixed point: (P, S) is the result.  2-cycle: x_{w+1} =
                // x_{w-1}, x_{w+2} = x_w, ..., so the count left picks one.
                const bool take = !fix && ((n - w) & 1) == 0;
                P = take ? Pn : P;
                S = take ? Sn : S;
                return;
            }
            Pq = P;
            Sq = S;
            P = Pn;
            S = Sn;
        }
    }
    for (; w <= n; w++) {
        const double S_e = (1 - S - swr) / d;
        double P_k = 0.0, PkS = 0.0;
        if (!(no_cap && S_e >= 0.0009765625)) {
            // the reference's powers; inf*0 = NaN propagates as there
            P_k = pd * pow(S_e, (-1.) / lam);
            const double AAA = pow(S_e, ((-1.) / lam) - 1.);
            PkS = AAA * pd / (lam * (1 - swr));
        }
        double Pn, Sn;
        newton_step(P, S, P_k, PkS, roS1, roS2, p, Pn, Sn);
        if (same_bits(Pn, P) && same_bits(Sn, S))
            return;
        P = Pn;
        S = Sn;
    }
}

// Source patch test of main.cpp:109 / cpu.cpp:346, generalised to bounds.
// Params: filt_params, or filt_params in the kernel argument segment (KArg).
template <class Params>
__host__ __device__ inline bool in_patch(int I, int k, const Params& p)
{
    return I >= p.src_x0 && I <= p.src_x1 && k >= p.src_z0 && k <= p.src_z1;
}

// main.cpp:145-158
__host__ __device__ inline int i_to_I(int i, int rank, int size, const filt_params& p)
{
    const int base = p.NX / size, rem = p.NX % size;
    if (rank <= rem)
        return rank == 0 ? i : (base + 1) * rank + i - 1;
    return (base + 1) * rank - (rank - rem) + i - 1;
}

// main.cpp:192-198
__host__ __device__ inline bool is_active(int i, int nx, int rank, int size)
{
    return !((rank != 0 && i == 0) || (rank != size - 1 && i == nx - 1));
}

// Interior test of cpu.cpp:220 / 295.
__host__ __device__ inline bool is_interior(int i, int j, int k, int nx, int ny, int nz)
{
    return i != 0 && i != nx - 1 && j != 0 && j != ny - 1 && ((k != 0 && k != nz - 1) || nz < 2);
}

}  // namespace filt

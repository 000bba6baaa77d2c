// filt_fused_tile.hpp -- the fused time step's kernel template: one launch
// per step for a batch of grids, three arithmetic forms.
//
// The fused step computes what the five defines.h stages compute for P1 and
// S2 (cpu.cpp:3-54 in the order of main.cpp:83-94) without materialising the
// intermediate arrays: state lives in ping-pong buffers (P, S in / out), the
// static fields (permeability, porosity, media) are read once per cell, and
// stage 1 (density, mobility) is evaluated on the fly for the 7-point
// stencil.  Per cell and step it moves 16 B of state in, 16 B out and 16 B of
// FP64 statics: 48 B against the 304 B of the stage path.
//
// Boundary conditions (cpu.cpp:44-54, 323-403) need no second launch: every
// shell cell's post-sweep value is either a copy of a value from the
// previous state, the source patch or P_atm (the "shell pass" blocks), or a
// copy of the new value of one interior cell, which the thread computing
// that cell writes as well:
//   interior (i,j,k) with value n also writes
//     (0,j,k)              if i == 1          [i = 0 face, interior source]
//     (N-1,j,k)            if i == N-2        [i = N-1 face]
//   and for column c in {i} and, if i == N-2, also {N-1}:
//     (c,M-1,k) = (n_S, n_P + ro1(i,1,0)*g*h2)   if j == M-2, plus k = 0 / L-1
//                                                 when k == 1 / L-2
//     (c,j,0) = n if k == 1,  (c,j,L-1) = n if k == L-2
//   shell pass: (0,j,k) = old (1,j,k) where (1,j,k) is not interior;
//     (i,0,k), i >= 1: P_atm and S = src(i',min(k,L-2)) for k >= 1,
//     old S(i',0,1) for k = 0, with i' = min(i, N-2)
// which is the closed form of the reference's i -> j -> k sweep.
//
// Launch: tiles of TX x TY columns covering the interior (CY columns per
// thread, stacked in y; x-tiles start at x = 0, on the rows' 128-byte lines)
// march over z-chunks of kz planes.  Per plane, the stage-1 values of the tile and
// its one-cell ring go to one of two LDS tiles; the raw inputs (state,
// statics) of the own column two planes ahead and of the ring one plane
// ahead are in flight while a plane is updated; the own column's plane k-1
// is read back from the thread's own slot of the other LDS tile, which no
// other thread writes; one barrier per plane orders both the writes before
// the reads and the reads of plane k-1 before the overwrite at plane k+1.
// The ring cells are evaluated by consecutive waves of the block starting at
// a wave chosen by the block index, so each SIMD pays for rings evenly.
// Shell-pass blocks follow each grid's tiles; a table maps each block to its
// grid, so grids of different shapes share one launch, and blocks are
// regrouped per XCD as in filt_tiles.hpp.
//
// Arrays are addressed by 32-bit byte offsets from the grid's uniform base
// pointers (global_load with a scalar base: one VGPR per address, no 64-bit
// address arithmetic); grids therefore keep every array below 4 GiB.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "../../include/filtration_hip.h"
#include "filt_device.hpp"
#include "filt_fused.hpp"
#include "filt_runtime.hpp"
#include "filt_zmarch.hpp"

namespace filt {
namespace fz {

// ------------------------------------------------------ memory access --
typedef __attribute__((address_space(1))) char gchar;

template <class T>
__device__ __forceinline__ T gld(const T* base, unsigned off)
{
    return *(const __attribute__((address_space(1))) T*)((const gchar*)base + off);
}

// Outputs of the step are not read again in the same launch: non-temporal
// stores keep them from displacing the L2 lines neighbouring tiles read.
__device__ __forceinline__ void gst(double* base, unsigned off, double v)
{
    __builtin_nontemporal_store(v, (__attribute__((address_space(1))) double*)((gchar*)base + off));
}

using zm::DROP;
using zm::out_rsrc;
using zm::Rsrc;
using zm::VMCNT0;
using zm::bstn;

// The media table in LDS, one row of FILT_MAX_MEDIA per quantity: the tile
// loop looks entries up by each cell's medium.
enum {
    M_LAM = 0,   // lambda
    M_SWR = 1,   // S_wr
    M_PD = 2,    // P_d
    M_M = 3,     // porosity
    M_K = 4,     // permeability
    M_INV1 = 5,  // 1 / (1 - S_wr), IEEE (fast and mixed forms)
    M_DY = 6,    // refined reciprocal of 1 - S_wr (DivBy, FP64 form)
    M_E1 = 7,    // (2 + 3 lambda) / lambda
    M_E2 = 8,    // (2 + lambda) / lambda
    M_ROWS = 9
};

// The grid's descriptor in the constant address space.
typedef const __attribute__((address_space(4))) FusedGrid CG;

// Block-uniform view of one grid.  Array pointers and parameters are read
// through G() where they are used: its value is hidden from the compiler, so
// each read is a scalar load (a scalar-cache hit) at the point of use rather
// than a value kept in SGPRs across the plane loop (the loop would otherwise
// spill them to VGPR lanes and read each back with v_readlane).
// FIELDS: every grid of the launch has per-cell K and porosity and one medium
// (no media lookups, no tests for absent fields).
template <bool FIELDS_>
struct Ctx {
    static constexpr bool FIELDS = FIELDS_;
    CG* g;
    const double* med;  // LDS media table
    const double* rh;   // LDS: refined reciprocals of h1, h2, h3, mu1, mu2
    int parity;

    __device__ __forceinline__ CG* G() const
    {
        CG* q = g;
        asm volatile("" : "+s"(q));
        return q;
    }
    __device__ double tab(int row, int md) const { return med[row * FILT_MAX_MEDIA + md]; }
    // the ping-pong buffers, chosen by a uniform select (an index by parity
    // would turn into a per-lane load of the pointer)
    __device__ const double* Pin() const { CG* q = G(); return parity ? q->P[1] : q->P[0]; }
    __device__ const double* Sin() const { CG* q = G(); return parity ? q->S[1] : q->S[0]; }
    __device__ double* Po() const { CG* q = G(); return parity ? q->P[0] : q->P[1]; }
    __device__ double* So() const { CG* q = G(); return parity ? q->S[0] : q->S[1]; }
    __device__ unsigned sz() const { CG* q = G(); return 8u * (unsigned)q->nx * (unsigned)q->ny; }
    __device__ unsigned bytes() const { CG* q = G(); return sz() * (unsigned)q->nz; }
    __device__ unsigned sy() const { return 8u * (unsigned)G()->nx; }
};

// Inputs of one cell: state, K (when per-cell), porosity (centre cells),
// the medium word as loaded (medium() gives the cell's medium).
struct Raw {
    double P, S, K, phi;
    int mdw;
};

// The cell's medium: 0 unless the grid has a media field with several media
// (the select is made where the medium is used, not where it is loaded, so
// that the load stays in flight).
template <class C>
__device__ __forceinline__ int medium(const C& c, const Raw& q)
{
    if constexpr (C::FIELDS)
        return 0;
    CG* G = c.G();
    return (G->media && G->p.n_media > 1) ? q.mdw : 0;
}

// The cell's permeability and porosity: per-cell where the grid has the
// field, the medium's otherwise.
template <class C>
__device__ __forceinline__ double perm_of(const C& c, const Raw& q, int md)
{
    if constexpr (C::FIELDS)
        return q.K;
    return c.G()->perm ? q.K : c.tab(M_K, md);
}

template <class C>
__device__ __forceinline__ double poro_of(const C& c, const Raw& q, int md)
{
    if constexpr (C::FIELDS)
        return q.phi;
    return c.G()->poro ? q.phi : c.tab(M_M, md);
}

// Every call issues the same loads whichever fields the grid has: a field
// that is absent is read from the pressure array instead and the value is not
// used.  No branch around a load, so the memory operations in flight are the
// same on every path through the plane loop and the compiler's waits for them
// stay exact (a skipped load makes it wait for everything outstanding).
template <class C>
__device__ __forceinline__ Raw load_raw(const C& c, unsigned off, bool centre)
{
    CG* G = c.G();
    const double* const Pin = c.parity ? G->P[1] : G->P[0];
    Raw r;
    if constexpr (C::FIELDS) {
        r.mdw = 0;
        r.P = gld(Pin, off);
        r.S = gld(c.parity ? G->S[1] : G->S[0], off);
        r.K = gld(G->perm, off);
        if (centre)
            r.phi = gld(G->poro, off);
        return r;
    }
    const bool media = G->media && G->p.n_media > 1;
    r.mdw = gld(media ? G->media : (const int*)Pin, off >> 1);
    r.P = gld(Pin, off);
    r.S = gld(c.parity ? G->S[1] : G->S[0], off);
    r.K = gld(G->perm ? G->perm : Pin, off);
    if (centre)
        r.phi = gld(G->poro ? G->poro : Pin, off);
    return r;
}

// ------------------------------------------------------------ FP64 form --
// Same expressions as the stage path (cpu.cpp:57-74, 218-291, 293-320 via
// filt_device.hpp), so the same bits: P2 = P1 + 0 = P1, densities from P,
// divisions by h1, h2, h3, mu1, mu2 and 1 - S_wr through DivBy (bit for bit
// the IEEE division).  LDS slot: P, Xi1, Xi2, ro1, ro2.
template <bool GEN>
struct FormFP64 {
    struct Val {
        double P, X1, X2;
    };
    static constexpr int NQ = 5;
    enum { QP = 0, QX1 = 1, QX2 = 2, QR1 = 3, QR2 = 4 };
    static constexpr int SLOT_BYTES = NQ * 8;
    static constexpr bool TOP64 = false;
    template <int SN, int TX>
    static constexpr int tile_bytes() { return SN * SLOT_BYTES; }
    // waves per SIMD the kernel is built for, with one or two columns per
    // thread
    static constexpr int WAVES = 4;
    static constexpr int WAVES2 = 3;
    // a plane's inputs are loaded at the top of its iteration: with them in
    // flight across the update this form needs more than the 128 registers
    // of four waves per SIMD (measured slower at three waves)
    static constexpr bool PIPELINED = false;

    template <class C>
    __device__ static Val eval(const C& c, const Raw& q)
    {
        const auto& p = c.G()->p;
        const int md = medium(c, q);
        const double K = perm_of(c, q, md);
        const double swr = c.tab(M_SWR, md);
        const DivBy d1{1. - swr, c.tab(M_DY, md)};
        const double S_e = d1(1. - q.S - swr);
        double k1, k2;
        if (GEN) {
            k1 = pow_spec(S_e, c.tab(M_E1, md));
            k2 = (1. - S_e) * (1. - S_e) * (1 - pow_spec(S_e, c.tab(M_E2, md)));
        } else {  // lambda = 2: exponents 4 and 2 exactly
            k1 = pow_spec(S_e, 4.0);
            k2 = (1. - S_e) * (1. - S_e) * (1 - S_e * S_e);
        }
        const DivBy dmu1{p.mu1, c.rh[3]}, dmu2{p.mu2, c.rh[4]};
        Val v;
        v.P = q.P;
        v.X1 = dmu1(-1. * K * k1);
        v.X2 = dmu2(-1. * K * k2);
        return v;
    }

    template <int SN, class C>
    __device__ static void put(char* T, int s, const C& c, const Val& v)
    {
        double* d = (double*)T;
        const auto& p = c.G()->p;
        d[QP * SN + s] = v.P;
        d[QX1 * SN + s] = v.X1;
        d[QX2 * SN + s] = v.X2;
        d[QR1 * SN + s] = ro1_of(v.P, p);
        d[QR2 * SN + s] = ro2_of(v.P, p);
    }

    // roS_calculation (cpu.cpp:218-291) then the Newton (cpu.cpp:293-320)
    // of an interior cell, expression for expression as k_step_zm.
    template <int SN, int SX, int TX, class C>
    __device__ static void update(const C& c, const Val& zc, const char* Tm, const Val& zp,
                                  const char* T, int s, const Raw& q, bool, int, double& nP,
                                  double& nS)
    {
        const double* L = (const double*)T;
        const double* M = (const double*)Tm;
        const auto& p = c.G()->p;
        const double g = p.g_const;
        const DivBy dx{p.h1, c.rh[0]}, dy{p.h2, c.rh[1]}, dz{p.h3, c.rh[2]};
        const double Pc = zc.P;
        const double r1c = ro1_of(Pc, p), r2c = ro2_of(Pc, p);
        const double S = q.S;
        const double roS1 = r1c * (1 - S);
        const double roS2 = r2c * S;
        const int xm = s - 1, xp = s + 1, ym = s - SX, yp = s + SX;
        double a2 = dz(-(zp.P - Pc));
        double a1 = dz(-(Pc - M[QP * SN + s]));
        const double r1p = ro1_of(zp.P, p), r2p = ro2_of(zp.P, p);
        const double f3 = upwind_div(a2, a1, zc.X1, r1c, M[QX1 * SN + s], M[QR1 * SN + s], zp.X1,
                                     r1p, dz);
        const double g3 = upwind_div(a2, a1, zc.X2, r2c, M[QX2 * SN + s], M[QR2 * SN + s], zp.X2,
                                     r2p, dz);
        a2 = dx(-(L[QP * SN + xp] - Pc));
        a1 = dx(-(Pc - L[QP * SN + xm]));
        const double f1 = upwind_div(a2, a1, zc.X1, r1c, L[QX1 * SN + xm], L[QR1 * SN + xm],
                                     L[QX1 * SN + xp], L[QR1 * SN + xp], dx);
        const double g1 = upwind_div(a2, a1, zc.X2, r2c, L[QX2 * SN + xm], L[QR2 * SN + xm],
                                     L[QX2 * SN + xp], L[QR2 * SN + xp], dx);
        const double dy2 = dy(-(L[QP * SN + yp] - Pc));
        const double dy1 = dy(-(Pc - L[QP * SN + ym]));
        const double f2 = upwind_div(dy2 + r1c * g, dy1 + r1c * g, zc.X1, r1c, L[QX1 * SN + ym],
                                     L[QR1 * SN + ym], L[QX1 * SN + yp], L[QR1 * SN + yp], dy);
        const double g2 = upwind_div(dy2 + r2c * g, dy1 + r2c * g, zc.X2, r2c, L[QX2 * SN + ym],
                                     L[QR2 * SN + ym], L[QX2 * SN + yp], L[QR2 * SN + yp], dy);
        const int md = medium(c, q);
        const double phi = poro_of(c, q, md);
        const double A1 = roS1 - (p.dt / phi) * (f1 + f2 + f3);
        const double A2 = roS2 - (p.dt / phi) * (g1 + g2 + g3);
        nP = Pc;
        nS = S;
        newton_cell<false>(nP, nS, A1, A2, c.tab(M_LAM, md), c.tab(M_SWR, md), c.tab(M_PD, md), p);
    }
};

// ------------------------------------------------------------ fast form --
// All FP64 with shortcuts that do not move the results beyond rounding: the
// Newton in closed form (capillary-free case) or with refined reciprocals;
// the Brooks-Corey powers as products for lambda = 2 and as exp / expm1 of
// one log1p otherwise; divisions by the uniform spacings, the viscosities and
// 1 - S_wr as multiplications by reciprocals.  LDS slot: P and the
// mobility-density products L1 = -Xi1*ro1, L2 = -Xi2*ro2.

// Mobility-density products of one cell.  1 - S_e = S/(1-S_wr) is formed
// directly, without the cancellation of 1 - S_e at small NAPL saturation.
template <bool GEN, class R, class C>
__device__ __forceinline__ void mobilities(const C& c, const Raw& q, R& L1, R& L2)
{
    CG& G = *c.G();
    const int md = medium(c, q);
    const double K = perm_of(c, q, md);
    const double P = q.P;
    double k1, k2d;
    const double lam = c.tab(M_LAM, md);
    const double oms = q.S * c.tab(M_INV1, md);
    if (!GEN || lam == 2.0) {  // Brooks-Corey exponents 4 and 2
        const double se = 1.0 - oms;
        const double s2 = se * se;
        k1 = s2 * s2;
        k2d = oms * oms * (oms * (2. - oms));  // (1-S_e)^2 (1-S_e^2)
    } else {
        // S_e^e1, and 1 - S_e^e2 = -expm1(e2*log1p(-oms)), accurate near S_e = 1
        const double l1p = log1p(-oms);
        k1 = exp(c.tab(M_E1, md) * l1p);
        k2d = oms * oms * -expm1(c.tab(M_E2, md) * l1p);
    }
    const double r1 = G.p.ro01 * (1. + G.p.beta1 * (P - G.p.P_atm));
    const double r2 = G.p.ro02 * (1. + G.p.beta2 * (P - G.p.P_atm));
    L1 = (R)(K * k1 * G.inv_mu1 * r1);
    L2 = (R)(K * k2d * G.inv_mu2 * r2);
}

// Square root of d > 0 from the hardware reciprocal square root and two
// coupled Newton-Raphson steps (s -> sqrt(d), h -> 1/(2 sqrt(d))), within an
// ulp or so.  No range scaling: 0, infinity and NaN give NaN, which the
// caller's finiteness test sends down its general path.
__device__ __forceinline__ double sqrt1(double d)
{
    const double y = __builtin_amdgcn_rsq(d);
    double s = d * y, h = 0.5 * y;
    const double r = fma(-s, h, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    const double e = fma(-s, s, d);
    return fma(e, h, s);
}

// Reciprocal from the hardware estimate with two Newton-Raphson refinements
// (within an ulp).  One refinement was measured no faster in the fused step.
__device__ __forceinline__ double rcp1(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = fma(fma(-d, r, 1.0), r, r);
    return fma(fma(-d, r, 1.0), r, r);
}

// The Newton of cpu.cpp:293-320 from the updated A1 = ro1*(1-S),
// A2 = ro2*S.  Without capillary pressure (P_k = 0) the two equations
//   ro01 (1 + b1 x)(1 - S) = A1,   ro02 (1 + b2 x) S = A2,   x = P - P_atm
// reduce to one quadratic al x^2 + be x + ga = 0.  Its root near the linear
// solution -ga/be is -2 ga / (be + sign(be) sqrt(be^2 - 4 al ga)), free of
// cancellation, and S = A2 / (ro02 (1 + b2 x)): the point the iteration
// converges to, for one square root and two reciprocals.  Outside that case
// (capillary terms, S_e < 2^-10, no finite real root) the iteration runs as
// the reference's, with refined reciprocals.
template <class C>
__device__ __forceinline__ void newton_fast(const C& c, int md, double P0, double S0,
                                            double A1, double A2, double& nP, double& nS)
{
    CG& G = *c.G();
    const auto& p = G.p;
    const double ro01 = p.ro01, ro02 = p.ro02, b1 = p.beta1, b2 = p.beta2, Pa = p.P_atm;
    const double swr = c.tab(M_SWR, md), pd = c.tab(M_PD, md);
    const double lam = c.tab(M_LAM, md), inv1 = c.tab(M_INV1, md);
    const bool no_cap = (pd == 0.0) && (lam >= 0.1);
    if (no_cap && p.newton_iterations > 0) {
        // the uniform products come with the grid's descriptor
        const double a1 = G.nq_a1, c1 = G.nq_c1;
        const double be = G.nq_be0 - A2 * a1 - A1 * c1;
        const double ga = G.nq_ga0 - A2 * ro01 - A1 * ro02;
        const double disc = be * be - G.nq_al4 * ga;
        // computed whatever the sign of disc (a negative one gives NaN) and
        // the tests combined bitwise: one branch instead of a chain
        const double x = -2.0 * ga * rcp1(be + copysign(sqrt1(disc), be));
        const double Sx = A2 * rcp1(ro02 + c1 * x);
        if ((int)(disc >= 0.0) & (int)isfinite(x) & (int)isfinite(Sx) &
            (int)((1 - Sx - swr) * inv1 >= 0.0009765625)) {
            nP = Pa + x;
            nS = Sx;
            return;
        }
    }
    double P = P0, Sv = S0;
    for (int w = 1; w <= p.newton_iterations; w++) {
        const double se = (1 - Sv - swr) * inv1;
        double P_k = 0.0, PkS = 0.0;
        if (!(no_cap && se >= 0.0009765625)) {
            P_k = pd * pow(se, (-1.) / lam);
            PkS = pow(se, ((-1.) / lam) - 1.) * pd / (lam * (1 - swr));
        }
        const double F1v = ro01 * (1. + b1 * (P - Pa)) * (1. - Sv) - A1;
        const double F2v = ro02 * (1. + b2 * (P + P_k - Pa)) * Sv - A2;
        const double F1P = ro01 * b1 * (1 - Sv);
        const double F2P = ro02 * b2 * Sv;
        const double F1S = -ro01 * (1 + b1 * (P - Pa));
        const double F2S = ro02 * (1 + b2 * (P + P_k - Pa + Sv * PkS));
        const double det = F1P * F2S - F1S * F2P;
        const double r = rcp1(det);
        const double Pn = P - r * (F2S * F1v - F1S * F2v);
        const double Sn = Sv - r * (F1P * F2v - F2P * F1v);
        if (same_bits(Pn, P) && same_bits(Sn, Sv))
            break;
        P = Pn;
        Sv = Sn;
    }
    nP = P;
    nS = Sv;
}

// dt / phi with an FP64 refined reciprocal
template <class C>
__device__ __forceinline__ double dt_over_phi(const C& c, const Raw& q)
{
    const double phi = poro_of(c, q, medium(c, q));
    return c.G()->p.dt * rcp1(phi);
}

// upwind flux with the face forces a2 (+ face) and a1 (- face)
template <class R>
__device__ __forceinline__ R uflux(R a2, R a1, R Lc, R Lm, R Lp)
{
    const R z = 0;
    return (fmax(a2, z) - fmin(a1, z)) * Lc - fmax(a1, z) * Lm + fmin(a2, z) * Lp;
}

template <bool GEN>
struct FormFast {
    struct Val {
        double P, L1, L2;
    };
    static constexpr int SLOT_BYTES = 24;
    static constexpr bool TOP64 = false;
    template <int SN, int TX>
    static constexpr int tile_bytes() { return SN * SLOT_BYTES; }
    static constexpr int WAVES = GEN ? 3 : 4;
    static constexpr int WAVES2 = GEN ? 2 : 3;
    static constexpr bool PIPELINED = true;

    template <class C>
    __device__ static Val eval(const C& c, const Raw& q)
    {
        Val v;
        v.P = q.P;
        mobilities<GEN>(c, q, v.L1, v.L2);
        return v;
    }

    template <int SN, class C>
    __device__ static void put(char* T, int s, const C&, const Val& v)
    {
        double* d = (double*)T;
        d[s] = v.P;
        d[SN + s] = v.L1;
        d[2 * SN + s] = v.L2;
    }

    template <int SN, int SX, int TX, class C>
    __device__ static void update(const C& c, const Val& zc, const char* Tm, const Val& zp,
                                  const char* T, int s, const Raw& q, bool, int, double& nP,
                                  double& nS)
    {
        const double* L = (const double*)T;
        const double* M = (const double*)Tm;
        CG& G = *c.G();
        const auto& p = G.p;
        const double S = q.S;
        const int xm = s - 1, xp = s + 1, ym = s - SX, yp = s + SX;
        const double p1 = zc.P;
        const double r1c = p.ro01 * (1. + p.beta1 * (p1 - p.P_atm));
        const double r2c = p.ro02 * (1. + p.beta2 * (p1 - p.P_atm));
        const double ihx = G.inv_h[0], ihy = G.inv_h[1], ihz = G.inv_h[2];
        const double ax2 = (p1 - L[xp]) * ihx, ax1 = (L[xm] - p1) * ihx;
        const double az2 = (p1 - zp.P) * ihz, az1 = (M[s] - p1) * ihz;
        const double dy2 = (p1 - L[yp]) * ihy, dy1 = (L[ym] - p1) * ihy;
        const double g = p.g_const;
        const double F1 = uflux(ax2, ax1, zc.L1, L[SN + xm], L[SN + xp]) * ihx +
                          uflux(dy2 + r1c * g, dy1 + r1c * g, zc.L1, L[SN + ym], L[SN + yp]) * ihy +
                          uflux(az2, az1, zc.L1, M[SN + s], zp.L1) * ihz;
        const double F2 = uflux(ax2, ax1, zc.L2, L[2 * SN + xm], L[2 * SN + xp]) * ihx +
                          uflux(dy2 + r2c * g, dy1 + r2c * g, zc.L2, L[2 * SN + ym],
                                L[2 * SN + yp]) * ihy +
                          uflux(az2, az1, zc.L2, M[2 * SN + s], zp.L2) * ihz;
        const double dtphi = dt_over_phi(c, q);
        const double A1 = r1c * (1 - S) - dtphi * F1;
        const double A2 = r2c * S - dtphi * F2;
        newton_fast(c, medium(c, q), p1, S, A1, A2, nP, nS);
    }
};

// ----------------------------------------------------------- mixed form --
// The upwind fluxes and the ro*S increment in FP32: the mobility-density
// products L1, L2 are FP32 (in LDS and registers), the six face forces are
// formed in FP64 (pressure differences, and for y the cancelling sum
// -dP/h + ro*g) and rounded to FP32 once, the flux products, their scaling
// by 1/h and the sum over the three axes are FP32.  FP64 keeps the pressure
// (LDS and registers), the face forces before rounding, the base
// ro(P)(1-S), the factor dt/phi and its product with the FP32 flux sum (a
// per-cell rounding of dt/phi would bias every increment of the cell alike),
// the Newton and the mass sums.  LDS slot: P (FP64), L1, L2 (FP32).
//
// One face keeps FP64 fluxes: the face between the top boundary row j = 0
// and the first interior row j = 1.  The top boundary holds a fixed state
// (P_atm, and S2_gr on the source patch), so the FP32 rounding of its
// mobility would be the same relative error in the inflow on every step, a
// bias that accumulates in the NAPL mass (measured 1e-8 relative on the
// configurations; with this face in FP64 the NAPL mass stays within 1e-11 of
// the all-FP64 result).  Interior x and z face fluxes telescope exactly in
// FP32 as well (both neighbours form each face flux from the same FP32
// operands), so their rounding moves no mass.  The j = 0 row's FP64
// mobilities are kept in an extra LDS row of the top tiles.
template <bool GEN>
struct FormMixed {
    struct Val {
        double P;
        float L1, L2;
    };
    static constexpr int SLOT_BYTES = 16;
    static constexpr int WAVES = GEN ? 3 : 4;
    static constexpr int WAVES2 = GEN ? 2 : 3;
    static constexpr bool PIPELINED = true;
    static constexpr bool TOP64 = true;
    // per tile: the slots, then FP64 L1 and L2 of the j = 0 ring row
    template <int SN, int TX>
    static constexpr int tile_bytes() { return SN * SLOT_BYTES + TX * 16; }

    // FP64 mobilities of the top boundary cell at tile column x
    template <int SN, int TX, class C>
    __device__ static void put_top(char* T, int x, const C& c, const Raw& q)
    {
        double* t = (double*)(T + SN * SLOT_BYTES);
        double L1, L2;
        mobilities<GEN>(c, q, L1, L2);
        t[x] = L1;
        t[TX + x] = L2;
    }

    template <class C>
    __device__ static Val eval(const C& c, const Raw& q)
    {
        Val v;
        v.P = q.P;
        mobilities<GEN>(c, q, v.L1, v.L2);
        return v;
    }

    template <int SN, class C>
    __device__ static void put(char* T, int s, const C&, const Val& v)
    {
        ((double*)T)[s] = v.P;
        float* f = (float*)(T + SN * 8);
        f[s] = v.L1;
        f[SN + s] = v.L2;
    }

    template <int SN, int SX, int TX, class C>
    __device__ static void update(const C& c, const Val& zc, const char* Tm, const Val& zp,
                                  const char* T, int s, const Raw& q, bool jtop, int tx,
                                  double& nP, double& nS)
    {
        const double* LP = (const double*)T;
        const double* MP = (const double*)Tm;
        const float* L = (const float*)(T + SN * 8);
        const float* M = (const float*)(Tm + SN * 8);
        CG& G = *c.G();
        const auto& p = G.p;
        const double S = q.S;
        const int xm = s - 1, xp = s + 1, ym = s - SX, yp = s + SX;
        const double p1 = zc.P;
        const double r1c = p.ro01 * (1. + p.beta1 * (p1 - p.P_atm));
        const double r2c = p.ro02 * (1. + p.beta2 * (p1 - p.P_atm));
        const double ihx = G.inv_h[0], ihy = G.inv_h[1], ihz = G.inv_h[2];
        const double g = p.g_const;
        // face forces in FP64, rounded once
        const float ax2 = (float)((p1 - LP[xp]) * ihx), ax1 = (float)((LP[xm] - p1) * ihx);
        const float az2 = (float)((p1 - zp.P) * ihz), az1 = (float)((MP[s] - p1) * ihz);
        const double dy2 = (p1 - LP[yp]) * ihy, dy1 = (LP[ym] - p1) * ihy;
        const float ay2w = (float)(dy2 + r1c * g), ay2n = (float)(dy2 + r2c * g);
        const float fx = (float)ihx, fy = (float)ihy, fzz = (float)ihz;
        const float X1 = uflux(ax2, ax1, zc.L1, L[xm], L[xp]) * fx +
                         uflux(az2, az1, zc.L1, M[s], zp.L1) * fzz;
        const float X2 = uflux(ax2, ax1, zc.L2, L[SN + xm], L[SN + xp]) * fx +
                         uflux(az2, az1, zc.L2, M[SN + s], zp.L2) * fzz;
        double F1, F2;
        if (!jtop) {
            const float ay1w = (float)(dy1 + r1c * g), ay1n = (float)(dy1 + r2c * g);
            F1 = (double)(X1 + uflux(ay2w, ay1w, zc.L1, L[ym], L[yp]) * fy);
            F2 = (double)(X2 + uflux(ay2n, ay1n, zc.L2, L[SN + ym], L[SN + yp]) * fy);
        } else {
            // + face in FP32; - face (to the top boundary) in FP64
            const float z = 0.f;
            const float pl1 = fmaxf(ay2w, z) * zc.L1 + fminf(ay2w, z) * L[yp];
            const float pl2 = fmaxf(ay2n, z) * zc.L2 + fminf(ay2n, z) * L[SN + yp];
            double L1c, L2c;
            mobilities<GEN>(c, q, L1c, L2c);
            const double* top = (const double*)(T + SN * SLOT_BYTES);
            const double a1w = dy1 + r1c * g, a1n = dy1 + r2c * g;
            const double mi1 = fmin(a1w, 0.) * L1c + fmax(a1w, 0.) * top[tx];
            const double mi2 = fmin(a1n, 0.) * L2c + fmax(a1n, 0.) * top[TX + tx];
            F1 = (double)X1 + ((double)pl1 - mi1) * ihy;
            F2 = (double)X2 + ((double)pl2 - mi2) * ihy;
        }
        const double dtphi = dt_over_phi(c, q);
        const double A1 = r1c * (1 - S) - dtphi * F1;
        const double A2 = r2c * S - dtphi * F2;
        newton_fast(c, medium(c, q), p1, S, A1, A2, nP, nS);
    }
};

// -------------------------------------------------------------- writing --

// The shell cells that copy interior cell (i,j,k) (byte offset off), whose
// own output the plane's buffer stores write.  i1: i == 1; lastcol: i ==
// nx-2; top: j == ny-2 (with Pt0 = ro1(i,1,0)*g*h2 of the input state); kl,
// kh: k == 1, k == nz-2.
template <class C>
__device__ __forceinline__ void write_copies(const C& c, unsigned off, bool i1, bool lastcol,
                                             bool top, double Pt0, bool kl, bool kh, double nP,
                                             double nS)
{
    double* const Po = c.Po();
    double* const So = c.So();
    if (i1) {
        gst(Po, off - 8, nP);
        gst(So, off - 8, nS);
    }
    if (lastcol) {
        gst(Po, off + 8, nP);
        gst(So, off + 8, nS);
    }
    if (!(top || kl || kh))
        return;
    const double Pt = nP + Pt0;
    const unsigned sz = c.sz(), sy = c.sy();
    const unsigned k0 = off - (kl ? sz : 0u);  // (i, j, 0) when kl
    const unsigned kL = off + (kh ? sz : 0u);  // (i, j, L-1) when kh
    for (int q = 0; q < (lastcol ? 2 : 1); q++) {
        const unsigned d = 8u * q;  // column i, then N-1
        if (top) {
            gst(Po, off + sy + d, Pt);
            gst(So, off + sy + d, nS);
            if (kl) {
                gst(Po, k0 + sy + d, Pt);
                gst(So, k0 + sy + d, nS);
            }
            if (kh) {
                gst(Po, kL + sy + d, Pt);
                gst(So, kL + sy + d, nS);
            }
        }
        if (kl) {
            gst(Po, k0 + d, nP);
            gst(So, k0 + d, nS);
        }
        if (kh) {
            gst(Po, kL + d, nP);
            gst(So, kL + d, nS);
        }
    }
}

// Shell cells that do not depend on this step's interior values.
__device__ inline void shell_cell(const FusedGrid& G, const double* Pin_, const double* Sin_,
                                  double* Po_, double* So_, long long q)
{
    typedef const __attribute__((address_space(1))) double* CG;
    typedef __attribute__((address_space(1))) double* WG;
    const CG Pin = (CG)Pin_, Sin = (CG)Sin_;
    const WG Po = (WG)Po_, So = (WG)So_;
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const long long sy = nx, sz = (long long)nx * ny;
    const long long ringA = 2LL * ny + 2LL * (nz - 2);
    if (q < ringA) {  // i = 0 face where (1, j, k) is a shell cell
        int j, k;
        if (q < ny) { j = (int)q; k = 0; }
        else if (q < 2LL * ny) { j = (int)(q - ny); k = nz - 1; }
        else if (q < 2LL * ny + (nz - 2)) { j = 0; k = 1 + (int)(q - 2LL * ny); }
        else { j = ny - 1; k = 1 + (int)(q - 2LL * ny - (nz - 2)); }
        const long long c = j * sy + k * sz;
        Po[c] = Pin[c + 1];
        So[c] = Sin[c + 1];
        return;
    }
    q -= ringA;
    if (q >= (long long)(nx - 1) * nz)
        return;
    const int i = 1 + (int)(q % (nx - 1));
    const int k = (int)(q / (nx - 1));
    const int ii = i < nx - 1 ? i : nx - 2;
    const long long c = i + k * sz;
    Po[c] = G.p.P_atm;
    So[c] = (k == 0) ? Sin[ii + sz]
                     : (in_patch(ii, k < nz - 1 ? k : nz - 2, G.p) ? G.p.S2_gr : 0.0);
}

// Update of a thread's CY cells of plane k and their outputs: the cells'
// own values by buffer stores every lane issues (DROP where a lane has no
// interior cell), the shell copies of edge cells after them.
template <class F, int SN, int SX, int TX, int CY, class C>
__device__ __forceinline__ void write_plane(const C& c, const typename F::Val* zc,
                                            const char* Lm, const typename F::Val* zp,
                                            const char* L, const int* s, const Raw* qc,
                                            const bool* jtop, int tx, const bool* inner,
                                            const unsigned* col, int k, int nz, bool i1,
                                            bool lastcol, const bool* top, const double* Pt0)
{
    double nP[CY], nS[CY];
#pragma unroll
    for (int q = 0; q < CY; q++)
        if (inner[q])
            F::template update<SN, SX, TX>(c, zc[q], Lm, zp[q], L, s[q], qc[q], jtop[q], tx,
                                           nP[q], nS[q]);
    const unsigned bytes = c.bytes(), plane = (unsigned)k * c.sz();
    const Rsrc rP = out_rsrc(c.Po(), bytes), rS = out_rsrc(c.So(), bytes);
#pragma unroll
    for (int q = 0; q < CY; q++) {
        const unsigned off = inner[q] ? col[q] + plane : DROP;
        bstn(rP, off, nP[q]);
        bstn(rS, off, nS[q]);
    }
    const bool kl = k == 1, kh = k == nz - 2;
#pragma unroll
    for (int q = 0; q < CY; q++)
        if (inner[q] && (i1 || lastcol || top[q] || kl || kh))
            write_copies(c, col[q] + plane, i1, lastcol, top[q], Pt0[q], kl, kh, nP[q], nS[q]);
}

// ------------------------------------------------------------- the tile --
// A tile of TX x TY columns; each thread owns CY of them, stacked in y
// (register blocking: the per-plane work common to a thread's cells -- loop
// control, descriptor reads, the barrier -- is paid once for CY cells, and
// their independent arithmetic interleaves).
template <class F, int TX, int TY, int CY, class C>
__device__ __forceinline__ void march(const C& c, const FusedGrid& G, unsigned lb, char* lds)
{
    constexpr int NT = TX * TY / CY;  // threads
    constexpr int NW = NT / 64;
    constexpr int SX = TX + 2;
    constexpr int SN = SX * (TY + 2);
    constexpr int RING = 2 * TX + 2 * TY;
    static_assert(NT % 64 == 0 && RING <= NT && TY % CY == 0, "tile shape");
    const unsigned nbx = (unsigned)G.nbx, nby = (unsigned)G.nby;
    const int bx = (int)(lb % nbx);
    const unsigned r = lb / nbx;
    const int by = (int)(r % nby);
    const int bz = (int)(r / nby);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    // x-tiles start at x = 0, so a tile's rows begin on 128-byte lines when
    // the grid's rows do: the column x = 0 is evaluated as a neighbour but
    // not updated
    const int i0 = bx * TX, j0 = 1 + by * TY;
    const int kb = 1 + bz * G.kz, ke = min(kb + G.kz, nz - 1);
    if (kb >= ke)
        return;  // uniform over the block
    const int t = threadIdx.x;
    const int tx = t % TX, ty = t / TX;  // this thread: rows CY*ty .. CY*ty+CY-1
    const int i = i0 + tx;
    bool in_grid[CY], inner[CY], top[CY], jtop[CY];
    unsigned col[CY];
    int s[CY];
    double Pt0[CY];
#pragma unroll
    for (int q = 0; q < CY; q++) {
        const int j = j0 + CY * ty + q;
        in_grid[q] = i < nx && j < ny;
        inner[q] = i >= 1 && i <= nx - 2 && j <= ny - 2;
        top[q] = j == ny - 2;   // shell copies of row M-1 (loop-invariant)
        jtop[q] = j == 1;       // the row next to the top boundary
        col[q] = 8u * (unsigned)(i + j * nx);
        s[q] = (CY * ty + q + 1) * SX + (tx + 1);
        Pt0[q] = 0.0;
        if (inner[q] && top[q]) {
            const double P10 = gld(c.Pin(), 8u * (unsigned)(i + nx));  // (i, 1, 0)
            CG& Gc = *c.G();
            Pt0[q] = ro1_of(P10, Gc.p) * Gc.p.g_const * Gc.p.h2;
        }
    }
    // ring cell h of this thread: waves w0, w0+1, ... (mod NW) take the ring
    const int w0 = (int)(blockIdx.x % NW);
    const int h = ((((t >> 6) - w0) + NW) % NW) * 64 + (t & 63);
    int rx = 0, ry = 0;
    if (h < TX) { rx = h; ry = -1; }
    else if (h < 2 * TX) { rx = h - TX; ry = TY; }
    else if (h < 2 * TX + TY) { rx = -1; ry = h - 2 * TX; }
    else if (h < RING) { rx = TX; ry = h - 2 * TX - TY; }
    const bool ring = h < RING && i0 + rx >= 0 && i0 + rx < nx && j0 + ry < ny;
    const unsigned rcol = 8u * (unsigned)((i0 + rx) + (j0 + ry) * nx);
    const int rs = (ry + 1) * SX + (rx + 1);
    const bool i1 = i == 1, lastcol = i == nx - 2;
    char* const T0 = lds;
    char* const T1 = lds + F::template tile_bytes<SN, TX>();
    // the top tiles' ring cells of row j = 0
    const bool top_ring = ring && ry == -1 && j0 == 1;
    using Val = typename F::Val;
    Val zc[CY], zp[CY];
    if constexpr (!F::PIPELINED) {
        // Loads of a plane at the top of its iteration: fewer registers live
        // across the update, for a form at the edge of an occupancy step.
        Raw qc[CY];
#pragma unroll
        for (int q = 0; q < CY; q++)
            if (in_grid[q]) {
                F::template put<SN>(T1, s[q], c,
                                    F::eval(c, load_raw(c, col[q] + (kb - 1) * c.sz(), false)));
                qc[q] = load_raw(c, col[q] + kb * c.sz(), true);
                zc[q] = F::eval(c, qc[q]);
            }
        for (int k = kb; k < ke; k++) {
            const bool odd = ((k - kb) & 1) != 0;
            char* const L = odd ? T1 : T0;
            const char* const Lm = odd ? T0 : T1;
            Raw qp[CY];
#pragma unroll
            for (int q = 0; q < CY; q++)
                if (in_grid[q]) {
                    qp[q] = load_raw(c, col[q] + (k + 1) * c.sz(), true);  // k+1 <= ke <= nz-1
                    F::template put<SN>(L, s[q], c, zc[q]);
                }
            if (ring) {
                const Raw rr = load_raw(c, rcol + k * c.sz(), false);
                F::template put<SN>(L, rs, c, F::eval(c, rr));
                if constexpr (F::TOP64)
                    if (top_ring)
                        F::template put_top<SN, TX>(L, rx, c, rr);
            }
#pragma unroll
            for (int q = 0; q < CY; q++)
                if (in_grid[q])
                    zp[q] = F::eval(c, qp[q]);
            __syncthreads();
            write_plane<F, SN, SX, TX, CY>(c, zc, Lm, zp, L, s, qc, jtop, tx, inner, col, k, nz,
                                           i1, lastcol, top, Pt0);
#pragma unroll
            for (int q = 0; q < CY; q++) {
                zc[q] = zp[q];
                qc[q] = qp[q];
            }
        }
        return;
    } else {
        // Pipelined: qp and rq hold loads in flight across the loop's back
        // edge; each is consumed at the top of the next iteration and only
        // then reloaded, never copied while in flight (a copy would make the
        // wave wait for every memory operation it has outstanding, stores
        // included, once per plane)
        Raw qc[CY], qp[CY], rq{};
#pragma unroll
        for (int q = 0; q < CY; q++)
            if (in_grid[q]) {
                // plane kb-1 goes to the own slot of the tile plane kb reads
                // as k-1
                F::template put<SN>(
                    T1, s[q], c, F::eval(c, load_raw(c, col[q] + (kb - 1) * c.sz(), false)));
                qc[q] = load_raw(c, col[q] + kb * c.sz(), true);
                zc[q] = F::eval(c, qc[q]);
            }
        // everything loaded so far has arrived before the loop's loads are
        // issued: the loop is entered with only those in flight (a field of qc
        // that eval does not read would otherwise still be pending at the loop
        // header, and the compiler's wait for it would then also cover loads
        // issued in every later plane)
        __builtin_amdgcn_s_waitcnt(VMCNT0);
#pragma unroll
        for (int q = 0; q < CY; q++)
            if (in_grid[q])
                qp[q] = load_raw(c, col[q] + (kb + 1) * c.sz(), true);  // kb+1 <= ke <= nz-1
        if (ring)
            rq = load_raw(c, rcol + kb * c.sz(), false);
        {
            // two stores that write nothing: the loop is entered with the same
            // memory operations behind the prefetch loads as at its back edge,
            // so the first plane's wait for them has the same count as later
            // planes' (without them, that count would be zero for all planes)
            const Rsrc rP = out_rsrc(c.Po(), c.bytes()), rS = out_rsrc(c.So(), c.bytes());
            bstn(rP, DROP, 0.0);
            bstn(rS, DROP, 0.0);
        }
        for (int k = kb; k < ke; k++) {
            const bool odd = ((k - kb) & 1) != 0;
            char* const L = odd ? T1 : T0;
            const char* const Lm = odd ? T0 : T1;
            Raw qnext[CY];  // the next plane's centre inputs, arrived
#pragma unroll
            for (int q = 0; q < CY; q++)
                if (in_grid[q]) {
                    zp[q] = F::eval(c, qp[q]);
                    qnext[q] = qp[q];
                    F::template put<SN>(L, s[q], c, zc[q]);
                }
            if (ring) {
                F::template put<SN>(L, rs, c, F::eval(c, rq));
                if constexpr (F::TOP64)
                    if (top_ring)
                        F::template put_top<SN, TX>(L, rx, c, rq);
            }
            if (k + 1 < ke) {  // the next plane is updated: fetch its inputs
#pragma unroll
                for (int q = 0; q < CY; q++)
                    if (in_grid[q])
                        qp[q] = load_raw(c, col[q] + (k + 2) * c.sz(), true);  // k+2 <= nz-1
                if (ring)
                    rq = load_raw(c, rcol + (k + 1) * c.sz(), false);
            }
            __syncthreads();
            write_plane<F, SN, SX, TX, CY>(c, zc, Lm, zp, L, s, qc, jtop, tx, inner, col, k, nz,
                                           i1, lastcol, top, Pt0);
#pragma unroll
            for (int q = 0; q < CY; q++) {
                zc[q] = zp[q];
                qc[q] = qnext[q];
            }
        }
    }
}

template <class F, int TX, int TY, int CY, bool FIELDS>
__global__ __launch_bounds__(TX* TY / CY) __attribute__((amdgpu_waves_per_eu(
    CY == 1 ? F::WAVES : F::WAVES2))) void k_fused(
    const FusedGrid* __restrict__ grids, const int* __restrict__ block_grid, long long total_blocks,
    int parity, int order)
{
    constexpr int NT = TX * TY / CY;
    constexpr int SN = (TX + 2) * (TY + 2);
    __shared__ __attribute__((aligned(16))) char lds[2 * F::template tile_bytes<SN, TX>()];
    __shared__ double med[M_ROWS * FILT_MAX_MEDIA];
    __shared__ double rh[5];
    long long b = blockIdx.x;
    if (order == 0 && total_blocks % 8 == 0)
        b = (b % 8) * (total_blocks / 8) + b / 8;
    const FusedGrid& G = grids[block_grid[b]];
    const long long lb = b - G.first_block;
    const int t = threadIdx.x;
    if (lb < G.n_interior_blocks) {  // uniform over the block
        if (t < M_ROWS * FILT_MAX_MEDIA) {
            const int row = t / FILT_MAX_MEDIA, q = t % FILT_MAX_MEDIA;
            const filt_params& p = G.p;
            double v;
            switch (row) {
            case M_LAM: v = p.lambda[q]; break;
            case M_SWR: v = p.S_wr[q]; break;
            case M_PD: v = p.P_d[q]; break;
            case M_M: v = p.m[q]; break;
            case M_K: v = p.K[q]; break;
            case M_INV1: v = G.inv_1mswr[q]; break;
            case M_DY: v = refined_reciprocal(1. - p.S_wr[q]); break;
            case M_E1: v = (2. + 3. * p.lambda[q]) / p.lambda[q]; break;
            default: v = (2. + p.lambda[q]) / p.lambda[q]; break;
            }
            med[t] = v;
        } else if (t >= 80 && t < 85) {
            const filt_params& p = G.p;
            const int q = t - 80;
            rh[q] = refined_reciprocal(q == 0 ? p.h1 : q == 1 ? p.h2 : q == 2 ? p.h3
                                       : q == 3 ? p.mu1 : p.mu2);
        }
        __syncthreads();
        Ctx<FIELDS> c;
        c.g = (CG*)(uintptr_t)&G;
        c.med = med;
        c.rh = rh;
        c.parity = parity;
        march<F, TX, TY, CY>(c, G, (unsigned)lb, lds);
        return;
    }
    const long long q = (lb - G.n_interior_blocks) * NT + t;
    shell_cell(G, G.P[parity], G.S[parity], G.P[parity ^ 1], G.S[parity ^ 1], q);
}

// Launch of form F at tile shape `tile`: for lambda = 2 only (with every
// grid's K and porosity per cell and one medium, or not), or general.
template <template <bool> class F>
int launch_form(const FusedGrid* g, const int* bg, long long total, int parity, bool general,
                bool fields, int tile, int order, hipStream_t s)
{
    const dim3 grid((unsigned)total);
    const FusedTile& ft = FUSED_TILES[tile];
#define FILT_FUSED_CASE(TXV, TYV, CYV)                                                     \
    if (ft.tx == TXV && ft.ty == TYV && ft.cy == CYV) {                                    \
        const dim3 block(TXV * TYV / CYV);                                                 \
        if (general)                                                                       \
            hipLaunchKernelGGL((k_fused<F<true>, TXV, TYV, CYV, false>), grid, block, 0, s, \
                               g, bg, total, parity, order);                               \
        else if (fields)                                                                   \
            hipLaunchKernelGGL((k_fused<F<false>, TXV, TYV, CYV, true>), grid, block, 0, s, \
                               g, bg, total, parity, order);                               \
        else                                                                               \
            hipLaunchKernelGGL((k_fused<F<false>, TXV, TYV, CYV, false>), grid, block, 0,   \
                               s, g, bg, total, parity, order);                            \
        FILT_LAUNCH_CHECK("k_fused");                                                      \
        return FILT_OK;                                                                    \
    }
    FILT_FUSED_CASE(16, 16, 1)
    FILT_FUSED_CASE(32, 8, 1)
    FILT_FUSED_CASE(64, 4, 1)
    FILT_FUSED_CASE(32, 16, 1)
    FILT_FUSED_CASE(64, 8, 1)
    FILT_FUSED_CASE(64, 8, 2)
    FILT_FUSED_CASE(32, 16, 2)
    FILT_FUSED_CASE(64, 16, 2)
    FILT_FUSED_CASE(128, 4, 1)
#undef FILT_FUSED_CASE
    return set_error(FILT_ERR_INVALID, "unknown fused tile %d", tile);
}

}  // namespace fz
}  // namespace filt

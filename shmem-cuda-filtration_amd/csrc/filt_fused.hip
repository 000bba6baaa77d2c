// filt_fused.hip -- one launch per time step for a batch of grids.
//
// The fused step computes what the five defines.h stages compute for P1 and
// S2 (cpu.cpp:3-54 in the order of main.cpp:83-94) without materialising the
// intermediate arrays: state lives in ping-pong buffers (P, S in / out), the
// static fields (permeability, porosity, media) are read once per cell, and
// stage 1 (density, mobility) is evaluated on the fly for the 7-point
// stencil.  Per cell and step it moves 16 B of state in, 16 B out and the
// statics (16 B FP64, 8 B FP32): 48 / 40 B against the 304 B of the stage
// path.  u_calculation is not on the P/S path; it runs on demand.
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
// Launch: interior tiles of TX x TY columns march over z-chunks (stage-1
// values of the tile plane and its one-cell ring in LDS, the own column's
// k-1 and k+1 in registers), shell-pass blocks follow each grid's tiles.  A
// table maps each block to its grid, so grids of different shapes share one
// launch; blocks are regrouped per XCD as in filt_tiles.hpp.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "../../include/filtration_hip.h"
#include "filt_device.hpp"
#include "filt_fused.hpp"
#include "filt_runtime.hpp"

namespace filt {

namespace {

constexpr int FTX = 16;
constexpr int FTY = 16;
constexpr int FT = FTX * FTY;
constexpr int RING = 2 * FTX + 2 * FTY;
constexpr int SX = FTX + 2;
constexpr int SN = SX * (FTY + 2);

__device__ __forceinline__ int slot(int x, int y) { return (y + 1) * SX + (x + 1); }

// The descriptor's device pointers are generic.  Accessed through gmem(),
// loads and stores are global_* instructions, counted by vmcnt alone; as
// flat_* instructions they would also count in lgkmcnt, so every wait for an
// LDS access in the tile loop would also wait for the HBM loads in flight.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gmem(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}

template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ int media_at(const FusedGrid& G, long long c)
{
    return (G.media && G.p.n_media > 1) ? gmem(G.media)[c] : 0;
}

// The media table in LDS, one row of FILT_MAX_MEDIA per quantity: the tile
// loop looks entries up by each cell's medium, and a per-lane load from the
// descriptor in global memory would make its wait also wait for the loads
// in flight for the next plane.
enum { M_LAM = 0, M_SWR = 1, M_PD = 2, M_M = 3, M_K = 4, M_INV1 = 5, M_ROWS = 6 };
__device__ __forceinline__ double mtab(const double* med, int row, int md)
{
    return med[row * FILT_MAX_MEDIA + md];
}

// ------------------------------------------------------------ FP64 form --
// Stage-1 values of one cell as the flux needs them (P2 = P1 + 0 = P1).
struct C1 {
    double P, X1, X2, r1, r2;
};

// Inputs of one cell, loaded a plane ahead of their use: state, K and, for
// the cell a thread updates, the porosity.
struct RawD {
    double P, S, K, phi;
    int md;
};

__device__ inline RawD fp64_raw(const FusedGrid& G, const double* med, const double* Pin,
                                const double* Sin, long long c, bool centre)
{
    RawD r;
    r.md = media_at(G, c);
    r.P = gmem(Pin)[c];
    r.S = gmem(Sin)[c];
    r.K = G.perm ? gmem(G.perm)[c] : mtab(med, M_K, r.md);
    r.phi = centre ? (G.poro ? gmem(G.poro)[c] : mtab(med, M_M, r.md)) : 0.0;
    return r;
}

__device__ inline C1 fp64_eval(const FusedGrid& G, const double* med, const RawD& q)
{
    const Stage1 r = stage1_cell(q.P, q.S, q.K, mtab(med, M_LAM, q.md), mtab(med, M_SWR, q.md),
                                 G.p);
    return C1{q.P, r.Xi1, r.Xi2, r.ro1, r.ro2};
}

// LDS planes of the FP64 form: P, Xi1, ro1, Xi2, ro2.
struct Fp64Lds {
    double* P;
    double* X1;
    double* R1;
    double* X2;
    double* R2;
    __device__ void put(int s, const C1& v) const
    {
        P[s] = v.P;
        X1[s] = v.X1;
        R1[s] = v.r1;
        X2[s] = v.X2;
        R2[s] = v.r2;
    }
};

// roS_calculation (cpu.cpp:218-291) then Newton (cpu.cpp:293-320) for an
// interior cell, expression for expression as in filt_stages.hip; neighbours
// in x / y from LDS slots, in z from the registers zm / zp.
__device__ inline void fp64_update(const FusedGrid& G, const double* med, const C1& cc,
                                   const C1& zm, const C1& zp, const Fp64Lds& L, int s,
                                   const RawD& q, double& nP, double& nS)
{
    const filt_params& p = G.p;
    const int xm = s - 1, xp = s + 1, ym = s - SX, yp = s + SX;
    const double p1 = cc.P, g = p.g_const;
    // face forces: P2 = P1 here, so both phases share them (y adds ro_c*g)
    const double az2 = -(zp.P - p1) / p.h3, az1 = -(p1 - zm.P) / p.h3;
    const double ax2 = -(L.P[xp] - p1) / p.h1, ax1 = -(p1 - L.P[xm]) / p.h1;
    const double dy2 = -(L.P[yp] - p1) / p.h2, dy1 = -(p1 - L.P[ym]) / p.h2;
    const int md = q.md;
    const double phi = q.phi, S = q.S;
    const double f3 = upwind(az2, az1, cc.X1, cc.r1, zm.X1, zm.r1, zp.X1, zp.r1, p.h3);
    const double f1 = upwind(ax2, ax1, cc.X1, cc.r1, L.X1[xm], L.R1[xm], L.X1[xp], L.R1[xp], p.h1);
    const double f2 = upwind(dy2 + cc.r1 * g, dy1 + cc.r1 * g, cc.X1, cc.r1, L.X1[ym], L.R1[ym],
                             L.X1[yp], L.R1[yp], p.h2);
    const double A1 = cc.r1 * (1 - S) - (p.dt / phi) * (f1 + f2 + f3);
    const double g3 = upwind(az2, az1, cc.X2, cc.r2, zm.X2, zm.r2, zp.X2, zp.r2, p.h3);
    const double g1 = upwind(ax2, ax1, cc.X2, cc.r2, L.X2[xm], L.R2[xm], L.X2[xp], L.R2[xp], p.h1);
    const double g2 = upwind(dy2 + cc.r2 * g, dy1 + cc.r2 * g, cc.X2, cc.r2, L.X2[ym], L.R2[ym],
                             L.X2[yp], L.R2[yp], p.h2);
    const double A2 = cc.r2 * S - (p.dt / phi) * (g1 + g2 + g3);
    nP = cc.P;
    nS = S;
    newton_cell<false>(nP, nS, A1, A2, mtab(med, M_LAM, md), mtab(med, M_SWR, md),
                       mtab(med, M_PD, md), p);
}

// ----------------------------------------------------------- mixed form --
// The fast form (FILT_PREC_MIXED, named for FP32 parts it no longer has).
// Its arithmetic differs from the FP64 form's where that buys speed without
// moving the results: the Newton in closed form (capillary-free case) or with
// refined reciprocals; the Brooks-Corey powers as products for lambda = 2 and
// as exp / expm1 of one log1p otherwise; divisions by the uniform spacings,
// the viscosities and 1 - S_wr as multiplications by reciprocals.  All of it
// is FP64, the mobility-density products L1 = -Xi1*ro1 and L2 = -Xi2*ro2 kept
// in LDS included.  On this GPU an unpacked FP32 operation costs the issue
// slot of an FP64 one: with L1 and the water fluxes in FP32 the form ran at
// the same speed and its NAPL mass drifted from the FP64 form's by up to 1e-9
// relative (through the pressure), with FP32 powers for lambda != 2 by 5e-8.
struct M1 {
    double P;
    double L2;
    double L1;
};

struct RawM {
    double P, S, K, phi;
    int md;
};

__device__ inline RawM mixed_raw(const FusedGrid& G, const double* med, const double* Pin,
                                 const double* Sin, long long c, bool centre)
{
    RawM r;
    r.md = media_at(G, c);
    r.P = gmem(Pin)[c];
    r.S = gmem(Sin)[c];
    // K by medium is looked up when the cell is evaluated, not here: a table
    // read indexed by the medium just loaded would wait for that load, and
    // with it for every load in flight
    r.K = G.perm ? gmem(G.perm)[c] : 0.0;
    r.phi = (centre && G.poro) ? gmem(G.poro)[c] : 0.0;
    return r;
}

// GENERAL: media with lambda != 2 occur in the batch.  Without it every
// medium has lambda = 2 (the launch checks) and the kernel carries no
// exp / log1p code, whose registers would halve its occupancy.
template <bool GENERAL>
__device__ inline M1 mixed_eval(const FusedGrid& G, const double* med, const RawM& q)
{
    const int md = q.md;
    const double K = G.perm ? q.K : mtab(med, M_K, md);
    const double P = q.P;
    // 1 - S_e = S/(1-S_wr) formed directly, without the cancellation of
    // 1 - S_e at small NAPL saturation
    double k1;
    double k2d;
    const double lamd = mtab(med, M_LAM, md);
    const double omsd = q.S * mtab(med, M_INV1, md);
    if (!GENERAL || lamd == 2.0) {  // Brooks-Corey exponents 4 and 2
        const double sed = 1.0 - omsd;
        const double s2 = sed * sed;
        k1 = s2 * s2;
        k2d = omsd * omsd * (omsd * (2. - omsd));  // (1-S_e)^2 (1-S_e^2)
    } else {
        // S_e^e1, and 1 - S_e^e2 = -expm1(e2*log1p(-oms)), accurate for S_e
        // near 1
        const double l1p = log1p(-omsd);
        k1 = exp((2. + 3. * lamd) / lamd * l1p);
        k2d = omsd * omsd * -expm1((2. + lamd) / lamd * l1p);
    }
    const double r1 = G.p.ro01 * (1. + G.p.beta1 * (P - G.p.P_atm));
    const double r2 = G.p.ro02 * (1. + G.p.beta2 * (P - G.p.P_atm));
    M1 v;
    v.P = P;
    v.L1 = K * k1 * G.inv_mu1 * r1;
    v.L2 = K * k2d * G.inv_mu2 * r2;
    return v;
}

struct MixedLds {
    double* P;
    double* L1;
    double* L2;
    __device__ void put(int s, const M1& v) const
    {
        P[s] = v.P;
        L1[s] = v.L1;
        L2[s] = v.L2;
    }
};

// upwind flux with the face forces a2 (+ face) and a1 (- face)
__device__ __forceinline__ double uflux64(double a2, double a1, double Lc, double Lm, double Lp)
{
    return (fmax(a2, 0.) - fmin(a1, 0.)) * Lc - fmax(a1, 0.) * Lm + fmin(a2, 0.) * Lp;
}

__device__ inline void mixed_update(const FusedGrid& G, const double* med, const M1& cc,
                                    const M1& zm, const M1& zp, const MixedLds& L, int s,
                                    const RawM& q, double& nP, double& nS)
{
    const double S = q.S;
    const filt_params& p = G.p;
    const int xm = s - 1, xp = s + 1, ym = s - SX, yp = s + SX;
    const double p1 = cc.P;
    const double r1c = p.ro01 * (1. + p.beta1 * (p1 - p.P_atm));
    const double r2c = p.ro02 * (1. + p.beta2 * (p1 - p.P_atm));
    const double ihx = G.inv_h[0], ihy = G.inv_h[1], ihz = G.inv_h[2];
    const double ax2d = (p1 - L.P[xp]) * ihx, ax1d = (L.P[xm] - p1) * ihx;
    const double az2d = (p1 - zp.P) * ihz, az1d = (zm.P - p1) * ihz;
    const double dy2 = (p1 - L.P[yp]) * ihy, dy1 = (L.P[ym] - p1) * ihy;
    const double g = p.g_const;
    const double F1 = uflux64(ax2d, ax1d, cc.L1, L.L1[xm], L.L1[xp]) * ihx +
                      uflux64(dy2 + r1c * g, dy1 + r1c * g, cc.L1, L.L1[ym], L.L1[yp]) * ihy +
                      uflux64(az2d, az1d, cc.L1, zm.L1, zp.L1) * ihz;
    const double F2 = uflux64(ax2d, ax1d, cc.L2, L.L2[xm], L.L2[xp]) * ihx +
                      uflux64(dy2 + r2c * g, dy1 + r2c * g, cc.L2, L.L2[ym], L.L2[yp]) * ihy +
                      uflux64(az2d, az1d, cc.L2, zm.L2, zp.L2) * ihz;
    const int md = q.md;
    // dt/phi in FP64: an FP32 factor would bias every increment alike
    const double phi = G.poro ? q.phi : mtab(med, M_M, md);
    double rphi = __builtin_amdgcn_rcp(phi);
    rphi = fma(fma(-phi, rphi, 1.0), rphi, rphi);
    const double dtphi = p.dt * rphi;
    const double A1 = r1c * (1 - S) - dtphi * F1;
    const double A2 = r2c * S - dtphi * F2;
    // Newton of cpu.cpp:293-320 in FP64 with reciprocals
    const double ro01 = p.ro01, ro02 = p.ro02, b1 = p.beta1, b2 = p.beta2, Pa = p.P_atm;
    const double swr = mtab(med, M_SWR, md), pd = mtab(med, M_PD, md);
    const double lam = mtab(med, M_LAM, md), inv1 = mtab(med, M_INV1, md);
    const bool no_cap = (pd == 0.0) && (lam >= 0.1);
    // Without capillary pressure (P_k = 0) the two equations
    //   ro01 (1 + b1 x)(1 - S) = A1,   ro02 (1 + b2 x) S = A2,   x = P - P_atm
    // reduce to one quadratic al x^2 + be x + ga = 0.  Its root near the
    // linear solution -ga/be is -2 ga / (be + sign(be) sqrt(be^2 - 4 al ga)),
    // free of cancellation, and S = A2 / (ro02 (1 + b2 x)): the point the
    // Newton iteration converges to, for one square root and two divisions.
    // Outside that case (capillary terms, S_e < 2^-10, no finite real root)
    // the Newton iteration runs as the reference's.
    if (no_cap && p.newton_iterations > 0) {
        const double a1 = ro01 * b1, c1 = ro02 * b2;
        const double al = a1 * c1;
        const double be = ro01 * c1 + a1 * ro02 - A2 * a1 - A1 * c1;
        const double ga = ro01 * ro02 - A2 * ro01 - A1 * ro02;
        const double disc = be * be - 4.0 * al * ga;
        if (disc >= 0.0) {
            const double x = -2.0 * ga / (be + copysign(sqrt(disc), be));
            const double Sx = A2 / (ro02 + c1 * x);
            if (isfinite(x) && isfinite(Sx) && (1 - Sx - swr) * inv1 >= 0.0009765625) {
                nP = Pa + x;
                nS = Sx;
                return;
            }
        }
    }
    double P = p1, Sv = S;
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
        double r = __builtin_amdgcn_rcp(det);
        r = fma(fma(-det, r, 1.0), r, r);
        const double Pn = P - r * (F2S * F1v - F1S * F2v);
        const double Sn = Sv - r * (F1P * F2v - F2P * F1v);
        // a fixed point ends the iteration exactly (newton_cell); the 2-cycle
        // exit costs this kernel an occupancy step (measured: slower)
        if (same_bits(Pn, P) && same_bits(Sn, Sv))
            break;
        P = Pn;
        Sv = Sn;
    }
    nP = P;
    nS = Sv;
}

// -------------------------------------------------------------- writing --

// Output of interior cell (i,j,k) and the shell cells that copy it.
__device__ inline void write_cell(const FusedGrid& G, double* Po_, double* So_,
                                  const double* Pin_, int i, int j, int k, double nP, double nS)
{
    const auto Po = gmem(Po_), So = gmem(So_);
    const auto Pin = gmem(Pin_);
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const long long sy = nx, sz = (long long)nx * ny;
    const long long c = i + j * sy + k * sz;
    Po[c] = nP;
    So[c] = nS;
    if (i == 1) {
        Po[c - 1] = nP;
        So[c - 1] = nS;
    }
    const bool lastcol = (i == nx - 2);
    if (lastcol) {
        Po[c + 1] = nP;
        So[c + 1] = nS;
    }
    const bool top = (j == ny - 2), kl = (k == 1), kh = (k == nz - 2);
    if (!(top || kl || kh))
        return;
    double Pt = 0;
    if (top) {
        const double P10 = Pin[i + sy];  // (i, 1, 0)
        const double ro1 = G.p.ro01 * (1. + G.p.beta1 * (P10 - G.p.P_atm));
        Pt = nP + ro1 * G.p.g_const * G.p.h2;
    }
    for (int q = 0; q < (lastcol ? 2 : 1); q++) {
        const long long col = (i + q) + j * sy;  // column i, then N-1
        if (top) {
            Po[col + sy + k * sz] = Pt;
            So[col + sy + k * sz] = nS;
            if (kl) {
                Po[col + sy] = Pt;
                So[col + sy] = nS;
            }
            if (kh) {
                Po[col + sy + (nz - 1) * sz] = Pt;
                So[col + sy + (nz - 1) * sz] = nS;
            }
        }
        if (kl) {
            Po[col] = nP;
            So[col] = nS;
        }
        if (kh) {
            Po[col + (nz - 1) * sz] = nP;
            So[col + (nz - 1) * sz] = nS;
        }
    }
}

// Shell cells that do not depend on this step's interior values.
__device__ inline void shell_cell(const FusedGrid& G, const double* Pin_, const double* Sin_,
                                  double* Po_, double* So_, long long q)
{
    const auto Pin = gmem(Pin_), Sin = gmem(Sin_);
    const auto Po = gmem(Po_), So = gmem(So_);
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
    So[c] = (k == 0) ? Sin[ii + sz] : (in_patch(ii, k < nz - 1 ? k : nz - 2, G.p) ? G.p.S2_gr : 0.0);
}

// ------------------------------------------------------------- the tiles --
// One tile: FTX x FTY columns, planes [kb, ke) of the interior.  Registers
// hold the stage-1 values of the own column at k-1, k, k+1.  In the mixed
// form the loads are software-pipelined: while plane k is updated, the raw
// inputs (state, K, porosity) of the own column at k+2 and of the ring cell
// at k+1 are in flight, and they are evaluated at the top of the next
// iteration.  The
// one-cell ring is evaluated by one wave of the block, chosen by the block
// index.  Two LDS tiles alternate, so one barrier per plane
// orders both the writes before the reads and the reads of plane k before
// the overwrite at plane k+2.
template <bool MIXED, bool GENERAL>
__device__ void interior_tile(const FusedGrid& G, long long lb, int parity, char* lds,
                              const double* med)
{
    using V = typename std::conditional<MIXED, M1, C1>::type;
    using Raw = typename std::conditional<MIXED, RawM, RawD>::type;
    using Lds = typename std::conditional<MIXED, MixedLds, Fp64Lds>::type;
    constexpr int BUF = MIXED ? SN * 24 : 5 * SN * 8;  // bytes of one LDS tile
    const double* Pin = G.P[parity];
    const double* Sin = G.S[parity];
    double* Po = G.P[parity ^ 1];
    double* So = G.S[parity ^ 1];
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int bx = (int)(lb % G.nbx);
    const long long r = lb / G.nbx;
    const int by = (int)(r % G.nby);
    const int bz = (int)(r / G.nby);
    const int i0 = 1 + bx * FTX, j0 = 1 + by * FTY;
    const int kb = 1 + bz * G.kz, ke = min(kb + G.kz, nz - 1);
    if (kb >= ke)
        return;  // uniform over the block
    const int tx = threadIdx.x, ty = threadIdx.y, t = ty * FTX + tx;
    const int i = i0 + tx, j = j0 + ty;
    const bool in_grid = i < nx && j < ny;
    const bool inner = i <= nx - 2 && j <= ny - 2;
    const long long sy = nx, sz = (long long)nx * ny;
    // ring cell h of this thread.  Mixed form: the 64 ring cells on one wave
    // of the block, a different wave from block to block, so each SIMD
    // evaluates one ring per four blocks (spread over all waves, every wave
    // would pay a full evaluation for 16 active lanes; measured 2.5% slower).
    // The FP64 form keeps them on wave 0 (measured faster there).
    const int h = MIXED ? ((t >> 6) == (int)(blockIdx.x & 3) ? (t & 63) : RING) : t;
    static_assert(RING == 64 && FT == 256, "one ring cell per lane of one wave");
    int rx = 0, ry = 0;
    if (h < FTX) { rx = h; ry = -1; }
    else if (h < 2 * FTX) { rx = h - FTX; ry = FTY; }
    else if (h < 2 * FTX + FTY) { rx = -1; ry = h - 2 * FTX; }
    else if (h < RING) { rx = FTX; ry = h - 2 * FTX - FTY; }
    const bool ring = h < RING && i0 + rx < nx && j0 + ry < ny;
    const long long rcol = (i0 + rx) + (j0 + ry) * sy;
    const long long col = i + j * sy;
    const int s = slot(tx, ty), rs = slot(rx, ry);

    auto tile = [&](int q) {
        char* base = lds + q * BUF;
        Lds L;
        if constexpr (MIXED) {
            L.P = (double*)base;
            L.L2 = (double*)base + SN;
            L.L1 = (double*)(base + 2 * SN * sizeof(double));
        } else {
            double* d = (double*)base;
            L.P = d;
            L.X1 = d + SN;
            L.R1 = d + 2 * SN;
            L.X2 = d + 3 * SN;
            L.R2 = d + 4 * SN;
        }
        return L;
    };
    auto raw = [&](long long c, bool centre) -> Raw {
        if constexpr (MIXED)
            return mixed_raw(G, med, Pin, Sin, c, centre);
        else
            return fp64_raw(G, med, Pin, Sin, c, centre);
    };
    auto eval = [&](const Raw& q) -> V {
        if constexpr (MIXED)
            return mixed_eval<GENERAL>(G, med, q);
        else
            return fp64_eval(G, med, q);
    };

    V zm{}, zc{}, zp{};
    if constexpr (!MIXED) {
        // The FP64 form loads and evaluates each plane at the top of its
        // iteration (registers for inputs in flight would cost it an
        // occupancy step) and keeps the own column's k-1 and k in the two
        // LDS tiles rather than in registers: the tile of plane k-1 is read
        // during plane k and overwritten only after the barrier that closes
        // plane k.  The registers this frees pay for the Newton's 2-cycle
        // exit.
        auto get = [&](const Lds& T, int q) { return C1{T.P[q], T.X1[q], T.X2[q], T.R1[q], T.R2[q]}; };
        if (in_grid) {
            tile(1).put(s, eval(raw(col + (kb - 1) * sz, false)));  // read as plane kb - 1
            zp = eval(raw(col + kb * sz, false));
        }
        for (int k = kb; k < ke; k++) {
            const Lds L = tile((k - kb) & 1), Lm = tile((k - kb + 1) & 1);
            if (in_grid) {
                L.put(s, zp);
                zp = eval(raw(col + (k + 1) * sz, false));  // k + 1 <= ke <= nz - 1
            }
            if (ring)
                L.put(rs, eval(raw(rcol + k * sz, false)));
            __syncthreads();
            if (in_grid && inner) {
                const long long c = col + k * sz;
                double nP, nS;
                fp64_update(G, med, get(L, s), get(Lm, s), zp, L, s, raw(c, true), nP, nS);
                write_cell(G, Po, So, Pin, i, j, k, nP, nS);
            }
            __syncthreads();
        }
        return;
    }
    Raw qc{}, qp{}, rq{};  // own column at k and (in flight) k+1; ring cell at k
    if (in_grid) {
        zm = eval(raw(col + (kb - 1) * sz, false));
        qc = raw(col + kb * sz, true);
        zc = eval(qc);
        qp = raw(col + (kb + 1) * sz, true);  // kb + 1 <= ke <= nz - 1
    }
    if (ring)
        rq = raw(rcol + kb * sz, false);
    for (int k = kb; k < ke; k++) {
        const Lds L = tile((k - kb) & 1);
        Raw qn = qp;
        if (in_grid) {
            zp = eval(qp);
            L.put(s, zc);
        }
        if (ring)
            L.put(rs, eval(rq));
        if (k + 1 < ke) {  // the next plane is updated: fetch its inputs
            if (in_grid)
                qp = raw(col + (k + 2) * sz, true);  // k + 2 <= ke <= nz - 1
            if (ring)
                rq = raw(rcol + (k + 1) * sz, false);
        }
        __syncthreads();
        if (in_grid && inner) {
            double nP, nS;
            if constexpr (MIXED)
                mixed_update(G, med, zc, zm, zp, L, s, qc, nP, nS);
            write_cell(G, Po, So, Pin, i, j, k, nP, nS);
        }
        zm = zc;
        zc = zp;
        qc = qn;
    }
}

constexpr size_t LDS_BYTES = 2 * 5 * SN * sizeof(double);  // two tiles of the FP64 layout

template <bool MIXED, bool GENERAL>
__global__ __launch_bounds__(FT) void k_fused(const FusedGrid* __restrict__ grids,
                                              const int* __restrict__ block_grid,
                                              long long total_blocks, int parity, int order)
{
    __shared__ __attribute__((aligned(16))) char lds[MIXED ? 2 * SN * 24 : LDS_BYTES];
    __shared__ double med[M_ROWS * FILT_MAX_MEDIA];
    long long b = blockIdx.x;
    if (order == 0 && total_blocks % 8 == 0)
        b = (b % 8) * (total_blocks / 8) + b / 8;
    const FusedGrid& G = grids[block_grid[b]];
    const long long lb = b - G.first_block;
    if (lb < G.n_interior_blocks) {  // uniform over the block
        const int t = threadIdx.y * FTX + threadIdx.x;
        if (t < M_ROWS * FILT_MAX_MEDIA) {
            const int row = t / FILT_MAX_MEDIA, q = t % FILT_MAX_MEDIA;
            const filt_params& p = G.p;
            med[t] = row == M_LAM ? p.lambda[q] : row == M_SWR ? p.S_wr[q] : row == M_PD ? p.P_d[q]
                   : row == M_M ? p.m[q] : row == M_K ? p.K[q] : G.inv_1mswr[q];
        }
        __syncthreads();
        interior_tile<MIXED, GENERAL>(G, lb, parity, lds, med);
        return;
    }
    const long long q = (lb - G.n_interior_blocks) * FT + threadIdx.y * FTX + threadIdx.x;
    shell_cell(G, G.P[parity], G.S[parity], G.P[parity ^ 1], G.S[parity ^ 1], q);
}

// ------------------------------------------------------------------ mass --
// Total water / NAPL mass over interior cells, deterministic: each block sums
// a fixed range of 4096 cells (fixed tree), a second pass sums the block
// partials in order.  Independent of any tiling.
constexpr int MASS_CHUNK = 4096;

__global__ __launch_bounds__(256) void k_mass_partial(const FusedGrid* __restrict__ Gp,
                                                      int parity, double* __restrict__ partial)
{
    const FusedGrid& G = *Gp;
    __shared__ double s1[256], s2[256];
    const long long n = (long long)G.nx * G.ny * G.nz;
    const long long base = (long long)blockIdx.x * MASS_CHUNK;
    const double V = G.p.h1 * G.p.h2 * G.p.h3;
    double a1 = 0, a2 = 0;
    for (int u = 0; u < MASS_CHUNK / 256; u++) {
        const long long c = base + u * 256 + threadIdx.x;
        if (c >= n)
            break;
        const int i = (int)(c % G.nx);
        const long long r = c / G.nx;
        const int j = (int)(r % G.ny), k = (int)(r / G.ny);
        if (!is_interior(i, j, k, G.nx, G.ny, G.nz))
            continue;
        const int md = media_at(G, c);
        const double phi = G.poro ? G.poro[c] : G.p.m[md];
        const double P = G.P[parity][c], S = G.S[parity][c];
        const double r1 = G.p.ro01 * (1. + G.p.beta1 * (P - G.p.P_atm));
        const double r2 = G.p.ro02 * (1. + G.p.beta2 * (P - G.p.P_atm));
        a1 += phi * r1 * (1. - S) * V;
        a2 += phi * r2 * S * V;
    }
    s1[threadIdx.x] = a1;
    s2[threadIdx.x] = a2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s1[0];
        partial[2 * blockIdx.x + 1] = s2[0];
    }
}

__global__ void k_mass_final(const double* __restrict__ partial, int nblocks,
                             double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double a = 0, b = 0;
    for (int q = 0; q < nblocks; q++) {
        a += partial[2 * q];
        b += partial[2 * q + 1];
    }
    out[0] = a;
    out[1] = b;
}

}  // namespace

// ------------------------------------------------------------ host side --

void fused_prepare(FusedGrid& G)
{
    const filt_params& p = G.p;
    G.inv_h[0] = 1.0 / p.h1;
    G.inv_h[1] = 1.0 / p.h2;
    G.inv_h[2] = 1.0 / p.h3;
    G.inv_mu1 = 1.0 / p.mu1;
    G.inv_mu2 = 1.0 / p.mu2;
    for (int q = 0; q < FILT_MAX_MEDIA; q++)
        G.inv_1mswr[q] = 1.0 / (1.0 - p.S_wr[q]);
    if (G.kz < 1)
        G.kz = 32;
    G.nbx = (G.nx - 2 + FTX - 1) / FTX;
    G.nby = (G.ny - 2 + FTY - 1) / FTY;
    G.nbz = (G.nz - 2 + G.kz - 1) / G.kz;
    G.n_interior_blocks = (long long)G.nbx * G.nby * G.nbz;
    const long long shell = 2LL * G.ny + 2LL * (G.nz - 2) + (long long)(G.nx - 1) * G.nz;
    G.n_shell_blocks = (shell + FT - 1) / FT;
}

bool fused_general_lambda(const FusedGrid& g)
{
    const int n = (g.media && g.p.n_media > 1) ? g.p.n_media : 1;
    for (int q = 0; q < n; q++)
        if (g.p.lambda[q] != 2.0)
            return true;
    return false;
}

int fused_launch(const FusedGrid* dev_grids, const int* dev_block_grid, long long total_blocks,
                 int parity, int precision, bool general_lambda, int order, hipStream_t s)
{
    if (total_blocks <= 0)
        return FILT_OK;
    if (total_blocks > 0x7fffffffLL)
        return set_error(FILT_ERR_INVALID, "batch too large for one launch");
    const dim3 grid((unsigned)total_blocks), block(FTX, FTY);
    if (precision == FILT_PREC_MIXED && general_lambda)
        hipLaunchKernelGGL((k_fused<true, true>), grid, block, 0, s, dev_grids, dev_block_grid,
                           total_blocks, parity, order);
    else if (precision == FILT_PREC_MIXED)
        hipLaunchKernelGGL((k_fused<true, false>), grid, block, 0, s, dev_grids, dev_block_grid,
                           total_blocks, parity, order);
    else
        hipLaunchKernelGGL((k_fused<false, false>), grid, block, 0, s, dev_grids, dev_block_grid,
                           total_blocks, parity, order);
    FILT_LAUNCH_CHECK("k_fused");
    return FILT_OK;
}

int fused_mass(const FusedGrid* dev_grid, const FusedGrid& G, int parity, double* dev_partial,
               double* dev_out, hipStream_t s)
{
    const long long n = (long long)G.nx * G.ny * G.nz;
    const long long nb = (n + MASS_CHUNK - 1) / MASS_CHUNK;
    hipLaunchKernelGGL(k_mass_partial, dim3((unsigned)nb), dim3(256), 0, s, dev_grid, parity,
                       dev_partial);
    FILT_LAUNCH_CHECK("k_mass_partial");
    hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(64), 0, s, dev_partial, (int)nb, dev_out);
    FILT_LAUNCH_CHECK("k_mass_final");
    return FILT_OK;
}

long long fused_mass_blocks(const FusedGrid& G)
{
    const long long n = (long long)G.nx * G.ny * G.nz;
    return (n + MASS_CHUNK - 1) / MASS_CHUNK;
}

}  // namespace filt

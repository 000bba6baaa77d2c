// filt_step.hip -- stages 2-4 of time_step_function (main.cpp:83-94) in one
// z-march: u_calculation (cpu.cpp:77-160), roS_calculation (cpu.cpp:218-291)
// and P1_S2_calculation (cpu.cpp:293-320), for one rank whose grid has every
// extent >= 3 and arrays below 4 GiB, after stage 1 has run on the state.
//
// Stage 1 writes P2 = P1 + P_k with P_k = 0 (cpu.cpp:63), so P2 holds P1's
// bits (for every P1 but -0.0, which no pressure takes).  The kernel reads
// P2 wherever the stages read P1 or P2; the Newton then rewrites P1 and S2
// of its own cell in place without any neighbour needing their old values
// (the stencils read P2, Xi and ro; S2 only at the centre).  ro1 and ro2
// are not read either: stage 1 computed them from P1 and P2 alone, and the
// same expressions on P2 give the same bits (ro1_of, ro2_of).  Every array
// is written with the values the three stages write one after another, so
// the state after the step is the same bit for bit, while HBM sees 136 B per
// cell instead of 240: the stencil operands are read once for u and roS, the
// densities not at all, roS is not read back by the Newton, and P1, S2 are
// not read twice.
//
// Structure as in filt_stencil.hip: TX x TY tiles of columns march over a
// chunk of planes; per plane the five stencil operands (P2, Xi1, Xi2, ro1,
// ro2) of the tile and its one-cell ring go to one of two LDS tiles, the ring
// and the point operands of plane k+1 and the own column of plane k+2 are in
// flight while plane k is computed, and one barrier orders each plane.  The
// own column's k and k+1 stay in registers (P2, Xi1, Xi2; the densities
// are evaluated from P2 as the plane is staged); its k-1 is read back from this
// thread's own slot of the other LDS tile, which no other thread writes.
// Array pointers and parameters are read from the kernel argument segment
// where they are used (filt_kernarg.hpp); the media table (lambda, S_wr,
// P_d, m) is copied to LDS, so its per-cell lookups are LDS reads rather
// than per-lane loads from the argument segment, whose waits would also wait
// for the loads in flight for the next plane.  The 15 divisions by h1, h2, h3
// (or their doubles) per cell go through DivBy (filt_device.hpp): three FP64
// operations each, bit for bit the IEEE division.  The kernel fits 97 VGPRs
// and runs 4 waves per SIMD.  u is computed on every plane of the grid; roS
// and the Newton on the interior.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "../../include/filtration_hip.h"
#include "filt_device.hpp"
#include "filt_kernarg.hpp"
#include "filt_runtime.hpp"
#include "filt_tiles.hpp"
#include "filt_zmarch.hpp"

namespace filt {

namespace {

using namespace zm;

constexpr int NA = 5;  // stencil operands in LDS: P2, Xi1, Xi2, ro1, ro2
constexpr int NL = 3;  // of which loaded: P2, Xi1, Xi2 (ro1, ro2 follow from P2)
enum { AP = 0, AX1 = 1, AX2 = 2, AR1 = 3, AR2 = 4 };

// The kernel's one argument: at offset 0 of the argument segment, so that
// the loop reads its members through kernarg<StepArgs>(0).
struct StepArgs {
    filt_fields f;
    filt_params p;
    int nx, kz;
    TileMap map;
};

__device__ __forceinline__ KArg<filt_fields>* fields_arg() { return &kernarg<StepArgs>(0)->f; }
__device__ __forceinline__ KArg<filt_params>* params_arg() { return &kernarg<StepArgs>(0)->p; }

// The boundary cells of the sweep's closed form (cpu.cpp:44-54, 323-403;
// k_bc_b in filt_stages.hip) that follow from interior cell (i, j, k)'s new
// P, S or from constants of its column, written by the cell's own thread:
//   (0, j, k) if i == 1;  (N-1, j, k) if i == N-2;
//   for column c in {i} and, if i == N-2, also {N-1}:
//     (c, M-1, k) = (S, P + Pt) if j == M-2, Pt = ro1(i,1,0)*g*h2, with
//       k = 0 / L-1 as well when k == 1 / L-2;
//     (c, j, 0) if k == 1, (c, j, L-1) if k == L-2;
//     (c, 0, k): P_atm and S = src(i, k) if j == 1, with (c, 0, 0) =
//       (P_atm, pre-sweep S(i, 0, 1)) when k == 1 and (c, 0, L-1) =
//       (P_atm, src(i, L-2)) when k == L-2.
// No other thread of the launch reads these cells (the stencils read P2, Xi
// and ro, and P1, S2 only at interior centres); the pre-sweep S(i, 0, 1) is
// read by the thread that then overwrites it.
template <class F, class Pp>
__device__ inline void boundary_from_cell(F* Fp, const Pp& p, unsigned c, unsigned sy,
                                          unsigned sz, int i, int j, int k, int nx, int ny,
                                          int nz, double P, double S, double Pt)
{
    double* const P1 = Fp->P1;
    double* const S2 = Fp->S2;
    if (i == 1) {
        stn(P1, c - 8, P);
        stn(S2, c - 8, S);
    }
    const bool last = i == nx - 2;
    if (last) {
        stn(P1, c + 8, P);
        stn(S2, c + 8, S);
    }
    const bool kl = k == 1, kh = k == nz - 2;
    double S01 = 0.0;
    if (j == 1) {
        if (kl)
            S01 = ld(S2, c - sy);  // pre-sweep S(i, 0, 1), before this thread rewrites it
    }
    const double Ssrc = j == 1 ? (in_patch(i, k, p) ? p.S2_gr : 0.0) : 0.0;
    for (int q = 0; q < (last ? 2 : 1); q++) {
        const unsigned d = 8u * q;  // column i, then N-1
        if (j == ny - 2) {
            const double Pb = P + Pt;
            stn(P1, c + sy + d, Pb);
            stn(S2, c + sy + d, S);
            if (kl) {
                stn(P1, c + sy - sz + d, Pb);
                stn(S2, c + sy - sz + d, S);
            }
            if (kh) {
                stn(P1, c + sy + sz + d, Pb);
                stn(S2, c + sy + sz + d, S);
            }
        }
        if (kl) {
            stn(P1, c - sz + d, P);
            stn(S2, c - sz + d, S);
        }
        if (kh) {
            stn(P1, c + sz + d, P);
            stn(S2, c + sz + d, S);
        }
        if (j == 1) {
            stn(P1, c - sy + d, p.P_atm);
            stn(S2, c - sy + d, Ssrc);
            if (kl) {
                stn(P1, c - sy - sz + d, p.P_atm);
                stn(S2, c - sy - sz + d, S01);
            }
            if (kh) {
                stn(P1, c - sy + sz + d, p.P_atm);
                stn(S2, c - sy + sz + d, Ssrc);
            }
        }
    }
}

template <int TX, int TY, bool WITH_U, bool BC>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(4))) void k_step_zm(
    StepArgs args)
{
    const filt_fields& f = args.f;
    const filt_params& p = args.p;
    const int nx = args.nx, kz = args.kz;
    const TileMap& map = args.map;
    using L = Lds<TX, TY>;
    __shared__ double shb[2][NA * L::SN];
    int bx, by, bz;
    map.decode(blockIdx.x, bx, by, bz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = bx * TX, j0 = by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int ny = p.Ny, nz = p.Nz;
    const bool valid = i < nx && j < ny;
    const bool inner = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
    const unsigned sy = 8u * nx, sz = 8u * nx * ny;  // byte strides
    // u on every plane; roS and the Newton on planes 1 .. nz-2
    const int kb = max(bz * kz, WITH_U ? 0 : 1);
    const int ke = min(bz * kz + kz, WITH_U ? nz : nz - 1);
    if (kb >= ke)
        return;  // uniform over the block
    const double* const A[NL] = {f.P2, f.Xi1, f.Xi2};
    const unsigned col = 8u * i + sy * j;
    int hslot;
    unsigned hcol;
    const bool hv = ring_slot<TX, TY>(tx, ty, i0, j0, nx, ny, sy, hslot, hcol);
    const bool per_cell_media = f.media && p.n_media > 1;
    // Own column at k-1, k, k+1 and (in flight) k+2; the ring cell and the
    // point operands (S2, porosity, medium) of the next plane.
    double zc[NL] = {}, zp[NL] = {}, zq[NL] = {}, hn[NL] = {};
    double Sc = 0, Sn = 0, phic = 0, phin = 0;
    int mdc = 0, mdn = 0;
    if (valid) {
        const unsigned c = col + kb * sz;
#pragma unroll
        for (int a = 0; a < NL; a++) {
            // plane kb-1 goes to the own slot of the tile that plane kb
            // reads as its predecessor
            if (kb >= 1)
                shb[1][L::at(a, ty + 1, tx + 1)] = ld(A[a], c - sz);
            zc[a] = ld(A[a], c);
            if (kb + 1 < nz)
                zp[a] = ld(A[a], c + sz);
        }
        if (kb >= 1) {
            const double Pm = shb[1][L::at(AP, ty + 1, tx + 1)];
            shb[1][L::at(AR1, ty + 1, tx + 1)] = ro1_of(Pm, p);
            shb[1][L::at(AR2, ty + 1, tx + 1)] = ro2_of(Pm, p);
        }
        Sc = ld(f.S2, c);
        if (f.poro)
            phic = ld(f.poro, c);
        if (per_cell_media)
            mdc = ldi(f.media, c >> 1);
    }
    if (hv)
#pragma unroll
        for (int a = 0; a < NL; a++) hn[a] = ld(A[a], hcol + kb * sz);
    const int cx = tx + 1, cy = ty + 1;
    const double g = p.g_const;
    // the bottom boundary's pressure offset ro1(i,1,0)*g*h2 (this step's
    // stage 1), for the cells of row M-2
    double Pt = 0.0;
    if (BC && valid && inner && j == ny - 2)
        Pt = ld(f.ro1, 8u * (unsigned)(i + nx)) * g * p.h2;
    // reciprocals of h1, h2, h3, 2h1, 2h2, 2h3 for DivBy; read after the
    // first barrier of the loop
    __shared__ double rh[6];
    // the media table (lambda, S_wr, P_d, m) for per-cell lookups: from LDS,
    // not from per-lane loads of the argument segment
    __shared__ double med[4][FILT_MAX_MEDIA];
    {
        const int t = tx + TX * ty;
        if (t < 4 * FILT_MAX_MEDIA) {
            const int q = t % FILT_MAX_MEDIA, w = t / FILT_MAX_MEDIA;
            med[w][q] = w == 0 ? p.lambda[q] : w == 1 ? p.S_wr[q] : w == 2 ? p.P_d[q] : p.m[q];
        }
    }
    if (tx == 0 && ty == 0) {
        rh[0] = refined_reciprocal(p.h1);
        rh[1] = refined_reciprocal(p.h2);
        rh[2] = refined_reciprocal(p.h3);
        rh[3] = refined_reciprocal(2. * p.h1);
        rh[4] = refined_reciprocal(2. * p.h2);
        rh[5] = refined_reciprocal(2. * p.h3);
    }
    for (int k = kb; k < ke; k++) {
        double* const sh = shb[(k - kb) & 1];
        const unsigned c = col + k * sz;
        // densities of the own cell and of the ring cell, from P2
        double r1c, r2c;
        {
            KArg<filt_params>* Q = params_arg();
            r1c = ro1_of(zc[AP], *Q);
            r2c = ro2_of(zc[AP], *Q);
            if (valid) {
#pragma unroll
                for (int a = 0; a < NL; a++) sh[L::at(a, cy, cx)] = zc[a];
                sh[L::at(AR1, cy, cx)] = r1c;
                sh[L::at(AR2, cy, cx)] = r2c;
            }
            if (hv) {
#pragma unroll
                for (int a = 0; a < NL; a++) sh[hslot + a * L::SN] = hn[a];
                sh[hslot + AR1 * L::SN] = ro1_of(hn[AP], *Q);
                sh[hslot + AR2 * L::SN] = ro2_of(hn[AP], *Q);
            }
        }
        if (k + 1 < ke) {  // the next plane is computed: fetch its operands
            KArg<filt_fields>* F = fields_arg();
            const double* const B[NL] = {F->P2, F->Xi1, F->Xi2};
            if (valid) {
                if (k + 2 < nz)
#pragma unroll
                    for (int a = 0; a < NL; a++) zq[a] = ld(B[a], c + 2 * sz);
                Sn = ld(F->S2, c + sz);
                if (F->poro)
                    phin = ld(F->poro, c + sz);
                if (per_cell_media)
                    mdn = ldi(F->media, (c + sz) >> 1);
            }
            if (hv)
#pragma unroll
                for (int a = 0; a < NL; a++) hn[a] = ld(B[a], hcol + (k + 1) * sz);
        }
        __syncthreads();
        const double* const shm = shb[(k - kb + 1) & 1];  // own slot: plane k-1
        const double Pc = zc[AP];
        if (WITH_U && valid) {
            // u_calculation, cpu.cpp:77-160 (P1 = P2)
            KArg<filt_fields>* F = fields_arg();
            KArg<filt_params>* Q = params_arg();
            // division by h and 2h, bit for bit as n / h (filt_device.hpp)
            const DivBy dx{Q->h1, rh[0]}, dy{Q->h2, rh[1]}, dz{Q->h3, rh[2]};
            const DivBy dx2{2. * Q->h1, rh[3]}, dy2{2. * Q->h2, rh[4]}, dz2{2. * Q->h3, rh[5]};
            const double X1 = zc[AX1], X2 = zc[AX2];
            double u1, u2;
            if (i == 0) {
                u1 = dx(X1 * (sh[L::at(AP, cy, cx + 1)] - Pc));
                u2 = dx(X2 * (sh[L::at(AP, cy, cx + 1)] - Pc));
            } else if (i == nx - 1) {
                u1 = dx(X1 * (Pc - sh[L::at(AP, cy, cx - 1)]));
                u2 = dx(X2 * (Pc - sh[L::at(AP, cy, cx - 1)]));
            } else {
                const double gx = dx2(sh[L::at(AP, cy, cx + 1)] - sh[L::at(AP, cy, cx - 1)]);
                u1 = X1 * gx;
                u2 = X2 * gx;
            }
            stn(F->u1x, c, u1);
            stn(F->u2x, c, u2);
            const double r1 = r1c, r2 = r2c;
            if (j == 0) {
                const double gy = dy(sh[L::at(AP, cy + 1, cx)] - Pc);
                u1 = X1 * (gy - r1 * g);
                u2 = X2 * (gy - r2 * g);
            } else if (j == ny - 1) {
                const double gy = dy(Pc - sh[L::at(AP, cy - 1, cx)]);
                u1 = X1 * (gy - r1 * g);
                u2 = X2 * (gy - r2 * g);
            } else {
                const double gy = dy2(sh[L::at(AP, cy + 1, cx)] - sh[L::at(AP, cy - 1, cx)]);
                u1 = X1 * (gy - r1 * g);
                u2 = X2 * (gy - r2 * g);
            }
            stn(F->u1y, c, u1);
            stn(F->u2y, c, u2);
            if (k == 0) {
                u1 = dz(X1 * (zp[AP] - Pc));
                u2 = dz(X2 * (zp[AP] - Pc));
            } else if (k == nz - 1) {
                u1 = dz(X1 * (Pc - shm[L::at(AP, cy, cx)]));
                u2 = dz(X2 * (Pc - shm[L::at(AP, cy, cx)]));
            } else {
                const double gz = dz2(zp[AP] - shm[L::at(AP, cy, cx)]);
                u1 = X1 * gz;
                u2 = X2 * gz;
            }
            stn(F->u1z, c, u1);
            stn(F->u2z, c, u2);
        }
        if (valid && inner && k >= 1 && k <= nz - 2) {
            // roS_calculation, cpu.cpp:218-291 (P1 = P2)
            KArg<filt_fields>* F = fields_arg();
            KArg<filt_params>* Q = params_arg();
            const DivBy dx{Q->h1, rh[0]}, dy{Q->h2, rh[1]}, dz{Q->h3, rh[2]};
            const double X1c = zc[AX1], X2c = zc[AX2];
            const double S = Sc;
            const double roS1 = r1c * (1 - S);
            const double roS2 = r2c * S;
            double a1, a2;
            a2 = dz(-(zp[AP] - Pc));
            a1 = dz(-(Pc - shm[L::at(AP, cy, cx)]));
            const double r1p = ro1_of(zp[AP], *Q), r2p = ro2_of(zp[AP], *Q);
            const double f3 = upwind_div(a2, a1, X1c, r1c, shm[L::at(AX1, cy, cx)],
                                         shm[L::at(AR1, cy, cx)], zp[AX1], r1p, dz);
            const double g3 = upwind_div(a2, a1, X2c, r2c, shm[L::at(AX2, cy, cx)],
                                         shm[L::at(AR2, cy, cx)], zp[AX2], r2p, dz);
            const int xp = cx + 1, xm = cx - 1, yp = cy + 1, ym = cy - 1;
            a2 = dx(-(sh[L::at(AP, cy, xp)] - Pc));
            a1 = dx(-(Pc - sh[L::at(AP, cy, xm)]));
            const double f1 = upwind_div(a2, a1, X1c, r1c, sh[L::at(AX1, cy, xm)],
                                         sh[L::at(AR1, cy, xm)], sh[L::at(AX1, cy, xp)],
                                         sh[L::at(AR1, cy, xp)], dx);
            const double g1 = upwind_div(a2, a1, X2c, r2c, sh[L::at(AX2, cy, xm)],
                                         sh[L::at(AR2, cy, xm)], sh[L::at(AX2, cy, xp)],
                                         sh[L::at(AR2, cy, xp)], dx);
            const double dy2 = dy(-(sh[L::at(AP, yp, cx)] - Pc));
            const double dy1 = dy(-(Pc - sh[L::at(AP, ym, cx)]));
            const double f2 = upwind_div(dy2 + r1c * g, dy1 + r1c * g, X1c, r1c,
                                         sh[L::at(AX1, ym, cx)], sh[L::at(AR1, ym, cx)],
                                         sh[L::at(AX1, yp, cx)], sh[L::at(AR1, yp, cx)], dy);
            const double g2 = upwind_div(dy2 + r2c * g, dy1 + r2c * g, X2c, r2c,
                                         sh[L::at(AX2, ym, cx)], sh[L::at(AR2, ym, cx)],
                                         sh[L::at(AX2, yp, cx)], sh[L::at(AR2, yp, cx)], dy);
            const int md = mdc;
            const double phi = F->poro ? phic : med[3][md];
            const double A1 = roS1 - (Q->dt / phi) * (f1 + f2 + f3);
            const double A2 = roS2 - (Q->dt / phi) * (g1 + g2 + g3);
            stn(F->roS1_old, c, roS1);
            stn(F->roS2_old, c, roS2);
            stn(F->roS1, c, A1);
            stn(F->roS2, c, A2);
            // P1_S2_calculation, cpu.cpp:293-320, on this cell's registers
            double P = Pc, Sv = S;
            newton_cell(P, Sv, A1, A2, med[0][md], med[1][md], med[2][md], *Q);
            stn(F->P1, c, P);
            stn(F->S2, c, Sv);
            if constexpr (BC)
                boundary_from_cell(F, *Q, c, sy, sz, i, j, k, nx, ny, nz, P, Sv, Pt);
        }
#pragma unroll
        for (int a = 0; a < NL; a++) {
            zc[a] = zp[a];
            zp[a] = zq[a];
        }
        Sc = Sn;
        phic = phin;
        mdc = mdn;
    }
}

// ---------------------------------------------------------------------------
// The whole time step but the boundary in one z-march (filt_time_step):
// stage 1 (ro_P2_Xi_calculation, cpu.cpp:3-15, 57-74) as well as stages 2-4.
// Each block reads P1, S2 (and K, porosity, media) of its columns and of the
// ring around its tile, evaluates stage 1 for all of them in registers and
// LDS, writes the stage-1 arrays of its own cells, and goes on as k_step_zm.
// Per cell it moves 32 B in and 136 B out instead of the 200 B of stage 1
// and k_step_zm apart.
//
// The new P1, S2 of a cell cannot always be written in place: other blocks
// read the old P1, S2 of the cells in their ring (the edge columns of the
// neighbouring tiles) and of the plane above and below their z-chunk (the
// first and last planes of the neighbouring chunks), and nothing orders
// those reads before this block's writes.  The Newton results of exactly
// those cells go to a stash (two arrays indexed like the grid); every other
// cell is written in place.  k_unstash then copies the stash into P1, S2 in
// the next launch, before the boundary kernels read them.  The stash carries
// the share of cells that are tile edges or chunk-boundary planes (about a
// quarter to a third of them at the usual tiles), 32 B per such cell.
struct FullArgs {
    StepArgs s;   // at offset 0: the kernarg helpers of k_step_zm apply
    double* stP;  // stash of the new P1 of the cells other blocks read
    double* stS;  // stash of the new S2
};

// Stage 1 (cpu.cpp:57-74) of one cell of a medium with lambda = 2, whose
// Brooks-Corey exponents (2+3*lambda)/lambda and (2+lambda)/lambda are 4 and
// 2 exactly: stage1_cell's expressions with the divisions by 1 - S_wr, mu1
// and mu2 through DivBy, bit for bit.  ro1, ro2 follow from P2 (ro1_of,
// ro2_of).
__device__ __forceinline__ void stage1_l2(double P1, double S2, double K, double swr,
                                          double dy_swr, const DivBy& dmu1, const DivBy& dmu2,
                                          double& P2, double& X1, double& X2)
{
    const DivBy d1{1. - swr, dy_swr};
    const double S_e = d1(1. - S2 - swr);
    const double k1 = pow_spec(S_e, 4.0);
    const double k2 = (1. - S_e) * (1. - S_e) * (1 - S_e * S_e);
    const double P_k = 0;  // capillary pressure disabled, cpu.cpp:63
    P2 = P1 + P_k;
    X1 = dmu1(-1. * K * k1);
    X2 = dmu2(-1. * K * k2);
}

// MEDIA: a per-cell media array is read (several media); without it every
// cell is medium 0 and the kernel keeps no medium indices in registers.
template <int TX, int TY, bool WITH_U, bool MEDIA>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(4))) void k_step_full(
    FullArgs args)
{
    const filt_fields& f = args.s.f;
    const filt_params& p = args.s.p;
    const int nx = args.s.nx, kz = args.s.kz;
    const TileMap& map = args.s.map;
    using L = Lds<TX, TY>;
    __shared__ double shb[2][NA * L::SN];
    int bx, by, bz;
    map.decode(blockIdx.x, bx, by, bz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = bx * TX, j0 = by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int ny = p.Ny, nz = p.Nz;
    const bool valid = i < nx && j < ny;
    const bool inner = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
    const unsigned sy = 8u * nx, sz = 8u * nx * ny;  // byte strides
    // stage 1 (and u) on every plane; roS and the Newton on planes 1 .. nz-2
    const int kb = bz * kz;
    const int ke = min(kb + kz, nz);
    if (kb >= ke)
        return;  // uniform over the block
    const unsigned col = 8u * i + sy * j;
    int hslot;
    unsigned hcol;
    const bool hv = ring_slot<TX, TY>(tx, ty, i0, j0, nx, ny, sy, hslot, hcol);
    // the tile's edge columns are other tiles' ring
    const bool edge = tx == 0 || tx == TX - 1 || ty == 0 || ty == TY - 1;
    // the media table (lambda, S_wr, P_d, m, K, refined 1/(1 - S_wr)) and the
    // refined reciprocals of h1, h2, h3, their doubles, mu1 and mu2
    __shared__ double med[6][FILT_MAX_MEDIA];
    __shared__ double rh[8];
    {
        const int t = tx + TX * ty;
        if (t < 6 * FILT_MAX_MEDIA) {
            const int q = t % FILT_MAX_MEDIA, w = t / FILT_MAX_MEDIA;
            med[w][q] = w == 0 ? p.lambda[q] : w == 1 ? p.S_wr[q] : w == 2 ? p.P_d[q]
                      : w == 3 ? p.m[q] : w == 4 ? p.K[q] : refined_reciprocal(1. - p.S_wr[q]);
        } else if (t >= 64 && t < 72) {
            const int q = t - 64;
            const double d = q == 0 ? p.h1 : q == 1 ? p.h2 : q == 2 ? p.h3 : q == 3 ? 2. * p.h1
                           : q == 4 ? 2. * p.h2 : q == 5 ? 2. * p.h3 : q == 6 ? p.mu1 : p.mu2;
            rh[q] = refined_reciprocal(d);
        }
    }
    __syncthreads();
    // Own column: stage-1 values at k and k+1; raw inputs of k+1 (read as
    // the next centre) and, in flight, of k+2; the ring cell's raw inputs of
    // the next plane.
    double zc[NL] = {}, zp[NL] = {};
    double Sc = 0, Sn = 0;
    int mdc = 0, mdn = 0;
    double rP = 0, rS = 0, rK = 0;  // raw of plane k+1 (then k+2)
    int rmd = 0;
    double hP = 0, hS = 0, hK = 0;  // ring raw of plane k (then k+1)
    int hmd = 0;
    // F and Q: the kernel argument segment, read at uniform points of the loop
    auto raw = [&](KArg<filt_fields>* F, unsigned c, double& P, double& S, double& K, int& md) {
        md = MEDIA ? ldi(F->media, c >> 1) : 0;
        P = ld(F->P1, c);
        S = ld(F->S2, c);
        K = F->perm ? ld(F->perm, c) : 0.0;
    };
    auto eval = [&](KArg<filt_fields>* F, KArg<filt_params>* Q, double P1v, double S2v, double Kv,
                    int md, double* z) {
        const DivBy dmu1{Q->mu1, rh[6]}, dmu2{Q->mu2, rh[7]};
        const double K = F->perm ? Kv : med[4][md];
        stage1_l2(P1v, S2v, K, med[1][md], med[5][md], dmu1, dmu2, z[AP], z[AX1], z[AX2]);
    };
    auto poro_at = [&](KArg<filt_fields>* F, unsigned c) {
        return F->poro ? ld(F->poro, c) : 0.0;
    };
    const int cx = tx + 1, cy = ty + 1;
    {
        KArg<filt_fields>* F = fields_arg();
        KArg<filt_params>* Q = params_arg();
        if (valid) {
            const unsigned c = col + kb * sz;
            if (kb >= 1) {
                // plane kb-1 goes to the own slot of the tile that plane kb
                // reads as its predecessor
                double P, S, K, zm[NL];
                int md;
                raw(F, c - sz, P, S, K, md);
                eval(F, Q, P, S, K, md, zm);
#pragma unroll
                for (int a = 0; a < NL; a++) shb[1][L::at(a, cy, cx)] = zm[a];
                shb[1][L::at(AR1, cy, cx)] = ro1_of(zm[AP], *Q);
                shb[1][L::at(AR2, cy, cx)] = ro2_of(zm[AP], *Q);
            }
            double K;
            raw(F, c, rP, Sc, K, mdc);
            eval(F, Q, rP, Sc, K, mdc, zc);
            if (kb + 1 < nz)
                raw(F, c + sz, rP, rS, rK, rmd);
        }
        if (hv)
            raw(F, hcol + kb * sz, hP, hS, hK, hmd);
    }
    const double g = p.g_const;
    for (int k = kb; k < ke; k++) {
        double* const sh = shb[(k - kb) & 1];
        const unsigned c = col + k * sz;
        double r1c, r2c, phic = 0;
        {
            KArg<filt_params>* Q = params_arg();
            KArg<filt_fields>* F = fields_arg();
            // the centre's porosity: only the update below reads it
            if (valid && inner && k >= 1 && k <= nz - 2)
                phic = poro_at(F, c);
            r1c = ro1_of(zc[AP], *Q);
            r2c = ro2_of(zc[AP], *Q);
            if (valid) {
                if (k + 1 < nz) {
                    eval(F, Q, rP, rS, rK, rmd, zp);
                    Sn = rS;
                    mdn = rmd;
                }
#pragma unroll
                for (int a = 0; a < NL; a++) sh[L::at(a, cy, cx)] = zc[a];
                sh[L::at(AR1, cy, cx)] = r1c;
                sh[L::at(AR2, cy, cx)] = r2c;
                // stage 1's outputs of this cell
                stn(F->P2, c, zc[AP]);
                stn(F->Xi1, c, zc[AX1]);
                stn(F->Xi2, c, zc[AX2]);
                stn(F->ro1, c, r1c);
                stn(F->ro2, c, r2c);
            }
            if (hv) {
                double hz[NL];
                eval(F, Q, hP, hS, hK, hmd, hz);
#pragma unroll
                for (int a = 0; a < NL; a++) sh[hslot + a * L::SN] = hz[a];
                sh[hslot + AR1 * L::SN] = ro1_of(hz[AP], *Q);
                sh[hslot + AR2 * L::SN] = ro2_of(hz[AP], *Q);
            }
        }
        if (k + 1 < ke) {  // the next plane is computed: fetch its inputs
            KArg<filt_fields>* F = fields_arg();
            if (valid && k + 2 < nz)
                raw(F, c + 2 * sz, rP, rS, rK, rmd);
            if (hv)
                raw(F, hcol + (k + 1) * sz, hP, hS, hK, hmd);
        }
        __syncthreads();
        const double* const shm = shb[(k - kb + 1) & 1];  // own slot: plane k-1
        const double Pc = zc[AP];
        if (WITH_U && valid) {
            // u_calculation, cpu.cpp:77-160 (P1 = P2)
            KArg<filt_fields>* F = fields_arg();
            KArg<filt_params>* Q = params_arg();
            const DivBy dx{Q->h1, rh[0]}, dy{Q->h2, rh[1]}, dz{Q->h3, rh[2]};
            const DivBy dx2{2. * Q->h1, rh[3]}, dy2{2. * Q->h2, rh[4]}, dz2{2. * Q->h3, rh[5]};
            const double X1 = zc[AX1], X2 = zc[AX2];
            double u1, u2;
            if (i == 0) {
                u1 = dx(X1 * (sh[L::at(AP, cy, cx + 1)] - Pc));
                u2 = dx(X2 * (sh[L::at(AP, cy, cx + 1)] - Pc));
            } else if (i == nx - 1) {
                u1 = dx(X1 * (Pc - sh[L::at(AP, cy, cx - 1)]));
                u2 = dx(X2 * (Pc - sh[L::at(AP, cy, cx - 1)]));
            } else {
                const double gx = dx2(sh[L::at(AP, cy, cx + 1)] - sh[L::at(AP, cy, cx - 1)]);
                u1 = X1 * gx;
                u2 = X2 * gx;
            }
            stn(F->u1x, c, u1);
            stn(F->u2x, c, u2);
            if (j == 0) {
                const double gy = dy(sh[L::at(AP, cy + 1, cx)] - Pc);
                u1 = X1 * (gy - r1c * g);
                u2 = X2 * (gy - r2c * g);
            } else if (j == ny - 1) {
                const double gy = dy(Pc - sh[L::at(AP, cy - 1, cx)]);
                u1 = X1 * (gy - r1c * g);
                u2 = X2 * (gy - r2c * g);
            } else {
                const double gy = dy2(sh[L::at(AP, cy + 1, cx)] - sh[L::at(AP, cy - 1, cx)]);
                u1 = X1 * (gy - r1c * g);
                u2 = X2 * (gy - r2c * g);
            }
            stn(F->u1y, c, u1);
            stn(F->u2y, c, u2);
            if (k == 0) {
                u1 = dz(X1 * (zp[AP] - Pc));
                u2 = dz(X2 * (zp[AP] - Pc));
            } else if (k == nz - 1) {
                u1 = dz(X1 * (Pc - shm[L::at(AP, cy, cx)]));
                u2 = dz(X2 * (Pc - shm[L::at(AP, cy, cx)]));
            } else {
                const double gz = dz2(zp[AP] - shm[L::at(AP, cy, cx)]);
                u1 = X1 * gz;
                u2 = X2 * gz;
            }
            stn(F->u1z, c, u1);
            stn(F->u2z, c, u2);
        }
        if (valid && inner && k >= 1 && k <= nz - 2) {
            // roS_calculation, cpu.cpp:218-291 (P1 = P2)
            KArg<filt_fields>* F = fields_arg();
            KArg<filt_params>* Q = params_arg();
            const DivBy dx{Q->h1, rh[0]}, dy{Q->h2, rh[1]}, dz{Q->h3, rh[2]};
            const double X1c = zc[AX1], X2c = zc[AX2];
            const double S = Sc;
            const double roS1 = r1c * (1 - S);
            const double roS2 = r2c * S;
            double a1, a2;
            a2 = dz(-(zp[AP] - Pc));
            a1 = dz(-(Pc - shm[L::at(AP, cy, cx)]));
            const double r1p = ro1_of(zp[AP], *Q), r2p = ro2_of(zp[AP], *Q);
            const double f3 = upwind_div(a2, a1, X1c, r1c, shm[L::at(AX1, cy, cx)],
                                         shm[L::at(AR1, cy, cx)], zp[AX1], r1p, dz);
            const double g3 = upwind_div(a2, a1, X2c, r2c, shm[L::at(AX2, cy, cx)],
                                         shm[L::at(AR2, cy, cx)], zp[AX2], r2p, dz);
            const int xp = cx + 1, xm = cx - 1, yp = cy + 1, ym = cy - 1;
            a2 = dx(-(sh[L::at(AP, cy, xp)] - Pc));
            a1 = dx(-(Pc - sh[L::at(AP, cy, xm)]));
            const double f1 = upwind_div(a2, a1, X1c, r1c, sh[L::at(AX1, cy, xm)],
                                         sh[L::at(AR1, cy, xm)], sh[L::at(AX1, cy, xp)],
                                         sh[L::at(AR1, cy, xp)], dx);
            const double g1 = upwind_div(a2, a1, X2c, r2c, sh[L::at(AX2, cy, xm)],
                                         sh[L::at(AR2, cy, xm)], sh[L::at(AX2, cy, xp)],
                                         sh[L::at(AR2, cy, xp)], dx);
            const double dy2 = dy(-(sh[L::at(AP, yp, cx)] - Pc));
            const double dy1 = dy(-(Pc - sh[L::at(AP, ym, cx)]));
            const double f2 = upwind_div(dy2 + r1c * g, dy1 + r1c * g, X1c, r1c,
                                         sh[L::at(AX1, ym, cx)], sh[L::at(AR1, ym, cx)],
                                         sh[L::at(AX1, yp, cx)], sh[L::at(AR1, yp, cx)], dy);
            const double g2 = upwind_div(dy2 + r2c * g, dy1 + r2c * g, X2c, r2c,
                                         sh[L::at(AX2, ym, cx)], sh[L::at(AR2, ym, cx)],
                                         sh[L::at(AX2, yp, cx)], sh[L::at(AR2, yp, cx)], dy);
            const int md = mdc;
            const double phi = F->poro ? phic : med[3][md];
            const double A1 = roS1 - (Q->dt / phi) * (f1 + f2 + f3);
            const double A2 = roS2 - (Q->dt / phi) * (g1 + g2 + g3);
            stn(F->roS1_old, c, roS1);
            stn(F->roS2_old, c, roS2);
            stn(F->roS1, c, A1);
            stn(F->roS2, c, A2);
            // P1_S2_calculation, cpu.cpp:293-320, on this cell's registers
            double P = Pc, Sv = S;
            newton_cell(P, Sv, A1, A2, med[0][md], med[1][md], med[2][md], *Q);
            if (edge || k == kb || k == ke - 1) {  // read by other blocks: stash
                KArg<FullArgs>* A = kernarg<FullArgs>(0);
                stn(A->stP, c, P);
                stn(A->stS, c, Sv);
            } else {
                stn(F->P1, c, P);
                stn(F->S2, c, Sv);
            }
        }
#pragma unroll
        for (int a = 0; a < NL; a++) zc[a] = zp[a];
        Sc = Sn;
        mdc = mdn;
    }
}

// Copies the stash of k_step_full into P1, S2: for each interior row (j, k)
// of the grid, the whole row on the first and last plane of a z-chunk and on
// the first and last row of a tile, else the tile-edge columns.  One wave
// per row.
__global__ __launch_bounds__(256) void k_unstash(FullArgs args, int tx, int ty)
{
    const filt_fields& f = args.s.f;
    const int nx = args.s.nx, ny = args.s.p.Ny, nz = args.s.p.Nz, kz = args.s.kz;
    const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long long rows = (long long)(ny - 2) * (nz - 2);
    if (row >= rows)
        return;
    const int j = 1 + (int)(row % (ny - 2)), k = 1 + (int)(row / (ny - 2));
    const int lane = threadIdx.x & 63;
    const unsigned base = 8u * (unsigned)nx * ((unsigned)j + (unsigned)ny * (unsigned)k);
    const int kk = k % kz, jj = j % ty;
    if (kk == 0 || kk == kz - 1 || jj == 0 || jj == ty - 1) {
        for (int i = 1 + lane; i <= nx - 2; i += 64) {
            const unsigned c = base + 8u * i;
            stn(f.P1, c, ld(args.stP, c));
            stn(f.S2, c, ld(args.stS, c));
        }
        return;
    }
    const int tiles = (nx + tx - 1) / tx;
    for (int q = lane; q < 2 * tiles; q += 64) {
        const int i = (q >> 1) * tx + ((q & 1) ? tx - 1 : 0);
        if (i < 1 || i > nx - 2)
            continue;
        const unsigned c = base + 8u * i;
        stn(f.P1, c, ld(args.stP, c));
        stn(f.S2, c, ld(args.stS, c));
    }
}

using StepFn = void (*)(StepArgs);

template <bool WITH_U, bool BC>
StepFn pick_step(int tx, int ty)
{
    if (tx == 64 && ty == 4) return k_step_zm<64, 4, WITH_U, BC>;
    if (tx == 32 && ty == 8) return k_step_zm<32, 8, WITH_U, BC>;
    if (tx == 16 && ty == 16) return k_step_zm<16, 16, WITH_U, BC>;
    if (tx == 64 && ty == 8) return k_step_zm<64, 8, WITH_U, BC>;
    if (tx == 32 && ty == 16) return k_step_zm<32, 16, WITH_U, BC>;
    if (tx == 64 && ty == 16) return k_step_zm<64, 16, WITH_U, BC>;
    if (tx == 128 && ty == 4) return k_step_zm<128, 4, WITH_U, BC>;
    if (tx == 128 && ty == 8) return k_step_zm<128, 8, WITH_U, BC>;
    if (tx == 256 && ty == 4) return k_step_zm<256, 4, WITH_U, BC>;
    return nullptr;
}

using FullFn = void (*)(FullArgs);

template <bool WITH_U, bool MEDIA>
FullFn pick_full(int tx, int ty)
{
    if (tx == 64 && ty == 4) return k_step_full<64, 4, WITH_U, MEDIA>;
    if (tx == 32 && ty == 8) return k_step_full<32, 8, WITH_U, MEDIA>;
    if (tx == 16 && ty == 16) return k_step_full<16, 16, WITH_U, MEDIA>;
    if (tx == 64 && ty == 8) return k_step_full<64, 8, WITH_U, MEDIA>;
    if (tx == 32 && ty == 16) return k_step_full<32, 16, WITH_U, MEDIA>;
    if (tx == 64 && ty == 16) return k_step_full<64, 16, WITH_U, MEDIA>;
    return nullptr;
}

// The stash of k_step_full: one pair of arrays per process, grown to the
// largest grid stepped and kept until the process ends.
std::mutex g_stash_mu;
double* g_stash = nullptr;
size_t g_stash_cells = 0;

}  // namespace

bool step_fusable(const filt_params& p, int nx, int rank, int size)
{
    return size == 1 && nx == p.NX && nx >= 3 && p.Ny >= 3 && p.Nz >= 3 &&
           tiled_fits(nx, p.Ny, p.Nz) && p.roS_scheme == FILT_SCHEME_UPWIND &&
           divby_ok(p.h1) && divby_ok(p.h2) && divby_ok(p.h3) &&
           tiled_enabled(2);
}

bool full_tile_ok(int tx, int ty) { return pick_full<true, false>(tx, ty) != nullptr; }
bool step_tile_ok(int tx, int ty) { return pick_step<true, false>(tx, ty) != nullptr; }

bool full_step_fusable(const filt_params& p, int nx, int rank, int size, bool per_cell_media)
{
    if (!step_fusable(p, nx, rank, size) || !tiled_enabled(3))
        return false;
    if (!divby_ok(p.mu1) || !divby_ok(p.mu2))
        return false;
    const int n = per_cell_media ? p.n_media : 1;
    if (n < 1 || n > FILT_MAX_MEDIA)
        return false;
    for (int q = 0; q < n; q++)
        if (p.lambda[q] != 2.0 || !divby_ok(1. - p.S_wr[q]))
            return false;
    return true;
}

int launch_step_full(const filt_fields& f, const filt_params& p, int nx, bool with_u,
                     hipEvent_t mid)
{
    const TileConfig t = full_tile_config(nx, p.Ny, p.Nz);
    FullArgs a;
    a.s.f = f;
    a.s.p = p;
    a.s.nx = nx;
    a.s.kz = t.kz;
    a.s.map = make_map(nx, p.Ny, p.Nz, t);
    const size_t cells = (size_t)nx * p.Ny * p.Nz;
    {
        std::lock_guard<std::mutex> lock(g_stash_mu);
        if (cells > g_stash_cells) {
            // launches already enqueued may still use the old stash
            FILT_HIP_TRY(hipStreamSynchronize(stream()));
            if (g_stash)
                (void)hipFree(g_stash);
            g_stash = nullptr;
            g_stash_cells = 0;
            FILT_HIP_TRY(hipMalloc((void**)&g_stash, 2 * cells * sizeof(double)));
            g_stash_cells = cells;
        }
        a.stP = g_stash;
        a.stS = g_stash + g_stash_cells;
    }
    const unsigned blocks = (unsigned)a.s.map.count();
    const bool media = f.media && p.n_media > 1;
    const FullFn fn = with_u ? (media ? pick_full<true, true>(t.tx, t.ty)
                                      : pick_full<true, false>(t.tx, t.ty))
                             : (media ? pick_full<false, true>(t.tx, t.ty)
                                      : pick_full<false, false>(t.tx, t.ty));
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(t.tx, t.ty), 0, stream(), a);
    FILT_LAUNCH_CHECK("k_step_full");
    if (mid)
        FILT_HIP_TRY(hipEventRecord(mid, stream()));
    const long long rows = (long long)(p.Ny - 2) * (p.Nz - 2);
    hipLaunchKernelGGL(k_unstash, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, stream(), a,
                       t.tx, t.ty);
    FILT_LAUNCH_CHECK("k_unstash");
    return FILT_OK;
}

// Launches of k_step_zm since the library was loaded (filt_step_zm_launches).
static std::atomic<long long> g_step_launches{0};

int launch_step_zm(const filt_fields& f, const filt_params& p, int nx, bool with_u, bool bc)
{
    return launch_step_zm_cfg(f, p, nx, with_u, bc, step_tile_config(nx, p.Ny, p.Nz));
}

int launch_step_zm_cfg(const filt_fields& f, const filt_params& p, int nx, bool with_u, bool bc,
                       const TileConfig& t)
{
    StepArgs a;
    a.f = f;
    a.p = p;
    a.nx = nx;
    a.kz = t.kz;
    a.map = make_map(nx, p.Ny, p.Nz, t);
    const unsigned blocks = (unsigned)a.map.count();
    const StepFn fn = with_u ? (bc ? pick_step<true, true>(t.tx, t.ty)
                                   : pick_step<true, false>(t.tx, t.ty))
                             : (bc ? pick_step<false, true>(t.tx, t.ty)
                                   : pick_step<false, false>(t.tx, t.ty));
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(t.tx, t.ty), 0, stream(), a);
    FILT_LAUNCH_CHECK("k_step_zm");
    g_step_launches.fetch_add(1, std::memory_order_relaxed);
    return FILT_OK;
}

}  // namespace filt

extern "C" long long filt_step_zm_launches(void)
{
    return filt::g_step_launches.load(std::memory_order_relaxed);
}

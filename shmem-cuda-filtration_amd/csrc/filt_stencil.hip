// filt_stencil.hip -- z-marching LDS-tiled kernels of the plain path's two
// stencil stages, u_calculation (cpu.cpp:77-160) and roS_calculation
// (cpu.cpp:218-291, and the divgrad scheme of cpu.cpp:162-216), for grids
// with NX, Ny, Nz >= 3 whose arrays are below 4 GiB each; their tile
// configurations, and the autotuners of these kernels and of the stages 2-4
// kernel (filt_autotune_stencil, filt_autotune_step).
//
// Each block owns a TX x TY tile of columns and a chunk of kz planes.  Per
// plane it writes its columns' values and a one-cell halo ring (read from
// global memory, where neighbouring tiles on the same XCD have just loaded
// them into L2) into one of two LDS tiles, synchronises once, and evaluates
// the cell from LDS (x, y neighbours) and registers (its own column's k-1,
// k, k+1).  The loads are software-pipelined: while plane k is evaluated,
// the halo ring and the point operands of plane k+1 and the own column of
// plane k+2 are in flight, so no load issued in an iteration is waited for
// in the same iteration; the two LDS tiles alternate, so one barrier per
// plane orders both the writes before the reads and the reads of plane k
// before the overwrite at plane k+2.  Halo cells are spread one per thread
// (ring cell h to thread h).  Every HBM element of the stencil arrays is
// read once per chunk.  Addresses are 32-bit byte offsets from the
// (uniform) array bases, which keeps them in one VGPR each.  The arithmetic
// is expression for expression the one of filt_stages.hip, so both paths
// give identical bits.  The roS kernel reads its own column's plane k-1 back
// from its own slot of the other LDS tile and divides by h1, h2, h3 through
// DivBy (filt_device.hpp, bit for bit the IEEE division), which fits it in
// 115 VGPRs at 4 waves per SIMD.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/filtration_hip.h"
#include "filt_device.hpp"
#include "filt_runtime.hpp"
#include "filt_tiles.hpp"
#include "filt_zmarch.hpp"
#include "../host/filt_guard.hpp"

namespace filt {

namespace {

using namespace zm;

// ---------------------------------------------------------------- roS ----
// cpu.cpp:218-291 for interior cells; arrays P1, P2, Xi1, Xi2, ro1, ro2 are
// the stencil operands (index 0..5 below).
template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(4))) void k_roS_zm(
    filt_fields f, filt_params p, int nx, TileMap map, int kz)
{
    using L = Lds<TX, TY>;
    __shared__ double shb[2][6 * L::SN];
    int bx, by, bz;
    map.decode(blockIdx.x, bx, by, bz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = bx * TX, j0 = by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int ny = p.Ny, nz = p.Nz;
    const bool valid = i < nx && j < ny;
    const bool inner = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
    const unsigned sy = 8u * nx, sz = 8u * nx * ny;  // byte strides
    const int kb = max(bz * kz, 1), ke = min(bz * kz + kz, nz - 1);
    if (kb >= ke)
        return;  // uniform over the block
    const double* const A[6] = {f.P1, f.P2, f.Xi1, f.Xi2, f.ro1, f.ro2};
    const unsigned col = 8u * i + sy * j;
    int hslot;
    unsigned hcol;
    const bool hv = ring_slot<TX, TY>(tx, ty, i0, j0, nx, ny, sy, hslot, hcol);
    // Porosity per cell, else by media index (one medium: index 0).
    const bool by_media = !f.poro && f.media && p.n_media > 1;
    // Pipeline registers: own column at k-1, k, k+1 and (in flight) k+2;
    // ring cell and point operands of the next plane.
    double zc[6] = {}, zp[6] = {}, zq[6] = {}, hn[6] = {};
    double Sc = 0, Sn = 0, phic = 0, phin = 0;
    int mdc = 0, mdn = 0;
    if (valid) {
        const unsigned c = col + kb * sz;
#pragma unroll
        for (int a = 0; a < 6; a++) {
            // plane kb-1 to the own slot of the tile plane kb reads as its
            // predecessor (no other thread writes that slot)
            shb[1][L::at(a, ty + 1, tx + 1)] = ld(A[a], c - sz);
            zc[a] = ld(A[a], c);
            zp[a] = ld(A[a], c + sz);  // kb + 1 <= ke <= nz - 1
        }
        Sc = ld(f.S2, c);
        if (f.poro)
            phic = ld(f.poro, c);
        else if (by_media)
            mdc = ldi(f.media, c >> 1);
    }
    if (hv)
#pragma unroll
        for (int a = 0; a < 6; a++) hn[a] = ld(A[a], hcol + kb * sz);
    const int cx = tx + 1, cy = ty + 1;
    const double g = p.g_const;
    __shared__ double rh[3];  // reciprocals of h1, h2, h3 for DivBy
    if (tx == 0 && ty == 0) {
        rh[0] = refined_reciprocal(p.h1);
        rh[1] = refined_reciprocal(p.h2);
        rh[2] = refined_reciprocal(p.h3);
    }
    for (int k = kb; k < ke; k++) {
        double* const sh = shb[(k - kb) & 1];
        const double* const shm = shb[(k - kb + 1) & 1];  // own slot: plane k-1
        const unsigned c = col + k * sz;
        if (valid)
#pragma unroll
            for (int a = 0; a < 6; a++) sh[L::at(a, cy, cx)] = zc[a];
        if (hv)
#pragma unroll
            for (int a = 0; a < 6; a++) sh[hslot + a * L::SN] = hn[a];
        if (k + 1 < ke) {  // next plane is evaluated: fetch its operands
            if (valid) {
#pragma unroll
                for (int a = 0; a < 6; a++) zq[a] = ld(A[a], c + 2 * sz);  // k+2 <= ke
                Sn = ld(f.S2, c + sz);
                if (f.poro)
                    phin = ld(f.poro, c + sz);
                else if (by_media)
                    mdn = ldi(f.media, (c + sz) >> 1);
            }
            if (hv)
#pragma unroll
                for (int a = 0; a < 6; a++) hn[a] = ld(A[a], hcol + (k + 1) * sz);
        }
        __syncthreads();
        if (valid && inner) {
            const DivBy dx{p.h1, rh[0]}, dy{p.h2, rh[1]}, dz{p.h3, rh[2]};
            double zm[6];
#pragma unroll
            for (int a = 0; a < 6; a++) zm[a] = shm[L::at(a, cy, cx)];
            const double p1 = zc[0], p2 = zc[1];
            const double X1c = zc[2], X2c = zc[3], r1c = zc[4], r2c = zc[5];
            const double S = Sc;
            const double roS1 = r1c * (1 - S);
            const double roS2 = r2c * S;
            double a1, a2;
            a2 = dz(-(zp[0] - p1));
            a1 = dz(-(p1 - zm[0]));
            const double f3 = upwind_div(a2, a1, X1c, r1c, zm[2], zm[4], zp[2], zp[4], dz);
            a2 = dz(-(zp[1] - p2));
            a1 = dz(-(p2 - zm[1]));
            const double g3 = upwind_div(a2, a1, X2c, r2c, zm[3], zm[5], zp[3], zp[5], dz);

            const int xp = cx + 1, xm = cx - 1, yp = cy + 1, ym = cy - 1;
            a2 = dx(-(sh[L::at(0, cy, xp)] - p1));
            a1 = dx(-(p1 - sh[L::at(0, cy, xm)]));
            const double f1 = upwind_div(a2, a1, X1c, r1c, sh[L::at(2, cy, xm)],
                                         sh[L::at(4, cy, xm)], sh[L::at(2, cy, xp)],
                                         sh[L::at(4, cy, xp)], dx);
            a2 = dy(-(sh[L::at(0, yp, cx)] - p1)) + r1c * g;
            a1 = dy(-(p1 - sh[L::at(0, ym, cx)])) + r1c * g;
            const double f2 = upwind_div(a2, a1, X1c, r1c, sh[L::at(2, ym, cx)],
                                         sh[L::at(4, ym, cx)], sh[L::at(2, yp, cx)],
                                         sh[L::at(4, yp, cx)], dy);

            a2 = dx(-(sh[L::at(1, cy, xp)] - p2));
            a1 = dx(-(p2 - sh[L::at(1, cy, xm)]));
            const double g1 = upwind_div(a2, a1, X2c, r2c, sh[L::at(3, cy, xm)],
                                         sh[L::at(5, cy, xm)], sh[L::at(3, cy, xp)],
                                         sh[L::at(5, cy, xp)], dx);
            a2 = dy(-(sh[L::at(1, yp, cx)] - p2)) + r2c * g;
            a1 = dy(-(p2 - sh[L::at(1, ym, cx)])) + r2c * g;
            const double g2 = upwind_div(a2, a1, X2c, r2c, sh[L::at(3, ym, cx)],
                                         sh[L::at(5, ym, cx)], sh[L::at(3, yp, cx)],
                                         sh[L::at(5, yp, cx)], dy);

            const double phi = f.poro ? phic : p.m[mdc];
            stn(f.roS1_old, c, roS1);
            stn(f.roS2_old, c, roS2);
            stn(f.roS1, c, roS1 - (p.dt / phi) * (f1 + f2 + f3));
            stn(f.roS2, c, roS2 - (p.dt / phi) * (g1 + g2 + g3));
        }
#pragma unroll
        for (int a = 0; a < 6; a++) {
            zc[a] = zp[a];
            zp[a] = zq[a];
        }
        Sc = Sn;
        phic = phin;
        mdc = mdn;
    }
}

// -------------------------------------------------------- roS, divgrad ----
// cpu.cpp:162-216 (assign_rS) for interior cells: the tau-relaxed scheme.
// Its x/y stencil operands are six products, each formed once per cell as
// the plane is staged: a1 = ro1*(1-S2) and a2 = ro2*S2 (diffusion), b1x =
// ro1*u1x and b2x = ro2*u2x (x transport), b1y = ro1*u1y and b2y = ro2*u2y
// (y transport).  Each is the product the reference forms at that neighbour,
// so the bits are the same.  Along z the own column's a1, a2 and b2y of plane
// k+1 are formed from registers; those of plane k-1, with b1x, are read back
// from the thread's own slots of the other LDS tile (no other thread writes
// them); the other term of the water z transport, ro1*u1x at (i+1, j, k+1),
// is loaded one plane ahead for every cell.  The divisions by h^2, 2h, dt^2
// and 2dt^2 go through DivBy, their divisors and reciprocals read from LDS
// where they are used (held in registers across the loop they would take 32
// VGPRs); the two by the cell's porosity are IEEE divisions.  Ring cells of
// the two rows (h < 2*TX) carry the y transport products, those of the two
// columns the x ones.  PORO: porosity per cell, else by media index.  At 4
// waves per SIMD one value is kept in scratch across the rare-numerator
// calls of DivBy.
template <int TX, int TY, bool PORO>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(4))) void k_roS_divgrad_zm(
    filt_fields f, filt_params p, int nx, TileMap map, int kz, double t)
{
    using L = Lds<TX, TY>;
    enum { A1, A2, B1X, B2X, B1Y, B2Y, NA };
    __shared__ double shb[2][NA * L::SN];
    // for DivBy: reciprocals of h1^2, h2^2, h3^2, 2h1, 2h2, 2h3, dt^2, 2dt^2,
    // then the divisors themselves
    __shared__ double rr[16];
    int bx, by, bz;
    map.decode(blockIdx.x, bx, by, bz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = bx * TX, j0 = by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int ny = p.Ny, nz = p.Nz;
    const bool valid = i < nx && j < ny;
    const bool inner = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
    const unsigned sy = 8u * nx, sz = 8u * nx * ny;  // byte strides
    const int kb = max(bz * kz, 1), ke = min(bz * kz + kz, nz - 1);
    if (kb >= ke)
        return;  // uniform over the block
    const unsigned col = 8u * i + sy * j;
    // (i+1, j), clamped into the grid for the lanes whose value is unused
    const unsigned scol = 8u * min(i + 1, nx - 1) + sy * j;
    int hslot;
    unsigned hcol;
    const bool hv = ring_slot<TX, TY>(tx, ty, i0, j0, nx, ny, sy, hslot, hcol);
    // a ring column cell (x transport), else a row cell (y transport); the
    // ring cells of one wave are all of one kind for tiles with TX >= 32
    const bool hx = tx + TX * ty >= 2 * TX;
    const int hb = hslot + (hx ? B1X : B1Y) * L::SN;  // its transport products
    const bool by_media = !PORO && f.media && p.n_media > 1;
    // Own column: ro1, ro2, S2, u2y at planes k, k+1 and (in flight) k+2;
    // u1x, u1y, u2x and roS1_old, roS2_old, porosity at k and (in flight)
    // k+1; ro1, u1x at (i+1, j, k+1) and (in flight) k+2.  Ring cell of the
    // next plane: ro1, ro2, S2 and its two u components.
    double zc[4] = {}, zp[4] = {}, zq[4] = {}, xc[3] = {}, xn[3] = {};
    double pc[3] = {}, pn[3] = {}, sc[2] = {}, sn[2] = {}, hn[5] = {};
    int mdc = 0, mdn = 0;
    if (valid) {
        const unsigned c = col + kb * sz;
        // plane kb-1's products to the own slots of the tile plane kb reads
        // as its predecessor
        const double r1m = ld(f.ro1, c - sz), r2m = ld(f.ro2, c - sz), Sm = ld(f.S2, c - sz);
        shb[1][L::at(A1, ty + 1, tx + 1)] = r1m * (1. - Sm);
        shb[1][L::at(A2, ty + 1, tx + 1)] = r2m * Sm;
        shb[1][L::at(B1X, ty + 1, tx + 1)] = r1m * ld(f.u1x, c - sz);
        shb[1][L::at(B2Y, ty + 1, tx + 1)] = r2m * ld(f.u2y, c - sz);
        zc[0] = ld(f.ro1, c), zc[1] = ld(f.ro2, c), zc[2] = ld(f.S2, c), zc[3] = ld(f.u2y, c);
        zp[0] = ld(f.ro1, c + sz), zp[1] = ld(f.ro2, c + sz), zp[2] = ld(f.S2, c + sz);
        zp[3] = ld(f.u2y, c + sz);  // kb + 1 <= ke <= nz - 1
        xc[0] = ld(f.u1x, c), xc[1] = ld(f.u1y, c), xc[2] = ld(f.u2x, c);
        pc[0] = ld(f.roS1_old, c), pc[1] = ld(f.roS2_old, c);
        if (PORO)
            pc[2] = ld(f.poro, c);
        else if (by_media)
            mdc = ldi(f.media, c >> 1);
        sc[0] = ld(f.ro1, scol + (kb + 1) * sz), sc[1] = ld(f.u1x, scol + (kb + 1) * sz);
    }
    if (hv) {
        const unsigned h = hcol + kb * sz;
        hn[0] = ld(f.ro1, h), hn[1] = ld(f.ro2, h), hn[2] = ld(f.S2, h);
        if (hx)
            hn[3] = ld(f.u1x, h), hn[4] = ld(f.u2x, h);
        else
            hn[3] = ld(f.u1y, h), hn[4] = ld(f.u2y, h);
    }
    if (tx == 0 && ty == 0) {
        const double dt = p.dt;
        rr[0] = refined_reciprocal(p.h1 * p.h1);
        rr[1] = refined_reciprocal(p.h2 * p.h2);
        rr[2] = refined_reciprocal(p.h3 * p.h3);
        rr[3] = refined_reciprocal(2. * p.h1);
        rr[4] = refined_reciprocal(2. * p.h2);
        rr[5] = refined_reciprocal(2. * p.h3);
        rr[6] = refined_reciprocal(dt * dt);
        rr[7] = refined_reciprocal(2. * dt * dt);
        rr[8] = p.h1 * p.h1, rr[9] = p.h2 * p.h2, rr[10] = p.h3 * p.h3;
        rr[11] = 2. * p.h1, rr[12] = 2. * p.h2, rr[13] = 2. * p.h3;
        rr[14] = dt * dt, rr[15] = 2. * dt * dt;
    }
    const int cx = tx + 1, cy = ty + 1;
    for (int k = kb; k < ke; k++) {
        double* const sh = shb[(k - kb) & 1];
        const double* const shm = shb[(k - kb + 1) & 1];  // own slots: plane k-1
        const unsigned c = col + k * sz;
        const double b1xs = sc[0] * sc[1];  // ro1*u1x at (i+1, j, k+1)
        if (valid) {
            sh[L::at(A1, cy, cx)] = zc[0] * (1. - zc[2]);
            sh[L::at(A2, cy, cx)] = zc[1] * zc[2];
            sh[L::at(B1X, cy, cx)] = zc[0] * xc[0];
            sh[L::at(B2X, cy, cx)] = zc[1] * xc[2];
            sh[L::at(B1Y, cy, cx)] = zc[0] * xc[1];
            sh[L::at(B2Y, cy, cx)] = zc[1] * zc[3];
        }
        if (hv) {
            sh[hslot + A1 * L::SN] = hn[0] * (1. - hn[2]);
            sh[hslot + A2 * L::SN] = hn[1] * hn[2];
            sh[hb] = hn[0] * hn[3];
            sh[hb + L::SN] = hn[1] * hn[4];
        }
        if (k + 1 < ke) {  // next plane is evaluated: fetch its operands
            if (valid) {
                const unsigned c1 = c + sz, c2 = c + 2 * sz;  // k+2 <= ke
                zq[0] = ld(f.ro1, c2), zq[1] = ld(f.ro2, c2), zq[2] = ld(f.S2, c2);
                zq[3] = ld(f.u2y, c2);
                xn[0] = ld(f.u1x, c1), xn[1] = ld(f.u1y, c1), xn[2] = ld(f.u2x, c1);
                pn[0] = ld(f.roS1_old, c1), pn[1] = ld(f.roS2_old, c1);
                if (PORO)
                    pn[2] = ld(f.poro, c1);
                else if (by_media)
                    mdn = ldi(f.media, c1 >> 1);
                sn[0] = ld(f.ro1, scol + (k + 2) * sz), sn[1] = ld(f.u1x, scol + (k + 2) * sz);
            }
            if (hv) {
                const unsigned h = hcol + (k + 1) * sz;
                hn[0] = ld(f.ro1, h), hn[1] = ld(f.ro2, h), hn[2] = ld(f.S2, h);
                if (hx)
                    hn[3] = ld(f.u1x, h), hn[4] = ld(f.u2x, h);
                else
                    hn[3] = ld(f.u1y, h), hn[4] = ld(f.u2y, h);
            }
        }
        __syncthreads();
        if (valid && inner) {
            // an index the compiler cannot see through keeps the LDS reads
            // of the divisors here rather than hoisted out of the loop
            int z0 = 0;
            asm volatile("" : "+v"(z0));
            const double* const q = rr + z0;
            const DivBy dxx{q[8], q[0]}, dyy{q[9], q[1]}, dzz{q[10], q[2]};
            const DivBy dx2{q[11], q[3]}, dy2{q[12], q[4]}, dz2{q[13], q[5]};
            const double r1c = zc[0], r2c = zc[1], Sc = zc[2];
            const double roS1 = r1c * (1 - Sc);
            const double roS2 = r2c * Sc;
            const double m = PORO ? pc[2] : p.m[mdc];
            const double k1 = m * p.l_w * p.c / 2.;
            const double k2 = m * p.l_n * p.c / 2.;
            const double t1 = 2 * r1c * (1. - Sc), t2 = 2 * r2c * Sc;
            double dg1 = dzz(k1 * (zp[0] * (1. - zp[2]) - t1 + shm[L::at(A1, cy, cx)]));
            double dg2 = dzz(k2 * (zp[1] * zp[2] - t2 + shm[L::at(A2, cy, cx)]));
            const double Tz1 = dz2(b1xs - shm[L::at(B1X, cy, cx)]);
            const double Tz2 = dz2(zp[1] * zp[3] - shm[L::at(B2Y, cy, cx)]);
            const int xp = cx + 1, xm = cx - 1, yp = cy + 1, ym = cy - 1;
            dg1 += k1 * (dxx(sh[L::at(A1, cy, xp)] - t1 + sh[L::at(A1, cy, xm)]) +
                         dyy(sh[L::at(A1, yp, cx)] - t1 + sh[L::at(A1, ym, cx)]));
            dg2 += k2 * (dxx(sh[L::at(A2, cy, xp)] - t2 + sh[L::at(A2, cy, xm)]) +
                         dyy(sh[L::at(A2, yp, cx)] - t2 + sh[L::at(A2, ym, cx)]));
            const double Tx1 = dx2(sh[L::at(B1X, cy, xp)] - sh[L::at(B1X, cy, xm)]);
            const double Ty1 = dy2(sh[L::at(B1Y, yp, cx)] - sh[L::at(B1Y, ym, cx)]);
            const double Tx2 = dx2(sh[L::at(B2X, cy, xp)] - sh[L::at(B2X, cy, xm)]);
            const double Ty2 = dy2(sh[L::at(B2Y, yp, cx)] - sh[L::at(B2Y, ym, cx)]);
            const double dt = p.dt, tau = p.tau;
            double A1v, A2v;
            if (t < 2 * dt) {
                A1v = roS1 + (dt / m) * (dg1 - Tx1 - Ty1 - Tz1);
                A2v = roS2 + (dt / m) * (dg2 - Tx2 - Ty2 - Tz2);
            } else {
                const DivBy dtt{q[14], q[6]}, d2tt{q[15], q[7]};
                const double w = (2. * dt * dt) / (m * (dt + 2. * tau));
                A1v = w * (dg1 - Tx1 - Ty1 - Tz1 + dtt(2. * roS1 * m * tau) +
                           d2tt(pc[0] * m * (dt - 2. * tau)));
                A2v = w * (dg2 - Tx2 - Ty2 - Tz2 + dtt(2. * roS2 * m * tau) +
                           d2tt(pc[1] * m * (dt - 2. * tau)));
            }
            stn(f.roS1_old, c, roS1);
            stn(f.roS2_old, c, roS2);
            stn(f.roS1, c, A1v);
            stn(f.roS2, c, A2v);
        }
#pragma unroll
        for (int a = 0; a < 4; a++) {
            zc[a] = zp[a];
            zp[a] = zq[a];
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            xc[a] = xn[a];
            pc[a] = pn[a];
        }
        sc[0] = sn[0], sc[1] = sn[1];
        mdc = mdn;
    }
}

// ------------------------------------------------------------------ u ----
// cpu.cpp:77-160 for every active cell; stencil operands P1, P2 only.
template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_u_zm(filt_fields f, filt_params p, int nx, int rank,
                                                 int size, TileMap map, int kz)
{
    using L = Lds<TX, TY>;
    __shared__ double shb[2][2 * L::SN];
    int bx, by, bz;
    map.decode(blockIdx.x, bx, by, bz);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = bx * TX, j0 = by * TY;
    const int i = i0 + tx, j = j0 + ty;
    const int ny = p.Ny, nz = p.Nz;
    const bool valid = i < nx && j < ny;
    const bool active = valid && is_active(i, nx, rank, size);
    const unsigned sy = 8u * nx, sz = 8u * nx * ny;
    const int kb = bz * kz, ke = min(bz * kz + kz, nz);
    if (kb >= ke)
        return;
    const unsigned col = 8u * i + sy * j;
    int hslot;
    unsigned hcol;
    const bool hv = ring_slot<TX, TY>(tx, ty, i0, j0, nx, ny, sy, hslot, hcol);
    // Pipeline registers: P1, P2 of the own column at k-1, k, k+1 and (in
    // flight) k+2; the ring cell and the point operands of the next plane.
    double zm0 = 0, zm1 = 0, zc0 = 0, zc1 = 0, zp0 = 0, zp1 = 0, zq0 = 0, zq1 = 0;
    double hn0 = 0, hn1 = 0;
    double X1c = 0, X2c = 0, r1c = 0, r2c = 0, X1n = 0, X2n = 0, r1n = 0, r2n = 0;
    if (valid) {
        const unsigned c = col + kb * sz;
        if (kb >= 1) {
            zm0 = ld(f.P1, c - sz);
            zm1 = ld(f.P2, c - sz);
        }
        zc0 = ld(f.P1, c);
        zc1 = ld(f.P2, c);
        if (kb + 1 < nz) {
            zp0 = ld(f.P1, c + sz);
            zp1 = ld(f.P2, c + sz);
        }
        X1c = ld(f.Xi1, c);
        X2c = ld(f.Xi2, c);
        r1c = ld(f.ro1, c);
        r2c = ld(f.ro2, c);
    }
    if (hv) {
        hn0 = ld(f.P1, hcol + kb * sz);
        hn1 = ld(f.P2, hcol + kb * sz);
    }
    const int cx = tx + 1, cy = ty + 1;
    const double g = p.g_const;
    for (int k = kb; k < ke; k++) {
        double* const sh = shb[(k - kb) & 1];
        const unsigned c = col + k * sz;
        if (valid) {
            sh[L::at(0, cy, cx)] = zc0;
            sh[L::at(1, cy, cx)] = zc1;
        }
        if (hv) {
            sh[hslot] = hn0;
            sh[hslot + L::SN] = hn1;
        }
        if (k + 1 < ke) {  // next plane is evaluated: fetch its operands
            if (valid) {
                if (k + 2 < nz) {
                    zq0 = ld(f.P1, c + 2 * sz);
                    zq1 = ld(f.P2, c + 2 * sz);
                }
                X1n = ld(f.Xi1, c + sz);
                X2n = ld(f.Xi2, c + sz);
                r1n = ld(f.ro1, c + sz);
                r2n = ld(f.ro2, c + sz);
            }
            if (hv) {
                hn0 = ld(f.P1, hcol + (k + 1) * sz);
                hn1 = ld(f.P2, hcol + (k + 1) * sz);
            }
        }
        __syncthreads();
        if (active) {
            const double X1 = X1c, X2 = X2c;
            const double P1c = zc0, P2c = zc1;
            double u1, u2;
            // x (p.NX > 2 holds: this kernel runs for NX >= 3 only)
            if (i == 0) {
                u1 = X1 * (sh[L::at(0, cy, cx + 1)] - P1c) / p.h1;
                u2 = X2 * (sh[L::at(1, cy, cx + 1)] - P2c) / p.h1;
            } else if (i == nx - 1) {
                u1 = X1 * (P1c - sh[L::at(0, cy, cx - 1)]) / p.h1;
                u2 = X2 * (P2c - sh[L::at(1, cy, cx - 1)]) / p.h1;
            } else {
                u1 = X1 * ((sh[L::at(0, cy, cx + 1)] - sh[L::at(0, cy, cx - 1)]) / (2. * p.h1));
                u2 = X2 * ((sh[L::at(1, cy, cx + 1)] - sh[L::at(1, cy, cx - 1)]) / (2. * p.h1));
            }
            stn(f.u1x, c, u1);
            stn(f.u2x, c, u2);
            // y
            const double r1 = r1c, r2 = r2c;
            if (j == 0) {
                u1 = X1 * ((sh[L::at(0, cy + 1, cx)] - P1c) / p.h2 - r1 * g);
                u2 = X2 * ((sh[L::at(1, cy + 1, cx)] - P2c) / p.h2 - r2 * g);
            } else if (j == ny - 1) {
                u1 = X1 * ((P1c - sh[L::at(0, cy - 1, cx)]) / p.h2 - r1 * g);
                u2 = X2 * ((P2c - sh[L::at(1, cy - 1, cx)]) / p.h2 - r2 * g);
            } else {
                u1 = X1 * ((sh[L::at(0, cy + 1, cx)] - sh[L::at(0, cy - 1, cx)]) / (2. * p.h2) - r1 * g);
                u2 = X2 * ((sh[L::at(1, cy + 1, cx)] - sh[L::at(1, cy - 1, cx)]) / (2. * p.h2) - r2 * g);
            }
            stn(f.u1y, c, u1);
            stn(f.u2y, c, u2);
            // z
            if (k == 0) {
                u1 = X1 * (zp0 - P1c) / p.h3;
                u2 = X2 * (zp1 - P2c) / p.h3;
            } else if (k == nz - 1) {
                u1 = X1 * (P1c - zm0) / p.h3;
                u2 = X2 * (P2c - zm1) / p.h3;
            } else {
                u1 = X1 * ((zp0 - zm0) / (2. * p.h3));
                u2 = X2 * ((zp1 - zm1) / (2. * p.h3));
            }
            stn(f.u1z, c, u1);
            stn(f.u2z, c, u2);
        }
        zm0 = zc0;
        zm1 = zc1;
        zc0 = zp0;
        zc1 = zp1;
        zp0 = zq0;
        zp1 = zq1;
        X1c = X1n;
        X2c = X2n;
        r1c = r1n;
        r2c = r2n;
    }
}

// Kernel variants by tile shape.
using RoSFn = void (*)(filt_fields, filt_params, int, TileMap, int);
using UFn = void (*)(filt_fields, filt_params, int, int, int, TileMap, int);

RoSFn pick_roS(int tx, int ty)
{
    if (tx == 64 && ty == 4) return k_roS_zm<64, 4>;
    if (tx == 32 && ty == 8) return k_roS_zm<32, 8>;
    if (tx == 16 && ty == 16) return k_roS_zm<16, 16>;
    if (tx == 64 && ty == 8) return k_roS_zm<64, 8>;
    if (tx == 128 && ty == 4) return k_roS_zm<128, 4>;
    if (tx == 128 && ty == 8) return k_roS_zm<128, 8>;
    return nullptr;
}

using DivgradFn = void (*)(filt_fields, filt_params, int, TileMap, int, double);

template <int TX, int TY>
DivgradFn divgrad_of(bool poro)
{
    return poro ? k_roS_divgrad_zm<TX, TY, true> : k_roS_divgrad_zm<TX, TY, false>;
}

DivgradFn pick_divgrad(int tx, int ty, bool poro)
{
    if (tx == 64 && ty == 4) return divgrad_of<64, 4>(poro);
    if (tx == 32 && ty == 8) return divgrad_of<32, 8>(poro);
    if (tx == 16 && ty == 16) return divgrad_of<16, 16>(poro);
    if (tx == 64 && ty == 8) return divgrad_of<64, 8>(poro);
    if (tx == 128 && ty == 4) return divgrad_of<128, 4>(poro);
    if (tx == 128 && ty == 8) return divgrad_of<128, 8>(poro);
    return nullptr;
}

UFn pick_u(int tx, int ty)
{
    if (tx == 64 && ty == 4) return k_u_zm<64, 4>;
    if (tx == 32 && ty == 8) return k_u_zm<32, 8>;
    if (tx == 16 && ty == 16) return k_u_zm<16, 16>;
    if (tx == 64 && ty == 8) return k_u_zm<64, 8>;
    if (tx == 128 && ty == 4) return k_u_zm<128, 4>;
    if (tx == 128 && ty == 8) return k_u_zm<128, 8>;
    return nullptr;
}

// Defaults: the fastest configurations of a sweep over every shape, kz in
// {16, 64} (the fused step: {8, 16, 32, 64}) and both orders on a 256^3 grid
// on MI355X (roS at 4 waves per SIMD: 64x8, kz 64).  The fused step's 64x8 tile runs two blocks of 8 waves per CU at
// 4 waves per SIMD; kz 64 timed within 1% of kz 16, which keeps more blocks
// for smaller grids.
constexpr TileConfig kRoSDefault = {64, 8, 64, 1};
constexpr TileConfig kUDefault = {64, 8, 64, 1};
constexpr TileConfig kStepDefault = {64, 8, 16, 1};
// The whole-step kernel is off by default (g_untiled[3]): on a 256^3 grid its
// best configuration (64x16, kz 64) ran 0.721 ms per step with its stash
// copy against 0.726 ms for stage 1 + the stages 2-4 kernel -- its 17 output
// streams run at 5.3 TB/s at best on MI355X against 5.7 and 6.2 for the two
// kernels' mixes, which eats the 32 B per cell it saves.
constexpr TileConfig kFullDefault = {64, 16, 64, 0};
TileConfig g_roS_cfg = kRoSDefault;
TileConfig g_u_cfg = kUDefault;
TileConfig g_step_cfg = kStepDefault;
TileConfig g_full_cfg = kFullDefault;
// Per kernel (0 roS, 1 u, 2 the fused step): run the one-cell-per-thread
// kernels of filt_stages.hip (for 2: the stages one by one) even where the
// tiled ones apply (filt_set_tile_config with tx = -1).
bool g_untiled[4] = {false, false, false, true};

// Autotuned configurations by grid shape (filt_autotune_stencil).
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, std::pair<TileConfig, TileConfig>> g_tuned;
// ... and of the stages 2-4 kernel (filt_autotune_step).
std::map<std::tuple<int, int, int>, TileConfig> g_step_tuned;

// The fused step offers the tile shapes of the roS kernel.
bool valid_cfg(const TileConfig& t, int kernel)
{
    if (!(t.kz >= 1 && t.kz <= 4096 && (t.order == 0 || t.order == 1)))
        return false;
    if (kernel == 3)
        return full_tile_ok(t.tx, t.ty);
    if (kernel == 2)
        return step_tile_ok(t.tx, t.ty);
    return kernel != 1 ? pick_roS(t.tx, t.ty) != nullptr : pick_u(t.tx, t.ty) != nullptr;
}

TileConfig cfg_for(int kernel, int nx, int ny, int nz)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (kernel == 2) {
        auto it = g_step_tuned.find(std::make_tuple(nx, ny, nz));
        return it != g_step_tuned.end() ? it->second : g_step_cfg;
    }
    if (kernel == 3)
        return g_full_cfg;
    auto it = g_tuned.find(std::make_tuple(nx, ny, nz));
    if (it != g_tuned.end())
        return kernel == 0 ? it->second.first : it->second.second;
    return kernel == 0 ? g_roS_cfg : g_u_cfg;
}

}  // namespace

// The tiled kernels address each array with 32-bit byte offsets.
bool tiled_fits(int nx, int ny, int nz)
{
    return (double)nx * ny * nz * 8.0 < 4294967296.0 - 65536.0;
}

bool tiled_enabled(int kernel)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return !g_untiled[kernel];
}

TileConfig step_tile_config(int nx, int ny, int nz) { return cfg_for(2, nx, ny, nz); }
TileConfig full_tile_config(int nx, int ny, int nz) { return cfg_for(3, nx, ny, nz); }

// Launchers used by filt_stages.hip (grids with NX, Ny, Nz >= 3).
int launch_roS_zm(const filt_fields& f, const filt_params& p, int nx)
{
    const TileConfig t = cfg_for(0, nx, p.Ny, p.Nz);
    const TileMap m = make_map(nx, p.Ny, p.Nz, t);
    hipLaunchKernelGGL(pick_roS(t.tx, t.ty), dim3((unsigned)m.count()), dim3(t.tx, t.ty), 0,
                       stream(), f, p, nx, m, t.kz);
    FILT_LAUNCH_CHECK("k_roS_zm");
    return FILT_OK;
}

bool divgrad_tiled_ok(const filt_params& p)
{
    const double dt = p.dt;
    return divby_ok(p.h1 * p.h1) && divby_ok(p.h2 * p.h2) && divby_ok(p.h3 * p.h3) &&
           divby_ok(2. * p.h1) && divby_ok(2. * p.h2) && divby_ok(2. * p.h3) &&
           divby_ok(dt * dt) && divby_ok(2. * dt * dt);
}

int launch_roS_divgrad_zm(const filt_fields& f, const filt_params& p, int nx, double t)
{
    const TileConfig c = cfg_for(0, nx, p.Ny, p.Nz);
    const TileMap m = make_map(nx, p.Ny, p.Nz, c);
    hipLaunchKernelGGL(pick_divgrad(c.tx, c.ty, f.poro != nullptr), dim3((unsigned)m.count()),
                       dim3(c.tx, c.ty), 0, stream(), f, p, nx, m, c.kz, t);
    FILT_LAUNCH_CHECK("k_roS_divgrad_zm");
    return FILT_OK;
}

int launch_u_zm(const filt_fields& f, const filt_params& p, int nx, int rank, int size)
{
    const TileConfig t = cfg_for(1, nx, p.Ny, p.Nz);
    const TileMap m = make_map(nx, p.Ny, p.Nz, t);
    hipLaunchKernelGGL(pick_u(t.tx, t.ty), dim3((unsigned)m.count()), dim3(t.tx, t.ty), 0,
                       stream(), f, p, nx, rank, size, m, t.kz);
    FILT_LAUNCH_CHECK("k_u_zm");
    return FILT_OK;
}

}  // namespace filt

extern "C" int filt_set_tile_config(int kernel, int tx, int ty, int kz, int order)
{
    return filt::guarded([&]() -> int {
        filt::TileConfig t = {tx, ty, kz, order};
        if (kernel < 0 || kernel > 3)
            return filt::set_error(FILT_ERR_INVALID,
                                   "kernel must be 0 (roS), 1 (u), 2 (stages 2-4) or 3 (step)");
        if (tx == -1) {  // the untiled kernels
            std::lock_guard<std::mutex> lk(filt::g_tune_mu);
            filt::g_untiled[kernel] = true;
            return FILT_OK;
        }
        if (tx == 0) {  // back to the built-in default
            t = kernel == 0 ? filt::kRoSDefault : kernel == 1 ? filt::kUDefault
              : kernel == 2 ? filt::kStepDefault : filt::kFullDefault;
            if (kernel == 3) {  // which is off for the whole-step kernel
                std::lock_guard<std::mutex> lk(filt::g_tune_mu);
                filt::g_full_cfg = t;
                filt::g_untiled[3] = true;
                return FILT_OK;
            }
        }
        if (!filt::valid_cfg(t, kernel))
            return filt::set_error(FILT_ERR_INVALID, "unsupported tile %dx%d kz=%d order=%d", tx, ty,
                                   kz, order);
        std::lock_guard<std::mutex> lk(filt::g_tune_mu);
        (kernel == 0 ? filt::g_roS_cfg : kernel == 1 ? filt::g_u_cfg
         : kernel == 2 ? filt::g_step_cfg : filt::g_full_cfg) = t;
        filt::g_untiled[kernel] = false;
        if (kernel < 2)
            filt::g_tuned.clear();
        if (kernel == 2)
            filt::g_step_tuned.clear();
        return FILT_OK;
    });
}

// Time every candidate tile configuration of the two stencil kernels on the
// given grid (reps launches each, median of the per-launch times) and keep
// the fastest per kernel for grids of this shape.  The kernels read the
// stage-1 outputs and rewrite only their own outputs (roS*, u*), which do
// not depend on earlier contents, so tuning on a live grid leaves the same
// values a single call would.
extern "C" int filt_autotune_stencil(const filt_fields* dev, const filt_params* p, int localNx,
                                     int reps, double* best_ms)
{
    return filt::guarded([&]() -> int {
        using namespace filt;
        int rc = check_shape(p, localNx);
        if (rc)
            return rc;
        if (!dev || !dev->P1)
            return set_error(FILT_ERR_INVALID, "null device fields");
        const int nx = localNx, ny = p->Ny, nz = p->Nz;
        if (nx < 3 || ny < 3 || nz < 3 || !tiled_fits(nx, ny, nz))
            return set_error(FILT_ERR_UNSUPPORTED, "grid does not use the tiled kernels");
        if (reps < 1)
            reps = 3;
        const bool divgrad = p->roS_scheme == FILT_SCHEME_DIVGRAD;
        if (divgrad && (!dev->u1x || !dev->u1y || !dev->u2x || !dev->u2y || !dev->roS1 ||
                        !dev->roS2 || !dev->roS1_old || !dev->roS2_old))
            return set_error(FILT_ERR_INVALID, "the divgrad scheme reads u and roS_old");
        static const int shapes[6][2] = {{64, 4}, {32, 8}, {16, 16}, {64, 8}, {128, 4}, {128, 8}};
        std::vector<int> kzs = {8, 16, 32, 64};
        if (nz > 64)
            kzs.push_back(nz);
        struct Ev {  // destroyed on every return
            hipEvent_t e = nullptr;
            ~Ev()
            {
                if (e) (void)hipEventDestroy(e);
            }
        } ev0, ev1;
        FILT_HIP_TRY(hipEventCreate(&ev0.e));
        FILT_HIP_TRY(hipEventCreate(&ev1.e));
        const hipEvent_t e0 = ev0.e, e1 = ev1.e;
        TileConfig best[2] = {cfg_for(0, nx, ny, nz), cfg_for(1, nx, ny, nz)};
        double best_t[2] = {1e30, 1e30};
        // the tiled roS kernels divide through DivBy: outside its divisor
        // range the stage runs untiled and only u is tuned
        const bool roS_ok = divby_ok(p->h1) && divby_ok(p->h2) && divby_ok(p->h3) &&
                            (!divgrad || divgrad_tiled_ok(*p));
        // The divgrad kernel reads roS1_old, roS2_old and rewrites them: all
        // four roS arrays are saved, the two inputs restored before each of
        // its launches, and all four restored at the end.
        const size_t bytes = sizeof(double) * (size_t)nx * ny * nz;
        double* const keep[4] = {dev->roS1_old, dev->roS2_old, dev->roS1, dev->roS2};
        struct Buf {
            double* p = nullptr;
            ~Buf()
            {
                if (p) {
                    (void)hipStreamSynchronize(stream());
                    (void)hipFree(p);
                }
            }
        } save;
        if (divgrad && roS_ok) {
            FILT_HIP_TRY(hipMalloc((void**)&save.p, 4 * bytes));
            for (int a = 0; a < 4; a++)
                FILT_HIP_TRY(hipMemcpyAsync((char*)save.p + a * bytes, keep[a], bytes,
                                            hipMemcpyDeviceToDevice, stream()));
        }
        // on an error return, the roS arrays go back to the state the call
        // found (declared after `save`, so it runs before `save` is freed)
        struct Restore {
            double* const* dst;
            const double* src;
            size_t bytes;
            bool armed;
            ~Restore()
            {
                if (armed)
                    for (int a = 0; a < 4; a++)
                        (void)hipMemcpyAsync(dst[a], (const char*)src + a * bytes, bytes,
                                             hipMemcpyDeviceToDevice, stream());
            }
        } restore{keep, save.p, bytes, save.p != nullptr};
        // the divgrad kernel is timed on its relaxed branch (t >= 2 dt)
        const double t_div = 2. * p->dt;
        for (int kern = roS_ok ? 0 : 1; kern < 2; kern++)
            for (auto& s : shapes)
                for (int kz : kzs)
                    for (int order = 0; order < 2; order++) {
                        const TileConfig t = {s[0], s[1], kz, order};
                        const TileMap m = make_map(nx, ny, nz, t);
                        std::vector<float> ms;
                        for (int r = 0; r <= reps; r++) {
                            if (kern == 0 && divgrad)
                                for (int a = 0; a < 2; a++)
                                    FILT_HIP_TRY(hipMemcpyAsync(keep[a], (char*)save.p + a * bytes,
                                                                bytes, hipMemcpyDeviceToDevice,
                                                                stream()));
                            FILT_HIP_TRY(hipEventRecord(e0, stream()));
                            if (kern == 0 && divgrad)
                                hipLaunchKernelGGL(pick_divgrad(t.tx, t.ty, dev->poro != nullptr),
                                                   dim3((unsigned)m.count()), dim3(t.tx, t.ty), 0,
                                                   stream(), *dev, *p, nx, m, t.kz, t_div);
                            else if (kern == 0)
                                hipLaunchKernelGGL(pick_roS(t.tx, t.ty), dim3((unsigned)m.count()),
                                                   dim3(t.tx, t.ty), 0, stream(), *dev, *p, nx, m, t.kz);
                            else
                                hipLaunchKernelGGL(pick_u(t.tx, t.ty), dim3((unsigned)m.count()),
                                                   dim3(t.tx, t.ty), 0, stream(), *dev, *p, nx, 0, 1,
                                                   m, t.kz);
                            FILT_LAUNCH_CHECK("autotune launch");
                            FILT_HIP_TRY(hipEventRecord(e1, stream()));
                            FILT_HIP_TRY(hipEventSynchronize(e1));
                            float v = 0;
                            FILT_HIP_TRY(hipEventElapsedTime(&v, e0, e1));
                            if (r > 0)  // the first launch warms caches and code
                                ms.push_back(v);
                        }
                        std::sort(ms.begin(), ms.end());
                        const double med = ms[ms.size() / 2];
                        if (med < best_t[kern]) {
                            best_t[kern] = med;
                            best[kern] = t;
                        }
                    }
        if (save.p) {
            restore.armed = false;
            for (int a = 0; a < 4; a++)
                FILT_HIP_TRY(hipMemcpyAsync(keep[a], (char*)save.p + a * bytes, bytes,
                                            hipMemcpyDeviceToDevice, stream()));
        }
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tuned[std::make_tuple(nx, ny, nz)] = std::make_pair(best[0], best[1]);
        }
        if (best_ms) {
            best_ms[0] = roS_ok ? best_t[0] : 0.0;
            best_ms[1] = best_t[1];
        }
        return FILT_OK;
    });
}

// Time every candidate tile configuration of the stages 2-4 kernel on the
// given grid and keep the fastest for grids of this shape.  The kernel reads
// S2 (the Newton's start) and rewrites it in place, so S2 is saved to a
// temporary array and restored before every timed launch; P1 is only
// written (the Newton starts from P2), and u, roS, roS_old follow from the
// inputs.  At the end S2 is restored once more and the chosen configuration
// runs once, so the grid holds what one filt_u_roS_P1_S2_calculation gives.
extern "C" int filt_autotune_step(const filt_fields* dev, const filt_params* p, int localNx,
                                  int reps, double* best_ms)
{
    return filt::guarded([&]() -> int {
        using namespace filt;
        int rc = check_shape(p, localNx);
        if (rc)
            return rc;
        if (!dev || !dev->P1 || !dev->S2)
            return set_error(FILT_ERR_INVALID, "null device fields");
        const int nx = localNx, ny = p->Ny, nz = p->Nz;
        if (!step_fusable(*p, nx, 0, 1))
            return set_error(FILT_ERR_UNSUPPORTED, "grid does not use the stages 2-4 kernel");
        if (reps < 1)
            reps = 3;
        const size_t bytes = sizeof(double) * (size_t)nx * ny * nz;
        struct Buf {
            double* p = nullptr;
            ~Buf()
            {
                if (p) {
                    (void)hipStreamSynchronize(stream());
                    (void)hipFree(p);
                }
            }
        } save;
        FILT_HIP_TRY(hipMalloc((void**)&save.p, bytes));
        FILT_HIP_TRY(hipMemcpyAsync(save.p, dev->S2, bytes, hipMemcpyDeviceToDevice, stream()));
        // on an error return, S2 goes back to the state the call found
        // (declared after `save`, so it runs before `save` is freed)
        struct Restore {
            double* dst;
            const double* src;
            size_t bytes;
            bool armed = true;
            ~Restore()
            {
                if (armed)
                    (void)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream());
            }
        } restore{dev->S2, save.p, bytes};
        struct Ev {
            hipEvent_t e = nullptr;
            ~Ev()
            {
                if (e) (void)hipEventDestroy(e);
            }
        } e0, e1;
        FILT_HIP_TRY(hipEventCreate(&e0.e));
        FILT_HIP_TRY(hipEventCreate(&e1.e));
        static const int shapes[][2] = {{64, 4},  {32, 8},  {16, 16}, {64, 8}, {32, 16},
                                        {64, 16}, {128, 4}, {128, 8}, {256, 4}};
        std::vector<int> kzs = {8, 16, 32, 64};
        if (nz > 64)
            kzs.push_back(nz);
        TileConfig best = cfg_for(2, nx, ny, nz);
        double best_t = 1e30;
        for (auto& s : shapes)
            for (int kz : kzs)
                for (int order = 0; order < 2; order++) {
                    const TileConfig t = {s[0], s[1], kz, order};
                    std::vector<float> ms;
                    for (int r = 0; r <= reps; r++) {
                        FILT_HIP_TRY(hipMemcpyAsync(dev->S2, save.p, bytes,
                                                    hipMemcpyDeviceToDevice, stream()));
                        FILT_HIP_TRY(hipEventRecord(e0.e, stream()));
                        rc = launch_step_zm_cfg(*dev, *p, nx, true, false, t);
                        if (rc)
                            return rc;
                        FILT_HIP_TRY(hipEventRecord(e1.e, stream()));
                        FILT_HIP_TRY(hipEventSynchronize(e1.e));
                        float v = 0;
                        FILT_HIP_TRY(hipEventElapsedTime(&v, e0.e, e1.e));
                        if (r > 0)  // the first launch warms caches and code
                            ms.push_back(v);
                    }
                    std::sort(ms.begin(), ms.end());
                    const double med = ms[ms.size() / 2];
                    if (med < best_t) {
                        best_t = med;
                        best = t;
                    }
                }
        FILT_HIP_TRY(hipMemcpyAsync(dev->S2, save.p, bytes, hipMemcpyDeviceToDevice, stream()));
        restore.armed = false;
        rc = launch_step_zm_cfg(*dev, *p, nx, true, false, best);
        if (rc)
            return rc;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_step_tuned[std::make_tuple(nx, ny, nz)] = best;
        }
        if (best_ms)
            *best_ms = best_t;
        return FILT_OK;
    });
}

extern "C" int filt_get_tile_config(int kernel, int nx, int ny, int nz, int* tx, int* ty, int* kz,
                                    int* order)
{
    if (kernel < 0 || kernel > 3)
        return filt::set_error(FILT_ERR_INVALID,
                               "kernel must be 0 (roS), 1 (u), 2 (stages 2-4) or 3 (step)");
    const filt::TileConfig t = filt::cfg_for(kernel, nx, ny, nz);
    if (tx) *tx = t.tx;
    if (ty) *ty = t.ty;
    if (kz) *kz = t.kz;
    if (order) *order = t.order;
    return FILT_OK;
}

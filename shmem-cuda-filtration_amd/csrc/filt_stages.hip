// filt_stages.hip -- the plain FP64 path: one kernel per defines.h stage.
//
// Drop-in semantics: every array of ptr_Arrays is materialised in HBM exactly
// as the reference's stage functions leave it (cpu.cpp:3-54), so a caller
// that reads any array between stages sees the reference's values.
//
// Launch shape of the cell kernels: blocks of 64 x 4 threads (one wavefront
// per x row of 64 cells, four rows), grid (ceil(nx/64), ceil(Ny/4), Nz).
// Loads are x-contiguous (coalesced); the stencil neighbours at j+-1 / k+-1 are
// served by L2 / MALL since neighbouring rows and planes are in flight at
// the same time.
#include <hip/hip_runtime.h>

#include "../../include/filtration_hip.h"
#include "filt_device.hpp"
#include "filt_runtime.hpp"

using namespace filt;

namespace {

constexpr int TX = 64;
constexpr int TY = 4;

struct Cell {
    int i, j, k;
    long long c;
    bool valid;
};

__device__ inline Cell cell_of_thread(int nx, int ny)
{
    Cell e;
    e.i = blockIdx.x * TX + threadIdx.x;
    e.j = blockIdx.y * TY + threadIdx.y;
    e.k = blockIdx.z;
    e.valid = e.i < nx && e.j < ny;
    e.c = idx3(e.i, e.j, e.k, nx, ny);
    return e;
}

// Medium index of a cell.  With a single medium in the table the media array
// is not read (4 bytes per cell and stage saved).
__device__ inline int media_of(const filt_fields& f, const filt_params& p, long long c)
{
    return (f.media && p.n_media > 1) ? f.media[c] : 0;
}

// ---------------------------------------------------------------- stage 1 --
// cpu.cpp:3-15, 57-74.  Bytes per cell: media 4 + S2 8 + P1 8 in,
// P2, Xi1, Xi2, ro1, ro2 40 out.
//
// BC_A (the step's boundary folded into its kernels, one rank, every extent
// >= 3): a cell (0, j, k) of the i = 0 face whose source (1, j, k) is a shell
// cell also takes its post-sweep value, the pre-sweep (1, j, k), after its
// own stage 1 has read its old value.  Nothing else in this launch writes
// (1, j, k), and the rest of the sweep, which reads only post-Newton values,
// is written by the stages 2-4 kernel (k_step_zm with BC).
template <bool BC_A>
__global__ __launch_bounds__(TX* TY) void k_ro_P2_Xi(filt_fields f, filt_params p, int nx,
                                                     int rank, int size)
{
    const Cell e = cell_of_thread(nx, p.Ny);
    if (!e.valid || !is_active(e.i, nx, rank, size))
        return;
    const int md = media_of(f, p, e.c);
    const double K = f.perm ? f.perm[e.c] : p.K[md];
    const Stage1 r = stage1_cell(f.P1[e.c], f.S2[e.c], K, p.lambda[md], p.S_wr[md], p);
    // non-temporal: the five arrays (40 B/cell) stream out past the caches
    __builtin_nontemporal_store(r.P2, f.P2 + e.c);
    __builtin_nontemporal_store(r.Xi1, f.Xi1 + e.c);
    __builtin_nontemporal_store(r.Xi2, f.Xi2 + e.c);
    __builtin_nontemporal_store(r.ro1, f.ro1 + e.c);
    __builtin_nontemporal_store(r.ro2, f.ro2 + e.c);
    if constexpr (BC_A) {
        if (e.i == 0 && !is_interior(1, e.j, e.k, nx, p.Ny, p.Nz)) {
            f.P1[e.c] = f.P1[e.c + 1];
            f.S2[e.c] = f.S2[e.c + 1];
        }
    }
}

// k_ro_P2_Xi for one medium (no media lookups) with 1 - S_wr, mu1 and mu2 in
// DivBy's range: the three divisions by them go through DivBy, and the two
// Brooks-Corey exponents (2 + 3 lam) / lam and (2 + lam) / lam are divided
// once per block; thread 0 puts them in LDS before the block's one barrier.
// The same expressions as stage1_cell, so the same bits.
template <bool BC_A>
__global__ __launch_bounds__(TX* TY) void k_ro_P2_Xi_one(filt_fields f, filt_params p, int nx,
                                                         int rank, int size)
{
    __shared__ double u[5];  // reciprocals of 1 - S_wr, mu1, mu2; the exponents
    const int t = threadIdx.x + TX * threadIdx.y;
    if (t == 0) {
        const double lam = p.lambda[0];
        u[0] = refined_reciprocal(1. - p.S_wr[0]);
        u[1] = refined_reciprocal(p.mu1);
        u[2] = refined_reciprocal(p.mu2);
        u[3] = (2. + 3. * lam) / lam;
        u[4] = (2. + lam) / lam;
    }
    __syncthreads();
    const Cell e = cell_of_thread(nx, p.Ny);
    if (!e.valid || !is_active(e.i, nx, rank, size))
        return;
    const double swr = p.S_wr[0];
    const double K = f.perm ? f.perm[e.c] : p.K[0];
    const double P1 = f.P1[e.c], S2 = f.S2[e.c];
    const DivBy dd{1. - swr, u[0]}, dmu1{p.mu1, u[1]}, dmu2{p.mu2, u[2]};
    const double S_e = dd(1. - S2 - swr);
    const double k1 = pow_spec(S_e, u[3]);
    const double k2 = (1. - S_e) * (1. - S_e) * (1 - pow_spec(S_e, u[4]));
    const double P_k = 0;  // capillary pressure disabled, cpu.cpp:63
    const double P2 = P1 + P_k;
    __builtin_nontemporal_store(P2, f.P2 + e.c);
    __builtin_nontemporal_store(dmu1(-1. * K * k1), f.Xi1 + e.c);
    __builtin_nontemporal_store(dmu2(-1. * K * k2), f.Xi2 + e.c);
    __builtin_nontemporal_store(ro1_of(P1, p), f.ro1 + e.c);
    __builtin_nontemporal_store(ro2_of(P2, p), f.ro2 + e.c);
    if constexpr (BC_A) {
        if (e.i == 0 && !is_interior(1, e.j, e.k, nx, p.Ny, p.Nz)) {
            f.P1[e.c] = f.P1[e.c + 1];
            f.S2[e.c] = f.S2[e.c + 1];
        }
    }
}

// Stage 1 takes k_ro_P2_Xi_one where it applies.
bool stage1_one(const filt_fields& f, const filt_params& p)
{
    return !(f.media && p.n_media > 1) && divby_ok(1. - p.S_wr[0]) && divby_ok(p.mu1) &&
           divby_ok(p.mu2);
}

// ---------------------------------------------------------------- stage 2 --
// cpu.cpp:17-24, 77-160.  Bytes per cell: Xi1, Xi2, ro1, ro2, P1, P2 in,
// u1x..u2z out (P neighbours from cache).
__global__ __launch_bounds__(TX* TY) void k_u(filt_fields f, filt_params p, int nx, int rank,
                                              int size)
{
    const Cell e = cell_of_thread(nx, p.Ny);
    if (!e.valid || !is_active(e.i, nx, rank, size))
        return;
    const int i = e.i, j = e.j, k = e.k, ny = p.Ny;
    const long long c = e.c;
    const double* __restrict__ P1 = f.P1;
    const double* __restrict__ P2 = f.P2;
    const double X1 = f.Xi1[c], X2 = f.Xi2[c];
    const double g = p.g_const;

    if (p.NX > 2) {
        if (i == 0) {
            const long long xp = c + 1;
            f.u1x[c] = X1 * (P1[xp] - P1[c]) / p.h1;
            f.u2x[c] = X2 * (P2[xp] - P2[c]) / p.h1;
        } else if (i == nx - 1) {
            const long long xm = c - 1;
            f.u1x[c] = X1 * (P1[c] - P1[xm]) / p.h1;
            f.u2x[c] = X2 * (P2[c] - P2[xm]) / p.h1;
        } else {
            f.u1x[c] = X1 * ((P1[c + 1] - P1[c - 1]) / (2. * p.h1));
            f.u2x[c] = X2 * ((P2[c + 1] - P2[c - 1]) / (2. * p.h1));
        }
    } else {
        f.u1x[c] = 0;
        f.u2x[c] = 0;
    }

    if (ny > 2) {
        const double r1 = f.ro1[c], r2 = f.ro2[c];
        const long long sy = nx;
        if (j == 0) {
            f.u1y[c] = X1 * ((P1[c + sy] - P1[c]) / p.h2 - r1 * g);
            f.u2y[c] = X2 * ((P2[c + sy] - P2[c]) / p.h2 - r2 * g);
        } else if (j == ny - 1) {
            f.u1y[c] = X1 * ((P1[c] - P1[c - sy]) / p.h2 - r1 * g);
            f.u2y[c] = X2 * ((P2[c] - P2[c - sy]) / p.h2 - r2 * g);
        } else {
            f.u1y[c] = X1 * ((P1[c + sy] - P1[c - sy]) / (2. * p.h2) - r1 * g);
            f.u2y[c] = X2 * ((P2[c + sy] - P2[c - sy]) / (2. * p.h2) - r2 * g);
        }
    } else {
        f.u1y[c] = 0;
        f.u2y[c] = 0;
    }

    if (p.Nz > 2) {
        const long long sz = (long long)nx * ny;
        if (k == 0) {
            f.u1z[c] = X1 * (P1[c + sz] - P1[c]) / p.h3;
            f.u2z[c] = X2 * (P2[c + sz] - P2[c]) / p.h3;
        } else if (k == p.Nz - 1) {
            f.u1z[c] = X1 * (P1[c] - P1[c - sz]) / p.h3;
            f.u2z[c] = X2 * (P2[c] - P2[c - sz]) / p.h3;
        } else {
            f.u1z[c] = X1 * ((P1[c + sz] - P1[c - sz]) / (2. * p.h3));
            f.u2z[c] = X2 * ((P2[c + sz] - P2[c - sz]) / (2. * p.h3));
        }
    } else {
        f.u1z[c] = 0;
        f.u2z[c] = 0;
    }
}

// ---------------------------------------------------------------- stage 3 --
// cpu.cpp:26-33, 218-291 (interior cells).  Bytes per cell: media, S2, P1,
// P2, Xi1, Xi2, ro1, ro2 in; roS1, roS2, roS1_old, roS2_old out.
__global__ __launch_bounds__(TX* TY) void k_roS(filt_fields f, filt_params p, int nx, int rank,
                                                int size)
{
    const Cell e = cell_of_thread(nx, p.Ny);
    if (!e.valid || !is_active(e.i, nx, rank, size))
        return;
    const int ny = p.Ny, nz = p.Nz;
    if (!is_interior(e.i, e.j, e.k, nx, ny, nz))
        return;
    const long long c = e.c;
    const int md = media_of(f, p, c);
    const double* __restrict__ P1 = f.P1;
    const double* __restrict__ P2 = f.P2;
    const double* __restrict__ X1 = f.Xi1;
    const double* __restrict__ X2 = f.Xi2;
    const double* __restrict__ r1 = f.ro1;
    const double* __restrict__ r2 = f.ro2;

    const double S = f.S2[c];
    const double roS1 = r1[c] * (1 - S);
    const double roS2 = r2[c] * S;
    const double p1 = P1[c], p2 = P2[c];
    const double X1c = X1[c], X2c = X2[c], r1c = r1[c], r2c = r2[c];
    double a1, a2, f3, g3;

    if (nz < 2) {
        f3 = 0;
        g3 = 0;
    } else {
        const long long sz = (long long)nx * ny;
        const long long zp = c + sz, zm = c - sz;
        a2 = -(P1[zp] - p1) / p.h3;
        a1 = -(p1 - P1[zm]) / p.h3;
        f3 = upwind(a2, a1, X1c, r1c, X1[zm], r1[zm], X1[zp], r1[zp], p.h3);
        a2 = -(P2[zp] - p2) / p.h3;
        a1 = -(p2 - P2[zm]) / p.h3;
        g3 = upwind(a2, a1, X2c, r2c, X2[zm], r2[zm], X2[zp], r2[zp], p.h3);
    }
    const long long xp = c + 1, xm = c - 1, yp = c + nx, ym = c - nx;
    const double g = p.g_const;

    a2 = -(P1[xp] - p1) / p.h1;
    a1 = -(p1 - P1[xm]) / p.h1;
    const double f1 = upwind(a2, a1, X1c, r1c, X1[xm], r1[xm], X1[xp], r1[xp], p.h1);
    a2 = -(P1[yp] - p1) / p.h2 + r1c * g;
    a1 = -(p1 - P1[ym]) / p.h2 + r1c * g;
    const double f2 = upwind(a2, a1, X1c, r1c, X1[ym], r1[ym], X1[yp], r1[yp], p.h2);

    a2 = -(P2[xp] - p2) / p.h1;
    a1 = -(p2 - P2[xm]) / p.h1;
    const double g1 = upwind(a2, a1, X2c, r2c, X2[xm], r2[xm], X2[xp], r2[xp], p.h1);
    a2 = -(P2[yp] - p2) / p.h2 + r2c * g;
    a1 = -(p2 - P2[ym]) / p.h2 + r2c * g;
    const double g2 = upwind(a2, a1, X2c, r2c, X2[ym], r2[ym], X2[yp], r2[yp], p.h2);

    const double phi = f.poro ? f.poro[c] : p.m[md];
    f.roS1_old[c] = roS1;
    f.roS2_old[c] = roS2;
    f.roS1[c] = roS1 - (p.dt / phi) * (f1 + f2 + f3);
    f.roS2[c] = roS2 - (p.dt / phi) * (g1 + g2 + g3);
}

// ------------------------------------------------- stage 3, divgrad scheme --
// cpu.cpp:162-216 (assign_rS), selected by p.roS_scheme: diffusion of ro*S
// with coefficient m*l*c/2, central transport of ro*u, and a tau relaxation
// through roS_old (plain Euler while t < 2 dt).  The z transport terms are
// the reference's own: water reads ro1*u1x at (i+1, j, k+1) and (i, j, k-1),
// NAPL reads ro2*u2y at k+-1.  One thread per cell; this scheme is not on the
// benchmarked path, so neighbours come straight from global memory (L2).
__global__ __launch_bounds__(TX* TY) void k_roS_divgrad(filt_fields f, filt_params p, int nx,
                                                        int rank, int size, double t)
{
    const Cell e = cell_of_thread(nx, p.Ny);
    if (!e.valid || !is_active(e.i, nx, rank, size))
        return;
    const int ny = p.Ny, nz = p.Nz;
    if (!is_interior(e.i, e.j, e.k, nx, ny, nz))
        return;
    const long long c = e.c;
    const int md = media_of(f, p, c);
    const double* __restrict__ r1 = f.ro1;
    const double* __restrict__ r2 = f.ro2;
    const double* __restrict__ S = f.S2;
    const double roS1 = r1[c] * (1 - S[c]);
    const double roS2 = r2[c] * S[c];
    const double m = f.poro ? f.poro[c] : p.m[md];
    const double k1 = m * p.l_w * p.c / 2.;
    const double k2 = m * p.l_n * p.c / 2.;
    double dg1, dg2, Tz1, Tz2;
    if (nz < 2) {
        dg1 = 0;
        dg2 = 0;
        Tz1 = 0;
        Tz2 = 0;
    } else {
        const long long sz = (long long)nx * ny;
        const long long zp = c + sz, zm = c - sz, xz = c + 1 + sz;
        dg1 = k1 * (r1[zp] * (1. - S[zp]) - 2 * r1[c] * (1. - S[c]) + r1[zm] * (1. - S[zm])) /
              (p.h3 * p.h3);
        dg2 = k2 * (r2[zp] * S[zp] - 2 * r2[c] * S[c] + r2[zm] * (S[zm])) / (p.h3 * p.h3);
        Tz1 = (r1[xz] * f.u1x[xz] - r1[zm] * f.u1x[zm]) / (2. * p.h3);
        Tz2 = (r2[zp] * f.u2y[zp] - r2[zm] * f.u2y[zm]) / (2. * p.h3);
    }
    const long long xp = c + 1, xm = c - 1, yp = c + nx, ym = c - nx;
    dg1 += k1 * ((r1[xp] * (1. - S[xp]) - 2 * r1[c] * (1. - S[c]) + r1[xm] * (1. - S[xm])) /
                     (p.h1 * p.h1) +
                 (r1[yp] * (1. - S[yp]) - 2 * r1[c] * (1. - S[c]) + r1[ym] * (1. - S[ym])) /
                     (p.h2 * p.h2));
    dg2 += k2 * ((r2[xp] * S[xp] - 2 * r2[c] * S[c] + r2[xm] * (S[xm])) / (p.h1 * p.h1) +
                 (r2[yp] * S[yp] - 2 * r2[c] * S[c] + r2[ym] * (S[ym])) / (p.h2 * p.h2));
    const double Tx1 = (r1[xp] * f.u1x[xp] - r1[xm] * f.u1x[xm]) / (2. * p.h1);
    const double Ty1 = (r1[yp] * f.u1y[yp] - r1[ym] * f.u1y[ym]) / (2. * p.h2);
    const double Tx2 = (r2[xp] * f.u2x[xp] - r2[xm] * f.u2x[xm]) / (2. * p.h1);
    const double Ty2 = (r2[yp] * f.u2y[yp] - r2[ym] * f.u2y[ym]) / (2. * p.h2);
    const double dt = p.dt, tau = p.tau;
    double A1, A2;
    if (t < 2 * dt) {
        A1 = roS1 + (dt / m) * (dg1 - Tx1 - Ty1 - Tz1);
        A2 = roS2 + (dt / m) * (dg2 - Tx2 - Ty2 - Tz2);
    } else {
        A1 = (2. * dt * dt) / (m * (dt + 2. * tau)) *
             (dg1 - Tx1 - Ty1 - Tz1 + (2. * roS1 * m * tau) / (dt * dt) +
              f.roS1_old[c] * m * (dt - 2. * tau) / (2. * dt * dt));
        A2 = (2. * dt * dt) / (m * (dt + 2. * tau)) *
             (dg2 - Tx2 - Ty2 - Tz2 + (2. * roS2 * m * tau) / (dt * dt) +
              f.roS2_old[c] * m * (dt - 2. * tau) / (2. * dt * dt));
    }
    f.roS1_old[c] = roS1;
    f.roS2_old[c] = roS2;
    f.roS1[c] = A1;
    f.roS2[c] = A2;
}

// ---------------------------------------------------------------- stage 4 --
// cpu.cpp:35-42, 293-320 (interior cells).  Bytes per cell: media, S2, P1,
// roS1, roS2 in; P1, S2 out.
__global__ __launch_bounds__(TX* TY) void k_newton(filt_fields f, filt_params p, int nx, int rank,
                                                   int size)
{
    const Cell e = cell_of_thread(nx, p.Ny);
    if (!e.valid || !is_active(e.i, nx, rank, size))
        return;
    if (!is_interior(e.i, e.j, e.k, nx, p.Ny, p.Nz))
        return;
    const long long c = e.c;
    const int md = media_of(f, p, c);
    double P = f.P1[c], S = f.S2[c];
    newton_cell(P, S, f.roS1[c], f.roS2[c], p.lambda[md], p.S_wr[md], p.P_d[md], p);
    __builtin_nontemporal_store(P, f.P1 + c);
    __builtin_nontemporal_store(S, f.S2 + c);
}

// ---------------------------------------------------------------- stage 5 --
// cpu.cpp:44-54, 323-403: an in-place sweep i -> j -> k where later shell
// cells read values earlier ones wrote.  For NX = localNx, Ny, Nz >= 3 on
// one rank the sweep has a closed form in terms of the post-Newton interior
// values n(.), the pre-sweep shell values, the source patch and ro1(i,1,0):
//
//   i = 0 face:            S, P = pre-sweep values at (1, j, k)
//   1 <= i <= N-2:
//     j = M-1:             S = n_S(i, M-2, c(k)),
//                          P = n_P(i, M-2, c(k)) + ro1(i,1,0)*g*h2
//     j = 0:               P = P_atm;  S = src(i,k) for 1 <= k <= L-2,
//                          src(i, L-2) at k = L-1, and at k = 0 the
//                          pre-sweep S(i, 0, 1) (the j = 0 rule falls
//                          through to the k = 0 copy, cpu.cpp:343-356)
//     1<=j<=M-2, k=0/L-1:  S, P = n(i, j, c(k))
//   i = N-1 face:          the post-sweep value of (N-2, j, k)
// with c(k) = clamp(k, 1, L-2).  Two kernels apply it without hazards:
// BC-A writes everything that reads pre-sweep shell values (the i = 0 face
// and S at (i, 0, 0), i >= 1), BC-B everything else from interior values.

__device__ inline double src_value(int i, int k, int rank, int size, const filt_params& p)
{
    return in_patch(i_to_I(i, rank, size, p), k, p) ? p.S2_gr : 0.0;
}

__global__ __launch_bounds__(256) void k_bc_a(filt_fields f, filt_params p, int nx)
{
    const int ny = p.Ny, nz = p.Nz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long face = (long long)ny * nz;
    double* __restrict__ S2 = f.S2;
    double* __restrict__ P1 = f.P1;
    if (t < face) {
        const int j = (int)(t % ny), k = (int)(t / ny);
        const long long c0 = idx3(0, j, k, nx, ny);
        if (j == 0 && k == 0) {
            // (0,0,0) copies the pre-sweep (1,0,0), which this thread then
            // rewrites with the pre-sweep (1,0,1).
            const double s10 = S2[c0 + 1], p10 = P1[c0 + 1];
            S2[c0] = s10;
            P1[c0] = p10;
            S2[c0 + 1] = S2[idx3(1, 0, 1, nx, ny)];
        } else {
            S2[c0] = S2[c0 + 1];
            P1[c0] = P1[c0 + 1];
        }
        return;
    }
    const long long u = t - face;  // i = 2 .. N-1 along the (i, 0, 0) line
    if (u >= nx - 2)
        return;
    const int i = 2 + (int)u;
    const int ii = i < nx - 1 ? i : nx - 2;
    S2[idx3(i, 0, 0, nx, ny)] = S2[idx3(ii, 0, 1, nx, ny)];
}

// Post-sweep value of shell cell (i, j, k), 1 <= i <= N-2, from the closed
// form.  Returns false for S at (i, 0, 0), which BC-A owns.
__device__ inline bool row_rule(const filt_fields& f, const filt_params& p, int nx, int i, int j,
                                int k, double& S, double& P)
{
    const int ny = p.Ny, nz = p.Nz;
    const int kc = k < 1 ? 1 : (k > nz - 2 ? nz - 2 : k);
    if (j == ny - 1) {
        const long long c = idx3(i, ny - 2, kc, nx, ny);
        S = f.S2[c];
        P = f.P1[c] + f.ro1[idx3(i, 1, 0, nx, ny)] * p.g_const * p.h2;
        return true;
    }
    if (j == 0) {
        P = p.P_atm;
        if (k == 0)
            return false;
        S = src_value(i, kc, 0, 1, p);
        return true;
    }
    const long long c = idx3(i, j, kc, nx, ny);
    S = f.S2[c];
    P = f.P1[c];
    return true;
}

__global__ __launch_bounds__(256) void k_bc_b(filt_fields f, filt_params p, int nx)
{
    const int ny = p.Ny, nz = p.Nz;
    const long long n1 = nx - 1, my = ny - 2, lz = nz - 2;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < 2 * n1 * nz) {  // j = 0 and j = M-1 faces, i >= 1
        j = t < n1 * nz ? 0 : ny - 1;
        if (j) t -= n1 * nz;
        i = 1 + (int)(t % n1);
        k = (int)(t / n1);
    } else if ((t -= 2 * n1 * nz) < 2 * n1 * my) {  // k = 0 and k = L-1, 1 <= j <= M-2
        k = t < n1 * my ? 0 : nz - 1;
        if (k) t -= n1 * my;
        i = 1 + (int)(t % n1);
        j = 1 + (int)(t / n1);
    } else if ((t -= 2 * n1 * my) < my * lz) {  // i = N-1 face, inner part
        i = nx - 1;
        j = 1 + (int)(t % my);
        k = 1 + (int)(t / my);
    } else {
        return;
    }
    double S = 0, P = 0;
    bool write_s;
    if (i == nx - 1) {
        if (is_interior(nx - 2, j, k, nx, ny, nz)) {
            const long long c = idx3(nx - 2, j, k, nx, ny);
            S = f.S2[c];
            P = f.P1[c];
            write_s = true;
        } else {
            write_s = row_rule(f, p, nx, nx - 2, j, k, S, P);
        }
    } else {
        write_s = row_rule(f, p, nx, i, j, k, S, P);
    }
    const long long c = idx3(i, j, k, nx, ny);
    if (write_s)
        f.S2[c] = S;
    f.P1[c] = P;
}

// Exact replay of the reference sweep on one thread, for grids the closed
// form does not cover (an extent below 3, or several ranks).  Visits the
// cells where a rule can fire, in i -> j -> k order.
__device__ void border_S2_dev(const filt_fields& f, const filt_params& p, int nx, int i, int j,
                              int k, int rank, int size)
{
    const int ny = p.Ny, nz = p.Nz;
    double* S2 = f.S2;
    const long long c = idx3(i, j, k, nx, ny);
    if (i == 0 && p.NX > 2) { S2[c] = S2[c + 1]; return; }
    if (i == nx - 1 && p.NX > 2) { S2[c] = S2[c - 1]; return; }
    if (j == ny - 1 && ny > 2) { S2[c] = S2[c - nx]; return; }
    if (j == 0 && ny > 2)
        S2[c] = in_patch(i_to_I(i, rank, size, p), k, p) ? p.S2_gr : 0.0;
    const long long sz = (long long)nx * ny;
    if (k == 0 && nz > 2) { S2[c] = S2[c + sz]; return; }
    if (k == nz - 1 && nz > 2) { S2[c] = S2[c - sz]; return; }
}

__device__ void border_P1_dev(const filt_fields& f, const filt_params& p, int nx, int i, int j,
                              int k)
{
    const int ny = p.Ny, nz = p.Nz;
    double* P1 = f.P1;
    const long long c = idx3(i, j, k, nx, ny);
    if (i == 0 && p.NX > 2) { P1[c] = P1[c + 1]; return; }
    if (i == nx - 1 && p.NX > 2) { P1[c] = P1[c - 1]; return; }
    if (j == ny - 1 && ny > 2) { P1[c] = P1[c - nx] + f.ro1[i + nx * 1] * p.g_const * p.h2; return; }
    if (j == 0 && ny > 2) { P1[c] = p.P_atm; return; }
    const long long sz = (long long)nx * ny;
    if (k == 0 && nz > 2) { P1[c] = P1[c + sz]; return; }
    if (k == nz - 1 && nz > 2) { P1[c] = P1[c - sz]; return; }
}

__global__ void k_bc_sequential(filt_fields f, filt_params p, int nx, int rank, int size)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    const int ny = p.Ny, nz = p.Nz;
    for (int i = 0; i < nx; i++) {
        if (!is_active(i, nx, rank, size))
            continue;
        for (int j = 0; j < ny; j++) {
            const bool line = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1);
            for (int k = 0; k < nz; k++) {
                if (!line && k != 0 && k != nz - 1)
                    continue;
                border_S2_dev(f, p, nx, i, j, k, rank, size);
                border_P1_dev(f, p, nx, i, j, k);
            }
        }
    }
}

int launch_check(const filt_fields* dev, const filt_params* p, int localNx)
{
    int rc = check_shape(p, localNx);
    if (rc)
        return rc;
    if (!dev || !dev->P1 || !dev->S2)
        return set_error(FILT_ERR_INVALID, "null device fields");
    if (p->Nz > 65535 || (p->Ny + TY - 1) / TY > 65535)
        return set_error(FILT_ERR_INVALID, "grid extent exceeds the launch grid (Ny/4, Nz <= 65535)");
    if (p->n_media < 0 || p->n_media > FILT_MAX_MEDIA)
        return set_error(FILT_ERR_INVALID, "n_media out of range");
    if (p->newton_iterations < 0 || p->newton_iterations > 1000)
        return set_error(FILT_ERR_INVALID, "newton_iterations out of range");
    if (p->roS_scheme != FILT_SCHEME_UPWIND && p->roS_scheme != FILT_SCHEME_DIVGRAD)
        return set_error(FILT_ERR_INVALID, "unknown roS_scheme %d", p->roS_scheme);
    return FILT_OK;
}

dim3 cell_grid(const filt_params* p, int nx)
{
    return dim3((nx + TX - 1) / TX, (p->Ny + TY - 1) / TY, p->Nz);
}

// The z-marching stencil kernels (filt_stencil.hip) need every extent >= 3;
// smaller grids take the one-cell-per-thread kernels above.
bool use_tiled(const filt_params* p, int nx, int kernel)
{
    // the tiled roS kernel divides by h1, h2, h3 through DivBy
    const bool div_ok = kernel != 0 || (divby_ok(p->h1) && divby_ok(p->h2) && divby_ok(p->h3));
    return nx >= 3 && p->Ny >= 3 && p->Nz >= 3 && tiled_fits(nx, p->Ny, p->Nz) &&
           tiled_enabled(kernel) && div_ok;
}

}  // namespace

extern "C" {

int filt_ro_P2_Xi_calculation(const filt_fields* dev, const filt_params* p, int localNx, int rank,
                              int size)
{
    int rc = launch_check(dev, p, localNx);
    if (rc) return rc;
    hipLaunchKernelGGL(stage1_one(*dev, *p) ? k_ro_P2_Xi_one<false> : k_ro_P2_Xi<false>,
                       cell_grid(p, localNx), dim3(TX, TY), 0, stream(), *dev, *p, localNx, rank,
                       size);
    FILT_LAUNCH_CHECK("k_ro_P2_Xi");
    return FILT_OK;
}

int filt_u_calculation(const filt_fields* dev, const filt_params* p, int localNx, int rank,
                       int size)
{
    int rc = launch_check(dev, p, localNx);
    if (rc) return rc;
    if (use_tiled(p, localNx, 1) && p->NX > 2)
        return launch_u_zm(*dev, *p, localNx, rank, size);
    hipLaunchKernelGGL(k_u, cell_grid(p, localNx), dim3(TX, TY), 0, stream(), *dev, *p, localNx,
                       rank, size);
    FILT_LAUNCH_CHECK("k_u");
    return FILT_OK;
}

int filt_roS_calculation(const filt_fields* dev, const filt_params* p, double t, int localNx,
                         int rank, int size)
{
    int rc = launch_check(dev, p, localNx);
    if (rc) return rc;
    if (p->roS_scheme == FILT_SCHEME_DIVGRAD) {
        if (!dev->u1x || !dev->u1y || !dev->u2x || !dev->u2y || !dev->roS1_old || !dev->roS2_old)
            return set_error(FILT_ERR_INVALID, "the divgrad scheme reads u and roS_old");
        if (use_tiled(p, localNx, 0) && divgrad_tiled_ok(*p))
            return launch_roS_divgrad_zm(*dev, *p, localNx, t);
        hipLaunchKernelGGL(k_roS_divgrad, cell_grid(p, localNx), dim3(TX, TY), 0, stream(), *dev,
                           *p, localNx, rank, size, t);
        FILT_LAUNCH_CHECK("k_roS_divgrad");
        return FILT_OK;
    }
    // t is unused by the upwind scheme, as by assign_rS_nr (cpu.cpp:218)
    if (use_tiled(p, localNx, 0))
        return launch_roS_zm(*dev, *p, localNx);
    hipLaunchKernelGGL(k_roS, cell_grid(p, localNx), dim3(TX, TY), 0, stream(), *dev, *p, localNx,
                       rank, size);
    FILT_LAUNCH_CHECK("k_roS");
    return FILT_OK;
}

int filt_P1_S2_calculation(const filt_fields* dev, const filt_params* p, int localNx, int rank,
                           int size)
{
    int rc = launch_check(dev, p, localNx);
    if (rc) return rc;
    hipLaunchKernelGGL(k_newton, cell_grid(p, localNx), dim3(TX, TY), 0, stream(), *dev, *p,
                       localNx, rank, size);
    FILT_LAUNCH_CHECK("k_newton");
    return FILT_OK;
}

int filt_boundary_conditions(const filt_fields* dev, const filt_params* p, int localNx, int rank,
                             int size)
{
    int rc = launch_check(dev, p, localNx);
    if (rc) return rc;
    const int nx = localNx, ny = p->Ny, nz = p->Nz;
    const bool closed = size == 1 && nx == p->NX && nx >= 3 && ny >= 3 && nz >= 3;
    if (!closed) {
        hipLaunchKernelGGL(k_bc_sequential, dim3(1), dim3(64), 0, stream(), *dev, *p, nx, rank,
                           size);
        FILT_LAUNCH_CHECK("k_bc_sequential");
        return FILT_OK;
    }
    const long long na = (long long)ny * nz + (nx - 2);
    hipLaunchKernelGGL(k_bc_a, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, stream(), *dev,
                       *p, nx);
    FILT_LAUNCH_CHECK("k_bc_a");
    const long long nb = 2LL * (nx - 1) * nz + 2LL * (nx - 1) * (ny - 2) +
                         (long long)(ny - 2) * (nz - 2);
    hipLaunchKernelGGL(k_bc_b, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, stream(), *dev,
                       *p, nx);
    FILT_LAUNCH_CHECK("k_bc_b");
    return FILT_OK;
}

int filt_u_roS_P1_S2_calculation(const filt_fields* dev, const filt_params* p, double t,
                                 int localNx, int rank, int size, int with_u)
{
    int rc = launch_check(dev, p, localNx);
    if (rc) return rc;
    if (step_fusable(*p, localNx, rank, size)) {
        if (!dev->P2 || !dev->Xi1 || !dev->Xi2 || !dev->ro1 || !dev->ro2 || !dev->roS1 ||
            !dev->roS2 || !dev->roS1_old || !dev->roS2_old ||
            (with_u && (!dev->u1x || !dev->u1y || !dev->u1z || !dev->u2x || !dev->u2y ||
                        !dev->u2z)))
            return set_error(FILT_ERR_INVALID, "null device fields");
        return launch_step_zm(*dev, *p, localNx, with_u != 0, false);
    }
    // the divgrad scheme reads this step's u (cpu.cpp:187-190)
    if ((with_u || p->roS_scheme == FILT_SCHEME_DIVGRAD) &&
        (rc = filt_u_calculation(dev, p, localNx, rank, size)))
        return rc;
    if ((rc = filt_roS_calculation(dev, p, t, localNx, rank, size))) return rc;
    return filt_P1_S2_calculation(dev, p, localNx, rank, size);
}

int filt_time_step_timed(const filt_fields* dev, const filt_params* p, double t, int localNx,
                         int rank, int size, int with_u, void* const* events, int* path)
{
    auto mark = [&](int q) -> int {
        if (events && events[q])
            FILT_HIP_TRY(hipEventRecord((hipEvent_t)events[q], stream()));
        return FILT_OK;
    };
    int rc;
    const bool full = dev && p && full_step_fusable(*p, localNx, rank, size,
                                                   dev->media && p->n_media > 1);
    if (path)
        *path = full ? 1 : 0;
    if ((rc = mark(0))) return rc;
    if (full) {
        // stage 1 and stages 2-4 in one launch, the stash copy, the boundary
        if ((rc = launch_check(dev, p, localNx))) return rc;
        if (!dev->P2 || !dev->Xi1 || !dev->Xi2 || !dev->ro1 || !dev->ro2 || !dev->roS1 ||
            !dev->roS2 || !dev->roS1_old || !dev->roS2_old ||
            (with_u && (!dev->u1x || !dev->u1y || !dev->u1z || !dev->u2x || !dev->u2y ||
                        !dev->u2z)))
            return set_error(FILT_ERR_INVALID, "null device fields");
        if ((rc = launch_step_full(*dev, *p, localNx, with_u != 0,
                                   events ? (hipEvent_t)events[1] : nullptr)))
            return rc;
    } else if (dev && p && step_fusable(*p, localNx, rank, size)) {
        // stage 1 with the boundary cells that read pre-sweep shell values,
        // then stages 2-4 with every other boundary cell: no boundary launch
        if ((rc = launch_check(dev, p, localNx))) return rc;
        if (!dev->P2 || !dev->Xi1 || !dev->Xi2 || !dev->ro1 || !dev->ro2 || !dev->roS1 ||
            !dev->roS2 || !dev->roS1_old || !dev->roS2_old ||
            (with_u && (!dev->u1x || !dev->u1y || !dev->u1z || !dev->u2x || !dev->u2y ||
                        !dev->u2z)))
            return set_error(FILT_ERR_INVALID, "null device fields");
        hipLaunchKernelGGL(stage1_one(*dev, *p) ? k_ro_P2_Xi_one<true> : k_ro_P2_Xi<true>,
                           cell_grid(p, localNx), dim3(TX, TY), 0, stream(), *dev, *p, localNx,
                           rank, size);
        FILT_LAUNCH_CHECK("k_ro_P2_Xi");
        if ((rc = mark(1))) return rc;
        if ((rc = launch_step_zm(*dev, *p, localNx, with_u != 0, true))) return rc;
        if ((rc = mark(2))) return rc;
        return mark(3);
    } else {
        if ((rc = filt_ro_P2_Xi_calculation(dev, p, localNx, rank, size))) return rc;
        if ((rc = mark(1))) return rc;
        if ((rc = filt_u_roS_P1_S2_calculation(dev, p, t, localNx, rank, size, with_u))) return rc;
    }
    if ((rc = mark(2))) return rc;
    if ((rc = filt_boundary_conditions(dev, p, localNx, rank, size))) return rc;
    return mark(3);
}

int filt_time_step(const filt_fields* dev, const filt_params* p, double t, int localNx, int rank,
                   int size, int with_u)
{
    return filt_time_step_timed(dev, p, t, localNx, rank, size, with_u, nullptr, nullptr);
}

}  // extern "C"

// filt_fused.hip -- host side of the fused step (the kernels are in
// filt_fused_tile.hpp, instantiated by filt_fused_fp64.hip,
// filt_fused_fast.hip and filt_fused_mixed.hip): descriptor preparation,
// the launch dispatch by form, and the deterministic mass sums.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "../../include/filtration_hip.h"
#include "filt_device.hpp"
#include "filt_fused.hpp"
#include "filt_runtime.hpp"

namespace filt {

namespace {

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
        const int md = (G.media && G.p.n_media > 1) ? G.media[c] : 0;
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

int fused_tile_index(int tx, int ty, int cy)
{
    for (int q = 0; q < N_FUSED_TILES; q++)
        if (FUSED_TILES[q].tx == tx && FUSED_TILES[q].ty == ty && FUSED_TILES[q].cy == cy)
            return q;
    return -1;
}

void fused_prepare(FusedGrid& G, int tile)
{
    const filt_params& p = G.p;
    G.inv_h[0] = 1.0 / p.h1;
    G.inv_h[1] = 1.0 / p.h2;
    G.inv_h[2] = 1.0 / p.h3;
    G.inv_mu1 = 1.0 / p.mu1;
    G.inv_mu2 = 1.0 / p.mu2;
    for (int q = 0; q < FILT_MAX_MEDIA; q++)
        G.inv_1mswr[q] = 1.0 / (1.0 - p.S_wr[q]);
    // the same products, in the same order, as the kernels would form (this
    // file is built without FP contraction)
    G.nq_a1 = p.ro01 * p.beta1;
    G.nq_c1 = p.ro02 * p.beta2;
    G.nq_al4 = 4.0 * (G.nq_a1 * G.nq_c1);
    G.nq_be0 = p.ro01 * G.nq_c1 + G.nq_a1 * p.ro02;
    G.nq_ga0 = p.ro01 * p.ro02;
    if (G.kz < 1)
        G.kz = 32;
    const int tx = FUSED_TILES[tile].tx, ty = FUSED_TILES[tile].ty;
    const int threads = tx * ty / FUSED_TILES[tile].cy;
    G.nbx = (G.nx - 1 + tx - 1) / tx;  // x-tiles from x = 0 over x <= nx - 2
    G.nby = (G.ny - 2 + ty - 1) / ty;
    G.nbz = (G.nz - 2 + G.kz - 1) / G.kz;
    G.n_interior_blocks = (long long)G.nbx * G.nby * G.nbz;
    const long long shell = 2LL * G.ny + 2LL * (G.nz - 2) + (long long)(G.nx - 1) * G.nz;
    G.n_shell_blocks = (shell + threads - 1) / threads;
}

bool fused_general_lambda(const FusedGrid& g)
{
    const int n = (g.media && g.p.n_media > 1) ? g.p.n_media : 1;
    for (int q = 0; q < n; q++)
        if (g.p.lambda[q] != 2.0)
            return true;
    return false;
}

bool fused_cell_fields(const FusedGrid& g)
{
    return g.perm && g.poro && !(g.media && g.p.n_media > 1);
}

bool fused_divisors_ok(const filt_params& p)
{
    if (!(divby_ok(p.h1) && divby_ok(p.h2) && divby_ok(p.h3) && divby_ok(p.mu1) &&
          divby_ok(p.mu2)))
        return false;
    const int n = p.n_media > 1 ? p.n_media : 1;
    for (int q = 0; q < n; q++)
        if (!divby_ok(1. - p.S_wr[q]))
            return false;
    return true;
}

int fused_launch(const FusedGrid* dev_grids, const int* dev_block_grid, long long total_blocks,
                 int parity, int precision, bool general_lambda, bool cell_fields, int tile,
                 int order, hipStream_t s)
{
    if (total_blocks <= 0)
        return FILT_OK;
    if (total_blocks > 0x7fffffffLL)
        return set_error(FILT_ERR_INVALID, "batch too large for one launch");
    if (tile < 0 || tile >= N_FUSED_TILES)
        return set_error(FILT_ERR_INVALID, "unknown fused tile %d", tile);
    switch (precision) {
    case FILT_PREC_FP64:
        return fused_launch_fp64(dev_grids, dev_block_grid, total_blocks, parity, general_lambda,
                                 cell_fields, tile, order, s);
    case FILT_PREC_FAST:
        return fused_launch_fast(dev_grids, dev_block_grid, total_blocks, parity, general_lambda,
                                 cell_fields, tile, order, s);
    case FILT_PREC_MIXED:
        return fused_launch_mixed(dev_grids, dev_block_grid, total_blocks, parity,
                                  general_lambda, cell_fields, tile, order, s);
    }
    return set_error(FILT_ERR_INVALID, "unknown precision %d", precision);
}

int fused_mass(const FusedGrid* dev_grid, const FusedGrid& G, int parity, double* dev_partial,
               double* dev_out, hipStream_t s)
{
    const long long nb = fused_mass_blocks(G);
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

// filt_fused.hpp -- device-side descriptor of one grid of a fused batch.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/filtration_hip.h"

namespace filt {

struct FusedGrid {
    double* P[2];         // ping-pong water pressure
    double* S[2];         // ping-pong NAPL saturation
    const double* perm;   // per-cell K (null: K[media])
    const double* poro;   // per-cell porosity (null: m[media])
    const int* media;     // null: medium 0
    int nx, ny, nz;
    int nbx, nby, nbz, kz;       // interior tiles
    long long n_interior_blocks;
    long long n_shell_blocks;
    long long first_block;       // global index of this grid's first block
    double inv_h[3];             // 1/h1, 1/h2, 1/h3 (mixed form)
    double inv_mu1, inv_mu2;     // mixed form
    double inv_1mswr[FILT_MAX_MEDIA];
    filt_params p;
};

// Fill the derived members (reciprocals, tile counts) from nx, ny, nz, kz, p.
void fused_prepare(FusedGrid& G);

// Whether a grid uses a medium with lambda != 2 (the fast form then needs
// its general Brooks-Corey powers).
bool fused_general_lambda(const FusedGrid& g);

// One step of every grid in the batch: blocks [0, total_blocks) map to grids
// through dev_block_grid; reads buffer `parity`, writes buffer parity ^ 1.
// general_lambda: some grid has fused_general_lambda.
int fused_launch(const FusedGrid* dev_grids, const int* dev_block_grid, long long total_blocks,
                 int parity, int precision, bool general_lambda, int order, hipStream_t s);

// Deterministic total water / NAPL mass of one grid into dev_out[2].
int fused_mass(const FusedGrid* dev_grid, const FusedGrid& G, int parity, double* dev_partial,
               double* dev_out, hipStream_t s);
long long fused_mass_blocks(const FusedGrid& G);

}  // namespace filt

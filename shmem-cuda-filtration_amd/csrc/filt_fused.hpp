// filt_fused.hpp -- descriptor of one grid of a fused batch and the host-side
// entry points of the fused step (filt_fused*.hip).
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
    int nbx, nby, nbz, kz;       // interior tiles of the launch's tile shape
    long long n_interior_blocks;
    long long n_shell_blocks;
    long long first_block;       // global index of this grid's first block
    double inv_h[3];             // 1/h1, 1/h2, 1/h3 (fast and mixed forms)
    double inv_mu1, inv_mu2;     // fast and mixed forms
    // the closed-form Newton's constant coefficients (fast and mixed forms):
    // ro01*beta1, ro02*beta2, 4*(ro01*beta1)*(ro02*beta2),
    // ro01*(ro02*beta2) + (ro01*beta1)*ro02, ro01*ro02
    double nq_a1, nq_c1, nq_al4, nq_be0, nq_ga0;
    double inv_1mswr[FILT_MAX_MEDIA];
    filt_params p;
};

// Tile shapes of the fused kernel: columns in x and y, columns per thread
// (stacked in y); each is a kernel instantiation, and the autotuner picks
// among them.
struct FusedTile {
    int tx, ty, cy;
};
constexpr FusedTile FUSED_TILES[] = {{16, 16, 1}, {32, 8, 1},  {64, 4, 1},  {32, 16, 1},
                                     {64, 8, 1},  {64, 8, 2},  {32, 16, 2}, {64, 16, 2},
                                     {128, 4, 1}};
constexpr int N_FUSED_TILES = sizeof(FUSED_TILES) / sizeof(FUSED_TILES[0]);
// Index of a tile shape in FUSED_TILES, -1 if none.
int fused_tile_index(int tx, int ty, int cy);

// Fill the derived members (reciprocals, tile counts) from nx, ny, nz, kz, p
// for tile shape tile (an index into FUSED_TILES).
void fused_prepare(FusedGrid& G, int tile);

// Whether a grid uses a medium with lambda != 2 (the kernels then carry the
// general Brooks-Corey powers).
bool fused_general_lambda(const FusedGrid& g);

// Whether a grid has per-cell K and porosity and a single medium (the kernels
// then read both fields unconditionally and look up no medium).
bool fused_cell_fields(const FusedGrid& g);

// Whether the FP64 form can divide by the grid's constant divisors (h1, h2,
// h3, mu1, mu2, 1 - S_wr) through DivBy: each in [2^-100, 2^101].
bool fused_divisors_ok(const filt_params& p);

// One step of every grid in the batch: blocks [0, total_blocks) map to grids
// through dev_block_grid; reads buffer `parity`, writes buffer parity ^ 1.
// general_lambda: some grid has fused_general_lambda; cell_fields: every grid
// has fused_cell_fields.
int fused_launch(const FusedGrid* dev_grids, const int* dev_block_grid, long long total_blocks,
                 int parity, int precision, bool general_lambda, bool cell_fields, int tile,
                 int order, hipStream_t s);

// Per-form launchers (filt_fused_fp64.hip, filt_fused_fast.hip,
// filt_fused_mixed.hip).
int fused_launch_fp64(const FusedGrid* g, const int* bg, long long total, int parity, bool general,
                      bool fields, int tile, int order, hipStream_t s);
int fused_launch_fast(const FusedGrid* g, const int* bg, long long total, int parity, bool general,
                      bool fields, int tile, int order, hipStream_t s);
int fused_launch_mixed(const FusedGrid* g, const int* bg, long long total, int parity,
                       bool general, bool fields, int tile, int order, hipStream_t s);

// Deterministic total water / NAPL mass of one grid into dev_out[2].
int fused_mass(const FusedGrid* dev_grid, const FusedGrid& G, int parity, double* dev_partial,
               double* dev_out, hipStream_t s);
long long fused_mass_blocks(const FusedGrid& G);

}  // namespace filt

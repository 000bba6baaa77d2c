// filt_tiles.hpp -- launch geometry of the z-marching stencil kernels.
//
// A block owns an x-y tile of TX x TY columns and marches along z over a
// chunk of kz planes, keeping its own column's k-1, k, k+1 values in
// registers and staging the plane's x/y neighbours through LDS.
//
// MI355X dispatches workgroup b to XCD b % 8 and each XCD has its own L2.
// TileMap renumbers blocks so that the tiles an XCD runs at the same time
// are neighbours: with order 0 the logical tile index L = xcd*C + b/8
// (C = tiles per XCD) runs x-tiles fastest, then y-tiles, then z-chunks, so
// an XCD works through whole z-chunks and the x/y halos its tiles read come
// from tiles on the same XCD.  order 1 keeps the hardware numbering.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/filtration_hip.h"

namespace filt {

struct TileMap {
    int nbx, nby, nbz;  // tiles along x, y and z-chunks along z
    int order;          // 0: XCD-grouped, 1: identity

    __host__ __device__ long long count() const { return (long long)nbx * nby * nbz; }

    __device__ inline void decode(long long b, int& bx, int& by, int& bz) const
    {
        const long long T = count();
        long long L = b;
        if (order == 0 && T % 8 == 0) {
            const long long C = T / 8;
            L = (b % 8) * C + b / 8;
        }
        bx = (int)(L % nbx);
        const long long r = L / nbx;
        by = (int)(r % nby);
        bz = (int)(r / nby);
    }
};

// Tile configuration of one stencil kernel.
struct TileConfig {
    int tx, ty;  // threads per block in x and y
    int kz;      // planes per z-chunk
    int order;   // TileMap order
};

// The fused step's configuration for a grid (filt_stencil.hip): autotuned
// for this grid shape (filt_autotune_step), else the set or built-in one.
TileConfig step_tile_config(int nx, int ny, int nz);
// The stages 2-4 kernel with an explicit tile configuration (filt_step.hip),
// one that step_tile_ok accepts.
int launch_step_zm_cfg(const filt_fields& f, const filt_params& p, int nx, bool with_u, bool bc,
                       const TileConfig& t);
// The whole-step kernel's configuration (stage 1 with stages 2-4).
TileConfig full_tile_config(int nx, int ny, int nz);

// Tiles and z-chunks covering an nx * ny * nz grid.
inline TileMap make_map(int nx, int ny, int nz, const TileConfig& t)
{
    TileMap m;
    m.nbx = (nx + t.tx - 1) / t.tx;
    m.nby = (ny + t.ty - 1) / t.ty;
    m.nbz = (nz + t.kz - 1) / t.kz;
    m.order = t.order;
    return m;
}

}  // namespace filt

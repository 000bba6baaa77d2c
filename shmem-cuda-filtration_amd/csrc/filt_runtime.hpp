// filt_runtime.hpp -- internal runtime helpers shared by the .hip sources.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/filtration_hip.h"

namespace filt {

// The stream all work is enqueued on (created on first use).
hipStream_t stream();

// Record an error message for filt_last_error() and return code.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Map a failed HIP call to FILT_ERR_HIP with its text.
int hip_fail(hipError_t e, const char* where);

// Clear the calling thread's error text.
void clear_error();

// Validate a grid's shape against params (non-null, positive extents).
int check_shape(const filt_params* p, int localNx);

// z-marching stencil launchers (filt_stencil.hip), for extents >= 3 and
// arrays below 4 GiB (tiled_fits).
bool tiled_fits(int nx, int ny, int nz);
// False after filt_set_tile_config(kernel, -1, ...): kernel 0 roS, 1 u,
// 2 the fused stages 2-4, 3 the whole step.
bool tiled_enabled(int kernel);
// The fused step (filt_step.hip): u, roS and the Newton in one z-march;
// false when the grid or the parameters need the stages one by one.
bool step_fusable(const filt_params& p, int nx, int rank, int size);
// bc: also write the boundary cells the closed form of the sweep takes from
// post-Newton values (after a stage 1 that wrote the others, k_ro_P2_Xi with
// BC_A in filt_stages.hip); the step then needs no boundary launch.
int launch_step_zm(const filt_fields& f, const filt_params& p, int nx, bool with_u, bool bc);
// The whole step but the boundary in one launch (filt_step.hip): stage 1
// with stages 2-4, then the copy of its stash; for grids step_fusable
// accepts whose media in use all have lambda = 2 (and DivBy-able mu1, mu2,
// 1 - S_wr), unless switched off with filt_set_tile_config(3, -1, ...).
bool full_step_fusable(const filt_params& p, int nx, int rank, int size, bool per_cell_media);
// Tile shapes of the whole-step kernel: 64x4, 32x8, 16x16, 64x8, 32x16, 64x16.
bool full_tile_ok(int tx, int ty);
// Tile shapes of the stages 2-4 kernel (the same six).
bool step_tile_ok(int tx, int ty);
// mid (may be null) is recorded between the two launches.
int launch_step_full(const filt_fields& f, const filt_params& p, int nx, bool with_u,
                     hipEvent_t mid = nullptr);
int launch_roS_zm(const filt_fields& f, const filt_params& p, int nx);
// The divgrad scheme's roS stage, z-marched with the roS kernel's tile
// configuration; divgrad_tiled_ok tells whether its eight uniform divisors
// (h^2, 2h, dt^2, 2dt^2) are in DivBy's range.
bool divgrad_tiled_ok(const filt_params& p);
int launch_roS_divgrad_zm(const filt_fields& f, const filt_params& p, int nx, double t);
int launch_u_zm(const filt_fields& f, const filt_params& p, int nx, int rank, int size);

// The no-allocation mass sum's scratch address on the current device
// (filt_mass.hip), looked up at set-up.
int mass_scratch_init();

}  // namespace filt

#define FILT_HIP_TRY(call)                                   \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return filt::hip_fail(e_, #call); \
    } while (0)

#define FILT_LAUNCH_CHECK(name)                              \
    do {                                                     \
        hipError_t e_ = hipGetLastError();                   \
        if (e_ != hipSuccess) return filt::hip_fail(e_, name); \
    } while (0)

/*
 * filt_oracle.h -- CPU restatement of the reference's FP64 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP backend is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  No product code links it.
 *
 * It restates cpu.cpp + no_communication.cpp + the driver pieces of main.cpp
 * of the reference, operation for operation (same parenthesisation, same
 * libm pow), so that built with -O2 -ffp-contract=off it reproduces the
 * compiled reference bit for bit on every input the reference can express.
 * It generalises the reference where the configurations need it: per-cell
 * permeability / porosity fields (filt_fields.perm / .poro), a media table
 * of up to FILT_MAX_MEDIA entries, explicit source-patch bounds and the
 * defines.h compile-time constants as runtime parameters.
 */
#ifndef FILT_ORACLE_H
#define FILT_ORACLE_H

#include "../include/filtration_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* main.cpp:145-158 */
int orc_i_to_I(int i, int rank, int size, const filt_params* p);
/* main.cpp:166-189 */
void orc_N_to_local(int* localNx, int* localNy, int size, int rank, const filt_params* p);
/* main.cpp:192-198 */
int orc_is_active_point(int i, int localNx, int rank, int size);

/* main.cpp:97-138 (x, y, z may be null; media set to 0 when non-null) */
void orc_initial_data(const filt_fields* f, const filt_params* p, int localNx,
                      int localNy, int rank, int size);

/* cpu.cpp:3-15 + 57-74 */
void orc_ro_P2_Xi_calculation(const filt_fields* f, const filt_params* p,
                              int localNx, int rank, int size);
/* cpu.cpp:17-24 + 77-160 */
void orc_u_calculation(const filt_fields* f, const filt_params* p, int localNx,
                       int rank, int size);
/* cpu.cpp:26-33 + 218-291 */
void orc_roS_calculation(const filt_fields* f, const filt_params* p, double t,
                         int localNx, int rank, int size);
/* cpu.cpp:35-42 + 293-320 */
void orc_P1_S2_calculation(const filt_fields* f, const filt_params* p,
                           int localNx, int rank, int size);
/* cpu.cpp:44-54 + 323-403 (sequential in-place sweep, i -> j -> k) */
void orc_boundary_conditions(const filt_fields* f, const filt_params* p,
                             int localNx, int rank, int size);
/* main.cpp:83-94 with the no_communication.cpp exchanges (no-ops).
 * with_u = 0 skips u_calculation, which nothing on the P/S path reads. */
void orc_time_step(const filt_fields* f, const filt_params* p, double t,
                   int localNx, int rank, int size, int with_u);
/* nsteps calls of orc_time_step with t = (step0 + s) * dt, s = 1..nsteps,
 * as main.cpp:41-43 passes t. */
void orc_run(const filt_fields* f, const filt_params* p, int localNx, int rank,
             int size, int step0, int nsteps, int with_u);

/* Total water / NAPL mass over interior cells:
 * M1 = sum phi*ro1(P1)*(1-S2)*V, M2 = sum phi*ro2(P1)*S2*V, V = h1*h2*h3,
 * with pairwise summation in cell-index order. */
void orc_mass(const filt_fields* f, const filt_params* p, int localNx,
              double out[2]);

/* Loop order of the stage sweeps: 0 = k, j, i (x innermost, cache friendly),
 * 1 = i, j, k as cpu.cpp writes it.  Stages 1-4 touch each cell
 * independently, so both orders give identical results; the boundary sweep is
 * always i, j, k. */
void orc_set_loop_order(int order);
/* Threads for stages 1-4 when built with OpenMP (1 = serial). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif

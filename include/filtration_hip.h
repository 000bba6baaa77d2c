/*
 * filtration_hip.h -- C ABI of the MI355X (gfx950) backend for the two-phase
 * filtration time step of shmem-cuda-filtration.
 *
 * The reference selects its compute backend at link time: gpu.cu or cpu.cpp
 * provides the functions declared in defines.h:66-116 (CUDA/CUDA.vcxproj:17-21,
 * CPU/CPU.vcxproj:81-83).  This header is the language-neutral layer under that
 * contract: extern "C", plain pointers and sizes, int status codes.  The C++
 * shims in shmem-cuda-filtration_amd/host/backend_defines.cpp implement the
 * defines.h functions (C++ linkage, same mangled names) on top of it, so the
 * reference driver main.cpp links against this backend unchanged.
 *
 * Every entry point that launches GPU work enqueues it on the library stream
 * (filt_set_stream, default: a stream owned by the library) and returns
 * without waiting, unless its comment says otherwise.  filt_synchronize()
 * waits for all of it.  Copies to and from the host are synchronous, as the
 * reference's cudaMemcpy calls are (gpu.cu:483-501).
 *
 * Grid layout (defines.h:41-47, cpu.cpp throughout): one array per field,
 * idx = i + j*localNx + k*localNx*Ny, x fastest, j vertical (j = 0 is the top
 * face, gravity acts toward +j), k slowest.  All physics is FP64; media is int32.
 */
#ifndef FILTRATION_HIP_H
#define FILTRATION_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by every int-returning entry point. */
#define FILT_OK 0
#define FILT_ERR_INVALID 1     /* null pointer, bad size or shape */
#define FILT_ERR_HIP 2         /* a HIP runtime call failed */
#define FILT_ERR_NOMEM 3       /* device allocation failed */
#define FILT_ERR_UNSTABLE 4    /* the explicit-scheme stability guard refused */
#define FILT_ERR_IO 5          /* file could not be opened or read */
#define FILT_ERR_UNSUPPORTED 6 /* valid request the library does not implement */
#define FILT_ERR_NOT_FOUND 7   /* unknown slot / handle */

/* Largest media table.  The reference has two media (defines.h:22-26). */
#define FILT_MAX_MEDIA 8

/*
 * Problem parameters: a superset of consts_tag (defines.h:50-61).
 * The first block keeps consts_tag's members and meaning (the per-media arrays
 * are widened from 2 to FILT_MAX_MEDIA entries); the second block holds what
 * defines.h:8-15 fixes at compile time; the third generalises the
 * (2*source+1)^2 source patch of main.cpp:109 / cpu.cpp:346 to explicit bounds.
 */
typedef struct filt_params {
    double K[FILT_MAX_MEDIA];      /* absolute permeability per medium, m^2 */
    double lambda[FILT_MAX_MEDIA]; /* Brooks-Corey pore-size index */
    double S_wr[FILT_MAX_MEDIA];   /* residual water saturation */
    double m[FILT_MAX_MEDIA];      /* porosity */
    double P_d[FILT_MAX_MEDIA];    /* entry (threshold) pressure, Pa */
    double h1, h2, h3;             /* node spacing in x, y, z, m */
    double dt;                     /* time step, s */
    double tau, l_w, l_n, c;       /* constants of the divgrad scheme (roS_scheme 1) */
    double beta1, beta2;           /* compressibility of water / NAPL, 1/Pa */
    int NX, Ny, Nz;                /* global grid */
    int source;                    /* half-width of the reference's source patch */
    int save_plots, print_screen;  /* driver output cadence */
    /* defines.h:8-15 */
    double mu1, mu2;               /* viscosity of water / NAPL, Pa*s */
    double ro01, ro02;             /* reference density of water / NAPL, kg/m^3 */
    double S2_gr;                  /* NAPL saturation held on the source patch */
    double P_atm;                  /* pressure on the top face, Pa */
    double g_const;                /* gravity, m/s^2 */
    int newton_iterations;         /* Newton iterations per cell and step */
    /* source patch on j = 0: src_x0 <= I <= src_x1 and src_z0 <= k <= src_z1 */
    int src_x0, src_x1, src_z0, src_z1;
    int n_media;                   /* entries in use in the media table */
    /* roS_calculation's scheme: FILT_SCHEME_UPWIND, the one the reference
     * calls (assign_rS_nr, cpu.cpp:218-291), or FILT_SCHEME_DIVGRAD, the
     * tau-relaxed scheme of assign_rS (cpu.cpp:162-216), which reads u and
     * roS*_old of the previous step */
    int roS_scheme;
} filt_params;

#define FILT_SCHEME_UPWIND 0
#define FILT_SCHEME_DIVGRAD 1

/*
 * Field arrays of one grid, host or device: the members of ptr_Arrays_tag
 * (defines.h:41-47) in the same order, plus two optional per-cell property
 * fields.  When perm (poro) is non-null it replaces K[media] (m[media]) cell
 * by cell; lambda, S_wr and P_d always come from the media table.
 */
typedef struct filt_fields {
    double *x, *y, *z;
    double *P1, *P2, *S2, *ro1, *ro2;
    double *u1x, *u1y, *u1z, *u2x, *u2y, *u2z;
    double *Xi1, *Xi2, *roS1, *roS1_old, *roS2, *roS2_old;
    int *media;
    double *perm;
    double *poro;
} filt_fields;

/* ---------------------------------------------------------------- setup -- */

/* Fill *p with the reference's compile-time defaults: defines.h:8-15 and
 * the media table of defines.h:22-26 (the ini-controlled members are zeroed,
 * as read_defines leaves them before parsing, main.cpp:601-612).
 * The source patch bounds are derived from NX, Nz, source by
 * filt_params_finalize. */
void filt_default_params(filt_params* p);

/* Derive src_x0..src_z1 from NX, Nz and source the way main.cpp:109 and
 * cpu.cpp:346 test the patch: NX/2-source <= I <= NX/2+source,
 * Nz/2-source <= k <= Nz/2+source.  Leaves explicit bounds alone when the
 * ini gave them (src_x0 >= 0). */
void filt_params_finalize(filt_params* p);

/* Parse a defines.ini file into *p (main.cpp:601-683): '#' lines are
 * skipped, a key runs to '=', a value to the first space, the last
 * occurrence wins, unknown keys are ignored.  Reads the 16 keys the reference
 * reads plus the keys defines.ini documents but the reference ignores
 * (K[i], LAMBDA[i], S_WR[i], M[i], P_D[i], MU1, MU2, RO_W, RO_N, S2_GR,
 * P_ATM, G_CONST) and this backend's extensions (NEWTON_ITERATIONS,
 * SRC_X0, SRC_X1, SRC_Z0, SRC_Z1, N_MEDIA, ROS_SCHEME).  Keys absent from the file keep
 * their value in *p.  Returns FILT_ERR_IO when the file cannot be opened
 * (the reference prints and continues, main.cpp:620-624). */
int filt_read_defines(const char* path, filt_params* p);

/* Stability guard of the explicit scheme: the amplification number
 * r = 4*dt*D*(1/h1^2+1/h2^2+1/h3^2) with the water diffusivity
 * D = K/(mu1*phi*beta1), maximised over the media table and, when given,
 * over the per-cell fields (perm, poro, n_cells long; either may be null).
 * The scheme is stable for r <= 2. */
double filt_stability_number(const filt_params* p, const double* perm,
                             const double* poro, size_t n_cells);

/* Seeded synthetic property fields (host arrays).  Draws come from a
 * counter-based generator (splitmix64 of seed, stream and cell index), so a
 * field does not depend on thread count.
 * filt_gen_uniform: out[i] = lo + (hi-lo)*U(0,1].
 * filt_gen_lognormal: K = exp(ln(k_mean) + sigma*z) on an nx*ny*nz grid
 * (x fastest), z = Gaussian white noise convolved with a separable 3D
 * Gaussian of std corr_cells/sqrt(2) cells (periodic wrap; corr_cells <= 0:
 * no filtering), renormalised to zero mean and unit variance and clipped to
 * [-clip_sigma, clip_sigma] when clip_sigma > 0.  threads <= 0: all cores
 * (at most 16). */
int filt_gen_uniform(double* out, size_t n, uint64_t seed, double lo, double hi);
int filt_gen_lognormal(double* out, int nx, int ny, int nz, uint64_t seed, double k_mean,
                       double sigma, double corr_cells, double clip_sigma, int threads);

/* ---------------------------------------------------- driver helpers -- */

/* i_to_I (main.cpp:145-158): local x index -> global x index. */
int filt_i_to_I(int i, int rank, int size, const filt_params* p);
/* N_to_local (main.cpp:166-189): local extents of rank's x slab. */
void filt_N_to_local(int* localNx, int* localNy, int size, int rank, const filt_params* p);
/* is_active_point (main.cpp:192-198): 0 for halo cells of inner ranks. */
int filt_is_active_point(int i, int localNx, int rank, int size);
/* initial_data (main.cpp:97-138) on host arrays: S2 = S2_gr on the source
 * patch of j = 0, else 0; P1 = P_atm + j*ro01*g*h2; x, y, z, media filled
 * when non-null. */
int filt_initial_data(const filt_fields* host, const filt_params* p, int localNx, int rank,
                      int size);

/* ---------------------------------------------------- device management -- */

/* Device_Initialize (defines.h:82, gpu.cu:565-610): select device
 * rank % device_count, report the launch-block counts the reference would use
 * (localNx/16, Ny/4, Nz/4 rounded up) and check that the grid fits. */
int filt_device_initialize(int rank, int localNx, const filt_params* p,
                           int* blocksX, int* blocksY, int* blocksZ);

/* device_memory_alloc (defines.h:68, gpu.cu:504-531): 20 FP64 arrays and one
 * int32 array of nX*Ny*Nz cells, plus a 2*Ny*Nz exchange buffer.  perm and
 * poro are set to null.  One allocation backs all arrays; each array starts
 * on a 256-byte boundary. */
int filt_device_memory_alloc(filt_fields* dev, double** dev_buffer, int nX,
                             const filt_params* p);

/* device_memory_free (defines.h:71, gpu.cu:534-560). */
int filt_device_memory_free(const filt_fields* dev, double* dev_buffer);

/* load_data_to_host / load_data_to_device / load_data_to_device_int
 * (defines.h:78-80, gpu.cu:483-501): synchronous copies of localNx*Ny*Nz
 * elements.  They wait for the library stream first. */
int filt_load_data_to_host(double* host, const double* dev, int localNx,
                           const filt_params* p);
int filt_load_data_to_device(const double* host, double* dev, int localNx,
                             const filt_params* p);
int filt_load_data_to_device_int(const int* host, int* dev, int localNx,
                                 const filt_params* p);

/* Raw synchronous copies of n bytes (direction 0: host->device,
 * 1: device->host, 2: device->device). */
int filt_memcpy(void* dst, const void* src, size_t n_bytes, int direction);

/* Device-to-device copy of n_bytes (a multiple of 16) by the library's own
 * streaming kernel, enqueued on the library stream.  Serves as the
 * reference point of the HBM roofline (its bytes moved are known exactly). */
int filt_device_copy(void* dst, const void* src, size_t n_bytes);

/* Allocate / free device memory for per-cell property fields and the like. */
int filt_device_malloc(void** ptr, size_t n_bytes);
int filt_device_free_ptr(void* ptr);

/* ---------------------------------------------------------- the stages -- */
/* Each stage enqueues its kernels and returns.  localNx is the local x
 * extent, rank/size the x-slab decomposition of main.cpp:145-198 (size 1 on
 * a single GPU). */

/* ro_P2_Xi_calculation (defines.h:106; cpu.cpp:3-15, 57-74): P2, Xi1, Xi2,
 * ro1, ro2 from P1, S2 and the media, in every active cell. */
int filt_ro_P2_Xi_calculation(const filt_fields* dev, const filt_params* p,
                              int localNx, int rank, int size);

/* u_calculation (defines.h:108; cpu.cpp:17-24, 77-160): Darcy velocities
 * u1x..u2z, one-sided at the faces, central inside. */
int filt_u_calculation(const filt_fields* dev, const filt_params* p,
                       int localNx, int rank, int size);

/* roS_calculation (defines.h:110; cpu.cpp:26-33, 218-291): explicit upwind
 * update of ro*S in interior cells; writes roS1, roS2, roS1_old, roS2_old.
 * t is accepted and unused, as in the reference. */
int filt_roS_calculation(const filt_fields* dev, const filt_params* p,
                         double t, int localNx, int rank, int size);

/* P1_S2_calculation (defines.h:111; cpu.cpp:35-42, 293-320): per-cell
 * Newton solve (newton_iterations iterations) for P1 and S2 from roS1, roS2,
 * in interior cells. */
int filt_P1_S2_calculation(const filt_fields* dev, const filt_params* p,
                           int localNx, int rank, int size);

/* boundary_conditions (defines.h:112; cpu.cpp:44-54, 323-403): the
 * reference's in-place i->j->k sweep over the shell cells, computed in
 * closed form by two kernels; grids with an extent below 3 use a sequential
 * kernel that replays the sweep. */
int filt_boundary_conditions(const filt_fields* dev, const filt_params* p,
                             int localNx, int rank, int size);

/* Stages 2-4 of time_step_function (main.cpp:86-90): u_calculation (when
 * with_u), roS_calculation, P1_S2_calculation, on a state stage 1 has just
 * been applied to (so P2 holds P1's bits, P2 = P1 + 0, cpu.cpp:63).  On one
 * rank whose grid has every extent >= 3 and arrays below 4 GiB, with the
 * upwind scheme, one z-marching kernel computes all three: it reads P2 where
 * the stages read P1 or P2, evaluates ro1, ro2 from P2 as stage 1 did, and
 * writes every array the three stages write, with the same bits, moving
 * 136 B per cell instead of 240.  Elsewhere (or
 * after filt_set_tile_config(2, -1, ...)) the three stages run in turn. */
int filt_u_roS_P1_S2_calculation(const filt_fields* dev, const filt_params* p,
                                 double t, int localNx, int rank, int size, int with_u);

/* Total water and NAPL mass of a device grid, FP64:
 * mass2[0] = sum phi*ro1(P1)*(1-S2)*V, mass2[1] = sum phi*ro2(P1)*S2*V over
 * the interior cells (cpu.cpp:220), V = h1*h2*h3, densities as stage 1
 * computes them (cpu.cpp:72-73), phi from poro or the medium.  No reference
 * counterpart: the mass-balance accumulator a driver reads step by step.
 * Deterministic (a fixed summation tree).  Waits for the library stream and
 * writes mass2 (host memory) before returning. */
int filt_mass(const filt_fields* dev, const filt_params* p, int localNx, double* mass2);

/* filt_mass without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Enqueues the same
 * kernels as filt_mass on the library stream and returns without waiting;
 * the two masses (water, NAPL) land in dev_mass2, device memory of 2
 * doubles the caller owns, bit for bit what filt_mass returns.  The partial
 * sums live in a fixed 1 MiB array in each GPU's memory (held until the
 * process ends), so grids of at most 268435456 cells (65536 partials of
 * 4096 cells); larger grids return FILT_ERR_INVALID.  Calls are
 * independent (nothing carries over, nothing needs clearing), but calls that
 * overlap on one GPU share that array and race: keep them on one stream. */
int filt_mass_noalloc(const filt_fields* dev, const filt_params* p, int localNx,
                      double* dev_mass2);

/* Number of launches of the one-kernel stages 2-4 (k_step_zm) since the
 * library was loaded, from any entry point.  A diagnostic with no
 * counterpart in the reference: it lets a caller see which path its calls
 * took (for example that the defines.h shims fused u, roS and the Newton). */
long long filt_step_zm_launches(void);

/* One whole time step in the order of time_step_function (main.cpp:83-94):
 * stage 1, filt_u_roS_P1_S2_calculation, boundary conditions; the exchanges
 * are no-ops on one GPU (no_communication.cpp:3-14).  with_u = 0 skips
 * u_calculation, which nothing on the P/S path reads.  After
 * filt_set_tile_config(3, tx, ty, kz, order), on one rank whose grid has
 * every extent >= 3, arrays below 4 GiB, the upwind scheme and media with
 * lambda = 2, stage 1 and stages 2-4 run as one z-marching kernel that reads
 * P1, S2 once and writes every array the four stages write (168 B per cell
 * instead of 200); the new P1, S2 of the cells other blocks read go through
 * a stash that a second kernel copies back.  Every array ends with the
 * same bits either way. */
int filt_time_step(const filt_fields* dev, const filt_params* p, double t,
                   int localNx, int rank, int size, int with_u);

/* filt_time_step, recording caller-supplied hipEvent_t between its launches
 * on the library stream (no reference counterpart: per-kernel timing for a
 * benchmark).  events (4 entries, any may be null): [0] before the step,
 * [1] after stage 1 with stages 2-4 in one kernel (*path = 1; the
 * lambda = 2 single-rank case) or after stage 1 alone (*path = 0), [2]
 * after the stash copy (path 1) or stages 2-4 (path 0), [3] after the
 * boundary conditions.  Same results as filt_time_step. */
int filt_time_step_timed(const filt_fields* dev, const filt_params* p, double t, int localNx,
                         int rank, int size, int with_u, void* const* events, int* path);

/* load_exchange_data / save_exchange_data (defines.h:115-116,
 * gpu.cu:613-653): pack x-planes i=1 and i=localNx-2 of dev_array into
 * dev_buffer and copy them to host_buffer (2*Ny*Nz doubles); unpack
 * host_buffer into planes i=0 (rank != 0) and i=localNx-1 (rank != size-1).
 * Synchronous.  Only a multi-rank driver calls them. */
int filt_load_exchange_data(const double* dev_array, double* dev_buffer,
                            double* host_buffer, int localNx,
                            const filt_params* p);
int filt_save_exchange_data(double* dev_array, double* dev_buffer,
                            const double* host_buffer, int localNx, int rank,
                            int size, const filt_params* p);

/* Tile shape of the z-marching stencil kernels of u_calculation (kernel 1),
 * roS_calculation (kernel 0; either scheme), the fused stages 2-4 (kernel 2)
 * and the whole step, stage 1 with stages 2-4 (kernel 3): tx*ty threads per
 * block (64x4, 32x8, 16x16 or 64x8; for kernels 0-2 also 128x4 and 128x8),
 * kz planes per z-chunk, order 0 = XCD-grouped tile numbering, 1 = hardware
 * order.  Used by grids with every extent >= 3.  The whole-step kernel also
 * takes 32x16 and 64x16 (1024 threads), and so does the stages 2-4 kernel,
 * which takes 256x4 too.  tx = 0 restores the built-in default
 * (roS 64x8, kz 64, order 1; u 64x8, kz 64, order 1; stages 2-4 64x8, kz 16,
 * order 1; the whole step off); tx = -1 switches the kernel off: the
 * one-cell-per-thread kernels for 0-1, the stages one by one for 2, stage 1
 * and stages 2-4 apart for 3 (grids with an extent below 3 or arrays of
 * 4 GiB and more take those anyway; results are the same).  Setting a tile
 * for kernel 3 switches the whole-step kernel on.  Process-wide. */
int filt_set_tile_config(int kernel, int tx, int ty, int kz, int order);
/* The configuration used for an nx*ny*nz grid (tuned or default). */
int filt_get_tile_config(int kernel, int nx, int ny, int nz, int* tx, int* ty, int* kz,
                         int* order);

/* Autotuner: times every candidate tile configuration (64x4, 32x8, 16x16,
 * 64x8, 128x4, 128x8 x kz in {8, 16, 32, 64, Nz} x both orders) of both
 * stencil kernels on the given device grid (after stage 1 has filled P2, Xi,
 * ro), reps launches each (median), and keeps the fastest per kernel for
 * grids of this shape.  Kernel 0 is the roS stage of the scheme p selects.
 * Upwind: rewrites only roS*, u* with the values one call would give.
 * Divgrad (its kernel reads roS*_old and rewrites them; timed on its t >= 2 dt
 * branch): the four roS arrays are saved to a temporary device array and
 * left as they were; u* holds the values one call gives.  best_ms[0] /
 * best_ms[1] receive the roS / u times when non-null.  The tiled roS kernels
 * need h1, h2, h3 in [2^-100, 2^100] (they divide through a refined
 * reciprocal; the divgrad kernel also h^2, 2h, dt^2 and 2 dt^2); outside
 * that range roS runs untiled, is not tuned, and best_ms[0] is 0.
 * filt_set_tile_config clears tuned entries. */
int filt_autotune_stencil(const filt_fields* dev, const filt_params* p, int localNx, int reps,
                          double* best_ms);

/* Autotuner of the stages 2-4 kernel (kernel 2 of filt_set_tile_config):
 * times its nine tile shapes (64x4, 32x8, 16x16, 64x8, 32x16, 64x16, 128x4,
 * 128x8, 256x4) x kz in {8, 16, 32, 64, Nz} x both block orders on the given
 * device grid (after stage 1 has filled P2, Xi, ro), reps launches each
 * (median), and keeps the fastest for grids of this shape; filt_time_step
 * and filt_u_roS_P1_S2_calculation then use it.  S2 is saved to a temporary
 * device array and restored before each launch; on return the grid holds
 * what one filt_u_roS_P1_S2_calculation call gives.  *best_ms (when
 * non-null) receives the chosen configuration's time.  FILT_ERR_UNSUPPORTED
 * where the kernel does not apply (several ranks, an extent below 3, arrays
 * of 4 GiB or more, the divgrad scheme, spacings outside [2^-100, 2^101], or
 * the fusion switched off).  filt_set_tile_config(2, ...) clears the tuned
 * entries.  No counterpart in the reference, whose launch shape is a fixed
 * 16x4x4 block (defines.h:18-20). */
int filt_autotune_step(const filt_fields* dev, const filt_params* p, int localNx, int reps,
                       double* best_ms);

/* -------------------------------------- fused / batched / mixed forms -- */
/*
 * No reference counterpart: these advance P1 and S2 by the same time step as
 * time_step_function (main.cpp:83-94) without materialising P2, ro, Xi, u,
 * roS, in one kernel launch per step for any number of grids.
 *
 * An ensemble holds member grids of any shapes (every extent >= 3, one rank),
 * each with its own parameters, optional per-cell permeability / porosity /
 * media and its state, all resident in HBM.  Results of a member do not
 * depend on the other members.  Three arithmetic forms:
 *   FILT_PREC_FP64  computes what the stage path computes (same expressions,
 *                   same bits as filt_time_step);
 *   FILT_PREC_FAST  all FP64, with algebraic shortcuts: the Newton in closed
 *                   form where the capillary pressure vanishes, the
 *                   Brooks-Corey powers as products for lambda = 2 and as
 *                   exp / expm1 of one log1p otherwise, divisions by uniform
 *                   quantities as multiplications by reciprocals;
 *   FILT_PREC_MIXED the upwind fluxes and the ro*S increment in FP32 (FP32
 *                   mobilities in LDS); the pressure differences and face
 *                   forces, the ro(P)(1-S) base, the Newton and the mass sums
 *                   in FP64.
 * Members must pass the stability guard (filt_stability_number <= 2 with
 * their fields), else filt_ensemble_add returns FILT_ERR_UNSTABLE.
 * Use an ensemble from one host thread at a time.
 */
#define FILT_PREC_FP64 0
#define FILT_PREC_FAST 1
#define FILT_PREC_MIXED 2

typedef struct filt_ensemble filt_ensemble;

int filt_ensemble_create(filt_ensemble** out, int precision);
int filt_ensemble_destroy(filt_ensemble* e);

/* What filt_ensemble_add would refuse for these parameters and fields
 * (perm, poro may be null), without touching the device: FILT_ERR_INVALID
 * (bad precision or media table), FILT_ERR_UNSUPPORTED (an extent below 3,
 * the divgrad scheme, an array of 4 GiB or more, or for the FP64 form a
 * divisor outside [2^-100, 2^101]), FILT_ERR_UNSTABLE (stability guard > 2);
 * else FILT_OK. */
int filt_ensemble_check(const filt_params* p, int precision, const double* perm,
                        const double* poro);

/* Add a member: copies the host arrays (NX*Ny*Nz cells each, x fastest) to
 * the device.  perm, poro, media may be null (media table entry 0 / the
 * media array).  *id receives the member's handle. */
int filt_ensemble_add(filt_ensemble* e, const filt_params* p, const double* P1, const double* S2,
                      const double* perm, const double* poro, const int* media, int* id);
int filt_ensemble_remove(filt_ensemble* e, int id);
int filt_ensemble_count(const filt_ensemble* e);
long long filt_ensemble_cells(const filt_ensemble* e);
/* Steps a member has advanced since it was added (-1: unknown id). */
long long filt_ensemble_steps(const filt_ensemble* e, int id);
/* Device bytes one member with these parameters occupies. */
size_t filt_ensemble_member_bytes(const filt_params* p, int precision, int with_perm,
                                  int with_poro, int with_media);
/* Device bytes held by all members. */
size_t filt_ensemble_device_bytes(const filt_ensemble* e);

/* Advance every member nsteps steps (one launch per step, enqueued). */
int filt_ensemble_step(filt_ensemble* e, int nsteps);
/* Record nsteps steps as a hipGraph; filt_ensemble_replay launches it.
 * Adding / removing members or changing the tiling drops the graph. */
int filt_ensemble_capture(filt_ensemble* e, int nsteps);
int filt_ensemble_replay(filt_ensemble* e);

/* Current state of a member to host arrays (either may be null). */
int filt_ensemble_get(filt_ensemble* e, int id, double* P1, double* S2);
/* Total water / NAPL mass over a member's interior cells,
 * sum phi*ro(P)*S*h1*h2*h3, FP64, fixed summation order. */
int filt_ensemble_mass(filt_ensemble* e, int id, double* out2);

/* Tiling of the fused kernel: tx x ty columns per tile and cy columns per
 * thread, stacked in y (16x16/1, 32x8/1, 64x4/1, 32x16/1, 64x8/1, 64x8/2,
 * 32x16/2, 64x16/2; 0x0: 16x16/1), kz planes per z-chunk, order 0 =
 * XCD-grouped, 1 = hardware block order.  Results do not depend on the
 * tiling. */
int filt_ensemble_set_tiling(filt_ensemble* e, int tx, int ty, int cy, int kz, int order);
int filt_ensemble_get_tiling(const filt_ensemble* e, int* tx, int* ty, int* cy, int* kz,
                             int* order);
/* Autotuner: times each (tile shape, kz, order) candidate -- every tile
 * shape above, kz in {4, 8, 16, 32, 64}, both orders -- over reps steps of
 * the current membership (the members advance by the steps run) and keeps
 * the fastest.  Outputs may be null. */
int filt_ensemble_autotune(filt_ensemble* e, int reps, int* best_tx, int* best_ty, int* best_cy,
                           int* best_kz, int* best_order, double* best_ms_per_step);

/* Free / total device memory (admission control). */
int filt_device_free_bytes(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------------------ serving loop -- */
/*
 * No reference counterpart.  Continuous batching of simulation requests: a
 * FIFO queue, admission while a slot (max_active) and the member's device
 * bytes plus reserve_bytes are free, all active requests advanced together
 * (one fused launch per step, up to max_round_steps steps per round and never
 * past the next completion), finished requests copied to host and freed.
 * State stays on the device between admission and retirement.  A request's
 * result equals the result of the same request run alone.  One host thread
 * at a time.
 */
typedef struct filt_server filt_server;

int filt_serve_create(filt_server** out, int precision, size_t reserve_bytes, int max_active,
                      int max_round_steps);
int filt_serve_destroy(filt_server* s);
/* Queue a request (arrays copied; perm, poro, media may be null).  Refuses
 * what filt_ensemble_check refuses, and with FILT_ERR_NOMEM a request whose
 * device bytes plus the reserve exceed the device's total memory. */
int filt_serve_submit(filt_server* s, const filt_params* p, const double* P1, const double* S2,
                      const double* perm, const double* poro, const int* media,
                      long long nsteps, int* id);
/* One round: admit, step, retire.  Counts of this round (may be null).
 * Returns FILT_ERR_NOMEM when nothing is active and the request at the head
 * of the queue still does not fit in the free device memory (it stays
 * queued). */
int filt_serve_round(filt_server* s, int* n_admitted, int* n_retired);
int filt_serve_status(const filt_server* s, int* queued, int* active, int* done,
                      long long* active_cells, long long* cell_updates);
int filt_serve_is_done(const filt_server* s, int id);
/* Result of a finished request (P1, S2: NX*Ny*Nz each; mass2: water, NAPL);
 * the server forgets it afterwards. */
int filt_serve_take(filt_server* s, int id, double* P1, double* S2, double* mass2);

/* ------------------------------------------- checkpoint, resume, plots -- */

/* save (main.cpp:485-537): append this rank's block
 *   [int j][P1][S2][x][y][z][roS1_old][roS2_old][media:int32]
 * (localNx*Ny*Nz values each, native byte order) to path, creating its
 * folder; truncate != 0 creates the file first (rank 0 of the reference).
 * Ranks call in rank order, as the reference's barrier loop orders them.
 * The device form copies P1, S2, roS1_old, roS2_old and media from the
 * device; x, y, z are the values initial_data gives (I*h1, j*h2, k*h3), as
 * in the _host form when its x, y or z is null. */
int filt_save(const char* path, const filt_fields* dev, const filt_params* p,
              int localNx, int rank, int size, int j, int truncate);
int filt_save_host(const char* path, const filt_fields* host, const filt_params* p,
                   int localNx, int rank, int size, int j, int truncate);

/* restore (main.cpp:540-569) then the uploads of Initialize
 * (main.cpp:217-228): read the blocks of ranks 0..rank in turn (each of its
 * own N_to_local width), keep rank's, and copy P1, S2, roS1_old, roS2_old,
 * media (and x, y, z where non-null) into the fields; *j receives the saved
 * step.  localNx must be rank's width.  FILT_ERR_IO when the file is
 * missing or short. */
int filt_restore(const char* path, const filt_fields* dev, const filt_params* p,
                 int localNx, int rank, int size, int* j);
int filt_restore_host(const char* path, const filt_fields* host, const filt_params* p,
                      int localNx, int rank, int size, int* j);

/* save_data_plots (main.cpp:319-482): Tecplot ASCII files
 * plot_S2/S2=%012.4f.dat, plot_P1/P1=..., plot_u/u=..., plot_S2_y/S2=...,
 * plot_S2_x/S2=... and media.dat under dir (NULL: the working folder, as
 * the reference), byte for byte as the reference writes them: rank 0
 * creates the folders and files with their headers, then each rank, in
 * rank order, appends its lines.  The device form copies P1, S2, u2x, u2y,
 * u2z and media; x, y, z as for filt_save.  u must be current: a driver
 * that skips u_calculation (filt_time_step with_u = 0) runs it on the
 * output steps, where the reference computes it in the same step. */
int filt_save_plots(const char* dir, const filt_fields* dev, const filt_params* p,
                    int localNx, int rank, int size, double t);
int filt_save_plots_host(const char* dir, const filt_fields* host, const filt_params* p,
                         int localNx, int rank, int size, double t);

/* Device-resident snapshots with asynchronous writers (no reference
 * counterpart: the reference stops to copy and format, main.cpp:50-55).
 * take: stream-ordered copy of the grid's P1, S2 (+ roS1_old, roS2_old with
 * FILT_SNAP_SAVE; + u2x, u2y, u2z with FILT_SNAP_PLOTS) and media into the
 * snapshot's own device buffer; returns at once and later steps may run.
 * write_async: a host thread copies the snapshot to pinned memory on a
 * second stream and writes the save block (save_path, or NULL) and / or the
 * plots (plot_dir, or NULL) as filt_save / filt_save_plots would have at
 * the moment of the take.  wait: joins the writer and returns its status.
 * restore: stream-ordered copy back into a grid (rollback, preemption).
 * A take or write_async waits for the snapshot's previous write.  A failed
 * write (bad path, full disk, short write) is reported once, by whichever
 * of take, write_async, wait or destroy comes next: that call returns the
 * write's status (take and write_async then do nothing). */
#define FILT_SNAP_SAVE 1
#define FILT_SNAP_PLOTS 2
typedef struct filt_snapshot filt_snapshot;
int filt_snapshot_create(filt_snapshot** out, const filt_params* p, int localNx,
                         int rank, int size, int what);
int filt_snapshot_destroy(filt_snapshot* s);
int filt_snapshot_take(filt_snapshot* s, const filt_fields* dev, int j, double t);
int filt_snapshot_write_async(filt_snapshot* s, const char* save_path, int truncate,
                              const char* plot_dir);
int filt_snapshot_wait(filt_snapshot* s);
int filt_snapshot_restore(filt_snapshot* s, const filt_fields* dev, int* j);

/* ------------------------------------------------------------- runtime -- */

/* Wait for all work the library enqueued. */
int filt_synchronize(void);

/* Use the given hipStream_t for all later work (null: the library's own
 * stream).  The caller keeps ownership of the stream. */
int filt_set_stream(void* hip_stream);

/* The hipStream_t the library currently enqueues on. */
void* filt_get_stream(void);

/* Text of the last error on the calling thread ("" when none). */
const char* filt_last_error(void);

/* Version / build string. */
const char* filt_build_info(void);

#ifdef __cplusplus
}
#endif

#endif /* FILTRATION_HIP_H */

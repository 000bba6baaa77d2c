/*
 * filt_oracle.c -- CPU restatement of the reference's FP64 time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP backend (see
 * filt_oracle.h).  Build with -O2 -ffp-contract=off and no -ffast-math; the
 * expressions below keep the reference's operation order and
 * parenthesisation so that the results match the compiled reference bit for
 * bit.  Each function cites the reference lines it restates.
 *
 * Parity status per function (how each is pinned; details in the tests):
 *   orc_ro_P2_Xi_calculation, orc_u_calculation, orc_roS_calculation,
 *   orc_P1_S2_calculation, orc_boundary_conditions, orc_time_step,
 *   orc_initial_data, orc_i_to_I, orc_N_to_local, orc_is_active_point:
 *     bit-exact against the reference compiled from its own sources
 *     (oracle/_ref, tests/test_oracle_vs_ref.py) and against committed golden
 *     vectors made from it (tests/golden/), plus closed-form checks
 *     (tests/test_oracle_math.py).
 *   per-cell perm/poro, patch bounds, runtime S2_gr: no reference counterpart;
 *     pinned by reduction (uniform fields equal the media-table path
 *     bit for bit; bounds from `source` equal the reference patch).
 *   orc_mass: no reference counterpart; pinned by closed forms on uniform
 *     states.
 *   orc_roS_calculation with roS_scheme = FILT_SCHEME_DIVGRAD (assign_rS,
 *     cpu.cpp:162-216): bit-exact against the reference's own assign_rS
 *     called through oracle/_ref, per stage and over trajectories
 *     (tests/test_divgrad.py).
 */
#include "filt_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_loop_order = 0;
static int g_threads = 1;

void orc_set_loop_order(int order) { g_loop_order = order; }
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* Flat index of the reference: i + j*localNx + k*localNx*Ny (cpu.cpp passim). */
#define IDX(i, j, k) ((ptrdiff_t)(i) + (ptrdiff_t)(j) * (ptrdiff_t)localNx + \
                      (ptrdiff_t)(k) * (ptrdiff_t)localNx * (ptrdiff_t)p->Ny)

/* ------------------------------------------------------ index helpers -- */

/* main.cpp:145-158: local x index -> global x index of the x-slab split. */
int orc_i_to_I(int i, int rank, int size, const filt_params* p)
{
    const int base = p->NX / size;
    const int rem = p->NX % size;
    if (rank <= rem)
        return rank == 0 ? i : (base + 1) * rank + i - 1;
    return (base + 1) * rank - (rank - rem) + i - 1;
}

/* main.cpp:166-189: local extent; slab owners below NX%size take one more
 * cell, and each neighbour adds one halo cell. */
void orc_N_to_local(int* localNx, int* localNy, int size, int rank, const filt_params* p)
{
    int n = p->NX / size;
    if (rank < p->NX % size)
        n++;
    if (size > 1)
        n += (rank == 0 || rank == size - 1) ? 1 : 2;
    *localNx = n;
    *localNy = p->Ny;
}

/* main.cpp:192-198: halo cells of interior ranks are not computed. */
int orc_is_active_point(int i, int localNx, int rank, int size)
{
    if ((rank != 0 && i == 0) || (rank != size - 1 && i == localNx - 1))
        return 0;
    return 1;
}

/* ------------------------------------------------------ initial state -- */

static int in_patch(int I, int k, const filt_params* p)
{
    return I >= p->src_x0 && I <= p->src_x1 && k >= p->src_z0 && k <= p->src_z1;
}

/* main.cpp:97-138: S2 = S2_gr on the source patch of the top face, else 0;
 * hydrostatic water pressure P_atm + j*ro01*g*h2; node coordinates; media 0. */
void orc_initial_data(const filt_fields* f, const filt_params* p, int localNx,
                      int localNy, int rank, int size)
{
    for (int i = 0; i < localNx; i++)
        for (int j = 0; j < localNy; j++)
            for (int k = 0; k < p->Nz; k++) {
                if (!orc_is_active_point(i, localNx, rank, size))
                    continue;
                const ptrdiff_t c = IDX(i, j, k);
                const int I = orc_i_to_I(i, rank, size, p);
                f->S2[c] = (j == 0 && in_patch(I, k, p)) ? p->S2_gr : 0;
                /* j*ro01 is an int product in the reference; exact here too */
                f->P1[c] = p->P_atm + (double)j * p->ro01 * p->g_const * p->h2;
                if (f->x) f->x[c] = I * p->h1;
                if (f->y) f->y[c] = j * p->h2;
                if (f->z) f->z[c] = k * p->h3;
                if (f->media) f->media[c] = 0;
            }
}

/* ------------------------------------------------------- cell kernels -- */

static inline double cell_K(const filt_fields* f, const filt_params* p, ptrdiff_t c, int md)
{
    return f->perm ? f->perm[c] : p->K[md];
}

static inline double cell_phi(const filt_fields* f, const filt_params* p, ptrdiff_t c, int md)
{
    return f->poro ? f->poro[c] : p->m[md];
}

/* Medium of a cell; a table with one entry makes every cell medium 0. */
static inline int cell_media_p(const filt_fields* f, const filt_params* p, ptrdiff_t c)
{
    return (f->media && p->n_media > 1) ? f->media[c] : 0;
}
#define cell_media(f, c) cell_media_p((f), p, (c))

/* cpu.cpp:57-74 (assign_P2_Xi1_Xi2 then assign_ro1_ro2). */
static void cell_ro_P2_Xi(const filt_fields* f, const filt_params* p, ptrdiff_t c)
{
    const int md = cell_media(f, c);
    const double lam = p->lambda[md];
    const double swr = p->S_wr[md];
    const double S_e = (1. - f->S2[c] - swr) / (1. - swr);
    const double k1 = pow(S_e, (2. + 3. * lam) / lam);
    const double k2 = (1. - S_e) * (1. - S_e) * (1 - pow(S_e, (2. + lam) / lam));
    const double P_k = 0; /* capillary pressure disabled, cpu.cpp:63 */
    const double K = cell_K(f, p, c, md);

    f->P2[c] = f->P1[c] + P_k;
    f->Xi1[c] = -1. * K * k1 / p->mu1;
    f->Xi2[c] = -1. * K * k2 / p->mu2;
    f->ro1[c] = p->ro01 * (1. + p->beta1 * (f->P1[c] - p->P_atm));
    f->ro2[c] = p->ro02 * (1. + p->beta2 * (f->P2[c] - p->P_atm));
}

/* cpu.cpp:77-160: one-sided differences on the faces, central inside, zero
 * along an axis of extent <= 2; the y components carry -ro*g. */
static void cell_u(const filt_fields* f, const filt_params* p, int i, int j, int k, int localNx)
{
    const ptrdiff_t c = IDX(i, j, k);
    const double* P1 = f->P1;
    const double* P2 = f->P2;
    const double g = p->g_const;

    if (p->NX > 2) {
        const ptrdiff_t xp = IDX(i + 1, j, k), xm = IDX(i - 1, j, k);
        if (i == 0) {
            f->u1x[c] = f->Xi1[c] * (P1[xp] - P1[c]) / p->h1;
            f->u2x[c] = f->Xi2[c] * (P2[xp] - P2[c]) / p->h1;
        } else if (i == localNx - 1) {
            f->u1x[c] = f->Xi1[c] * (P1[c] - P1[xm]) / p->h1;
            f->u2x[c] = f->Xi2[c] * (P2[c] - P2[xm]) / p->h1;
        } else {
            f->u1x[c] = f->Xi1[c] * ((P1[xp] - P1[xm]) / (2. * p->h1));
            f->u2x[c] = f->Xi2[c] * ((P2[xp] - P2[xm]) / (2. * p->h1));
        }
    } else {
        f->u1x[c] = 0;
        f->u2x[c] = 0;
    }

    if (p->Ny > 2) {
        const ptrdiff_t yp = IDX(i, j + 1, k), ym = IDX(i, j - 1, k);
        if (j == 0) {
            f->u1y[c] = f->Xi1[c] * ((P1[yp] - P1[c]) / p->h2 - f->ro1[c] * g);
            f->u2y[c] = f->Xi2[c] * ((P2[yp] - P2[c]) / p->h2 - f->ro2[c] * g);
        } else if (j == p->Ny - 1) {
            f->u1y[c] = f->Xi1[c] * ((P1[c] - P1[ym]) / p->h2 - f->ro1[c] * g);
            f->u2y[c] = f->Xi2[c] * ((P2[c] - P2[ym]) / p->h2 - f->ro2[c] * g);
        } else {
            f->u1y[c] = f->Xi1[c] * ((P1[yp] - P1[ym]) / (2. * p->h2) - f->ro1[c] * g);
            f->u2y[c] = f->Xi2[c] * ((P2[yp] - P2[ym]) / (2. * p->h2) - f->ro2[c] * g);
        }
    } else {
        f->u1y[c] = 0;
        f->u2y[c] = 0;
    }

    if (p->Nz > 2) {
        const ptrdiff_t zp = IDX(i, j, k + 1), zm = IDX(i, j, k - 1);
        if (k == 0) {
            f->u1z[c] = f->Xi1[c] * (P1[zp] - P1[c]) / p->h3;
            f->u2z[c] = f->Xi2[c] * (P2[zp] - P2[c]) / p->h3;
        } else if (k == p->Nz - 1) {
            f->u1z[c] = f->Xi1[c] * (P1[c] - P1[zm]) / p->h3;
            f->u2z[c] = f->Xi2[c] * (P2[c] - P2[zm]) / p->h3;
        } else {
            f->u1z[c] = f->Xi1[c] * ((P1[zp] - P1[zm]) / (2. * p->h3));
            f->u2z[c] = f->Xi2[c] * ((P2[zp] - P2[zm]) / (2. * p->h3));
        }
    } else {
        f->u1z[c] = 0;
        f->u2z[c] = 0;
    }
}

/* Interior test shared by cpu.cpp:220 and cpu.cpp:295. */
static inline int is_interior(int i, int j, int k, int localNx, const filt_params* p)
{
    return i != 0 && i != localNx - 1 && j != 0 && j != p->Ny - 1 &&
           ((k != 0 && k != p->Nz - 1) || p->Nz < 2);
}

/* Upwind flux difference along one axis, cpu.cpp:242-244 (and its five
 * siblings): a2 / a1 are the forces on the + / - faces; lc, lm, lp the
 * mobility*density products Xi*ro at the centre, - and + neighbours. */
static inline double upwind(double a2, double a1, double Xc, double rc, double Xm,
                            double rm, double Xp, double rp, double h)
{
    return (((a2 + fabs(a2)) / 2.0 - (a1 - fabs(a1)) / 2.0) * (-1) * Xc * rc -
            (a1 + fabs(a1)) / 2.0 * (-1) * Xm * rm +
            (a2 - fabs(a2)) / 2.0 * (-1) * Xp * rp) / h;
}

/* cpu.cpp:218-291 (assign_rS_nr). */
static void cell_roS(const filt_fields* f, const filt_params* p, int i, int j, int k, int localNx)
{
    if (!is_interior(i, j, k, localNx, p))
        return;
    const ptrdiff_t c = IDX(i, j, k);
    const int md = cell_media(f, c);
    const double* P1 = f->P1;
    const double* P2 = f->P2;
    const double* X1 = f->Xi1;
    const double* X2 = f->Xi2;
    const double* r1 = f->ro1;
    const double* r2 = f->ro2;

    f->roS1[c] = r1[c] * (1 - f->S2[c]);
    f->roS2[c] = r2[c] * f->S2[c];

    const double p1 = P1[c];
    const double p2 = P2[c];
    double a1, a2, f1, f2, f3, g1, g2, g3;

    if (p->Nz < 2) {
        f3 = 0;
        g3 = 0;
    } else {
        const ptrdiff_t zp = IDX(i, j, k + 1), zm = IDX(i, j, k - 1);
        a2 = -(P1[zp] - p1) / p->h3;
        a1 = -(p1 - P1[zm]) / p->h3;
        f3 = upwind(a2, a1, X1[c], r1[c], X1[zm], r1[zm], X1[zp], r1[zp], p->h3);
        a2 = -(P2[zp] - p2) / p->h3;
        a1 = -(p2 - P2[zm]) / p->h3;
        g3 = upwind(a2, a1, X2[c], r2[c], X2[zm], r2[zm], X2[zp], r2[zp], p->h3);
    }

    const ptrdiff_t xp = IDX(i + 1, j, k), xm = IDX(i - 1, j, k);
    const ptrdiff_t yp = IDX(i, j + 1, k), ym = IDX(i, j - 1, k);
    const double g = p->g_const;

    a2 = -(P1[xp] - p1) / p->h1;
    a1 = -(p1 - P1[xm]) / p->h1;
    f1 = upwind(a2, a1, X1[c], r1[c], X1[xm], r1[xm], X1[xp], r1[xp], p->h1);
    /* gravity enters both y faces with the centre density, cpu.cpp:257-258 */
    a2 = -(P1[yp] - p1) / p->h2 + r1[c] * g;
    a1 = -(p1 - P1[ym]) / p->h2 + r1[c] * g;
    f2 = upwind(a2, a1, X1[c], r1[c], X1[ym], r1[ym], X1[yp], r1[yp], p->h2);

    a2 = -(P2[xp] - p2) / p->h1;
    a1 = -(p2 - P2[xm]) / p->h1;
    g1 = upwind(a2, a1, X2[c], r2[c], X2[xm], r2[xm], X2[xp], r2[xp], p->h1);
    a2 = -(P2[yp] - p2) / p->h2 + r2[c] * g;
    a1 = -(p2 - P2[ym]) / p->h2 + r2[c] * g;
    g2 = upwind(a2, a1, X2[c], r2[c], X2[ym], r2[ym], X2[yp], r2[yp], p->h2);

    const double phi = cell_phi(f, p, c, md);
    const double A1 = f->roS1[c] - (p->dt / phi) * (f1 + f2 + f3);
    const double A2 = f->roS2[c] - (p->dt / phi) * (g1 + g2 + g3);
    f->roS1_old[c] = f->roS1[c];
    f->roS2_old[c] = f->roS2[c];
    f->roS1[c] = A1;
    f->roS2[c] = A2;
}

/* cpu.cpp:162-216 (assign_rS): the tau-relaxed "divgrad" scheme, selected by
 * p->roS_scheme.  It adds a capillary-like diffusion of ro*S with
 * coefficient m*l*c/2 and central transport terms of ro*u, and relaxes in
 * time with tau through roS_old (plain Euler on the first step, t < 2 dt).
 * The reference's own z transport terms are kept as written: water reads
 * ro1*u1x at (i+1, j, k+1) and (i, j, k-1), NAPL reads ro2*u2y at k+-1
 * (cpu.cpp:184-185). */
static void cell_roS_divgrad(const filt_fields* f, const filt_params* p, int i, int j, int k,
                             int localNx, double t)
{
    if (!is_interior(i, j, k, localNx, p))
        return;
    const ptrdiff_t c = IDX(i, j, k);
    const int md = cell_media(f, c);
    const double* r1 = f->ro1;
    const double* r2 = f->ro2;
    const double* S = f->S2;

    f->roS1[c] = r1[c] * (1 - S[c]);
    f->roS2[c] = r2[c] * S[c];

    const double m = cell_phi(f, p, c, md);
    const double k1 = m * p->l_w * p->c / 2.;
    const double k2 = m * p->l_n * p->c / 2.;
    double dg1, dg2, Tz1, Tz2;
    if (p->Nz < 2) {
        dg1 = 0;
        dg2 = 0;
        Tz1 = 0;
        Tz2 = 0;
    } else {
        const ptrdiff_t zp = IDX(i, j, k + 1), zm = IDX(i, j, k - 1), xz = IDX(i + 1, j, k + 1);
        dg1 = k1 * (r1[zp] * (1. - S[zp]) - 2 * r1[c] * (1. - S[c]) + r1[zm] * (1. - S[zm])) /
              (p->h3 * p->h3);
        dg2 = k2 * (r2[zp] * S[zp] - 2 * r2[c] * S[c] + r2[zm] * (S[zm])) / (p->h3 * p->h3);
        Tz1 = (r1[xz] * f->u1x[xz] - r1[zm] * f->u1x[zm]) / (2. * p->h3);
        Tz2 = (r2[zp] * f->u2y[zp] - r2[zm] * f->u2y[zm]) / (2. * p->h3);
    }
    const ptrdiff_t xp = IDX(i + 1, j, k), xm = IDX(i - 1, j, k);
    const ptrdiff_t yp = IDX(i, j + 1, k), ym = IDX(i, j - 1, k);
    dg1 += k1 * ((r1[xp] * (1. - S[xp]) - 2 * r1[c] * (1. - S[c]) + r1[xm] * (1. - S[xm])) /
                     (p->h1 * p->h1) +
                 (r1[yp] * (1. - S[yp]) - 2 * r1[c] * (1. - S[c]) + r1[ym] * (1. - S[ym])) /
                     (p->h2 * p->h2));
    dg2 += k2 * ((r2[xp] * S[xp] - 2 * r2[c] * S[c] + r2[xm] * (S[xm])) / (p->h1 * p->h1) +
                 (r2[yp] * S[yp] - 2 * r2[c] * S[c] + r2[ym] * (S[ym])) / (p->h2 * p->h2));
    const double Tx1 = (r1[xp] * f->u1x[xp] - r1[xm] * f->u1x[xm]) / (2. * p->h1);
    const double Ty1 = (r1[yp] * f->u1y[yp] - r1[ym] * f->u1y[ym]) / (2. * p->h2);
    const double Tx2 = (r2[xp] * f->u2x[xp] - r2[xm] * f->u2x[xm]) / (2. * p->h1);
    const double Ty2 = (r2[yp] * f->u2y[yp] - r2[ym] * f->u2y[ym]) / (2. * p->h2);

    const double dt = p->dt, tau = p->tau;
    double A1, A2;
    if (t < 2 * dt) {
        A1 = f->roS1[c] + (dt / m) * (dg1 - Tx1 - Ty1 - Tz1);
        A2 = f->roS2[c] + (dt / m) * (dg2 - Tx2 - Ty2 - Tz2);
    } else {
        A1 = (2. * dt * dt) / (m * (dt + 2. * tau)) *
             (dg1 - Tx1 - Ty1 - Tz1 + (2. * f->roS1[c] * m * tau) / (dt * dt) +
              f->roS1_old[c] * m * (dt - 2. * tau) / (2. * dt * dt));
        A2 = (2. * dt * dt) / (m * (dt + 2. * tau)) *
             (dg2 - Tx2 - Ty2 - Tz2 + (2. * f->roS2[c] * m * tau) / (dt * dt) +
              f->roS2_old[c] * m * (dt - 2. * tau) / (2. * dt * dt));
    }
    f->roS1_old[c] = f->roS1[c];
    f->roS2_old[c] = f->roS2[c];
    f->roS1[c] = A1;
    f->roS2[c] = A2;
}

/* cpu.cpp:293-320: Newton iterations on
 *   F1 = ro01(1+b1(P-Pa))(1-S) - roS1,  F2 = ro02(1+b2(P+Pk-Pa))S - roS2
 * with the 2x2 Jacobian inverted through its adjugate. */
static void cell_newton(const filt_fields* f, const filt_params* p, int i, int j, int k, int localNx)
{
    if (!is_interior(i, j, k, localNx, p))
        return;
    const ptrdiff_t c = IDX(i, j, k);
    const int md = cell_media(f, c);
    const double lam = p->lambda[md];
    const double swr = p->S_wr[md];
    const double pd = p->P_d[md];
    const double roS1 = f->roS1[c];
    const double roS2 = f->roS2[c];
    const double ro01 = p->ro01, ro02 = p->ro02, b1 = p->beta1, b2 = p->beta2, Pa = p->P_atm;
    double P = f->P1[c];
    double S = f->S2[c];

    for (int w = 1; w <= p->newton_iterations; w++) {
        const double S_e = (1 - S - swr) / (1 - swr);
        const double P_k = pd * pow(S_e, (-1.) / lam);
        const double AAA = pow(S_e, ((-1.) / lam) - 1.);
        const double F1 = ro01 * (1. + b1 * (P - Pa)) * (1. - S) - roS1;
        const double F2 = ro02 * (1. + b2 * (P + P_k - Pa)) * S - roS2;
        const double PkS = AAA * pd / (lam * (1 - swr));
        const double F1P = ro01 * b1 * (1 - S);
        const double F2P = ro02 * b2 * S;
        const double F1S = (-1) * ro01 * (1 + b1 * (P - Pa));
        const double F2S = ro02 * (1 + b2 * (P + P_k - Pa + (S * PkS)));
        const double det = F1P * F2S - F1S * F2P;
        const double Pn = P - (1 / det) * (F2S * F1 - F1S * F2);
        const double Sn = S - (1 / det) * (F1P * F2 - F2P * F1);
        P = Pn;
        S = Sn;
    }
    f->P1[c] = P;
    f->S2[c] = S;
}

/* cpu.cpp:323-363 (Border_S2).  Note the j == 0 rule has no return: the
 * k == 0 / k == Nz-1 rules still apply after it. */
static void cell_border_S2(const filt_fields* f, const filt_params* p, int i, int j, int k,
                           int localNx, int rank, int size)
{
    double* S2 = f->S2;
    const ptrdiff_t c = IDX(i, j, k);
    if (i == 0 && p->NX > 2) {
        S2[c] = S2[IDX(i + 1, j, k)];
        return;
    }
    if (i == localNx - 1 && p->NX > 2) {
        S2[c] = S2[IDX(i - 1, j, k)];
        return;
    }
    if (j == p->Ny - 1 && p->Ny > 2) {
        S2[c] = S2[IDX(i, j - 1, k)];
        return;
    }
    if (j == 0 && p->Ny > 2) {
        const int I = orc_i_to_I(i, rank, size, p);
        S2[c] = in_patch(I, k, p) ? p->S2_gr : 0;
    }
    if (k == 0 && p->Nz > 2) {
        S2[c] = S2[IDX(i, j, k + 1)];
        return;
    }
    if (k == p->Nz - 1 && p->Nz > 2) {
        S2[c] = S2[IDX(i, j, k - 1)];
        return;
    }
}

/* cpu.cpp:366-403 (Border_P1).  The bottom face adds the hydrostatic
 * increment with the water density of cell (i, 1, 0), a fixed index in the
 * reference (cpu.cpp:382). */
static void cell_border_P1(const filt_fields* f, const filt_params* p, int i, int j, int k,
                           int localNx)
{
    double* P1 = f->P1;
    const ptrdiff_t c = IDX(i, j, k);
    if (i == 0 && p->NX > 2) {
        P1[c] = P1[IDX(i + 1, j, k)];
        return;
    }
    if (i == localNx - 1 && p->NX > 2) {
        P1[c] = P1[IDX(i - 1, j, k)];
        return;
    }
    if (j == p->Ny - 1 && p->Ny > 2) {
        P1[c] = P1[IDX(i, j - 1, k)] + f->ro1[i + localNx * 1] * p->g_const * p->h2;
        return;
    }
    if (j == 0 && p->Ny > 2) {
        P1[c] = p->P_atm;
        return;
    }
    if (k == 0 && p->Nz > 2) {
        P1[c] = P1[IDX(i, j, k + 1)];
        return;
    }
    if (k == p->Nz - 1 && p->Nz > 2) {
        P1[c] = P1[IDX(i, j, k - 1)];
        return;
    }
}

/* ------------------------------------------------------------- sweeps -- */

typedef void (*cell_fn)(const filt_fields*, const filt_params*, int, int, int, int);

/* Apply fn to every active cell.  Stages 1-4 write only the cell they visit
 * and read only arrays the same stage does not write, so order and threading
 * cannot change the result. */
static void sweep(cell_fn fn, const filt_fields* f, const filt_params* p, int localNx,
                  int rank, int size)
{
    if (g_loop_order == 1) {
        for (int i = 0; i < localNx; i++)
            for (int j = 0; j < p->Ny; j++)
                for (int k = 0; k < p->Nz; k++)
                    if (orc_is_active_point(i, localNx, rank, size))
                        fn(f, p, i, j, k, localNx);
        return;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
#endif
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < localNx; i++)
                if (orc_is_active_point(i, localNx, rank, size))
                    fn(f, p, i, j, k, localNx);
}

static void cell_ro_P2_Xi_ijk(const filt_fields* f, const filt_params* p, int i, int j, int k,
                              int localNx)
{
    cell_ro_P2_Xi(f, p, IDX(i, j, k));
}

/* cpu.cpp:3-15 */
void orc_ro_P2_Xi_calculation(const filt_fields* f, const filt_params* p, int localNx,
                              int rank, int size)
{
    sweep(cell_ro_P2_Xi_ijk, f, p, localNx, rank, size);
}

/* cpu.cpp:17-24 */
void orc_u_calculation(const filt_fields* f, const filt_params* p, int localNx, int rank,
                       int size)
{
    sweep(cell_u, f, p, localNx, rank, size);
}

/* cpu.cpp:26-33 (t is unused by assign_rS_nr); the divgrad scheme reads
 * only ro, S2, u and roS_old and writes only roS, roS_old of its own cell,
 * so its cells are independent as well. */
void orc_roS_calculation(const filt_fields* f, const filt_params* p, double t, int localNx,
                         int rank, int size)
{
    if (p->roS_scheme != FILT_SCHEME_DIVGRAD) {
        sweep(cell_roS, f, p, localNx, rank, size);
        return;
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
#endif
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < localNx; i++)
                if (orc_is_active_point(i, localNx, rank, size))
                    cell_roS_divgrad(f, p, i, j, k, localNx, t);
}

/* cpu.cpp:35-42 */
void orc_P1_S2_calculation(const filt_fields* f, const filt_params* p, int localNx, int rank,
                           int size)
{
    sweep(cell_newton, f, p, localNx, rank, size);
}

/* cpu.cpp:44-54: in place, strictly in i -> j -> k order (later cells read
 * values earlier cells of the same sweep wrote). */
void orc_boundary_conditions(const filt_fields* f, const filt_params* p, int localNx,
                             int rank, int size)
{
    for (int i = 0; i < localNx; i++)
        for (int j = 0; j < p->Ny; j++)
            for (int k = 0; k < p->Nz; k++)
                if (orc_is_active_point(i, localNx, rank, size)) {
                    cell_border_S2(f, p, i, j, k, localNx, rank, size);
                    cell_border_P1(f, p, i, j, k, localNx);
                }
}

/* main.cpp:83-94 */
void orc_time_step(const filt_fields* f, const filt_params* p, double t, int localNx,
                   int rank, int size, int with_u)
{
    orc_ro_P2_Xi_calculation(f, p, localNx, rank, size);
    if (with_u || p->roS_scheme == FILT_SCHEME_DIVGRAD)  /* divgrad reads u */
        orc_u_calculation(f, p, localNx, rank, size);
    orc_roS_calculation(f, p, t, localNx, rank, size);
    orc_P1_S2_calculation(f, p, localNx, rank, size);
    orc_boundary_conditions(f, p, localNx, rank, size);
}

/* main.cpp:41-43 */
void orc_run(const filt_fields* f, const filt_params* p, int localNx, int rank, int size,
             int step0, int nsteps, int with_u)
{
    for (int s = step0 + 1; s <= step0 + nsteps; s++)
        orc_time_step(f, p, s * p->dt, localNx, rank, size, with_u);
}

/* ---------------------------------------------------------------- mass -- */

/* Pairwise sum of n terms produced by term(idx) in index order. */
typedef struct {
    const filt_fields* f;
    const filt_params* p;
    int localNx;
} mass_ctx;

static void mass_terms(const mass_ctx* m, ptrdiff_t c, int i, int j, int k, double out[2])
{
    const filt_fields* f = m->f;
    const filt_params* p = m->p;
    const int localNx = m->localNx;
    out[0] = out[1] = 0;
    if (!is_interior(i, j, k, localNx, p))
        return;
    const int md = cell_media(f, c);
    const double phi = cell_phi(f, p, c, md);
    const double V = p->h1 * p->h2 * p->h3;
    const double S = f->S2[c];
    const double P = f->P1[c];
    const double r1 = p->ro01 * (1. + p->beta1 * (P - p->P_atm));
    const double r2 = p->ro02 * (1. + p->beta2 * (P - p->P_atm));
    out[0] = phi * r1 * (1. - S) * V;
    out[1] = phi * r2 * S * V;
}

void orc_mass(const filt_fields* f, const filt_params* p, int localNx, double out[2])
{
    /* Row sums (x lines) accumulated sequentially, then a pairwise tree over
     * the Ny*Nz row sums. */
    const ptrdiff_t rows = (ptrdiff_t)p->Ny * p->Nz;
    double* r = (double*)malloc(sizeof(double) * 2 * (size_t)(rows > 0 ? rows : 1));
    mass_ctx m = {f, p, localNx};
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++) {
            double s0 = 0, s1 = 0, t[2];
            for (int i = 0; i < localNx; i++) {
                mass_terms(&m, IDX(i, j, k), i, j, k, t);
                s0 += t[0];
                s1 += t[1];
            }
            const ptrdiff_t row = (ptrdiff_t)j + (ptrdiff_t)k * p->Ny;
            r[2 * row] = s0;
            r[2 * row + 1] = s1;
        }
    ptrdiff_t n = rows;
    while (n > 1) {
        const ptrdiff_t half = (n + 1) / 2;
        for (ptrdiff_t q = 0; q < n / 2; q++) {
            r[2 * q] = r[4 * q] + r[4 * q + 2];
            r[2 * q + 1] = r[4 * q + 1] + r[4 * q + 3];
        }
        if (n % 2) {
            r[2 * (n / 2)] = r[2 * (n - 1)];
            r[2 * (n / 2) + 1] = r[2 * (n - 1) + 1];
        }
        n = half;
    }
    out[0] = rows > 0 ? r[0] : 0;
    out[1] = rows > 0 ? r[1] : 0;
    free(r);
}

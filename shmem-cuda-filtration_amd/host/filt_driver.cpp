// filt_driver.cpp -- driver pieces of the reference's main.cpp that sit on
// the hot path's setup: slab decomposition helpers and the initial state.
#include <stddef.h>

#include "../../include/filtration_hip.h"

extern "C" {

// main.cpp:145-158: local x index -> global x index of the x-slab split.
int filt_i_to_I(int i, int rank, int size, const filt_params* p)
{
    const int base = p->NX / size, rem = p->NX % size;
    if (rank <= rem)
        return rank == 0 ? i : (base + 1) * rank + i - 1;
    return (base + 1) * rank - (rank - rem) + i - 1;
}

// main.cpp:166-189: owners below NX%size take one more cell; every slab
// with a neighbour carries one halo cell per neighbour.
void filt_N_to_local(int* localNx, int* localNy, int size, int rank, const filt_params* p)
{
    int n = p->NX / size;
    if (rank < p->NX % size)
        n++;
    if (size > 1)
        n += (rank == 0 || rank == size - 1) ? 1 : 2;
    if (localNx)
        *localNx = n;
    if (localNy)
        *localNy = p->Ny;
}

// main.cpp:192-198
int filt_is_active_point(int i, int localNx, int rank, int size)
{
    return ((rank != 0 && i == 0) || (rank != size - 1 && i == localNx - 1)) ? 0 : 1;
}

// main.cpp:97-138 on host arrays: S2 = S2_gr on the source patch of the top
// face (j = 0), 0 elsewhere; P1 hydrostatic, P_atm + j*ro01*g*h2; node
// coordinates x = I*h1, y = j*h2, z = k*h3 and media 0 where those arrays
// are given (x, y, z, media may be null).
int filt_initial_data(const filt_fields* host, const filt_params* p, int localNx, int rank,
                      int size)
{
    if (!host || !p || !host->P1 || !host->S2 || localNx < 1 || p->Ny < 1 || p->Nz < 1)
        return FILT_ERR_INVALID;
    const long long sx = 1, sy = localNx, sz = (long long)localNx * p->Ny;
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < localNx; i++) {
                if (!filt_is_active_point(i, localNx, rank, size))
                    continue;
                const long long c = i * sx + j * sy + k * sz;
                const int I = filt_i_to_I(i, rank, size, p);
                const bool src = j == 0 && I >= p->src_x0 && I <= p->src_x1 && k >= p->src_z0 &&
                                 k <= p->src_z1;
                host->S2[c] = src ? p->S2_gr : 0;
                host->P1[c] = p->P_atm + (double)j * p->ro01 * p->g_const * p->h2;
                if (host->x) host->x[c] = I * p->h1;
                if (host->y) host->y[c] = j * p->h2;
                if (host->z) host->z[c] = k * p->h3;
                if (host->media) host->media[c] = 0;
            }
    return FILT_OK;
}

}  // extern "C"

// ref_harness.cpp -- drives the reference's own CPU path from Python.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
// with the unmodified reference sources main.cpp (its main renamed by
// -Dmain=reference_main), cpu.cpp and no_communication.cpp, found under the
// reference tree given by REF, into oracle/_ref/libref.so.  The functions
// below only marshal arguments: every number comes from the reference's code.
// Nothing here is shipped.
#include "defines.h"  // the reference's contract (defines.h:41-116), via -I$(REF)

#include <unistd.h>

extern double* HostBuffer;  // main.cpp:6

#ifndef REF_HARNESS_NO_CPU
// assign_rS (cpu.cpp:162-216): defined in cpu.cpp, not declared in defines.h.
// Left out of the build that links another backend in place of cpu.cpp.
void assign_rS(ptr_Arrays HostArraysPtr, double t, int i, int j, int k, int localNx, consts def);
#endif

extern "C" {

size_t ref_sizeof_consts(void) { return sizeof(consts); }
size_t ref_sizeof_ptr_arrays(void) { return sizeof(ptr_Arrays); }

// Reference compile-time constants (defines.h:8-15), for the Python side to
// check that it mirrors them.
void ref_constants(double* out)
{
    out[0] = mu1;
    out[1] = mu2;
    out[2] = ro01;
    out[3] = ro02;
    out[4] = S2_gr;
    out[5] = P_atm;
    out[6] = g_const;
    out[7] = NEWTON_ITERATIONS;
}

// One stage of time_step_function (main.cpp:83-94) on host arrays:
// 1 ro_P2_Xi_calculation, 2 u_calculation, 3 roS_calculation,
// 4 P1_S2_calculation, 5 boundary_conditions.
int ref_stage(int stage, const ptr_Arrays* host, const consts* def, int localNx, int rank,
              int size, double t)
{
    ptr_Arrays h = *host;
    ptr_Arrays d = *host;  // the CPU backend ignores the device bundle
    switch (stage) {
    case 1: ro_P2_Xi_calculation(h, d, *def, localNx, rank, size, 0, 0, 0); return 0;
    case 2: u_calculation(h, d, localNx, rank, size, 0, 0, 0, *def); return 0;
    case 3: roS_calculation(h, d, *def, t, localNx, rank, size, 0, 0, 0); return 0;
    case 4: P1_S2_calculation(h, d, *def, localNx, rank, size, 0, 0, 0); return 0;
    case 5: boundary_conditions(h, d, localNx, rank, size, 0, 0, 0, *def); return 0;
    }
    return 1;
}

#ifndef REF_HARNESS_NO_CPU
// roS_calculation's loop (cpu.cpp:26-33) with assign_rS in place of
// assign_rS_nr.
static void roS_divgrad(const ptr_Arrays& h, const consts& def, double t, int localNx, int rank,
                        int size)
{
    for (int i = 0; i < localNx; i++)
        for (int j = 0; j < def.Ny; j++)
            for (int k = 0; k < def.Nz; k++)
                if (is_active_point(i, localNx, rank, size))
                    assign_rS(h, t, i, j, k, localNx, def);
}

void ref_roS_divgrad(const ptr_Arrays* host, const consts* def, double t, int localNx, int rank,
                     int size)
{
    roS_divgrad(*host, *def, t, localNx, rank, size);
}

// time_step_function (main.cpp:83-94) with the divgrad roS stage, nsteps
// times from step0, t = s*dt as main.cpp:43.
void ref_run_divgrad(const ptr_Arrays* host, const consts* def, int localNx, int rank, int size,
                     int step0, int nsteps)
{
    ptr_Arrays h = *host;
    for (int s = step0 + 1; s <= step0 + nsteps; s++) {
        const double t = s * def->dt;
        ro_P2_Xi_calculation(h, h, *def, localNx, rank, size, 0, 0, 0);
        u_calculation(h, h, localNx, rank, size, 0, 0, 0, *def);
        roS_divgrad(h, *def, t, localNx, rank, size);
        P1_S2_calculation(h, h, *def, localNx, rank, size, 0, 0, 0);
        boundary_conditions(h, h, localNx, rank, size, 0, 0, 0, *def);
    }
}
#endif

// time_step_function (main.cpp:83-94), nsteps times, t = s*dt as main.cpp:43.
void ref_run(const ptr_Arrays* host, const consts* def, int localNx, int localNy, int rank,
             int size, int step0, int nsteps)
{
    ptr_Arrays h = *host;
    for (int s = step0 + 1; s <= step0 + nsteps; s++)
        time_step_function(h, h, HostBuffer, *def, s * def->dt, localNx, localNy, rank, size,
                           0, 0, 0);
}

// initial_data (main.cpp:97-138).
void ref_initial_data(const ptr_Arrays* host, const consts* def, int localNx, int localNy,
                      int rank, int size)
{
    int t = 0;
    initial_data(*host, &t, localNx, localNy, rank, size, *def);
}

int ref_i_to_I(int i, int rank, int size, const consts* def) { return i_to_I(i, rank, size, *def); }

void ref_N_to_local(int* localNx, int* localNy, int size, int rank, const consts* def)
{
    N_to_local(localNx, localNy, size, rank, *def);
}

int ref_is_active_point(int i, int localNx, int rank, int size)
{
    return is_active_point(i, localNx, rank, size);
}

// read_defines (main.cpp:601-683).  The reference opens the fixed relative
// path DEFINES_FILE (defines.h:5); the caller lays the file out under
// workdir and this changes into workdir for the call.
int ref_read_defines(const char* workdir, consts* out)
{
    char old[4096];
    if (!getcwd(old, sizeof old)) return 1;
    if (chdir(workdir) != 0) return 2;
    char arg0[] = "ref";
    char* argv[] = {arg0, nullptr};
    read_defines(1, argv, out);
    return chdir(old) != 0 ? 3 : 0;
}

const char* ref_defines_file(void) { return DEFINES_FILE; }

// load_exchange_data / save_exchange_data of the linked backend on host
// arrays (cpu.cpp:436-465 for the CPU backend): which = 0 packs array into
// buffer, which = 1 unpacks buffer into array.
void ref_exchange(int which, double* array, double* buffer, int localNx, int rank, int size,
                  const consts* def)
{
    if (which == 0)
        load_exchange_data(array, array, buffer, buffer, localNx, 0, 0, rank, size, *def);
    else
        save_exchange_data(array, array, buffer, buffer, localNx, 0, 0, rank, size, *def);
}

// The reference driver's file output, run in workdir (its paths are
// relative): save (main.cpp:485-537) writes save/save.dat, restore
// (main.cpp:540-569) reads it, save_data_plots (main.cpp:319-347) writes the
// plot_* folders and media.dat.  The CPU backend's load_data_to_host is a
// no-op, so the host arrays are what they write.
int ref_save(const char* workdir, const ptr_Arrays* host, const consts* def, int j, int localNx,
             int rank, int size)
{
    char old[4096];
    if (!getcwd(old, sizeof old) || chdir(workdir) != 0) return 1;
    save(*host, *host, j, rank, size, localNx, *def);
    return chdir(old) != 0 ? 3 : 0;
}

int ref_restore(const char* workdir, const ptr_Arrays* host, const consts* def, int* j,
                int localNx, int rank, int size)
{
    char old[4096];
    if (!getcwd(old, sizeof old) || chdir(workdir) != 0) return 1;
    restore(*host, j, rank, size, localNx, *def);
    return chdir(old) != 0 ? 3 : 0;
}

int ref_save_data_plots(const char* workdir, const ptr_Arrays* host, const consts* def, double t,
                        int localNx, int rank, int size)
{
    char old[4096];
    if (!getcwd(old, sizeof old) || chdir(workdir) != 0) return 1;
    save_data_plots(*host, *host, t, size, rank, localNx, *def);
    return chdir(old) != 0 ? 3 : 0;
}

extern double* DevBuffer;  // main.cpp:7

// The reference driver end to end with whichever backend this library links:
// Initialize (main.cpp:204-231: Device_Initialize, memory_alloc,
// initial_data, load_data_to_device), nsteps calls of time_step_function
// (main.cpp:41-43), load_data_to_host of P1 and S2 (main.cpp:322-323),
// Finalize.  Copies localNx*Ny*Nz values of P1 and S2 to the outputs.
int ref_drive(const consts* def, int nsteps, double* P1_out, double* S2_out, size_t capacity)
{
    ptr_Arrays H, D;
    memset(&H, 0, sizeof H);
    memset(&D, 0, sizeof D);
    int j = 0, localNx = 0, localNy = 0, size = 0, rank = 0, bx = 0, by = 0, bz = 0;
    char arg0[] = "ref";
    char* argv[] = {arg0, nullptr};
    Initialize(&H, &D, &j, &localNx, &localNy, &size, &rank, &bx, &by, &bz, 1, argv, *def);
    for (int s = 1; s <= nsteps; s++)
        time_step_function(H, D, DevBuffer, *def, s * def->dt, localNx, localNy, rank, size, bx,
                           by, bz);
    load_data_to_host(H.P1, D.P1, localNx, *def);
    load_data_to_host(H.S2, D.S2, localNx, *def);
    const size_t n = (size_t)localNx * def->Ny * def->Nz;
    if (n <= capacity) {
        memcpy(P1_out, H.P1, n * sizeof(double));
        memcpy(S2_out, H.S2, n * sizeof(double));
    }
    Finalize(H, D, DevBuffer);
    return n <= capacity ? 0 : 1;
}

}  // extern "C"

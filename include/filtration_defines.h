// filtration_defines.h -- the reference's backend contract, restated.
//
// shmem-cuda-filtration swaps its compute backend at link time against the
// declarations of defines.h (defines.h:41-116).  This header restates that
// contract with identical struct tags, member order and types (so the
// by-value structs have the same layout and every function the same mangled
// C++ name) and the same compile-time constants (defines.h:8-15).  It uses
// the reference's include guard on purpose: a translation unit that already
// saw the reference's defines.h skips this one, so both can never disagree.
//
// The HIP backend (lib/libfiltration_defines.so, source
// shmem-cuda-filtration_amd/host/backend_defines.cpp) defines the 13
// functions marked [backend] below; the reference's driver (main.cpp) and
// comm layer (no_communication.cpp / mpi.cpp) define the rest.
#ifndef DEFINES_H
#define DEFINES_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

// defines.h:8-15
#define mu1 0.001
#define mu2 0.0009
#define ro01 1000
#define ro02 1460
#define S2_gr 0.4
#define P_atm 100000
#define g_const 9.8
#define NEWTON_ITERATIONS 7

// defines.h:41-47: field arrays of one subdomain (host or device).
struct ptr_Arrays_tag {
    double *x, *y, *z;
    double *P1, *P2, *S2, *ro1, *ro2, *u1x, *u1y, *u1z, *u2x, *u2y, *u2z, *Xi1, *Xi2, *roS1,
        *roS1_old, *roS2, *roS2_old;
    int* media;
};
typedef struct ptr_Arrays_tag ptr_Arrays;

// defines.h:50-61: problem parameters (two media).
struct consts_tag {
    double K[2];
    double lambda[2];
    double S_wr[2];
    double m[2];
    double P_d[2];
    double h1, h2, h3, dt, tau, l_w, l_n, c, beta1, beta2;
    int NX, Ny, Nz;
    int source, save_plots, print_screen;
};
typedef struct consts_tag consts;

// ---- [backend] defines.h:68-82, 106-116 --------------------------------
void device_memory_alloc(ptr_Arrays* ArraysPtr, double** DevBuffer, int nX, consts def);
void device_memory_free(ptr_Arrays DevArraysPtr, double* DevBuffer);
void load_data_to_host(double* HostArrayPtr, double* DevArrayPtr, int localNx, consts def);
void load_data_to_device(double* HostArrayPtr, double* DevArrayPtr, int localNx, consts def);
void load_data_to_device_int(int* HostArrayPtr, int* DevArrayPtr, int localNx, consts def);
void Device_Initialize(int rank, int* blocksX, int* blocksY, int* blocksZ, int localNx,
                       consts def);
void ro_P2_Xi_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def,
                          int localNx, int rank, int size, int blocksX, int blocksY,
                          int blocksZ);
void u_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, int localNx, int rank,
                   int size, int blocksX, int blocksY, int blocksZ, consts def);
void roS_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def, double t,
                     int localNx, int rank, int size, int blocksX, int blocksY, int blocksZ);
void P1_S2_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def,
                       int localNx, int rank, int size, int blocksX, int blocksY, int blocksZ);
void boundary_conditions(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, int localNx,
                         int rank, int size, int blocksX, int blocksY, int blocksZ,
                         consts def);
void load_exchange_data(double* HostArrayPtr, double* DevArrayPtr, double* HostBuffer,
                        double* DevBuffer, int localNx, int blocksY, int blocksZ, int rank,
                        int size, consts def);
void save_exchange_data(double* HostArrayPtr, double* DevArrayPtr, double* HostBuffer,
                        double* DevBuffer, int localNx, int blocksY, int blocksZ, int rank,
                        int size, consts def);

// ---- driver and comm layer, defines.h:63-104, 107-113 --------------------
void time_step_function(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, double* DevBuffer,
                        consts def, double t, int localNx, int localNy, int rank, int size,
                        int blocksX, int blocksY, int blocksZ);
void Initialize(ptr_Arrays* HostArraysPtr, ptr_Arrays* DevArraysPtr, int* j, int* localNx,
                int* localNy, int* size, int* rank, int* blocksX, int* blocksY, int* blocksZ,
                int argc, char* argv[], consts def);
void Finalize(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, double* DevBuffer);
void memory_alloc(ptr_Arrays* HostArraysPtr, ptr_Arrays* DevArraysPtr, int localNx, int nY,
                  consts def);
void host_memory_alloc(ptr_Arrays* ArraysPtr, int nX, int nY, consts def);
void memory_free(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr);
void host_memory_free(ptr_Arrays HostArraysPtr);
void save_data_plots(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, double t, int size,
                     int rank, int localNx, consts def);
void initial_data(ptr_Arrays HostArraysPtr, int* t, int localNx, int localNy, int rank,
                  int size, consts def);
void Communication_Initialize(int argc, char* argv[], int* size, int* rank, consts def);
void Communication_Finalize(void);
void N_to_local(int* localNx, int* localNy, int size, int rank, consts def);
int is_active_point(int i, int localNx, int rank, int size);
void print_plots_top(double t, consts def);
void print_plots(ptr_Arrays HostArraysPtr, double t, int rank, int size, int localNx,
                 consts def);
void Barrier(void);
void restore(ptr_Arrays HostArraysPtr, int* j, int rank, int size, int localNx, consts def);
void save(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, int j, int rank, int size,
          int localNx, consts def);
void read_defines(int argc, char* argv[], consts* def);
void test_correct_P1_S2(ptr_Arrays HostArraysPtr, int nX, int rank, consts def);
int i_to_I(int i, int rank, int size, consts def);
void P2_ro_Xi_exchange(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, double* HostBuffer,
                       double* DevBuffer, int localNx, int blocksY, int blocksZ, int rank,
                       int size, consts def);
void u_exchange(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, double* HostBuffer,
                double* DevBuffer, int localNx, int blocksY, int blocksZ, int rank, int size,
                consts def);
void P1_S2_exchange(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, double* HostBuffer,
                    double* DevBuffer, int localNx, int blocksY, int blocksZ, int rank,
                    int size, consts def);

#endif  // DEFINES_H

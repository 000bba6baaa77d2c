// backend_defines.cpp -- the defines.h compute backend over the HIP C ABI.
//
// Defines, with C++ linkage and the reference's signatures, the 13 functions
// gpu.cu provides (defines.h:68-82, 106-116), so that the reference driver
// main.cpp links against lib/libfiltration_defines.so in place of gpu.cu.
// Compile against the reference's own defines.h with
//   -DFILT_DEFINES_HEADER='"defines.h"' -I<reference tree>
// or, by default, against include/filtration_defines.h (same layout).
//
// Behaviour kept from the reference backend: each call is complete when it
// returns (gpu.cu synchronises after every launch, gpu.cu:7-27); errors are
// printed and execution continues.  Set FILT_ABORT_ON_ERROR=1 to abort
// instead.
//
// One difference, invisible through this interface: time_step_function
// (main.cpp:86-91) calls u_calculation, roS_calculation and P1_S2_calculation
// right after ro_P2_Xi_calculation on the same arrays (the exchanges between
// them do nothing on one rank).  The shim defers u and roS and runs the three
// stages in P1_S2_calculation as one z-marching launch
// (filt_u_roS_P1_S2_calculation), which writes every array the three stages
// write with the same bits and moves 136 instead of 240 bytes per cell.  The
// fused kernel relies on stage 1 having just run (P2 then holds P1's bits),
// so the shim only defers u when the previous call was ro_P2_Xi_calculation
// with the same arrays and parameters.  Every other entry point first runs
// whatever is deferred, so data read back through load_data_to_host or the
// exchange functions is what the separate calls would have left.  Device
// memory read by other means between u_calculation and P1_S2_calculation
// would not hold u yet.  FILT_FUSE_STAGES=0 turns the deferral off.  The
// deferred call is process-wide state: call the shims from one thread at a
// time, as the reference driver does.
#ifdef FILT_DEFINES_HEADER
#include FILT_DEFINES_HEADER
#else
#include "filtration_defines.h"
#endif

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

// The header's compile-time constants (defines.h:8-15) are macros whose names
// collide with filt_params members: take their values, then drop the macros.
namespace {
const double k_mu1 = mu1, k_mu2 = mu2, k_ro01 = ro01, k_ro02 = ro02;
const double k_S2_gr = S2_gr, k_P_atm = P_atm, k_g_const = g_const;
const int k_newton_iterations = NEWTON_ITERATIONS;
}  // namespace
#undef mu1
#undef mu2
#undef ro01
#undef ro02
#undef S2_gr
#undef P_atm
#undef g_const

#include "filtration_hip.h"

namespace {

filt_params to_params(const consts& def)
{
    filt_params p;
    filt_default_params(&p);
    for (int q = 0; q < 2; q++) {
        p.K[q] = def.K[q];
        p.lambda[q] = def.lambda[q];
        p.S_wr[q] = def.S_wr[q];
        p.m[q] = def.m[q];
        p.P_d[q] = def.P_d[q];
    }
    p.h1 = def.h1;
    p.h2 = def.h2;
    p.h3 = def.h3;
    p.dt = def.dt;
    p.tau = def.tau;
    p.l_w = def.l_w;
    p.l_n = def.l_n;
    p.c = def.c;
    p.beta1 = def.beta1;
    p.beta2 = def.beta2;
    p.NX = def.NX;
    p.Ny = def.Ny;
    p.Nz = def.Nz;
    p.source = def.source;
    p.save_plots = def.save_plots;
    p.print_screen = def.print_screen;
    // the header's compile-time constants, defines.h:8-15
    p.mu1 = k_mu1;
    p.mu2 = k_mu2;
    p.ro01 = k_ro01;
    p.ro02 = k_ro02;
    p.S2_gr = k_S2_gr;
    p.P_atm = k_P_atm;
    p.g_const = k_g_const;
    p.newton_iterations = k_newton_iterations;
    p.n_media = 2;
    p.src_x0 = -1;
    // The reference switches roS_calculation to assign_rS (cpu.cpp:162-216)
    // only by editing cpu.cpp:31; FILT_ROS_SCHEME=1 selects it here.
    const char* scheme = getenv("FILT_ROS_SCHEME");
    p.roS_scheme = scheme ? atoi(scheme) : FILT_SCHEME_UPWIND;
    filt_params_finalize(&p);
    return p;
}

filt_fields to_fields(const ptr_Arrays& a)
{
    filt_fields f;
    f.x = a.x;
    f.y = a.y;
    f.z = a.z;
    f.P1 = a.P1;
    f.P2 = a.P2;
    f.S2 = a.S2;
    f.ro1 = a.ro1;
    f.ro2 = a.ro2;
    f.u1x = a.u1x;
    f.u1y = a.u1y;
    f.u1z = a.u1z;
    f.u2x = a.u2x;
    f.u2y = a.u2y;
    f.u2z = a.u2z;
    f.Xi1 = a.Xi1;
    f.Xi2 = a.Xi2;
    f.roS1 = a.roS1;
    f.roS1_old = a.roS1_old;
    f.roS2 = a.roS2;
    f.roS2_old = a.roS2_old;
    f.media = a.media;
    f.perm = nullptr;
    f.poro = nullptr;
    return f;
}

void report(int rc, const char* where)
{
    if (rc == FILT_OK)
        return;
    printf("HIP backend error: %s (at %s)\n", filt_last_error(), where);
    fflush(stdout);
    const char* a = getenv("FILT_ABORT_ON_ERROR");
    if (a && a[0] == '1')
        abort();
}

// Launch, then wait: a stage is complete on return (gpu.cu:7-27).
void finish(int rc, const char* where)
{
    report(rc, where);
    if (rc == FILT_OK)
        report(filt_synchronize(), where);
}

// The stages 2-4 chain of time_step_function: what has run or been deferred
// since the last stage-1 call, and that call's arguments.
struct Chain {
    enum { kNone, kStage1, kU, kURoS } state = kNone;
    ptr_Arrays dev;
    consts def;
    int localNx, rank, size;
    double t;
};
Chain g_chain;

bool fuse_enabled()
{
    const char* e = getenv("FILT_FUSE_STAGES");
    return !(e && e[0] == '0');
}

// ptr_Arrays (pointers only) and consts (20 doubles, 6 ints) have no padding,
// so equal bytes mean equal arguments.
bool same_call(const ptr_Arrays& dev, const consts& def, int localNx, int rank, int size)
{
    return memcmp(&dev, &g_chain.dev, sizeof dev) == 0 &&
           memcmp(&def, &g_chain.def, sizeof def) == 0 && localNx == g_chain.localNx &&
           rank == g_chain.rank && size == g_chain.size;
}

// Runs the deferred stages, in the order the caller made the calls.
void flush()
{
    const int s = g_chain.state;
    g_chain.state = Chain::kNone;
    if (s < Chain::kU)
        return;
    filt_params p = to_params(g_chain.def);
    filt_fields f = to_fields(g_chain.dev);
    finish(filt_u_calculation(&f, &p, g_chain.localNx, g_chain.rank, g_chain.size),
           "u_calculation");
    if (s == Chain::kURoS)
        finish(filt_roS_calculation(&f, &p, g_chain.t, g_chain.localNx, g_chain.rank,
                                    g_chain.size),
               "roS_calculation");
}

}  // namespace

void device_memory_alloc(ptr_Arrays* ArraysPtr, double** DevBuffer, int nX, consts def)
{
    flush();
    filt_params p = to_params(def);
    filt_fields f;
    const int rc = filt_device_memory_alloc(&f, DevBuffer, nX, &p);
    report(rc, "device_memory_alloc");
    if (rc == FILT_OK && ArraysPtr) {
        ArraysPtr->x = f.x;
        ArraysPtr->y = f.y;
        ArraysPtr->z = f.z;
        ArraysPtr->P1 = f.P1;
        ArraysPtr->P2 = f.P2;
        ArraysPtr->S2 = f.S2;
        ArraysPtr->ro1 = f.ro1;
        ArraysPtr->ro2 = f.ro2;
        ArraysPtr->u1x = f.u1x;
        ArraysPtr->u1y = f.u1y;
        ArraysPtr->u1z = f.u1z;
        ArraysPtr->u2x = f.u2x;
        ArraysPtr->u2y = f.u2y;
        ArraysPtr->u2z = f.u2z;
        ArraysPtr->Xi1 = f.Xi1;
        ArraysPtr->Xi2 = f.Xi2;
        ArraysPtr->roS1 = f.roS1;
        ArraysPtr->roS1_old = f.roS1_old;
        ArraysPtr->roS2 = f.roS2;
        ArraysPtr->roS2_old = f.roS2_old;
        ArraysPtr->media = f.media;
    }
}

void device_memory_free(ptr_Arrays DevArraysPtr, double* DevBuffer)
{
    flush();
    filt_fields f = to_fields(DevArraysPtr);
    report(filt_device_memory_free(&f, DevBuffer), "device_memory_free");
}

void load_data_to_host(double* HostArrayPtr, double* DevArrayPtr, int localNx, consts def)
{
    flush();
    filt_params p = to_params(def);
    report(filt_load_data_to_host(HostArrayPtr, DevArrayPtr, localNx, &p), "load_data_to_host");
}

void load_data_to_device(double* HostArrayPtr, double* DevArrayPtr, int localNx, consts def)
{
    flush();
    filt_params p = to_params(def);
    report(filt_load_data_to_device(HostArrayPtr, DevArrayPtr, localNx, &p),
           "load_data_to_device");
}

void load_data_to_device_int(int* HostArrayPtr, int* DevArrayPtr, int localNx, consts def)
{
    flush();
    filt_params p = to_params(def);
    report(filt_load_data_to_device_int(HostArrayPtr, DevArrayPtr, localNx, &p),
           "load_data_to_device_int");
}

void Device_Initialize(int rank, int* blocksX, int* blocksY, int* blocksZ, int localNx,
                       consts def)
{
    flush();
    filt_params p = to_params(def);
    const int rc = filt_device_initialize(rank, localNx, &p, blocksX, blocksY, blocksZ);
    if (rc == FILT_OK) {
        printf("HIP initialized: %s\n", filt_build_info());
        fflush(stdout);
    }
    report(rc, "Device_Initialize");
}

void ro_P2_Xi_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def,
                          int localNx, int rank, int size, int blocksX, int blocksY, int blocksZ)
{
    flush();
    filt_params p = to_params(def);
    filt_fields f = to_fields(DevArraysPtr);
    const int rc = filt_ro_P2_Xi_calculation(&f, &p, localNx, rank, size);
    finish(rc, "ro_P2_Xi_calculation");
    if (rc == FILT_OK && fuse_enabled()) {
        g_chain.state = Chain::kStage1;
        g_chain.dev = DevArraysPtr;
        g_chain.def = def;
        g_chain.localNx = localNx;
        g_chain.rank = rank;
        g_chain.size = size;
    }
}

void u_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, int localNx, int rank,
                   int size, int blocksX, int blocksY, int blocksZ, consts def)
{
    if (g_chain.state == Chain::kStage1 && same_call(DevArraysPtr, def, localNx, rank, size)) {
        g_chain.state = Chain::kU;  // runs in P1_S2_calculation or at the next flush
        return;
    }
    flush();
    filt_params p = to_params(def);
    filt_fields f = to_fields(DevArraysPtr);
    finish(filt_u_calculation(&f, &p, localNx, rank, size), "u_calculation");
}

void roS_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def, double t,
                     int localNx, int rank, int size, int blocksX, int blocksY, int blocksZ)
{
    if (g_chain.state == Chain::kU && same_call(DevArraysPtr, def, localNx, rank, size)) {
        g_chain.state = Chain::kURoS;
        g_chain.t = t;
        return;
    }
    flush();
    filt_params p = to_params(def);
    filt_fields f = to_fields(DevArraysPtr);
    finish(filt_roS_calculation(&f, &p, t, localNx, rank, size), "roS_calculation");
}

void P1_S2_calculation(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, consts def,
                       int localNx, int rank, int size, int blocksX, int blocksY, int blocksZ)
{
    filt_params p = to_params(def);
    filt_fields f = to_fields(DevArraysPtr);
    if (g_chain.state == Chain::kURoS && same_call(DevArraysPtr, def, localNx, rank, size)) {
        g_chain.state = Chain::kNone;
        finish(filt_u_roS_P1_S2_calculation(&f, &p, g_chain.t, localNx, rank, size, 1),
               "P1_S2_calculation");
        return;
    }
    flush();
    finish(filt_P1_S2_calculation(&f, &p, localNx, rank, size), "P1_S2_calculation");
}

void boundary_conditions(ptr_Arrays HostArraysPtr, ptr_Arrays DevArraysPtr, int localNx,
                         int rank, int size, int blocksX, int blocksY, int blocksZ, consts def)
{
    flush();
    filt_params p = to_params(def);
    filt_fields f = to_fields(DevArraysPtr);
    finish(filt_boundary_conditions(&f, &p, localNx, rank, size), "boundary_conditions");
}

void load_exchange_data(double* HostArrayPtr, double* DevArrayPtr, double* HostBuffer,
                        double* DevBuffer, int localNx, int blocksY, int blocksZ, int rank,
                        int size, consts def)
{
    flush();
    filt_params p = to_params(def);
    report(filt_load_exchange_data(DevArrayPtr, DevBuffer, HostBuffer, localNx, &p),
           "load_exchange_data");
}

void save_exchange_data(double* HostArrayPtr, double* DevArrayPtr, double* HostBuffer,
                        double* DevBuffer, int localNx, int blocksY, int blocksZ, int rank,
                        int size, consts def)
{
    flush();
    filt_params p = to_params(def);
    report(filt_save_exchange_data(DevArrayPtr, DevBuffer, HostBuffer, localNx, rank, size, &p),
           "save_exchange_data");
}

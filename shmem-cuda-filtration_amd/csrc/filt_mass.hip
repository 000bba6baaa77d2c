// filt_mass.hip -- total water and NAPL mass of a grid on the device.
//
// M1 = sum phi*ro1(P1)*(1-S2)*V and M2 = sum phi*ro2(P1)*S2*V over the
// interior cells (cpu.cpp:220's test), V = h1*h2*h3, with the densities of
// stage 1 (cpu.cpp:72-73) and the porosity of the cell (per-cell field or
// medium).  The sum is deterministic: block b sums cells
// [b*CHUNK, (b+1)*CHUNK), each thread a fixed stride of them in order and
// the block a fixed tree; one block then sums the partials, each thread a
// fixed stride in order and a fixed tree.  The result does not depend on the
// launch or on other work.  It is the accumulator the mass-balance series of
// a run reads step by step without copying the state to the host.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "../../include/filtration_hip.h"
#include "filt_device.hpp"
#include "filt_runtime.hpp"
#include "../host/filt_guard.hpp"

namespace filt {

namespace {

constexpr int CHUNK = 4096;  // cells per partial
constexpr int NT = 256;      // threads per block

struct MassIn {
    const double* P1;
    const double* S2;
    const double* poro;
    const int* media;
    int nx, ny, nz;
    int n_media;
    double m[FILT_MAX_MEDIA];
    double ro01, ro02, beta1, beta2, P_atm, V;
};

__device__ __forceinline__ void block_tree(double a, double b, double* out)
{
    __shared__ double s1[NT], s2[NT];
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s1[0];
        out[1] = s2[0];
    }
}

__global__ __launch_bounds__(NT) void k_grid_mass_partial(MassIn in, double* __restrict__ partial)
{
    const long long n = (long long)in.nx * in.ny * in.nz;
    const long long base = (long long)blockIdx.x * CHUNK;
    double a1 = 0, a2 = 0;
    for (int u = 0; u < CHUNK / NT; u++) {
        const long long c = base + u * NT + threadIdx.x;
        if (c >= n)
            break;
        const int i = (int)(c % in.nx);
        const long long r = c / in.nx;
        const int j = (int)(r % in.ny), k = (int)(r / in.ny);
        if (!is_interior(i, j, k, in.nx, in.ny, in.nz))
            continue;
        const int md = (in.media && in.n_media > 1) ? in.media[c] : 0;
        const double phi = in.poro ? in.poro[c] : in.m[md];
        const double P = in.P1[c], S = in.S2[c];
        const double r1 = in.ro01 * (1. + in.beta1 * (P - in.P_atm));
        const double r2 = in.ro02 * (1. + in.beta2 * (P - in.P_atm));
        a1 += phi * r1 * (1. - S) * in.V;
        a2 += phi * r2 * S * in.V;
    }
    block_tree(a1, a2, partial + 2 * blockIdx.x);
}

__global__ __launch_bounds__(NT) void k_grid_mass_final(const double* __restrict__ partial,
                                                        long long nb, double* __restrict__ out)
{
    double a = 0, b = 0;
    for (long long q = threadIdx.x; q < nb; q += NT) {
        a += partial[2 * q];
        b += partial[2 * q + 1];
    }
    block_tree(a, b, out);
}

MassIn mass_in(const filt_fields& f, const filt_params& p, int nx)
{
    MassIn in;
    in.P1 = f.P1;
    in.S2 = f.S2;
    in.poro = f.poro;
    in.media = f.media;
    in.nx = nx;
    in.ny = p.Ny;
    in.nz = p.Nz;
    in.n_media = p.n_media;
    for (int q = 0; q < FILT_MAX_MEDIA; q++)
        in.m[q] = p.m[q];
    in.ro01 = p.ro01;
    in.ro02 = p.ro02;
    in.beta1 = p.beta1;
    in.beta2 = p.beta2;
    in.P_atm = p.P_atm;
    in.V = p.h1 * p.h2 * p.h3;
    return in;
}

// Partials of filt_mass_noalloc: a fixed array in each GPU's memory, held
// until the process ends.  Host code never passes its name to a kernel or a
// copy (in host code it stands for a host-side shadow, not a device
// address): it uses the device address hipGetSymbolAddress returns for the
// current device, looked up once per device (scratch_on).
constexpr long long NOALLOC_MAX_PARTIALS = 65536;  // 2 doubles each: 1 MiB
constexpr long long NOALLOC_MAX_CELLS = NOALLOC_MAX_PARTIALS * CHUNK;
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_PARTIALS];

std::mutex g_scratch_mutex;
std::vector<double*> g_scratch_ptr;  // per device; null until looked up

// The current device's address of `scratch` (looked up on the first call
// with that device current).
int scratch_on(double** out)
{
    int dev = 0;
    FILT_HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= dev)
        g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                            "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = (double*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return FILT_OK;
}

// Partials of the mass kernels: one buffer per process, grown on demand and
// kept until the process ends (calls are serialised by the mutex).
std::mutex g_mass_mutex;
double* g_mass_buf = nullptr;
size_t g_mass_cap = 0;  // doubles

}  // namespace

// Looks up this device's scratch address at set-up (filt_device_initialize),
// so that a later no-allocation call finds it ready.
int mass_scratch_init()
{
    double* p = nullptr;
    return scratch_on(&p);
}

long long mass_partials(long long cells) { return (cells + CHUNK - 1) / CHUNK; }

// Enqueue the two mass kernels: partials (2 per CHUNK cells) in `partial`
// (device), the result in out2 (device).
int launch_grid_mass(const filt_fields& f, const filt_params& p, int nx, double* partial,
                     double* out2, hipStream_t s)
{
    const long long n = (long long)nx * p.Ny * p.Nz;
    const long long nb = mass_partials(n);
    if (nb <= 0 || nb > 0x7fffffffLL)
        return set_error(FILT_ERR_INVALID, "grid size out of range for the mass sum");
    hipLaunchKernelGGL(k_grid_mass_partial, dim3((unsigned)nb), dim3(NT), 0, s, mass_in(f, p, nx),
                       partial);
    FILT_LAUNCH_CHECK("k_grid_mass_partial");
    hipLaunchKernelGGL(k_grid_mass_final, dim3(1), dim3(NT), 0, s, partial, nb, out2);
    FILT_LAUNCH_CHECK("k_grid_mass_final");
    return FILT_OK;
}

}  // namespace filt

extern "C" int filt_mass(const filt_fields* dev, const filt_params* p, int localNx, double* mass2)
{
    using namespace filt;
    return guarded([&]() -> int {
        if (!dev || !mass2 || !dev->P1 || !dev->S2)
            return set_error(FILT_ERR_INVALID, "null argument");
        int rc = check_shape(p, localNx);
        if (rc)
            return rc;
        const long long nb = mass_partials((long long)localNx * p->Ny * p->Nz);
        std::lock_guard<std::mutex> lock(g_mass_mutex);
        const size_t need = 2 * (size_t)nb + 2;
        if (need > g_mass_cap) {
            // work enqueued earlier may still read the old buffer
            FILT_HIP_TRY(hipStreamSynchronize(stream()));
            if (g_mass_buf)
                (void)hipFree(g_mass_buf);
            g_mass_buf = nullptr;
            g_mass_cap = 0;
            FILT_HIP_TRY(hipMalloc((void**)&g_mass_buf, need * sizeof(double)));
            g_mass_cap = need;
        }
        double* out = g_mass_buf + 2 * nb;
        rc = launch_grid_mass(*dev, *p, localNx, g_mass_buf, out, stream());
        if (rc)
            return rc;
        FILT_HIP_TRY(hipMemcpyAsync(mass2, out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                    stream()));
        FILT_HIP_TRY(hipStreamSynchronize(stream()));
        return FILT_OK;
    });
}

extern "C" int filt_mass_noalloc(const filt_fields* dev, const filt_params* p, int localNx,
                                 double* dev_mass2)
{
    using namespace filt;
    return guarded([&]() -> int {
        if (!dev || !dev_mass2 || !dev->P1 || !dev->S2)
            return set_error(FILT_ERR_INVALID, "null argument");
        int rc = check_shape(p, localNx);
        if (rc)
            return rc;
        const long long cells = (long long)localNx * p->Ny * p->Nz;
        if (cells > NOALLOC_MAX_CELLS)
            return set_error(FILT_ERR_INVALID,
                             "filt_mass_noalloc: %lld cells exceed its fixed scratch of %lld "
                             "partials (at most %lld cells); use filt_mass",
                             cells, NOALLOC_MAX_PARTIALS, NOALLOC_MAX_CELLS);
        double* sp = nullptr;
        rc = scratch_on(&sp);
        if (rc)
            return rc;
        // every partial the final pass reads is written by this call's first
        // kernel, so nothing needs clearing first
        return launch_grid_mass(*dev, *p, localNx, sp, dev_mass2, stream());
    });
}

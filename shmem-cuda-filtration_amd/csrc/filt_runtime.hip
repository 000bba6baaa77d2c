// filt_runtime.hip -- stream, errors, device memory, copies, halo packing.
//
// C-ABI entry points for the device-management half of the defines.h
// contract (defines.h:68-82, 115-116; gpu.cu:483-653).
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/filtration_hip.h"
#include "filt_device.hpp"
#include "filt_runtime.hpp"

namespace filt {

static thread_local std::string t_error;
static hipStream_t g_user_stream = nullptr;
static hipStream_t g_own_stream = nullptr;
static bool g_use_user = false;
static std::mutex g_stream_mu;

hipStream_t stream()
{
    if (g_use_user)
        return g_user_stream;
    if (!g_own_stream) {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        if (!g_own_stream) {
            hipStream_t s = nullptr;
            if (hipStreamCreate(&s) == hipSuccess)  // default priority
                g_own_stream = s;
        }
    }
    return g_own_stream;
}

int set_error(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_error = buf;
    return code;
}

void clear_error() { t_error.clear(); }

int hip_fail(hipError_t e, const char* where)
{
    return set_error(e == hipErrorOutOfMemory ? FILT_ERR_NOMEM : FILT_ERR_HIP, "%s: %s (%d)",
                     where, hipGetErrorString(e), (int)e);
}

int check_shape(const filt_params* p, int localNx)
{
    if (!p)
        return set_error(FILT_ERR_INVALID, "null params");
    if (localNx < 1 || p->Ny < 1 || p->Nz < 1 || p->NX < 1)
        return set_error(FILT_ERR_INVALID, "bad grid localNx=%d NX=%d Ny=%d Nz=%d", localNx,
                         p->NX, p->Ny, p->Nz);
    if ((long long)localNx * p->Ny * p->Nz > (1LL << 40))
        return set_error(FILT_ERR_INVALID, "grid too large");
    return FILT_OK;
}

// Byte size of one array of the grid, rounded up to 256 bytes.
static size_t array_bytes(long long cells, size_t elem)
{
    size_t b = (size_t)cells * elem;
    return (b + 255) & ~(size_t)255;
}

}  // namespace filt

using namespace filt;

extern "C" {

const char* filt_last_error(void) { return t_error.c_str(); }

const char* filt_build_info(void)
{
    return "filtration_hip gfx950 FP64 (" __DATE__ ")";
}

int filt_set_stream(void* s)
{
    g_user_stream = (hipStream_t)s;
    g_use_user = (s != nullptr);
    return FILT_OK;
}

void* filt_get_stream(void) { return (void*)stream(); }

int filt_synchronize(void)
{
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    return FILT_OK;
}

int filt_device_malloc(void** ptr, size_t n)
{
    if (!ptr)
        return set_error(FILT_ERR_INVALID, "null ptr");
    *ptr = nullptr;
    FILT_HIP_TRY(hipMalloc(ptr, n ? n : 1));
    return FILT_OK;
}

int filt_device_free_ptr(void* ptr)
{
    if (ptr)
        FILT_HIP_TRY(hipFree(ptr));
    return FILT_OK;
}

int filt_memcpy(void* dst, const void* src, size_t n, int direction)
{
    if ((!dst || !src) && n)
        return set_error(FILT_ERR_INVALID, "null pointer");
    static const hipMemcpyKind kinds[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                                           hipMemcpyDeviceToDevice};
    if (direction < 0 || direction > 2)
        return set_error(FILT_ERR_INVALID, "bad direction %d", direction);
    if (!n)
        return FILT_OK;
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    FILT_HIP_TRY(hipMemcpyAsync(dst, src, n, kinds[direction], stream()));
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    return FILT_OK;
}

// gpu.cu:565-610: device rank % count, reference block counts, capacity check.
int filt_device_initialize(int rank, int localNx, const filt_params* p, int* bx, int* by, int* bz)
{
    int rc = check_shape(p, localNx);
    if (rc)
        return rc;
    int count = 0;
    FILT_HIP_TRY(hipGetDeviceCount(&count));
    if (count < 1)
        return set_error(FILT_ERR_HIP, "no HIP device");
    FILT_HIP_TRY(hipSetDevice(rank % count));
    rc = mass_scratch_init();  // filt_mass_noalloc's scratch on this device
    if (rc)
        return rc;
    if (bx) *bx = (localNx + 15) / 16;
    if (by) *by = (p->Ny + 3) / 4;
    if (bz) *bz = (p->Nz + 3) / 4;
    size_t free_b = 0, total_b = 0;
    FILT_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const double need = (double)(localNx + 2) * p->Ny * p->Nz * (20 * 8 + 4);
    if (need > (double)free_b)
        return set_error(FILT_ERR_NOMEM, "grid needs %.3g bytes, %.3g free", need, (double)free_b);
    return FILT_OK;
}

// gpu.cu:504-531.  One allocation backs the 21 arrays (x first, so that
// freeing x frees all); the exchange buffer is separate.
int filt_device_memory_alloc(filt_fields* dev, double** dev_buffer, int nX, const filt_params* p)
{
    if (!dev)
        return set_error(FILT_ERR_INVALID, "null fields");
    int rc = check_shape(p, nX);
    if (rc)
        return rc;
    memset(dev, 0, sizeof *dev);
    const long long cells = (long long)nX * p->Ny * p->Nz;
    const size_t ad = array_bytes(cells, sizeof(double));
    const size_t ai = array_bytes(cells, sizeof(int));
    char* base = nullptr;
    FILT_HIP_TRY(hipMalloc((void**)&base, 20 * ad + ai));
    double** slots[20] = {&dev->x,   &dev->y,   &dev->z,   &dev->P1,  &dev->P2,
                          &dev->S2,  &dev->ro1, &dev->ro2, &dev->u1x, &dev->u1y,
                          &dev->u1z, &dev->u2x, &dev->u2y, &dev->u2z, &dev->Xi1,
                          &dev->Xi2, &dev->roS1, &dev->roS1_old, &dev->roS2, &dev->roS2_old};
    for (int a = 0; a < 20; a++)
        *slots[a] = (double*)(base + a * ad);
    dev->media = (int*)(base + 20 * ad);
    hipError_t e = hipMemsetAsync(base, 0, 20 * ad + ai, stream());
    if (e == hipSuccess)
        e = hipStreamSynchronize(stream());
    if (e != hipSuccess) {
        (void)hipFree(base);
        memset(dev, 0, sizeof *dev);
        return hip_fail(e, "clear grid");
    }
    if (dev_buffer) {
        *dev_buffer = nullptr;
        e = hipMalloc((void**)dev_buffer, sizeof(double) * 2 * (size_t)p->Ny * p->Nz);
        if (e != hipSuccess) {
            (void)hipFree(base);
            memset(dev, 0, sizeof *dev);
            return hip_fail(e, "hipMalloc exchange buffer");
        }
    }
    return FILT_OK;
}

int filt_device_memory_free(const filt_fields* dev, double* dev_buffer)
{
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    if (dev && dev->x)
        FILT_HIP_TRY(hipFree(dev->x));
    if (dev_buffer)
        FILT_HIP_TRY(hipFree(dev_buffer));
    return FILT_OK;
}

static int copy_cells(void* dst, const void* src, int localNx, const filt_params* p, size_t elem,
                      hipMemcpyKind kind)
{
    int rc = check_shape(p, localNx);
    if (rc)
        return rc;
    if (!dst || !src)
        return set_error(FILT_ERR_INVALID, "null array");
    const size_t n = (size_t)localNx * p->Ny * p->Nz * elem;
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    FILT_HIP_TRY(hipMemcpyAsync(dst, src, n, kind, stream()));
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    return FILT_OK;
}

int filt_load_data_to_host(double* host, const double* dev, int localNx, const filt_params* p)
{
    return copy_cells(host, dev, localNx, p, sizeof(double), hipMemcpyDeviceToHost);
}

int filt_load_data_to_device(const double* host, double* dev, int localNx, const filt_params* p)
{
    return copy_cells(dev, host, localNx, p, sizeof(double), hipMemcpyHostToDevice);
}

int filt_load_data_to_device_int(const int* host, int* dev, int localNx, const filt_params* p)
{
    return copy_cells(dev, host, localNx, p, sizeof(int), hipMemcpyHostToDevice);
}

}  // extern "C"

// ------------------------------------------------------------ copy kernel --
// 16-byte vector accesses; each thread issues its four loads before its four
// stores (a block moves 16 KiB); one pass over src and dst.  Loads and stores
// carry the non-temporal hint: nothing is read twice, and with it the copy
// moves 6.2-6.3 TB/s on MI355X against 5.6-5.7 TB/s without (1 GiB each way).
typedef unsigned v4u __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(const v4u* __restrict__ src,
                                                v4u* __restrict__ dst, size_t n16)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    v4u v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < n16)
            v[u] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < n16)
            __builtin_nontemporal_store(v[u], dst + i);
    }
}

extern "C" int filt_device_copy(void* dst, const void* src, size_t n_bytes)
{
    if ((!dst || !src) && n_bytes)
        return set_error(FILT_ERR_INVALID, "null pointer");
    if (n_bytes % 16)
        return set_error(FILT_ERR_INVALID, "n_bytes must be a multiple of 16");
    const size_t n16 = n_bytes / 16;
    if (!n16)
        return FILT_OK;
    const size_t blocks = (n16 + 1023) / 1024;
    if (blocks > 0x7fffffffULL)
        return set_error(FILT_ERR_INVALID, "copy too large");
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, stream(),
                       (const v4u*)src, (v4u*)dst, n16);
    FILT_LAUNCH_CHECK("k_copy16");
    return FILT_OK;
}

// ----------------------------------------------------------- halo planes --
// gpu.cu:613-653: x-planes i = 1 and i = localNx-2 packed as buffer[j + Ny*k]
// and buffer[Ny*Nz + j + Ny*k]; unpacked into i = 0 and i = localNx-1.

__global__ void k_pack_x_planes(const double* __restrict__ a, double* __restrict__ buf, int nx,
                                int ny, int nz)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= ny || k >= nz)
        return;
    const long long plane = (long long)ny * nz;
    buf[j + (long long)ny * k] = a[filt::idx3(1, j, k, nx, ny)];
    buf[plane + j + (long long)ny * k] = a[filt::idx3(nx - 2, j, k, nx, ny)];
}

__global__ void k_unpack_x_planes(double* __restrict__ a, const double* __restrict__ buf, int nx,
                                  int ny, int nz, int left, int right)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= ny || k >= nz)
        return;
    const long long plane = (long long)ny * nz;
    if (left)
        a[filt::idx3(0, j, k, nx, ny)] = buf[j + (long long)ny * k];
    if (right)
        a[filt::idx3(nx - 1, j, k, nx, ny)] = buf[plane + j + (long long)ny * k];
}

extern "C" {

int filt_load_exchange_data(const double* dev_array, double* dev_buffer, double* host_buffer,
                            int localNx, const filt_params* p)
{
    int rc = check_shape(p, localNx);
    if (rc)
        return rc;
    if (!dev_array || !dev_buffer || !host_buffer || localNx < 3)
        return set_error(FILT_ERR_INVALID, "exchange needs arrays and localNx >= 3");
    if (p->Nz > 65535)
        return set_error(FILT_ERR_INVALID, "Nz too large for the halo kernel grid");
    dim3 grid((p->Ny + 63) / 64, p->Nz);
    hipLaunchKernelGGL(k_pack_x_planes, grid, dim3(64), 0, stream(), dev_array, dev_buffer,
                       localNx, p->Ny, p->Nz);
    FILT_LAUNCH_CHECK("k_pack_x_planes");
    const size_t n = sizeof(double) * 2 * (size_t)p->Ny * p->Nz;
    FILT_HIP_TRY(hipMemcpyAsync(host_buffer, dev_buffer, n, hipMemcpyDeviceToHost, stream()));
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    return FILT_OK;
}

int filt_save_exchange_data(double* dev_array, double* dev_buffer, const double* host_buffer,
                            int localNx, int rank, int size, const filt_params* p)
{
    int rc = check_shape(p, localNx);
    if (rc)
        return rc;
    if (!dev_array || !dev_buffer || !host_buffer || localNx < 2)
        return set_error(FILT_ERR_INVALID, "exchange needs arrays and localNx >= 2");
    if (p->Nz > 65535)
        return set_error(FILT_ERR_INVALID, "Nz too large for the halo kernel grid");
    const size_t n = sizeof(double) * 2 * (size_t)p->Ny * p->Nz;
    FILT_HIP_TRY(hipMemcpyAsync(dev_buffer, host_buffer, n, hipMemcpyHostToDevice, stream()));
    dim3 grid((p->Ny + 63) / 64, p->Nz);
    hipLaunchKernelGGL(k_unpack_x_planes, grid, dim3(64), 0, stream(), dev_array, dev_buffer,
                       localNx, p->Ny, p->Nz, rank != 0, rank != size - 1);
    FILT_LAUNCH_CHECK("k_unpack_x_planes");
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    return FILT_OK;
}

}  // extern "C"

// stream_mix.hip -- HBM bandwidth of pure streaming kernels with the same
// array mix as each plain-path kernel.
//
// A kernel that reads I FP64 arrays and writes O FP64 arrays, one cell per
// thread, with no arithmetic beyond a sum, is the fastest any kernel with that
// access mix can be: it is the roofline of the mix on this GPU.  Mixes with
// many output streams move fewer bytes per second than a copy, so each
// plain-path kernel is compared with the stream of its own mix:
//
//   copy          1 in,  1 out   (2^27 cells: 1 GiB each way)
//   k_ro_P2_Xi    3 in,  5 out   S2, P1, perm -> P2, Xi1, Xi2, ro1, ro2
//   k_u_zm        6 in,  6 out   Xi1, Xi2, ro1, ro2, P1, P2 -> six u
//   k_roS_zm      8 in,  4 out   S2, P1, P2, Xi1, Xi2, ro1, ro2, poro -> four roS
//   k_roS_divgrad_zm 10 in, 4 out  ro1, ro2, S2, poro, u1x, u1y, u2x, u2y,
//                                roS1_old, roS2_old -> roS1_old, roS2_old, roS1, roS2
//   k_newton      4 in,  2 out   S2, P1, roS1, roS2 -> P1, S2
//   k_step_zm     5 in, 12 out   P2, Xi1, Xi2, S2, poro -> six u, four roS, P1, S2
//   k_step_full   4 in, 17 out   P1, S2, perm, poro -> P2, Xi1, Xi2, ro1, ro2,
//                                six u, four roS, P1, S2
//
// Every mix runs with plain and with non-temporal stores.  Prints one JSON
// line per mix and store kind.  Usage: stream_mix [cells [mix]] (default
// 256^3 cells, every mix; a mix name runs that mix alone).
//
// Tiled mode: stream_mix tiled NX NY NZ TX TY KZ [mix] streams the same
// mixes with the z-marching kernels' access pattern instead of a flat one:
// blocks of TX x TY columns (tiles from x = 0) each march over KZ planes of
// an NX x NY x NZ grid, one cell per thread per plane, non-temporal stores.
// No halo and no arithmetic: the fastest a z-march over those arrays with
// that tiling can run.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            exit(1);                                                                   \
        }                                                                              \
    } while (0)

namespace {

constexpr int kMaxIn = 10, kMaxOut = 17;

struct Streams {
    const double* in[kMaxIn];
    double* out[kMaxOut];
};

template <int I, int O, bool NT>
__global__ __launch_bounds__(256) void k_mix(Streams s, size_t n)
{
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n)
        return;
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < I; a++)
        v += s.in[a][c];
#pragma unroll
    for (int b = 0; b < O; b++) {
        if (NT)
            __builtin_nontemporal_store(v + b, s.out[b] + c);
        else
            s.out[b][c] = v + b;
    }
}

template <int I, int O>
__global__ __launch_bounds__(1024) void k_tiled(Streams s, int nx, int ny, int nz, int tx, int ty,
                                                int kz, int nbx, int nby)
{
    const int b = blockIdx.x;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int i = bx * tx + (int)threadIdx.x % tx, j = by * ty + (int)threadIdx.x / tx;
    if (i >= nx || j >= ny)
        return;
    const int k0 = bz * kz, k1 = min(k0 + kz, nz);
    const size_t sz = (size_t)nx * ny;
    size_t c = (size_t)i + (size_t)j * nx + (size_t)k0 * sz;
    for (int k = k0; k < k1; k++, c += sz) {
        double v = 0.0;
#pragma unroll
        for (int a = 0; a < I; a++)
            v += s.in[a][c];
#pragma unroll
        for (int o = 0; o < O; o++)
            __builtin_nontemporal_store(v + o, s.out[o] + c);
    }
}

struct Tiling {
    int nx, ny, nz, tx, ty, kz;
};

template <int I, int O>
void run_tiled(const char* name, const Streams& s, const Tiling& t)
{
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    const int nbx = (t.nx + t.tx - 1) / t.tx, nby = (t.ny + t.ty - 1) / t.ty;
    const int nbz = (t.nz + t.kz - 1) / t.kz;
    const dim3 grid((unsigned)(nbx * nby * nbz)), block((unsigned)(t.tx * t.ty));
    const int reps = 20;
    for (int w = 0; w < 3; w++)
        hipLaunchKernelGGL((k_tiled<I, O>), grid, block, 0, 0, s, t.nx, t.ny, t.nz, t.tx, t.ty,
                           t.kz, nbx, nby);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r3 = 0; r3 < 3; r3++) {
        CHECK(hipEventRecord(a));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL((k_tiled<I, O>), grid, block, 0, 0, s, t.nx, t.ny, t.nz, t.tx,
                               t.ty, t.kz, nbx, nby);
        CHECK(hipEventRecord(b));
        CHECK(hipEventSynchronize(b));
        float ms;
        CHECK(hipEventElapsedTime(&ms, a, b));
        best = ms / reps < best ? ms / reps : best;
    }
    CHECK(hipGetLastError());
    const double n = (double)t.nx * t.ny * t.nz;
    const double bytes = 8.0 * (I + O) * n;
    printf("{\"mix\": \"%s\", \"in\": %d, \"out\": %d, \"nontemporal_stores\": 1, "
           "\"tiled\": [%d, %d, %d], \"grid\": [%d, %d, %d], \"cells\": %.0f, "
           "\"bytes_per_cell\": %d, \"ms\": %.4f, \"TBps\": %.3f}\n",
           name, I, O, t.tx, t.ty, t.kz, t.nx, t.ny, t.nz, n, 8 * (I + O), best,
           bytes / (best * 1e-3) / 1e12);
    fflush(stdout);
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
}

__global__ void k_fill(double* a, size_t n, double base)
{
    const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (c < n)
        a[c] = base + (double)(c & 1023);
}

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

template <int I, int O, bool NT>
void run(const char* name, const Streams& s, size_t n)
{
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    const int reps = 20;
    for (int w = 0; w < 3; w++)
        hipLaunchKernelGGL((k_mix<I, O, NT>), dim3(blocks_for(n)), dim3(256), 0, 0, s, n);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int t = 0; t < 3; t++) {
        CHECK(hipEventRecord(a));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL((k_mix<I, O, NT>), dim3(blocks_for(n)), dim3(256), 0, 0, s, n);
        CHECK(hipEventRecord(b));
        CHECK(hipEventSynchronize(b));
        float ms;
        CHECK(hipEventElapsedTime(&ms, a, b));
        best = ms / reps < best ? ms / reps : best;
    }
    CHECK(hipGetLastError());
    const double bytes = 8.0 * (I + O) * (double)n;
    printf("{\"mix\": \"%s\", \"in\": %d, \"out\": %d, \"nontemporal_stores\": %d, "
           "\"cells\": %zu, \"bytes_per_cell\": %d, \"ms\": %.4f, \"TBps\": %.3f}\n",
           name, I, O, NT ? 1 : 0, n, 8 * (I + O), best, bytes / (best * 1e-3) / 1e12);
    fflush(stdout);
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
}

template <int I, int O>
void both(const char* name, const Streams& s, size_t n)
{
    run<I, O, false>(name, s, n);
    run<I, O, true>(name, s, n);
}

}  // namespace

int main(int argc, char** argv)
{
    const bool tiled = argc > 1 && strcmp(argv[1], "tiled") == 0;
    Tiling tl = {0, 0, 0, 0, 0, 0};
    if (tiled) {
        if (argc < 8) {
            fprintf(stderr, "usage: stream_mix tiled NX NY NZ TX TY KZ [mix]\n");
            return 2;
        }
        tl = {atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]),
              atoi(argv[7])};
        const double cells = (double)tl.nx * tl.ny * tl.nz;
        if (tl.nx < 1 || tl.ny < 1 || tl.nz < 1 || tl.kz < 1 || tl.tx < 1 || tl.ty < 1 ||
            tl.tx * tl.ty > 1024 || (tl.tx * tl.ty) % 64 != 0 || cells > (double)(1u << 31)) {
            fprintf(stderr, "tiled: extents >= 1, tile of 64..1024 threads, cells <= 2^31\n");
            return 2;
        }
    }
    const size_t n = tiled ? (size_t)tl.nx * tl.ny * tl.nz
                           : argc > 1 ? (size_t)atoll(argv[1]) : (size_t)256 * 256 * 256;
    const char* only = tiled ? (argc > 8 ? argv[8] : nullptr) : argc > 2 ? argv[2] : nullptr;
    auto want = [only](const char* name) { return !only || strcmp(only, name) == 0; };
    if (n == 0 || n > ((size_t)1 << 31)) {
        fprintf(stderr, "cells must be in [1, 2^31]\n");
        return 2;
    }
    Streams s;
    for (int a = 0; a < kMaxIn; a++) {
        double* p;
        CHECK(hipMalloc(&p, n * sizeof(double)));
        hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(256), 0, 0, p, n, 1.0 + a);
        s.in[a] = p;
    }
    for (int b = 0; b < kMaxOut; b++)
        CHECK(hipMalloc(&s.out[b], n * sizeof(double)));
    CHECK(hipDeviceSynchronize());

    if (tiled) {
        if (want("k_ro_P2_Xi")) run_tiled<3, 5>("k_ro_P2_Xi", s, tl);
        if (want("k_u_zm")) run_tiled<6, 6>("k_u_zm", s, tl);
        if (want("k_roS_zm")) run_tiled<8, 4>("k_roS_zm", s, tl);
        if (want("k_roS_divgrad_zm")) run_tiled<10, 4>("k_roS_divgrad_zm", s, tl);
        if (want("k_newton")) run_tiled<4, 2>("k_newton", s, tl);
        if (want("k_step_zm")) run_tiled<5, 12>("k_step_zm", s, tl);
        if (want("k_step_full")) run_tiled<4, 17>("k_step_full", s, tl);
        if (want("k_fused")) run_tiled<4, 2>("k_fused", s, tl);
    } else {
        if (want("copy")) {  // 1 GiB each way, past the 256 MB last-level cache
            const size_t nc = (size_t)1 << 27;
            Streams c = s;
            double *src, *dst;
            CHECK(hipMalloc(&src, nc * sizeof(double)));
            CHECK(hipMalloc(&dst, nc * sizeof(double)));
            hipLaunchKernelGGL(k_fill, dim3(blocks_for(nc)), dim3(256), 0, 0, src, nc, 1.0);
            c.in[0] = src;
            c.out[0] = dst;
            both<1, 1>("copy", c, nc);
            CHECK(hipFree(src));
            CHECK(hipFree(dst));
        }
        if (want("k_ro_P2_Xi")) both<3, 5>("k_ro_P2_Xi", s, n);
        if (want("k_u_zm")) both<6, 6>("k_u_zm", s, n);
        if (want("k_roS_zm")) both<8, 4>("k_roS_zm", s, n);
        if (want("k_roS_divgrad_zm")) both<10, 4>("k_roS_divgrad_zm", s, n);
        if (want("k_newton")) both<4, 2>("k_newton", s, n);
        if (want("k_step_zm")) both<5, 12>("k_step_zm", s, n);
        if (want("k_step_full")) both<4, 17>("k_step_full", s, n);
    }

    for (int a = 0; a < kMaxIn; a++)
        CHECK(hipFree((void*)s.in[a]));
    for (int b = 0; b < kMaxOut; b++)
        CHECK(hipFree(s.out[b]));
    return 0;
}

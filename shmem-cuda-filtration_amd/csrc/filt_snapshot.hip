// filt_snapshot.hip -- device-resident snapshots of a grid's state and the
// asynchronous file writers behind them.
//
// filt_snapshot_take copies, in stream order after the work already
// enqueued, the arrays an output needs into the snapshot's own HBM buffer and
// returns: the library stream can go on stepping at once, since later steps
// overwrite only the grid's arrays.  filt_snapshot_write_async hands the
// snapshot to a host thread, which waits for that copy, moves the buffer to
// pinned host memory on a second stream (default priority) and writes the
// save block (main.cpp:485-537) and / or the Tecplot plots (main.cpp:319-482)
// with the writers of filt_io.cpp.  filt_snapshot_restore copies a snapshot
// back into a grid in stream order: a rollback or a preempted grid resumed
// without touching a file.
#include <hip/hip_runtime.h>

#include <string>
#include <thread>

#include "../../include/filtration_hip.h"
#include "../host/filt_io.hpp"
#include "filt_runtime.hpp"
#include "../host/filt_guard.hpp"

namespace {

// Snapshot slots, in buffer order.
enum Slot { P1, S2, R1, R2, UX, UY, UZ, NSLOT };

// Zero fill of a media slot (a grid without media is one medium, index 0);
// a kernel rather than a memset node, so that a captured take replays right.
__global__ void k_zero_int(int* a, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        a[i] = 0;
}

}  // namespace

struct filt_snapshot {
    filt_params p;
    int nx = 0, rank = 0, size = 1, what = 0, device = 0;
    size_t n = 0;
    bool has[NSLOT] = {};
    char* dbuf = nullptr;      // device: NSLOT FP64 slabs then media (int32)
    char* hbuf = nullptr;      // pinned host mirror
    size_t off[NSLOT + 1] = {};  // byte offsets; off[NSLOT] = media
    size_t bytes = 0;
    hipEvent_t taken = nullptr;
    hipStream_t io = nullptr;
    std::thread writer;
    bool busy = false;
    bool valid = false;        // a take has happened
    int status = FILT_OK;
    std::string err;           // the writer's error text
    int j = 0;
    double t = 0.0;
};

namespace {

double* dev_slot(const filt_fields* f, int s)
{
    switch (s) {
    case P1: return f->P1;
    case S2: return f->S2;
    case R1: return f->roS1_old;
    case R2: return f->roS2_old;
    case UX: return f->u2x;
    case UY: return f->u2y;
    case UZ: return f->u2z;
    }
    return nullptr;
}

int join(filt_snapshot* s)
{
    if (s->busy) {
        s->writer.join();
        s->busy = false;
    }
    return s->status;
}

// The previous write's failure, reported once: by whichever of take,
// write_async, wait or destroy comes first after it.
int report_previous(filt_snapshot* s, const char* what)
{
    const int rc = join(s);
    if (rc == FILT_OK)
        return FILT_OK;
    s->status = FILT_OK;
    return filt::set_error(rc, "%s: the snapshot's previous write failed: %s", what,
                           s->err.c_str());
}

// Writer thread body: pinned copy of the snapshot, then the files.
int write_files(filt_snapshot* s, const std::string& save_path, bool truncate,
                const std::string& plot_dir)
{
    if (hipSetDevice(s->device) != hipSuccess)
        return filt::set_error(FILT_ERR_HIP, "hipSetDevice in the snapshot writer");
    if (hipStreamWaitEvent(s->io, s->taken, 0) != hipSuccess ||
        hipMemcpyAsync(s->hbuf, s->dbuf, s->bytes, hipMemcpyDeviceToHost, s->io) != hipSuccess ||
        hipStreamSynchronize(s->io) != hipSuccess)
        return filt::set_error(FILT_ERR_HIP, "snapshot copy to host");
    const double* h[NSLOT];
    for (int q = 0; q < NSLOT; q++)
        h[q] = s->has[q] ? (const double*)(s->hbuf + s->off[q]) : nullptr;
    const int* media = (const int*)(s->hbuf + s->off[NSLOT]);
    filt_fields hf = {};
    hf.P1 = (double*)h[P1];
    hf.S2 = (double*)h[S2];
    hf.roS1_old = (double*)h[R1];
    hf.roS2_old = (double*)h[R2];
    hf.u2x = (double*)h[UX];
    hf.u2y = (double*)h[UY];
    hf.u2z = (double*)h[UZ];
    hf.media = (int*)media;
    int rc = FILT_OK;
    if (!save_path.empty())
        rc = filt_save_host(save_path.c_str(), &hf, &s->p, s->nx, s->rank, s->size, s->j,
                            truncate ? 1 : 0);
    if (!rc && !plot_dir.empty())
        rc = filt_save_plots_host(plot_dir.c_str(), &hf, &s->p, s->nx, s->rank, s->size, s->t);
    return rc;
}

}  // namespace

extern "C" {

int filt_snapshot_create(filt_snapshot** out, const filt_params* p, int localNx, int rank,
                         int size, int what)
{
    return filt::guarded([&]() -> int {
        using namespace filt;
        if (!out)
            return set_error(FILT_ERR_INVALID, "null handle");
        *out = nullptr;
        int rc = check_shape(p, localNx);
        if (rc)
            return rc;
        if (!(what & (FILT_SNAP_SAVE | FILT_SNAP_PLOTS)) || (what & ~(FILT_SNAP_SAVE | FILT_SNAP_PLOTS)))
            return set_error(FILT_ERR_INVALID, "what must combine FILT_SNAP_SAVE and FILT_SNAP_PLOTS");
        if (size < 1 || rank < 0 || rank >= size)
            return set_error(FILT_ERR_INVALID, "bad rank / size");
        filt_snapshot* s = new filt_snapshot();
        s->p = *p;
        s->nx = localNx, s->rank = rank, s->size = size, s->what = what;
        s->n = (size_t)localNx * p->Ny * p->Nz;
        s->has[P1] = s->has[S2] = true;
        s->has[R1] = s->has[R2] = (what & FILT_SNAP_SAVE) != 0;
        s->has[UX] = s->has[UY] = s->has[UZ] = (what & FILT_SNAP_PLOTS) != 0;
        size_t o = 0;
        for (int q = 0; q < NSLOT; q++) {
            s->off[q] = o;
            if (s->has[q])
                o += (s->n * sizeof(double) + 255) & ~(size_t)255;
        }
        s->off[NSLOT] = o;
        s->bytes = o + s->n * sizeof(int);
        hipError_t e = hipGetDevice(&s->device);
        if (e == hipSuccess)
            e = hipMalloc((void**)&s->dbuf, s->bytes);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&s->hbuf, s->bytes, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&s->taken, hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipStreamCreate(&s->io);  // default priority
        if (e != hipSuccess) {
            filt_snapshot_destroy(s);
            return e == hipErrorOutOfMemory ? set_error(FILT_ERR_NOMEM, "snapshot buffers")
                                            : hip_fail(e, "filt_snapshot_create");
        }
        *out = s;
        return FILT_OK;
    });
}

int filt_snapshot_destroy(filt_snapshot* s)
{
    if (!s)
        return FILT_OK;
    const int rc = report_previous(s, "filt_snapshot_destroy");
    if (s->dbuf) {
        (void)hipStreamSynchronize(filt::stream());  // a take may still be copying
        (void)hipFree(s->dbuf);
    }
    if (s->hbuf)
        (void)hipHostFree(s->hbuf);
    if (s->taken)
        (void)hipEventDestroy(s->taken);
    if (s->io)
        (void)hipStreamDestroy(s->io);
    delete s;
    return rc;
}

int filt_snapshot_take(filt_snapshot* s, const filt_fields* dev, int j, double t)
{
    using namespace filt;
    if (!s || !dev)
        return set_error(FILT_ERR_INVALID, "null argument");
    // the previous write reads the buffer until it ends; its failure is
    // returned here rather than lost (the take is not made)
    int prev = report_previous(s, "filt_snapshot_take");
    if (prev)
        return prev;
    for (int q = 0; q < NSLOT; q++) {
        if (!s->has[q])
            continue;
        const double* src = dev_slot(dev, q);
        if (!src)
            return set_error(FILT_ERR_INVALID, "the grid lacks an array this snapshot holds");
        FILT_HIP_TRY(hipMemcpyAsync(s->dbuf + s->off[q], src, s->n * sizeof(double),
                                    hipMemcpyDeviceToDevice, stream()));
    }
    if (dev->media)
        FILT_HIP_TRY(hipMemcpyAsync(s->dbuf + s->off[NSLOT], dev->media, s->n * sizeof(int),
                                    hipMemcpyDeviceToDevice, stream()));
    else {
        hipLaunchKernelGGL(k_zero_int, dim3((unsigned)((s->n + 255) / 256)), dim3(256), 0,
                           stream(), (int*)(s->dbuf + s->off[NSLOT]), s->n);
        FILT_LAUNCH_CHECK("k_zero_int");
    }
    FILT_HIP_TRY(hipEventRecord(s->taken, stream()));
    s->j = j;
    s->t = t;
    s->valid = true;
    s->status = FILT_OK;
    return FILT_OK;
}

int filt_snapshot_write_async(filt_snapshot* s, const char* save_path, int truncate,
                              const char* plot_dir)
{
    return filt::guarded([&]() -> int {
        using namespace filt;
        if (!s)
            return set_error(FILT_ERR_INVALID, "null snapshot");
        if (!s->valid)
            return set_error(FILT_ERR_INVALID, "nothing taken yet");
        if (save_path && !(s->what & FILT_SNAP_SAVE))
            return set_error(FILT_ERR_INVALID, "snapshot made without FILT_SNAP_SAVE");
        if (plot_dir && !(s->what & FILT_SNAP_PLOTS))
            return set_error(FILT_ERR_INVALID, "snapshot made without FILT_SNAP_PLOTS");
        int prev = report_previous(s, "filt_snapshot_write_async");
        if (prev)
            return prev;
        const std::string sp = save_path ? save_path : "", pd = plot_dir ? plot_dir : "";
        const bool tr = truncate != 0;
        s->status = FILT_OK;
        s->busy = true;
        s->writer = std::thread([s, sp, tr, pd] {
            s->status = filt::guarded([&] { return write_files(s, sp, tr, pd); });
            if (s->status)
                s->err = filt_last_error();
        });
        return FILT_OK;
    });
}

int filt_snapshot_wait(filt_snapshot* s)
{
    if (!s)
        return filt::set_error(FILT_ERR_INVALID, "null snapshot");
    return report_previous(s, "filt_snapshot_wait");
}

int filt_snapshot_restore(filt_snapshot* s, const filt_fields* dev, int* j)
{
    using namespace filt;
    if (!s || !dev)
        return set_error(FILT_ERR_INVALID, "null argument");
    if (!s->valid)
        return set_error(FILT_ERR_INVALID, "nothing taken yet");
    for (int q = 0; q < NSLOT; q++) {
        double* dst = dev_slot(dev, q);
        if (s->has[q] && dst)
            FILT_HIP_TRY(hipMemcpyAsync(dst, s->dbuf + s->off[q], s->n * sizeof(double),
                                        hipMemcpyDeviceToDevice, stream()));
    }
    if (dev->media)
        FILT_HIP_TRY(hipMemcpyAsync(dev->media, s->dbuf + s->off[NSLOT], s->n * sizeof(int),
                                    hipMemcpyDeviceToDevice, stream()));
    if (j)
        *j = s->j;
    return FILT_OK;
}

}  // extern "C"

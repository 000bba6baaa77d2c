// filt_io.cpp -- the reference driver's file output, byte for byte:
// checkpoint / resume in the save/save.dat layout (save and restore,
// main.cpp:485-569) and the Tecplot ASCII plots of save_data_plots
// (main.cpp:319-482).
//
// The *_host entry points take the host arrays the reference's driver holds.
// The device entry points copy what the reference copies to the host
// (load_data_to_host of P1, S2, u2*, roS*_old; media too) and take x, y, z
// from the formulas of initial_data (x = I*h1, y = j*h2, z = k*h3), which are
// the values the reference's host arrays hold.  Plot text is formatted by
// several threads, each into its own buffer over a range of i, and written
// in the reference's line order; save blocks are written with one fwrite per
// array as the reference does.
#include <errno.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "../../include/filtration_hip.h"
#include "filt_io.hpp"
#include "filt_guard.hpp"

namespace filt {
int set_error(int code, const char* fmt, ...);
}

namespace filt {
namespace io {

namespace {

// mkdir -p
bool make_dir(const std::string& d)
{
    for (size_t pos = 1; pos <= d.size(); pos++) {
        if (pos < d.size() && d[pos] != '/')
            continue;
        const std::string part = d.substr(0, pos);
        if (mkdir(part.c_str(), 0777) != 0 && errno != EEXIST)
            return false;
    }
    return true;
}

std::string parent_of(const char* path)
{
    const char* s = strrchr(path, '/');
    return s ? std::string(path, (size_t)(s - path)) : std::string();
}

size_t cells(const filt_params* p, int localNx) { return (size_t)localNx * p->Ny * p->Nz; }

// x, y, z of initial_data (main.cpp:113-116) for every cell of the slab.
void coordinates(const filt_params* p, int localNx, int rank, int size, std::vector<double>& x,
                 std::vector<double>& y, std::vector<double>& z)
{
    const size_t n = cells(p, localNx);
    x.resize(n);
    y.resize(n);
    z.resize(n);
    for (int k = 0; k < p->Nz; k++)
        for (int j = 0; j < p->Ny; j++)
            for (int i = 0; i < localNx; i++) {
                const size_t c = i + (size_t)localNx * (j + (size_t)p->Ny * k);
                x[c] = filt_i_to_I(i, rank, size, p) * p->h1;
                y[c] = j * p->h2;
                z[c] = k * p->h3;
            }
}

// printf-append into a string.
void appendf(std::string& s, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void appendf(std::string& s, const char* fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    const int m = vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (m > 0)
        s.append(buf, (size_t)std::min(m, (int)sizeof buf - 1));
}

int io_threads()
{
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 32u));
}

// Run fn(lo, hi, t) over [0, n) split into contiguous ranges, one per thread.
template <class Fn>
void parallel_ranges(int n, int T, Fn fn)
{
    std::vector<std::thread> th;
    std::atomic<bool> failed{false};  // a worker's exception, rethrown below
    for (int t = 0; t < T; t++) {
        const int lo = (int)((long long)n * t / T), hi = (int)((long long)n * (t + 1) / T);
        th.emplace_back([=, &failed] {
            try {
                fn(lo, hi, t);
            } catch (...) {
                failed = true;
            }
        });
    }
    for (auto& x : th)
        x.join();
    if (failed)
        throw std::bad_alloc();
}

struct PlotFiles {
    FILE *S2 = nullptr, *P1 = nullptr, *u = nullptr, *S2y = nullptr, *S2x = nullptr,
         *media = nullptr;
    bool open(const std::string& dir, double t, const char* mode)
    {
        char name[4][64];
        snprintf(name[0], sizeof name[0], "plot_S2/S2=%012.4f.dat", t);
        snprintf(name[1], sizeof name[1], "plot_P1/P1=%012.4f.dat", t);
        snprintf(name[2], sizeof name[2], "plot_u/u=%012.4f.dat", t);
        snprintf(name[3], sizeof name[3], "plot_S2_y/S2=%012.4f.dat", t);
        char name_x[64];
        snprintf(name_x, sizeof name_x, "plot_S2_x/S2=%012.4f.dat", t);
        S2 = fopen((dir + "/" + name[0]).c_str(), mode);
        P1 = fopen((dir + "/" + name[1]).c_str(), mode);
        u = fopen((dir + "/" + name[2]).c_str(), mode);
        S2y = fopen((dir + "/" + name[3]).c_str(), mode);
        S2x = fopen((dir + "/" + name_x).c_str(), mode);
        media = fopen((dir + "/media.dat").c_str(), mode);
        return S2 && P1 && u && S2y && S2x && media;
    }
    void close()
    {
        for (FILE* f : {S2, P1, u, S2y, S2x, media})
            if (f)
                fclose(f);
        S2 = P1 = u = S2y = S2x = media = nullptr;
    }
};

bool put(FILE* f, const std::string& s) { return fwrite(s.data(), 1, s.size(), f) == s.size(); }

}  // namespace

int write_save(const char* path, const View& v, size_t n, int j, bool truncate)
{
    const std::string dir = parent_of(path);
    if (!dir.empty() && !make_dir(dir))
        return set_error(FILT_ERR_IO, "cannot create directory %s", dir.c_str());
    FILE* f = fopen(path, truncate ? "wb" : "ab");
    if (!f)
        return set_error(FILT_ERR_IO, "cannot open %s for writing", path);
    bool ok = fwrite(&j, sizeof(int), 1, f) == 1;
    for (const double* a : {v.P1, v.S2, v.x, v.y, v.z, v.roS1_old, v.roS2_old})
        ok = ok && fwrite(a, sizeof(double), n, f) == n;
    ok = ok && fwrite(v.media, sizeof(int), n, f) == n;
    ok = (fclose(f) == 0) && ok;
    return ok ? FILT_OK : set_error(FILT_ERR_IO, "short write to %s", path);
}

int read_save(const char* path, const filt_params* p, int rank, int size, const Out& o, int* j)
{
    FILE* f = fopen(path, "rb");
    if (!f)
        return set_error(FILT_ERR_IO, "cannot open %s", path);
    int rc = FILT_OK;
    for (int q = 0; q <= rank && rc == FILT_OK; q++) {
        int lnx = 0;
        filt_N_to_local(&lnx, nullptr, size, q, p);
        const size_t n = cells(p, lnx);
        int jj = 0;
        if (fread(&jj, sizeof(int), 1, f) != 1) {
            rc = set_error(FILT_ERR_IO, "%s ends before the block of rank %d", path, q);
            break;
        }
        if (j)
            *j = jj;
        double* const dst[7] = {o.P1, o.S2, o.x, o.y, o.z, o.roS1_old, o.roS2_old};
        for (int a = 0; a < 8 && rc == FILT_OK; a++) {
            const size_t bytes = n * (a < 7 ? sizeof(double) : sizeof(int));
            void* to = q < rank ? nullptr : (a < 7 ? (void*)dst[a] : (void*)o.media);
            if (to) {
                if (fread(to, 1, bytes, f) != bytes)
                    rc = set_error(FILT_ERR_IO, "%s: short block of rank %d", path, q);
            } else if (fseek(f, (long)bytes, SEEK_CUR) != 0) {
                rc = set_error(FILT_ERR_IO, "%s: cannot skip the block of rank %d", path, q);
            }
        }
    }
    fclose(f);
    return rc;
}

// print_plots_top (main.cpp:350-419): create the plot folders, truncate the
// six files and write their Tecplot headers.
int write_plots_top(const std::string& dir, const filt_params* p, double t)
{
    if (!make_dir(dir))
        return set_error(FILT_ERR_IO, "cannot create %s", dir.c_str());
    for (const char* d : {"plot_P1", "plot_S2", "plot_u", "plot_S2_y", "plot_S2_x"})
        if (!make_dir(dir + "/" + d))
            return set_error(FILT_ERR_IO, "cannot create %s/%s", dir.c_str(), d);
    PlotFiles F;
    if (!F.open(dir, t, "wt")) {
        F.close();
        return set_error(FILT_ERR_IO, "cannot open the plot files under %s", dir.c_str());
    }
    const int NX = p->NX, Ny = p->Ny, Nz = p->Nz;
    fprintf(F.S2, "TITLE =  \"Saturation of DNALP in time=%5.2f\" \n", t);
    fprintf(F.S2, "VARIABLES = \"X\",\"Y\",\"Z\",\"S2\",\"P1\",\"u_x\", \"u_y\", \"u_z\" \n");
    fprintf(F.S2, "ZONE T = \"BIG ZONE\", K=%d,J=%d,I=%d, F = POINT\n", NX, Nz, Ny);
    fprintf(F.P1, "TITLE =  \"Pressure of water in time=%5.2f\" \n", t);
    fprintf(F.P1, "VARIABLES = \"X\",\"Y\",\"P1\" \n");
    fprintf(F.P1, "ZONE T = \"BIG ZONE\", J=%d,I=%d, F = POINT\n", NX, Ny);
    fprintf(F.u, "TITLE =  \"Velocity of DNALP in time=%5.2f\" \n", t);
    fprintf(F.u, "VARIABLES = \"X\",\"Y\",\"u_x\", \"u_y\" \n");
    fprintf(F.u, "ZONE T = \"BIG ZONE\", J=%d,I=%d, F = POINT\n", NX - 2, Ny - 2);
    fprintf(F.S2y, "TITLE =  \"Saturation of DNALP in time=%5.2f\" \n", t);
    fprintf(F.S2y, "VARIABLES = \"Y\",\"S2\" \n");
    fprintf(F.S2y, "ZONE T = \"BIG ZONE\", J=%d, F = POINT\n", Ny);
    fprintf(F.S2x, "TITLE =  \"Saturation of DNALP in time=%5.2f\" \n", t);
    fprintf(F.S2x, "VARIABLES = \"X\",\"S2\" \n");
    fprintf(F.S2x, "ZONE T = \"BIG ZONE\", J=%d, F = POINT\n", NX);
    fprintf(F.media, "TITLE =  \"Porous media\" \n");
    fprintf(F.media, "VARIABLES = \"X\",\"Y\",\"media\" \n");
    fprintf(F.media, "ZONE T = \"BIG ZONE\", J=%d,I=%d, F = POINT\n", NX, Ny);
    F.close();
    return FILT_OK;
}

// print_plots (main.cpp:422-482): append this rank's lines.  The reference's
// own index choices are kept: the S2-along-y profile reads column localNx/2
// of every (i, k) whose global I is NX/2, and the S2-along-x profile reads
// offset i + (localNx*Ny)/2 + k*localNx*Ny.
int write_plots_body(const std::string& dir, const View& v, const filt_params* p, int localNx,
                     int rank, int size, double t)
{
    PlotFiles F;
    if (!F.open(dir, t, "at")) {
        F.close();
        return set_error(FILT_ERR_IO, "cannot open the plot files under %s", dir.c_str());
    }
    const int Ny = p->Ny, Nz = p->Nz;
    const size_t sj = (size_t)localNx, sk = (size_t)localNx * Ny;
    const double top = Ny * p->h2;  // (def.Ny)*(def.h2)
    const int T = io_threads();
    bool ok = true;
    // cells in i-major, then j, then k order; i in rounds of a few per thread
    // so that the text of one round stays small
    const int per_round = std::max(1, T * 2);
    std::vector<std::string> bS2(T), bP1(T), bMedia(T), bU(T);
    for (int i0 = 0; i0 < localNx && ok; i0 += per_round) {
        const int i1 = std::min(localNx, i0 + per_round);
        parallel_ranges(i1 - i0, T, [&](int lo, int hi, int t_) {
            std::string &s2 = bS2[t_], &p1 = bP1[t_], &md = bMedia[t_], &u = bU[t_];
            s2.clear(), p1.clear(), md.clear(), u.clear();
            for (int i = i0 + lo; i < i0 + hi; i++) {
                if (!filt_is_active_point(i, localNx, rank, size))
                    continue;
                for (int j = 0; j < Ny; j++)
                    for (int k = 0; k < Nz; k++) {
                        const size_t c = i + j * sj + k * sk;
                        appendf(s2, "%.2e %.2e %.2e %.3e %.3e %.3e %.3e %.3e\n", v.x[c], v.z[c],
                                top - v.y[c], v.S2[c], v.P1[c], v.u2x[c], v.u2z[c],
                                (-1) * v.u2y[c]);
                        appendf(p1, "%d %d %d %.3e %.3e %.3e %.3e %.3e\n", i, j, k, v.S2[c],
                                v.P1[c], v.u2x[c], v.u2z[c], (-1) * v.u2y[c]);
                        appendf(md, "%.2e %.2e %d\n", v.x[c], top - v.y[c], v.media[c]);
                        appendf(u, "%.2e %.2e %.3e %.3e\n", v.x[c], top - v.y[c], v.u2x[c],
                                (-1) * v.u2y[c]);
                    }
            }
        });
        for (int t_ = 0; t_ < T; t_++)
            ok = ok && put(F.S2, bS2[t_]) && put(F.P1, bP1[t_]) && put(F.media, bMedia[t_]) &&
                 put(F.u, bU[t_]);
    }
    // second pass of the velocity file over the interior (main.cpp:460-464)
    for (int i0 = 1; i0 < localNx - 1 && ok; i0 += per_round) {
        const int i1 = std::min(localNx - 1, i0 + per_round);
        parallel_ranges(i1 - i0, T, [&](int lo, int hi, int t_) {
            std::string& u = bU[t_];
            u.clear();
            for (int i = i0 + lo; i < i0 + hi; i++)
                for (int j = 1; j < Ny - 1; j++)
                    for (int k = 1; k < Nz - 1; k++) {
                        if ((rank != 0 && i == 0) || (rank != size - 1 && i == localNx - 1))
                            continue;
                        const size_t c = i + j * sj + k * sk;
                        appendf(u, "%.2e %.2e %.3e %.3e\n", v.x[c], top - v.y[c], v.u2x[c],
                                (-1) * v.u2y[c]);
                    }
        });
        for (int t_ = 0; t_ < T; t_++)
            ok = ok && put(F.u, bU[t_]);
    }
    std::string line;
    for (int i = 0; i < localNx && ok; i++)
        for (int k = 0; k < Nz; k++)
            if (filt_i_to_I(i, rank, size, p) == p->NX / 2 &&
                filt_is_active_point(i, localNx, rank, size)) {
                line.clear();
                for (int j = 0; j < Ny; j++) {
                    const size_t c = localNx / 2 + j * sj + k * sk;
                    appendf(line, "%.2e %.3e\n", v.y[c], v.S2[c]);
                }
                ok = ok && put(F.S2y, line);
            }
    const size_t mid = ((size_t)localNx * Ny) / 2;
    for (int i = 0; i < localNx && ok; i++) {
        line.clear();
        for (int k = 0; k < Nz; k++) {
            const size_t c = i + mid + k * sk;
            appendf(line, "%.2e %.3e\n", v.x[c], v.S2[c]);
        }
        ok = ok && put(F.S2x, line);
    }
    F.close();
    return ok ? FILT_OK : set_error(FILT_ERR_IO, "short write under %s", dir.c_str());
}

int write_plots(const char* dir, const View& v, const filt_params* p, int localNx, int rank,
                int size, double t)
{
    const std::string d = (dir && *dir) ? std::string(dir) : std::string(".");
    if (rank == 0) {
        const int rc = write_plots_top(d, p, t);
        if (rc)
            return rc;
    }
    return write_plots_body(d, v, p, localNx, rank, size, t);
}

// Host copies of the device arrays a writer needs (filt_memcpy is
// synchronous on the library stream).
int fetch(const double* dev, std::vector<double>& h, size_t n, const char* name)
{
    h.resize(n);
    if (!dev)
        return set_error(FILT_ERR_INVALID, "device array %s is null", name);
    return filt_memcpy(h.data(), dev, n * sizeof(double), 1);
}

int fetch_int(const int* dev, std::vector<int>& h, size_t n)
{
    h.assign(n, 0);
    if (!dev)
        return FILT_OK;  // no media array: one medium, index 0
    return filt_memcpy(h.data(), dev, n * sizeof(int), 1);
}

int check_args(const void* path, const void* fields, const filt_params* p, int localNx,
               int rank, int size)
{
    if (!path || !fields || !p)
        return set_error(FILT_ERR_INVALID, "null argument");
    if (localNx < 1 || p->Ny < 1 || p->Nz < 1 || size < 1 || rank < 0 || rank >= size)
        return set_error(FILT_ERR_INVALID, "bad shape or rank");
    return FILT_OK;
}

}  // namespace io
}  // namespace filt

using namespace filt::io;

extern "C" {

int filt_save_host(const char* path, const filt_fields* h, const filt_params* p, int localNx,
                   int rank, int size, int j, int truncate)
{
    return filt::guarded([&]() -> int {
        int rc = check_args(path, h, p, localNx, rank, size);
        if (rc)
            return rc;
        if (!h->P1 || !h->S2 || !h->roS1_old || !h->roS2_old || !h->media)
            return filt::set_error(FILT_ERR_INVALID, "P1, S2, roS1_old, roS2_old and media are needed");
        const size_t n = (size_t)localNx * p->Ny * p->Nz;
        std::vector<double> x, y, z;
        if (!h->x || !h->y || !h->z)
            coordinates(p, localNx, rank, size, x, y, z);
        View v{};
        v.P1 = h->P1, v.S2 = h->S2, v.roS1_old = h->roS1_old, v.roS2_old = h->roS2_old;
        v.x = h->x ? h->x : x.data();
        v.y = h->y ? h->y : y.data();
        v.z = h->z ? h->z : z.data();
        v.media = h->media;
        return write_save(path, v, n, j, truncate != 0);
    });
}

int filt_restore_host(const char* path, const filt_fields* h, const filt_params* p, int localNx,
                      int rank, int size, int* j)
{
    return filt::guarded([&]() -> int {
        int rc = check_args(path, h, p, localNx, rank, size);
        if (rc)
            return rc;
        int lnx = 0;
        filt_N_to_local(&lnx, nullptr, size, rank, p);
        if (lnx != localNx)
            return filt::set_error(FILT_ERR_INVALID, "localNx %d is not rank %d's %d", localNx, rank, lnx);
        if (!h->P1 || !h->S2 || !h->roS1_old || !h->roS2_old || !h->media)
            return filt::set_error(FILT_ERR_INVALID, "P1, S2, roS1_old, roS2_old and media are needed");
        Out o{h->P1, h->S2, h->x, h->y, h->z, h->roS1_old, h->roS2_old, h->media};
        return read_save(path, p, rank, size, o, j);
    });
}

int filt_save_plots_host(const char* dir, const filt_fields* h, const filt_params* p, int localNx,
                         int rank, int size, double t)
{
    return filt::guarded([&]() -> int {
        int rc = check_args(dir ? dir : ".", h, p, localNx, rank, size);
        if (rc)
            return rc;
        if (!h->P1 || !h->S2 || !h->u2x || !h->u2y || !h->u2z || !h->media)
            return filt::set_error(FILT_ERR_INVALID, "P1, S2, u2x, u2y, u2z and media are needed");
        std::vector<double> x, y, z;
        if (!h->x || !h->y || !h->z)
            coordinates(p, localNx, rank, size, x, y, z);
        View v{};
        v.P1 = h->P1, v.S2 = h->S2, v.u2x = h->u2x, v.u2y = h->u2y, v.u2z = h->u2z;
        v.x = h->x ? h->x : x.data();
        v.y = h->y ? h->y : y.data();
        v.z = h->z ? h->z : z.data();
        v.media = h->media;
        return write_plots(dir, v, p, localNx, rank, size, t);
    });
}

int filt_save(const char* path, const filt_fields* dev, const filt_params* p, int localNx,
              int rank, int size, int j, int truncate)
{
    return filt::guarded([&]() -> int {
        int rc = check_args(path, dev, p, localNx, rank, size);
        if (rc)
            return rc;
        const size_t n = (size_t)localNx * p->Ny * p->Nz;
        std::vector<double> P1, S2, r1, r2, x, y, z;
        std::vector<int> media;
        if ((rc = fetch(dev->P1, P1, n, "P1")) || (rc = fetch(dev->S2, S2, n, "S2")) ||
            (rc = fetch(dev->roS1_old, r1, n, "roS1_old")) ||
            (rc = fetch(dev->roS2_old, r2, n, "roS2_old")) || (rc = fetch_int(dev->media, media, n)))
            return rc;
        coordinates(p, localNx, rank, size, x, y, z);
        View v{};
        v.P1 = P1.data(), v.S2 = S2.data(), v.roS1_old = r1.data(), v.roS2_old = r2.data();
        v.x = x.data(), v.y = y.data(), v.z = z.data(), v.media = media.data();
        return write_save(path, v, n, j, truncate != 0);
    });
}

int filt_restore(const char* path, const filt_fields* dev, const filt_params* p, int localNx,
                 int rank, int size, int* j)
{
    return filt::guarded([&]() -> int {
        int rc = check_args(path, dev, p, localNx, rank, size);
        if (rc)
            return rc;
        int lnx = 0;
        filt_N_to_local(&lnx, nullptr, size, rank, p);
        if (lnx != localNx)
            return filt::set_error(FILT_ERR_INVALID, "localNx %d is not rank %d's %d", localNx, rank, lnx);
        if (!dev->P1 || !dev->S2 || !dev->roS1_old || !dev->roS2_old)
            return filt::set_error(FILT_ERR_INVALID, "P1, S2, roS1_old and roS2_old are needed");
        const size_t n = (size_t)localNx * p->Ny * p->Nz;
        std::vector<double> a[7];
        for (auto& v : a)
            v.resize(n);
        std::vector<int> media(n);
        Out o{a[0].data(), a[1].data(), a[2].data(), a[3].data(), a[4].data(),
              a[5].data(), a[6].data(), media.data()};
        if ((rc = read_save(path, p, rank, size, o, j)))
            return rc;
        // the arrays Initialize uploads after restore (main.cpp:224-228), then
        // x, y, z where the device bundle has them
        double* const to[7] = {dev->P1, dev->S2, dev->x, dev->y, dev->z, dev->roS1_old, dev->roS2_old};
        for (int q = 0; q < 7; q++)
            if (to[q] && (rc = filt_memcpy(to[q], a[q].data(), n * sizeof(double), 0)))
                return rc;
        if (dev->media && (rc = filt_memcpy(dev->media, media.data(), n * sizeof(int), 0)))
            return rc;
        return FILT_OK;
    });
}

int filt_save_plots(const char* dir, const filt_fields* dev, const filt_params* p, int localNx,
                    int rank, int size, double t)
{
    return filt::guarded([&]() -> int {
        int rc = check_args(dir ? dir : ".", dev, p, localNx, rank, size);
        if (rc)
            return rc;
        const size_t n = (size_t)localNx * p->Ny * p->Nz;
        std::vector<double> P1, S2, ux, uy, uz, x, y, z;
        std::vector<int> media;
        if ((rc = fetch(dev->P1, P1, n, "P1")) || (rc = fetch(dev->S2, S2, n, "S2")) ||
            (rc = fetch(dev->u2x, ux, n, "u2x")) || (rc = fetch(dev->u2y, uy, n, "u2y")) ||
            (rc = fetch(dev->u2z, uz, n, "u2z")) || (rc = fetch_int(dev->media, media, n)))
            return rc;
        coordinates(p, localNx, rank, size, x, y, z);
        View v{};
        v.P1 = P1.data(), v.S2 = S2.data(), v.u2x = ux.data(), v.u2y = uy.data(), v.u2z = uz.data();
        v.x = x.data(), v.y = y.data(), v.z = z.data(), v.media = media.data();
        return write_plots(dir, v, p, localNx, rank, size, t);
    });
}

}  // extern "C"

// filt_ensemble.hip -- C ABI of the fused / batched / mixed forms.
//
// An ensemble holds any number of grids ("members"), each with its own shape,
// parameters, property fields and state, resident in HBM.  One step of the
// whole ensemble is one launch of k_fused over the concatenated tile lists;
// filt_ensemble_capture records a run of steps as a hipGraph that
// filt_ensemble_replay launches as a unit.  Members can be added and removed
// between steps (continuous batching); either invalidates a captured graph.
// An ensemble is used from one host thread at a time.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <vector>

#include "../../include/filtration_hip.h"
#include "filt_fused.hpp"
#include "filt_runtime.hpp"
#include "../host/filt_guard.hpp"

using namespace filt;

struct filt_ensemble {
    int precision = FILT_PREC_FP64;
    int tile = 0;  // index into FUSED_TILES
    int kz = 32;
    int order = 0;
    int parity = 0;
    int next_id = 1;
    struct Member {
        int id;
        FusedGrid g;        // host copy (device pointers inside)
        void* block = nullptr;  // one allocation for state and statics
        size_t bytes = 0;
        long long steps = 0;
    };
    std::vector<Member> members;
    bool dirty = true;
    FusedGrid* d_grids = nullptr;
    size_t d_grids_cap = 0;
    int* d_block_grid = nullptr;
    size_t d_block_cap = 0;
    long long total_blocks = 0;
    bool general_lambda = false;  // some member has a medium with lambda != 2
    bool cell_fields = false;     // every member has per-cell K, porosity, one medium
    double* d_partial = nullptr;
    size_t d_partial_cap = 0;
    double* d_mass = nullptr;
    FusedGrid* d_one = nullptr;  // single-grid descriptor for the mass kernel
    // one graph per starting buffer parity (the kernels' parity argument is
    // baked into the captured launches)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int graph_steps = 0;
};

namespace {

void drop_graph(filt_ensemble* e)
{
    for (int q = 0; q < 2; q++) {
        if (e->exec[q]) (void)hipGraphExecDestroy(e->exec[q]);
        if (e->graph[q]) (void)hipGraphDestroy(e->graph[q]);
        e->exec[q] = nullptr;
        e->graph[q] = nullptr;
    }
    e->graph_steps = 0;
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Upload the grid table and the block -> grid map when membership changed.
int sync_tables(filt_ensemble* e)
{
    if (!e->dirty)
        return FILT_OK;
    std::vector<FusedGrid> gs;
    std::vector<int> bg;
    long long first = 0;
    for (size_t q = 0; q < e->members.size(); q++) {
        FusedGrid g = e->members[q].g;
        g.kz = e->kz;
        fused_prepare(g, e->tile);
        g.first_block = first;
        const long long nb = g.n_interior_blocks + g.n_shell_blocks;
        first += nb;
        e->members[q].g = g;
        gs.push_back(g);
        bg.insert(bg.end(), (size_t)nb, (int)q);
    }
    e->total_blocks = first;
    e->general_lambda = false;
    e->cell_fields = !gs.empty();
    for (const FusedGrid& g : gs) {
        e->general_lambda = e->general_lambda || fused_general_lambda(g);
        e->cell_fields = e->cell_fields && fused_cell_fields(g);
    }
    // launches already enqueued read the current tables: let them finish
    // before the tables are rewritten or reallocated
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    if (gs.size() > e->d_grids_cap) {
        if (e->d_grids) (void)hipFree(e->d_grids);
        e->d_grids = nullptr;
        const size_t cap = std::max<size_t>(gs.size() * 2, 16);
        FILT_HIP_TRY(hipMalloc((void**)&e->d_grids, cap * sizeof(FusedGrid)));
        e->d_grids_cap = cap;
    }
    if (bg.size() > e->d_block_cap) {
        if (e->d_block_grid) (void)hipFree(e->d_block_grid);
        e->d_block_grid = nullptr;
        const size_t cap = std::max<size_t>(bg.size() * 2, 1024);
        FILT_HIP_TRY(hipMalloc((void**)&e->d_block_grid, cap * sizeof(int)));
        e->d_block_cap = cap;
    }
    if (!gs.empty()) {
        FILT_HIP_TRY(hipMemcpyAsync(e->d_grids, gs.data(), gs.size() * sizeof(FusedGrid),
                                    hipMemcpyHostToDevice, stream()));
        FILT_HIP_TRY(hipMemcpyAsync(e->d_block_grid, bg.data(), bg.size() * sizeof(int),
                                    hipMemcpyHostToDevice, stream()));
        FILT_HIP_TRY(hipStreamSynchronize(stream()));
    }
    e->dirty = false;
    return FILT_OK;
}

filt_ensemble::Member* find(filt_ensemble* e, int id)
{
    for (auto& m : e->members)
        if (m.id == id)
            return &m;
    return nullptr;
}

}  // namespace

extern "C" {

int filt_ensemble_create(filt_ensemble** out, int precision)
{
    return filt::guarded([&]() -> int {
        if (!out)
            return set_error(FILT_ERR_INVALID, "null out");
        if (precision != FILT_PREC_FP64 && precision != FILT_PREC_FAST &&
            precision != FILT_PREC_MIXED)
            return set_error(FILT_ERR_INVALID,
                             "precision must be FILT_PREC_FP64, FILT_PREC_FAST or FILT_PREC_MIXED");
        filt_ensemble* e = new filt_ensemble;
        e->precision = precision;
        hipError_t r = hipMalloc((void**)&e->d_mass, 2 * sizeof(double));
        if (r == hipSuccess)
            r = hipMalloc((void**)&e->d_one, sizeof(FusedGrid));
        if (r != hipSuccess) {
            if (e->d_mass) (void)hipFree(e->d_mass);
            delete e;
            return hip_fail(r, "filt_ensemble_create");
        }
        *out = e;
        return FILT_OK;
    });
}

int filt_ensemble_destroy(filt_ensemble* e)
{
    if (!e)
        return FILT_OK;
    (void)hipStreamSynchronize(stream());
    drop_graph(e);
    for (auto& m : e->members)
        if (m.block) (void)hipFree(m.block);
    if (e->d_grids) (void)hipFree(e->d_grids);
    if (e->d_block_grid) (void)hipFree(e->d_block_grid);
    if (e->d_partial) (void)hipFree(e->d_partial);
    if (e->d_mass) (void)hipFree(e->d_mass);
    if (e->d_one) (void)hipFree(e->d_one);
    delete e;
    return FILT_OK;
}

int filt_ensemble_set_tiling(filt_ensemble* e, int tx, int ty, int cy, int kz, int order)
{
    const int tile = (tx == 0 && ty == 0) ? 0 : fused_tile_index(tx, ty, cy);
    if (!e || tile < 0 || kz < 1 || kz > 4096 || (order != 0 && order != 1))
        return set_error(FILT_ERR_INVALID, "bad tiling %dx%d/%d kz=%d order=%d", tx, ty, cy, kz,
                         order);
    e->tile = tile;
    e->kz = kz;
    e->order = order;
    e->dirty = true;
    drop_graph(e);
    return FILT_OK;
}

size_t filt_ensemble_member_bytes(const filt_params* p, int precision, int with_perm,
                                  int with_poro, int with_media)
{
    if (!p)
        return 0;
    const size_t n = (size_t)p->NX * p->Ny * p->Nz;
    (void)precision;  // every form keeps K and porosity in FP64
    return 4 * round256(n * 8) + (with_perm ? round256(n * 8) : 0) +
           (with_poro ? round256(n * 8) : 0) + (with_media ? round256(n * 4) : 0);
}

int filt_ensemble_check(const filt_params* p, int precision, const double* perm,
                        const double* poro)
{
    return filt::guarded([&]() -> int {
        if (!p)
            return set_error(FILT_ERR_INVALID, "null params");
        if (precision != FILT_PREC_FP64 && precision != FILT_PREC_FAST &&
            precision != FILT_PREC_MIXED)
            return set_error(FILT_ERR_INVALID, "unknown precision %d", precision);
        if (p->NX < 3 || p->Ny < 3 || p->Nz < 3)
            return set_error(FILT_ERR_UNSUPPORTED, "the fused step needs every extent >= 3");
        if (p->roS_scheme != FILT_SCHEME_UPWIND)
            return set_error(FILT_ERR_UNSUPPORTED,
                             "the fused step implements the upwind roS scheme only (the divgrad "
                             "scheme runs on the stage path)");
        // the kernels address arrays by 32-bit byte offsets
        if ((double)p->NX * p->Ny * p->Nz * 8.0 >= 4294967296.0)
            return set_error(FILT_ERR_UNSUPPORTED,
                             "the fused step keeps every array below 4 GiB (%dx%dx%d)", p->NX,
                             p->Ny, p->Nz);
        if (p->n_media < 0 || p->n_media > FILT_MAX_MEDIA || p->newton_iterations < 0)
            return set_error(FILT_ERR_INVALID, "bad n_media / newton_iterations");
        if (precision == FILT_PREC_FP64 && !fused_divisors_ok(*p))
            return set_error(FILT_ERR_UNSUPPORTED,
                             "the FP64 form needs h1, h2, h3, mu1, mu2 and 1 - S_wr in "
                             "[2^-100, 2^101]");
        // the guard maximises over the media table and the per-cell fields
        const size_t n = (size_t)p->NX * p->Ny * p->Nz;
        const double r = filt_stability_number(p, perm, poro, n);
        if (!(r <= 2.0))
            return set_error(FILT_ERR_UNSTABLE,
                             "stability guard: amplification number %.3g > 2 (explicit scheme "
                             "unstable for these fields and dt)", r);
        return FILT_OK;
    });
}

int filt_ensemble_add(filt_ensemble* e, const filt_params* p, const double* P1, const double* S2,
                      const double* perm, const double* poro, const int* media, int* id)
{
    return filt::guarded([&]() -> int {
        if (!e || !p || !P1 || !S2)
            return set_error(FILT_ERR_INVALID, "null argument");
        int rc = filt_ensemble_check(p, e->precision, perm, poro);
        if (rc)
            return rc;
        const size_t n = (size_t)p->NX * p->Ny * p->Nz;
        const size_t bytes = filt_ensemble_member_bytes(p, e->precision, perm != nullptr,
                                                        poro != nullptr, media != nullptr);
        char* base = nullptr;
        FILT_HIP_TRY(hipMalloc((void**)&base, bytes));
        filt_ensemble::Member m;
        m.id = e->next_id++;
        m.block = base;
        m.bytes = bytes;
        memset(&m.g, 0, sizeof m.g);
        FusedGrid& g = m.g;
        g.p = *p;
        g.nx = p->NX;
        g.ny = p->Ny;
        g.nz = p->Nz;
        size_t off = 0;
        for (int b = 0; b < 2; b++) {
            g.P[b] = (double*)(base + off);
            off += round256(n * 8);
            g.S[b] = (double*)(base + off);
            off += round256(n * 8);
        }
        hipError_t r = hipSuccess;
        auto put = [&](void* dst, const void* src, size_t nb) {
            if (r == hipSuccess)
                r = hipMemcpy(dst, src, nb, hipMemcpyHostToDevice);
        };
        put(g.P[e->parity], P1, n * 8);
        put(g.S[e->parity], S2, n * 8);
        for (int which = 0; which < 2; which++) {
            const double* src = which == 0 ? perm : poro;
            if (!src)
                continue;
            double* d = (double*)(base + off);
            put(d, src, n * 8);
            (which == 0 ? g.perm : g.poro) = d;
            off += round256(n * 8);
        }
        if (media) {
            int* d = (int*)(base + off);
            put(d, media, n * 4);
            g.media = d;
            off += round256(n * 4);
        }
        if (r != hipSuccess) {
            (void)hipFree(base);
            return hip_fail(r, "filt_ensemble_add upload");
        }
        e->members.push_back(m);
        e->dirty = true;
        drop_graph(e);
        if (id)
            *id = m.id;
        return FILT_OK;
    });
}

int filt_ensemble_remove(filt_ensemble* e, int id)
{
    return filt::guarded([&]() -> int {
        if (!e)
            return set_error(FILT_ERR_INVALID, "null ensemble");
        for (size_t q = 0; q < e->members.size(); q++)
            if (e->members[q].id == id) {
                FILT_HIP_TRY(hipStreamSynchronize(stream()));
                (void)hipFree(e->members[q].block);
                e->members.erase(e->members.begin() + q);
                e->dirty = true;
                drop_graph(e);
                return FILT_OK;
            }
        return set_error(FILT_ERR_NOT_FOUND, "no member %d", id);
    });
}

int filt_ensemble_count(const filt_ensemble* e) { return e ? (int)e->members.size() : 0; }

long long filt_ensemble_cells(const filt_ensemble* e)
{
    long long c = 0;
    if (e)
        for (auto& m : e->members)
            c += (long long)m.g.nx * m.g.ny * m.g.nz;
    return c;
}

int filt_ensemble_step(filt_ensemble* e, int nsteps)
{
    return filt::guarded([&]() -> int {
        if (!e || nsteps < 0)
            return set_error(FILT_ERR_INVALID, "bad arguments");
        int rc = sync_tables(e);
        if (rc)
            return rc;
        for (int s = 0; s < nsteps; s++) {
            rc = fused_launch(e->d_grids, e->d_block_grid, e->total_blocks, e->parity, e->precision,
                              e->general_lambda, e->cell_fields, e->tile, e->order, stream());
            if (rc)
                return rc;
            e->parity ^= 1;
        }
        for (auto& m : e->members)
            m.steps += nsteps;
        return FILT_OK;
    });
}

int filt_ensemble_capture(filt_ensemble* e, int nsteps)
{
    return filt::guarded([&]() -> int {
        if (!e || nsteps < 1)
            return set_error(FILT_ERR_INVALID, "bad arguments");
        int rc = sync_tables(e);
        if (rc)
            return rc;
        drop_graph(e);
        FILT_HIP_TRY(hipStreamSynchronize(stream()));
        for (int start = 0; start < 2; start++) {
            FILT_HIP_TRY(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
            int par = start;
            for (int s = 0; s < nsteps && rc == FILT_OK; s++) {
                rc = fused_launch(e->d_grids, e->d_block_grid, e->total_blocks, par, e->precision,
                                  e->general_lambda, e->cell_fields, e->tile, e->order, stream());
                par ^= 1;
            }
            hipGraph_t g = nullptr;
            hipError_t r = hipStreamEndCapture(stream(), &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                drop_graph(e);
                return rc;
            }
            if (r != hipSuccess) {
                drop_graph(e);
                return hip_fail(r, "hipStreamEndCapture");
            }
            r = hipGraphInstantiate(&e->exec[start], g, nullptr, nullptr, 0);
            e->graph[start] = g;
            if (r != hipSuccess) {
                drop_graph(e);
                return hip_fail(r, "hipGraphInstantiate");
            }
        }
        e->graph_steps = nsteps;
        return FILT_OK;
    });
}

int filt_ensemble_replay(filt_ensemble* e)
{
    if (!e || !e->exec[e->parity])
        return set_error(FILT_ERR_INVALID, "no captured graph (membership changed?)");
    FILT_HIP_TRY(hipGraphLaunch(e->exec[e->parity], stream()));
    if (e->graph_steps & 1)
        e->parity ^= 1;
    for (auto& m : e->members)
        m.steps += e->graph_steps;
    return FILT_OK;
}

int filt_ensemble_get(filt_ensemble* e, int id, double* P1, double* S2)
{
    if (!e)
        return set_error(FILT_ERR_INVALID, "null ensemble");
    filt_ensemble::Member* m = find(e, id);
    if (!m)
        return set_error(FILT_ERR_NOT_FOUND, "no member %d", id);
    const size_t n = (size_t)m->g.nx * m->g.ny * m->g.nz * 8;
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    if (P1) FILT_HIP_TRY(hipMemcpy(P1, m->g.P[e->parity], n, hipMemcpyDeviceToHost));
    if (S2) FILT_HIP_TRY(hipMemcpy(S2, m->g.S[e->parity], n, hipMemcpyDeviceToHost));
    return FILT_OK;
}

long long filt_ensemble_steps(const filt_ensemble* e, int id)
{
    if (!e)
        return -1;
    for (auto& m : e->members)
        if (m.id == id)
            return m.steps;
    return -1;
}

int filt_ensemble_mass(filt_ensemble* e, int id, double* out)
{
    if (!e || !out)
        return set_error(FILT_ERR_INVALID, "null argument");
    int rc = sync_tables(e);
    if (rc)
        return rc;
    filt_ensemble::Member* m = find(e, id);
    if (!m)
        return set_error(FILT_ERR_NOT_FOUND, "no member %d", id);
    const size_t nb = (size_t)fused_mass_blocks(m->g);
    if (nb > e->d_partial_cap) {
        if (e->d_partial) (void)hipFree(e->d_partial);
        e->d_partial = nullptr;
        FILT_HIP_TRY(hipMalloc((void**)&e->d_partial, 2 * nb * sizeof(double)));
        e->d_partial_cap = nb;
    }
    FILT_HIP_TRY(hipMemcpyAsync(e->d_one, &m->g, sizeof(FusedGrid), hipMemcpyHostToDevice,
                                stream()));
    rc = fused_mass(e->d_one, m->g, e->parity, e->d_partial, e->d_mass, stream());
    if (rc)
        return rc;
    FILT_HIP_TRY(hipMemcpyAsync(out, e->d_mass, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                stream()));
    FILT_HIP_TRY(hipStreamSynchronize(stream()));
    return FILT_OK;
}

size_t filt_ensemble_device_bytes(const filt_ensemble* e)
{
    size_t b = 0;
    if (e)
        for (auto& m : e->members)
            b += m.bytes;
    return b;
}

int filt_device_free_bytes(size_t* free_b, size_t* total_b)
{
    size_t f = 0, t = 0;
    FILT_HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_b) *free_b = f;
    if (total_b) *total_b = t;
    return FILT_OK;
}

int filt_ensemble_get_tiling(const filt_ensemble* e, int* tx, int* ty, int* cy, int* kz,
                             int* order)
{
    if (!e)
        return set_error(FILT_ERR_INVALID, "null ensemble");
    if (tx) *tx = FUSED_TILES[e->tile].tx;
    if (ty) *ty = FUSED_TILES[e->tile].ty;
    if (cy) *cy = FUSED_TILES[e->tile].cy;
    if (kz) *kz = e->kz;
    if (order) *order = e->order;
    return FILT_OK;
}

// Time every (tile shape, kz, order) for the current membership: reps steps
// each after one untimed step (the state advances by the steps run), keep
// the fastest.
int filt_ensemble_autotune(filt_ensemble* e, int reps, int* best_tx, int* best_ty, int* best_cy,
                           int* best_kz, int* best_order, double* best_ms)
{
    return filt::guarded([&]() -> int {
        if (!e || reps < 1)
            return set_error(FILT_ERR_INVALID, "bad arguments");
        static const int kzs[] = {4, 8, 16, 32, 64};
        hipEvent_t a = nullptr, b = nullptr;
        int rc = FILT_OK;
        hipError_t he = hipEventCreate(&a);
        if (he == hipSuccess)
            he = hipEventCreate(&b);
        double best = 1e30;
        int bt = e->tile, bk = e->kz, bo = e->order;
        for (int tile = 0; tile < N_FUSED_TILES && rc == FILT_OK && he == hipSuccess; tile++)
            for (int kz : kzs)
                for (int order = 0; order < 2 && rc == FILT_OK && he == hipSuccess; order++) {
                    rc = filt_ensemble_set_tiling(e, FUSED_TILES[tile].tx, FUSED_TILES[tile].ty,
                                                  FUSED_TILES[tile].cy, kz, order);
                    if (!rc) rc = filt_ensemble_step(e, 1);  // warm
                    if (!rc) he = hipEventRecord(a, stream());
                    if (!rc && he == hipSuccess) rc = filt_ensemble_step(e, reps);
                    if (!rc && he == hipSuccess) he = hipEventRecord(b, stream());
                    if (!rc && he == hipSuccess) he = hipEventSynchronize(b);
                    float ms = 0;
                    if (!rc && he == hipSuccess) he = hipEventElapsedTime(&ms, a, b);
                    if (!rc && he == hipSuccess && ms / reps < best) {
                        best = ms / reps;
                        bt = tile;
                        bk = kz;
                        bo = order;
                    }
                }
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (he != hipSuccess)
            return hip_fail(he, "filt_ensemble_autotune");
        if (rc)
            return rc;
        rc = filt_ensemble_set_tiling(e, FUSED_TILES[bt].tx, FUSED_TILES[bt].ty,
                                      FUSED_TILES[bt].cy, bk, bo);
        if (best_tx) *best_tx = FUSED_TILES[bt].tx;
        if (best_ty) *best_ty = FUSED_TILES[bt].ty;
        if (best_cy) *best_cy = FUSED_TILES[bt].cy;
        if (best_kz) *best_kz = bk;
        if (best_order) *best_order = bo;
        if (best_ms) *best_ms = best;
        return rc;
    });
}

}  // extern "C"

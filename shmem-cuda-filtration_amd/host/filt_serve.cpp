// filt_serve.cpp -- continuous-batching serving loop over an ensemble.
//
// Requests (a grid, its parameters, fields, initial state and a step count)
// wait in a FIFO queue.  Each scheduling round:
//   1. admits queued requests, in order, while a slot is free and the
//      device has the member's bytes plus a reserve free (hipMemGetInfo);
//   2. advances every active request together, one fused launch per step,
//      for as many steps as remain until the next request completes (at most
//      max_round_steps, so that arrivals join within a few steps);
//   3. retires requests that reached their step count: copies P1, S2 and the
//      mass to host memory and frees their device memory.
// State stays on the device from admission to retirement.  A request's
// arithmetic never touches another's data, so its result is the same as when
// it runs alone.  Use a server from one host thread at a time.
#include <stdint.h>
#include <string.h>

#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "../../include/filtration_hip.h"
#include "filt_guard.hpp"

namespace {

struct Request {
    int id = 0;
    filt_params p{};
    std::vector<double> P1, S2, perm, poro;
    std::vector<int> media;
    long long nsteps = 0;
    int member = 0;  // ensemble member id while active
    long long steps_at_admission = 0;
    double mass[2] = {0, 0};
};

size_t cells_of(const filt_params& p) { return (size_t)p.NX * p.Ny * p.Nz; }

}  // namespace

struct filt_server {
    int precision = FILT_PREC_MIXED;
    size_t reserve = 0;
    int max_active = 64;
    int max_round_steps = 8;
    int next_id = 1;
    filt_ensemble* ens = nullptr;
    long long ens_steps = 0;  // steps the ensemble has taken
    std::deque<std::unique_ptr<Request>> queue;
    std::map<int, std::unique_ptr<Request>> active;
    std::map<int, std::unique_ptr<Request>> done;
    long long cell_updates = 0;
};

extern "C" {

int filt_serve_create(filt_server** out, int precision, size_t reserve_bytes, int max_active,
                      int max_round_steps)
{
    return filt::guarded([&]() -> int {
        if (!out || max_active < 1 || max_round_steps < 1)
            return FILT_ERR_INVALID;
        filt_server* s = new filt_server;
        s->precision = precision;
        s->reserve = reserve_bytes;
        s->max_active = max_active;
        s->max_round_steps = max_round_steps;
        int rc = filt_ensemble_create(&s->ens, precision);
        if (rc) {
            delete s;
            return rc;
        }
        *out = s;
        return FILT_OK;
    });
}

int filt_serve_destroy(filt_server* s)
{
    if (!s)
        return FILT_OK;
    filt_ensemble_destroy(s->ens);
    delete s;
    return FILT_OK;
}

int filt_serve_submit(filt_server* s, const filt_params* p, const double* P1, const double* S2,
                      const double* perm, const double* poro, const int* media, long long nsteps,
                      int* id)
{
    return filt::guarded([&]() -> int {
        if (!s || !p || !P1 || !S2 || nsteps < 0)
            return FILT_ERR_INVALID;
        int rc = filt_ensemble_check(p, s->precision, perm, poro);
        if (rc)
            return rc;
        // a request that can never fit would block the FIFO queue for good
        const size_t need = filt_ensemble_member_bytes(p, s->precision, perm != nullptr,
                                                       poro != nullptr, media != nullptr);
        size_t free_b = 0, total_b = 0;
        rc = filt_device_free_bytes(&free_b, &total_b);
        if (rc)
            return rc;
        if (need + s->reserve > total_b)
            return filt::set_error(FILT_ERR_NOMEM,
                                   "request needs %zu device bytes plus a reserve of %zu; the "
                                   "device has %zu", need, s->reserve, total_b);
        auto r = std::make_unique<Request>();
        r->id = s->next_id++;
        r->p = *p;
        r->nsteps = nsteps;
        const size_t n = cells_of(*p);
        r->P1.assign(P1, P1 + n);
        r->S2.assign(S2, S2 + n);
        if (perm) r->perm.assign(perm, perm + n);
        if (poro) r->poro.assign(poro, poro + n);
        if (media) r->media.assign(media, media + n);
        if (id)
            *id = r->id;
        s->queue.push_back(std::move(r));
        return FILT_OK;
    });
}

// One scheduling round (see the file comment).  *n_admitted / *n_retired
// receive the counts of this round (either may be null).
int filt_serve_round(filt_server* s, int* n_admitted, int* n_retired)
{
    return filt::guarded([&]() -> int {
        if (!s)
            return FILT_ERR_INVALID;
        int admitted = 0, retired = 0;
        // 1. admission, FIFO
        while (!s->queue.empty() && (int)s->active.size() < s->max_active) {
            Request& r = *s->queue.front();
            const size_t need = filt_ensemble_member_bytes(&r.p, s->precision, !r.perm.empty(),
                                                           !r.poro.empty(), !r.media.empty());
            size_t free_b = 0, total_b = 0;
            int rc = filt_device_free_bytes(&free_b, &total_b);
            if (rc)
                return rc;
            if (free_b < need + s->reserve) {
                if (s->active.empty())  // nothing will retire and free memory
                    return filt::set_error(FILT_ERR_NOMEM,
                                           "request %d needs %zu device bytes plus a reserve of "
                                           "%zu; %zu are free and no request is active",
                                           r.id, need, s->reserve, free_b);
                break;
            }
            rc = filt_ensemble_add(s->ens, &r.p, r.P1.data(), r.S2.data(),
                                   r.perm.empty() ? nullptr : r.perm.data(),
                                   r.poro.empty() ? nullptr : r.poro.data(),
                                   r.media.empty() ? nullptr : r.media.data(), &r.member);
            if (rc)
                return rc;
            r.steps_at_admission = s->ens_steps;
            // inputs now live on the device
            std::vector<double>().swap(r.P1);
            std::vector<double>().swap(r.S2);
            std::vector<double>().swap(r.perm);
            std::vector<double>().swap(r.poro);
            std::vector<int>().swap(r.media);
            std::unique_ptr<Request> own = std::move(s->queue.front());
            s->queue.pop_front();
            s->active[own->id] = std::move(own);
            admitted++;
        }
        // 2. advance everyone until the next completion (bounded)
        long long k = s->max_round_steps;
        long long cells = 0;
        for (auto& kv : s->active) {
            const Request& r = *kv.second;
            const long long left = r.nsteps - (s->ens_steps - r.steps_at_admission);
            if (left < k)
                k = left;
            cells += (long long)cells_of(r.p);
        }
        if (!s->active.empty() && k > 0) {
            int rc = filt_ensemble_step(s->ens, (int)k);
            if (rc)
                return rc;
            s->ens_steps += k;
            s->cell_updates += cells * k;
        }
        // 3. retirement
        for (auto it = s->active.begin(); it != s->active.end();) {
            Request& r = *it->second;
            if (s->ens_steps - r.steps_at_admission < r.nsteps) {
                ++it;
                continue;
            }
            const size_t n = cells_of(r.p);
            r.P1.resize(n);
            r.S2.resize(n);
            int rc = filt_ensemble_get(s->ens, r.member, r.P1.data(), r.S2.data());
            if (!rc)
                rc = filt_ensemble_mass(s->ens, r.member, r.mass);
            if (!rc)
                rc = filt_ensemble_remove(s->ens, r.member);
            if (rc)
                return rc;
            s->done[r.id] = std::move(it->second);
            it = s->active.erase(it);
            retired++;
        }
        if (n_admitted) *n_admitted = admitted;
        if (n_retired) *n_retired = retired;
        return FILT_OK;
    });
}

int filt_serve_status(const filt_server* s, int* queued, int* active, int* done,
                      long long* active_cells, long long* cell_updates)
{
    if (!s)
        return FILT_ERR_INVALID;
    long long ac = 0;
    for (auto& kv : s->active)
        ac += (long long)cells_of(kv.second->p);
    if (queued) *queued = (int)s->queue.size();
    if (active) *active = (int)s->active.size();
    if (done) *done = (int)s->done.size();
    if (active_cells) *active_cells = ac;
    if (cell_updates) *cell_updates = s->cell_updates;
    return FILT_OK;
}

// 1 when request id has finished (its result waits in the server), else 0.
int filt_serve_is_done(const filt_server* s, int id)
{
    return (s && s->done.count(id)) ? 1 : 0;
}

// Copy a finished request's P1, S2 (NX*Ny*Nz each, either may be null) and
// water / NAPL mass (mass2, may be null) out and forget it.
int filt_serve_take(filt_server* s, int id, double* P1, double* S2, double* mass2)
{
    return filt::guarded([&]() -> int {
        if (!s)
            return FILT_ERR_INVALID;
        auto it = s->done.find(id);
        if (it == s->done.end())
            return FILT_ERR_NOT_FOUND;
        Request& r = *it->second;
        if (P1) memcpy(P1, r.P1.data(), r.P1.size() * sizeof(double));
        if (S2) memcpy(S2, r.S2.data(), r.S2.size() * sizeof(double));
        if (mass2) {
            mass2[0] = r.mass[0];
            mass2[1] = r.mass[1];
        }
        s->done.erase(it);
        return FILT_OK;
    });
}

}  // extern "C"

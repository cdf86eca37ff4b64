/* Request server for SES meshes on one GPU.
 *
 * Surface requests (atoms, radii, voxel size, Laplacian passes) arrive at any time through
 * qs_serve_submit.  Worker threads, each owning an engine with its own HIP stream and
 * workspace, take the pending requests in arrival order while their estimated device memory
 * fits the worker's share of the budget, mesh them as one batch (one launch sequence), and
 * copy each mesh into a device-resident result arena where it stays until the client
 * releases it.  While one worker's batch runs, requests that arrive are taken by an idle
 * worker, so batches overlap on the device (continuous batching across the workers' slots).
 * With an express threshold set, the last worker takes only requests of at most that many grid
 * points (oldest first), so that a small request arriving while the other workers run large
 * batches starts at once on its own stream; its kernels then share the device with theirs.
 *
 * Every result equals the mesh of the same request computed alone (the batched kernels give
 * each molecule exactly its single-molecule mesh).  The reference has no counterpart: its
 * API_computeSES is one synchronous call per molecule (CudaSurf.cu:787-842).
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/quickses.h"
#include "qs_engine.h"

using namespace qs;

namespace {

using Clock = std::chrono::steady_clock;

enum ReqState { PENDING = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

struct Request {
    int64_t id = 0;
    std::vector<float> pos, rad;
    float reso = 0.5f;
    int smooth = 1;
    BBox bb{};
    uint64_t est_bytes = 0;
    uint64_t voxels = 0; /* g^3: the request's share of a batch's work */
    Clock::time_point t_submit, t_start, t_done;
    ReqState state = PENDING;
    char err[256] = "";
    uint32_t nv = 0, nt = 0;
    uint64_t arena_off = 0, arena_bytes = 0; /* vertices, then triangles, then atom ids */
    int batch_size = 0;
};

/* first-fit allocator over one device buffer (offsets 256-byte aligned) */
struct Arena {
    DevBuf buf;
    uint64_t size = 0, used = 0;
    std::map<uint64_t, uint64_t> free_; /* offset -> bytes */

    int init(uint64_t bytes)
    {
        bytes = (bytes + 255) & ~255ull;
        QS_TRY(ensure(buf, bytes));
        size = bytes;
        used = 0;
        free_.clear();
        free_[0] = bytes;
        return 0;
    }
    bool alloc(uint64_t bytes, uint64_t &off)
    {
        bytes = std::max<uint64_t>((bytes + 255) & ~255ull, 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used += bytes;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        bytes = std::max<uint64_t>((bytes + 255) & ~255ull, 256);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;
    std::map<int64_t, Request *> reqs;
    std::vector<std::thread> threads;
    std::vector<Engine *> engines;
    Arena arena;
    uint64_t budget_per_worker = 0;
    int max_batch = 256;
    uint64_t max_batch_voxels = 0; /* 0: no cap */
    uint64_t express_voxels = 0;   /* > 0 (and 2 workers or more): the last worker's request size cap */
    int device = 0;
    bool closing = false;
    int64_t next_id = 0;
    uint64_t batches = 0, batched_requests = 0, largest_batch = 0, busy_workers = 0, max_busy = 0, express_batches = 0;
    int ready_workers = 0;
};

Server *g_srv = nullptr;

/* device bytes of one request in an engine's workspace: grid arrays per voxel (field 4, bit
   masks and per-word counts about 1.4), atom arrays, neighbour cells, and the mesh arrays for a
   surface of about 6 g^2 vertices */
uint64_t estimate_bytes(const QsGrid &G, uint32_t n)
{
    const uint64_t g = (uint64_t)G.g, gn = (uint64_t)G.gn;
    return g * g * g * 11 / 2 + (uint64_t)n * 64 + gn * gn * gn * 20 + 6 * g * g * 140 + 4096;
}

void finish(Server *S, std::vector<Request *> &batch, const char *err)
{
    std::lock_guard<std::mutex> lk(S->mu);
    const auto now = Clock::now();
    for (Request *r : batch) {
        if (err) {
            r->state = FAILED;
            snprintf(r->err, sizeof(r->err), "%s", err);
        } else if (r->state == RUNNING) {
            r->state = DONE;
        }
        r->t_done = now;
    }
    S->busy_workers--;
    S->cv_done.notify_all();
}

/* mesh one batch on engine E and move each mesh into the result arena */
int run_batch_requests(Server *S, Engine &E, std::vector<Request *> &batch)
{
    const int nb = (int)batch.size();
    std::vector<MolIn> in(nb);
    std::vector<BBox> bbs(nb);
    uint64_t n = 0;
    for (int i = 0; i < nb; i++) {
        in[i].atom_off = (uint32_t)n;
        in[i].n = (uint32_t)batch[i]->rad.size();
        in[i].reso = batch[i]->reso;
        bbs[i] = batch[i]->bb;
        n += in[i].n;
    }
    QS_TRY(ensure(E.in_pos, std::max<uint64_t>(n, 1) * 12));
    QS_TRY(ensure(E.in_rad, std::max<uint64_t>(n, 1) * 4));
    QS_TRY(E.upload_begin());
    for (int i = 0; i < nb; i++) {
        const size_t k = batch[i]->rad.size();
        if (!k) continue;
        QS_TRY(E.upload_input(as<float>(E.in_pos) + 3 * (size_t)in[i].atom_off, batch[i]->pos.data(), k * 12));
        QS_TRY(E.upload_input(as<float>(E.in_rad) + in[i].atom_off, batch[i]->rad.data(), k * 4));
    }
    QS_TRY(E.run_batch(as<float>(E.in_pos), as<float>(E.in_rad), in, batch[0]->smooth, &bbs));
    if (E.nt) {
        QS_TRY(ensure(E.tri_loc, (size_t)E.nt * 12));
        qs_launch_tri_localize(E.res_tri, E.nt, as<uint32_t>(E.bounds), as<uint32_t>(E.d_off) + (nb + 1), nb,
                               as<int>(E.tri_loc), E.stream);
        QS_HIP(hipGetLastError());
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (int i = 0; i < nb; i++) {
            Request *r = batch[i];
            r->nv = E.vert_off[i + 1] - E.vert_off[i];
            r->nt = E.tri_off[i + 1] - E.tri_off[i];
            r->batch_size = nb;
            const uint64_t vb = ((uint64_t)r->nv * 12 + 255) & ~255ull, tb = ((uint64_t)r->nt * 12 + 255) & ~255ull;
            r->arena_bytes = std::max<uint64_t>(vb + tb + (uint64_t)r->nv * 4, 256); /* never 0: release keys on it */
            if (!S->arena.alloc(r->arena_bytes, r->arena_off)) {
                r->state = FAILED;
                r->arena_bytes = 0;
                snprintf(r->err, sizeof(r->err), "result arena full (%llu of %llu bytes in use): release results",
                         (unsigned long long)S->arena.used, (unsigned long long)S->arena.size);
            }
        }
    }
    uint8_t *base = as<uint8_t>(S->arena.buf);
    for (int i = 0; i < nb; i++) {
        Request *r = batch[i];
        if (r->state == FAILED) continue;
        const uint64_t vb = ((uint64_t)r->nv * 12 + 255) & ~255ull, tb = ((uint64_t)r->nt * 12 + 255) & ~255ull;
        uint8_t *dst = base + r->arena_off;
        if (r->nv) {
            QS_HIP(hipMemcpyAsync(dst, E.res_vert + 3 * (size_t)E.vert_off[i], (size_t)r->nv * 12, hipMemcpyDeviceToDevice,
                                  E.stream));
            QS_HIP(hipMemcpyAsync(dst + vb + tb, E.res_atom + E.vert_off[i], (size_t)r->nv * 4, hipMemcpyDeviceToDevice,
                                  E.stream));
        }
        if (r->nt)
            QS_HIP(hipMemcpyAsync(dst + vb, as<int>(E.tri_loc) + 3 * (size_t)E.tri_off[i], (size_t)r->nt * 12,
                                  hipMemcpyDeviceToDevice, E.stream));
    }
    QS_HIP(hipStreamSynchronize(E.stream));
    return 0;
}

void worker_main(Server *S, int w)
{
    if (hipSetDevice(S->device) != hipSuccess) (void)hipGetLastError(); /* the engine's first call reports it */
    g_grow_shift = 0; /* grow by doubling: a reallocation (hipFree) would stall the other workers */
    Engine &E = *S->engines[w];
    int init_rc = E.init();
    /* workspace for batches of about a twelfth of the budget in voxels, so that growth is rare */
    const uint64_t vox = S->budget_per_worker / 24;
    if (!init_rc) init_rc = E.reserve(vox, vox / 64, vox / 64, vox / 24, S->max_batch);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->ready_workers++;
    }
    S->cv_done.notify_all();
    const int nworkers = (int)S->engines.size();
    for (;;) {
        std::vector<Request *> batch;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            /* the express worker waits for a pending request within its size cap */
            auto express = [&] { return S->express_voxels > 0 && nworkers >= 2 && w == nworkers - 1; };
            auto small_pending = [&] {
                for (Request *r : S->pending)
                    if (r->voxels <= S->express_voxels) return true;
                return false;
            };
            S->cv_work.wait(lk, [&] { return S->closing || (express() ? small_pending() : !S->pending.empty()); });
            const bool ex = express();
            /* closing, and nothing left that this worker takes (the express worker leaves larger
               requests to the others, which finish every pending request before they stop) */
            if (ex ? !small_pending() : S->pending.empty()) return;
            /* arrival order; the batch ends at the first request that does not fit the budget
               (a request larger than the budget runs alone) or asks for another pass count.  The
               express worker takes the requests within its cap, skipping the others. */
            uint64_t acc = 0, accv = 0;
            int smooth = -1;
            const auto now = Clock::now();
            for (auto it = S->pending.begin(); it != S->pending.end() && (int)batch.size() < S->max_batch;) {
                Request *r = *it;
                if (ex && r->voxels > S->express_voxels) {
                    ++it;
                    continue;
                }
                if (smooth < 0) smooth = r->smooth;
                if (r->smooth != smooth) break;
                if (!batch.empty() && acc + r->est_bytes > S->budget_per_worker) break;
                /* work cap: a batch stays short, so requests arriving meanwhile wait less */
                if (!batch.empty() && S->max_batch_voxels && accv + r->voxels > S->max_batch_voxels) break;
                it = S->pending.erase(it);
                r->state = RUNNING;
                r->t_start = now;
                acc += r->est_bytes;
                accv += r->voxels;
                batch.push_back(r);
            }
            if (batch.empty()) continue; /* (another worker took them) */
            S->batches++;
            if (ex) S->express_batches++;
            S->batched_requests += batch.size();
            S->largest_batch = std::max<uint64_t>(S->largest_batch, batch.size());
            S->busy_workers++;
            S->max_busy = std::max(S->max_busy, S->busy_workers);
        }
        int rc = init_rc ? init_rc : run_batch_requests(S, E, batch);
        finish(S, batch, rc ? g_errmsg : nullptr);
    }
}

int check_server()
{
    if (!g_srv) return set_error(6, "server not open (qs_serve_open)");
    return 0;
}

}  // namespace

extern "C" {

int qs_serve_open(uint64_t hbm_budget_bytes, uint64_t result_bytes, int workers)
{
    g_err = 0;
    g_errmsg[0] = 0;
    if (g_srv) return set_error(6, "server already open");
    if (workers < 1 || workers > 8) return set_error(3, "workers must be 1..8");
    if (hbm_budget_bytes < (64ull << 20) || result_bytes < (1ull << 20)) return set_error(3, "budget too small");
    Server *S = new Server();
    if (hipGetDevice(&S->device) != hipSuccess) {
        delete S;
        return set_error(2, "no HIP device");
    }
    S->budget_per_worker = hbm_budget_bytes / (uint64_t)workers;
    if (S->arena.init(result_bytes)) {
        delete S;
        return g_err;
    }
    for (int w = 0; w < workers; w++) S->engines.push_back(new Engine());
    for (int w = 0; w < workers; w++) S->threads.emplace_back(worker_main, S, w);
    {
        /* the workers have set up their engines and workspaces */
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] { return S->ready_workers == workers; });
    }
    g_srv = S;
    return 0;
}

int64_t qs_serve_submit(const float *pos, const float *rad, unsigned int N, float reso, int doSmoothing)
{
    g_err = 0;
    g_errmsg[0] = 0;
    if (check_server()) return -g_err;
    if (N > 1 && (!pos || !rad)) return -set_error(3, "null atom arrays");
    if (N > (1u << 30)) return -set_error(3, "too many atoms (%u)", N);
    if (!(reso > 0.0f) || !std::isfinite(reso)) return -set_error(3, "voxel size must be a positive finite number");
    Request *r = new Request();
    r->reso = reso;
    r->smooth = doSmoothing;
    if (N > 1) {
        int nonfinite = 0;
        bbox_host(pos, rad, N, r->bb, nonfinite);
        if (nonfinite) {
            delete r;
            return -set_error(3, "atom positions or radii contain NaN or Inf");
        }
        for (int c = 0; c < 3; c++)
            if (!(r->bb.vmin[c] > -100000.0f && r->bb.vmax[c] < 100000.0f)) {
                delete r;
                return -set_error(3, "atom coordinates must lie strictly within (-1e5, 1e5) Angstrom");
            }
        QsGrid G;
        grid_from_bbox(r->bb, reso, G);
        if ((uint64_t)G.g * G.g * G.g >= (1ull << 32)) {
            delete r;
            return -set_error(4, "grid of %d^3 points", G.g);
        }
        r->est_bytes = estimate_bytes(G, N);
        r->voxels = (uint64_t)G.g * G.g * G.g;
        r->pos.assign(pos, pos + 3 * (size_t)N);
        r->rad.assign(rad, rad + N);
    } else {
        /* fewer than 2 atoms: an empty mesh, as the reference (CudaSurf.cu:539-543) */
        r->est_bytes = 4096;
    }
    Server *S = g_srv;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        r->id = S->next_id++;
        r->t_submit = Clock::now();
        S->reqs[r->id] = r;
        S->pending.push_back(r);
    }
    /* every worker re-checks: the express worker takes only small requests */
    S->cv_work.notify_all();
    return r->id;
}

int qs_serve_poll(int64_t id, int wait_ms, unsigned int *nvert, unsigned int *ntri, float **d_vert, int **d_tri,
                  int **d_atom_id, double *latency_ms)
{
    if (check_server()) return -g_err;
    Server *S = g_srv;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return -set_error(3, "unknown request %lld", (long long)id);
    Request *r = it->second;
    auto ready = [&] { return r->state == DONE || r->state == FAILED; };
    if (!ready() && wait_ms != 0) {
        if (wait_ms < 0) S->cv_done.wait(lk, ready);
        else S->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    }
    if (!ready()) return 1;
    if (r->state == FAILED) return -set_error(5, "request %lld: %s", (long long)id, r->err);
    const uint64_t vb = ((uint64_t)r->nv * 12 + 255) & ~255ull, tb = ((uint64_t)r->nt * 12 + 255) & ~255ull;
    uint8_t *base = as<uint8_t>(S->arena.buf) + r->arena_off;
    if (nvert) *nvert = r->nv;
    if (ntri) *ntri = r->nt;
    if (d_vert) *d_vert = (float *)base;
    if (d_tri) *d_tri = (int *)(base + vb);
    if (d_atom_id) *d_atom_id = (int *)(base + vb + tb);
    if (latency_ms) {
        latency_ms[0] = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
        latency_ms[1] = std::chrono::duration<double, std::milli>(r->t_start - r->t_submit).count();
        latency_ms[2] = (double)r->batch_size;
    }
    return 0;
}

int qs_serve_fetch(int64_t id, float *vert, int *tri, int *atom_id)
{
    unsigned nv = 0, nt = 0;
    float *dv = nullptr;
    int *dt = nullptr, *da = nullptr;
    const int rc = qs_serve_poll(id, -1, &nv, &nt, &dv, &dt, &da, nullptr);
    if (rc) return rc;
    if (nv && vert) QS_HIP(hipMemcpy(vert, dv, (size_t)nv * 12, hipMemcpyDeviceToHost));
    if (nt && tri) QS_HIP(hipMemcpy(tri, dt, (size_t)nt * 12, hipMemcpyDeviceToHost));
    if (nv && atom_id) QS_HIP(hipMemcpy(atom_id, da, (size_t)nv * 4, hipMemcpyDeviceToHost));
    return 0;
}

int qs_serve_release(int64_t id)
{
    if (check_server()) return -g_err;
    Server *S = g_srv;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return -set_error(3, "unknown request %lld", (long long)id);
    Request *r = it->second;
    if (r->state == PENDING || r->state == RUNNING) return -set_error(3, "request %lld is not finished", (long long)id);
    if (r->arena_bytes) S->arena.release(r->arena_off, r->arena_bytes);
    S->reqs.erase(it);
    delete r;
    return 0;
}

int qs_serve_set(const char *key, uint64_t value)
{
    if (check_server()) return -g_err;
    Server *S = g_srv;
    std::lock_guard<std::mutex> lk(S->mu);
    const std::string k = key ? key : "";
    if (k == "max_batch_voxels") S->max_batch_voxels = value;
    else if (k == "express_voxels") S->express_voxels = value;
    else if (k == "max_batch") S->max_batch = (int)std::max<uint64_t>(1, std::min<uint64_t>(value, 256));
    else return -set_error(3, "unknown server setting %s", k.c_str());
    return 0;
}

int qs_serve_stats(uint64_t *out, int n)
{
    if (check_server()) return -g_err;
    Server *S = g_srv;
    std::lock_guard<std::mutex> lk(S->mu);
    const uint64_t v[8] = {S->batches, S->batched_requests, S->largest_batch, S->max_busy, S->arena.used, S->arena.size,
                           (uint64_t)S->pending.size(), S->express_batches};
    for (int i = 0; i < n && i < 8; i++) out[i] = v[i];
    return 8;
}

int qs_serve_close(void)
{
    if (check_server()) return -g_err;
    Server *S = g_srv;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->closing = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->threads) t.join();
    for (Engine *E : S->engines) {
        E->release_all();
        delete E;
    }
    for (auto &kv : S->reqs) delete kv.second;
    release(S->arena.buf);
    delete S;
    g_srv = nullptr;
    return 0;
}

}  // extern "C"

// tpsc_batch.cpp -- many independent parameter points on one GPU.
//
// One worker (the default the tuner usually picks): every point is queued without a host wait
// (tpsc_queue_solve: the root searches read and write device memory only), one context per grid,
// its results copied into a pinned slot of its own; the host then drains the streams and reads
// the slots.  Small points (N <= 128) are launch- and round-trip-bound, so queuing them back to
// back keeps the GPU busy where one point at a time leaves it mostly idle.  A point whose root
// search has not settled in the queued passes is queued again with longer searches, a
// mixed-precision point that needs the FP64 chi0 stage again with it (as tpsc_solve redoes
// it); one that the queue cannot finish as tpsc_solve would (a failed search, a mu0 search
// without a root) is solved again with tpsc_solve.
// Several workers: host threads, each with one context per grid re-targeted point by point
// (tpsc_set_params) before tpsc_solve.  Points are taken in decreasing cost N^2 Nw (longest
// first).  The worker count is either given or chosen by tpsc_batch_tune, which times a probe
// subset at 1, 2, 4 and 8 workers.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <random>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/tpsc_capi.h"
#include "tpsc_queue.hpp"

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// contexts of one worker, one per (N, Nw, precision), re-targeted point by point
struct Pool {
    std::map<std::tuple<int, int, int>, tpsc_ctx*> ctx;
    ~Pool()
    {
        for (auto& kv : ctx) tpsc_destroy(kv.second);
    }
    // the context of p's grid, created at p when missing (not re-targeted)
    int ctx_for(const tpsc_params& p, tpsc_ctx** out)
    {
        const auto key = std::make_tuple(p.nk0, p.Nw0, p.precision);
        auto it = ctx.find(key);
        if (it != ctx.end()) {
            *out = it->second;
            return TPSC_OK;
        }
        tpsc_ctx* c = nullptr;
        const int rc = tpsc_create(&p, &c);
        if (rc) {
            tpsc_destroy(c);
            return rc;
        }
        ctx[key] = c;
        *out = c;
        return TPSC_OK;
    }
    // the context of p's grid, re-targeted to p
    int get(const tpsc_params& p, tpsc_ctx** out)
    {
        const bool fresh = !ctx.count(std::make_tuple(p.nk0, p.Nw0, p.precision));
        const int rc = ctx_for(p, out);
        if (rc || fresh) return rc;
        return tpsc_set_params(*out, &p);
    }
};

void solve_point(Pool& pool, const tpsc_params& p, tpsc_batch_result& r, double* jj, int ldc)
{
    const double t0 = now_s();
    std::memset(&r, 0, sizeof r);
    tpsc_ctx* c = nullptr;
    int rc = pool.get(p, &c);
    if (!rc) rc = tpsc_solve(c);
    if (c) {
        tpsc_get_scalar(c, TPSC_S_MU0, &r.mu0);
        tpsc_get_scalar(c, TPSC_S_USP, &r.Usp);
        tpsc_get_scalar(c, TPSC_S_UCH, &r.Uch);
        tpsc_get_scalar(c, TPSC_S_MU, &r.mu);
        tpsc_get_scalar(c, TPSC_S_DBLOCC, &r.dblOcc);
        tpsc_get_scalar(c, TPSC_S_KX0, &r.kx0);
        if (!rc && jj && ldc > 0) {
            const int nwb = p.Nw0 / 2 + 1;
            std::vector<double> a((size_t)nwb), b((size_t)nwb);
            if (!tpsc_get(c, TPSC_F_CHIJJ1, a.data(), a.size() * sizeof(double)) &&
                !tpsc_get(c, TPSC_F_CHIJJ2, b.data(), b.size() * sizeof(double)))
                for (int n = 0; n < std::min(nwb, ldc); n++) jj[n] = a[n] + b[n];
        }
    }
    r.status = rc;
    r.seconds = now_s() - t0;
}

std::vector<int> by_cost(const tpsc_params* pts, int n)
{
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return (double)pts[a].nk0 * pts[a].nk0 * pts[a].Nw0 > (double)pts[b].nk0 * pts[b].nk0 * pts[b].Nw0;
    });
    return order;
}

// one host thread: queue every point, drain, read the slots; queue again, with longer root
// searches or the FP64 chi0 stage, the points that need them; solve what is left with tpsc_solve
// The queued path's contexts (one per grid and precision) and its pinned slots, kept across
// calls: a context costs ~1-2 ms of allocations and stream set-up, more than a small point's
// solve.  Calls on the queued path take turns on the mutex; tpsc_batch_release frees them.
struct QueuedState {
    Pool pool;
    double* slots = nullptr;  // pinned, mapped host memory
    double* slots_dev = nullptr;
    size_t cap = 0;
    ~QueuedState()
    {
        if (slots) (void)hipHostFree(slots);
    }
};
std::mutex g_queued_mu;
QueuedState* g_queued = nullptr;  // never freed at exit: the HIP runtime may be gone by then

void run_queued(const tpsc_params* pts, const std::vector<int>& order, tpsc_batch_result* out, double* jj, int ldc)
{
    const double t0 = now_s();
    std::lock_guard<std::mutex> lock(g_queued_mu);
    if (!g_queued) g_queued = new QueuedState;
    Pool& pool = g_queued->pool;
    const size_t n = order.size();
    std::vector<size_t> off(n + 1, 0);
    std::vector<tpsc_ctx*> ctxs(n, nullptr);
    for (size_t k = 0; k < n; k++) {
        const int i = order[k];
        std::memset(&out[i], 0, sizeof out[i]);
        if (pool.ctx_for(pts[i], &ctxs[k])) ctxs[k] = nullptr;
        off[k + 1] = off[k] + (ctxs[k] ? tpsc_queue_slot_doubles(ctxs[k]) : 0);
    }
    // the slots: pinned host memory that the kernels write through its device address
    QueuedState& q = *g_queued;
    if (off[n] > q.cap) {
        if (q.slots) (void)hipHostFree(q.slots);
        q.slots = q.slots_dev = nullptr;
        q.cap = 0;
        if (hipHostMalloc((void**)&q.slots, off[n] * sizeof(double), hipHostMallocMapped) == hipSuccess &&
            hipHostGetDevicePointer((void**)&q.slots_dev, q.slots, 0) == hipSuccess)
            q.cap = off[n];
        else if (q.slots) {
            (void)hipHostFree(q.slots);
            q.slots = nullptr;
        }
    }
    double *slots = q.cap ? q.slots : nullptr, *slots_dev = q.cap ? q.slots_dev : nullptr;
    struct Item {
        size_t k;
        bool fp64, longer;
    };
    std::vector<Item> todo;
    std::vector<size_t> redo;
    for (size_t k = 0; k < n; k++)
        if (slots && ctxs[k])
            todo.push_back({k, false, false});
        else
            redo.push_back(k);
    size_t done = 0;
    while (!todo.empty()) {  // each item's flags only turn on: at most three rounds
        std::vector<int> qrc(todo.size());
        for (size_t t = 0; t < todo.size(); t++) {
            const Item& it = todo[t];
            qrc[t] = tpsc_queue_solve(ctxs[it.k], &pts[order[it.k]], slots_dev + off[it.k], it.fp64, it.longer);
        }
        for (auto& kv : pool.ctx) tpsc_queue_sync(kv.second);
        std::vector<Item> next;
        for (size_t t = 0; t < todo.size(); t++) {
            Item it = todo[t];
            const int i = order[it.k];
            const int rc = qrc[t] == TPSC_OK ? tpsc_queue_result(&pts[i], slots + off[it.k], it.fp64, &out[i],
                                                                 jj ? jj + (size_t)i * ldc : nullptr, ldc)
                                             : qrc[t];
            if (rc == TPSC_OK) {
                out[i].queued = 1;
                done++;
            } else if (rc == kQueuePromote) {
                it.fp64 = true;
                next.push_back(it);
            } else if (rc == kQueueLonger && !it.longer) {
                it.longer = true;
                next.push_back(it);
            } else
                redo.push_back(it.k);
        }
        todo.swap(next);
    }
    const double per = done ? (now_s() - t0) / done : 0.0;
    for (size_t k = 0; k < n; k++)
        if (out[order[k]].queued) out[order[k]].seconds = per;
    for (size_t k : redo) {
        const int i = order[k];
        solve_point(pool, pts[i], out[i], jj ? jj + (size_t)i * ldc : nullptr, ldc);
    }
}

void run(const tpsc_params* pts, const std::vector<int>& order, int workers, tpsc_batch_result* out, double* jj,
         int ldc)
{
    if (workers <= 1) {
        run_queued(pts, order, out, jj, ldc);
        return;
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
        Pool pool;
        for (size_t k; (k = next.fetch_add(1)) < order.size();) {
            const int i = order[k];
            solve_point(pool, pts[i], out[i], jj ? jj + (size_t)i * ldc : nullptr, ldc);
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < workers; w++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

}  // namespace

extern "C" int tpsc_batch_tune(const tpsc_params* pts, int npts, int* workers, double* rates)
{
    if (!pts || npts <= 0 || !workers) return TPSC_E_INVALID;
    std::vector<int> order = by_cost(pts, npts);
    order.resize((size_t)std::min(npts, 8));
    std::vector<tpsc_batch_result> tmp((size_t)npts);
    const int cand[4] = {1, 2, 4, 8};
    run(pts, order, 1, tmp.data(), nullptr, 0);  // warm-up: code objects and allocator pools
    int best = 1;
    double best_rate = 0.0;
    for (int k = 0; k < 4; k++) {
        const double t0 = now_s();
        run(pts, order, cand[k], tmp.data(), nullptr, 0);
        const double rate = order.size() / (now_s() - t0);
        if (rates) rates[k] = rate;
        if (rate > best_rate) {
            best_rate = rate;
            best = cand[k];
        }
    }
    *workers = best;
    return TPSC_OK;
}

extern "C" int tpsc_solve_batch(const tpsc_params* pts, int npts, int workers, tpsc_batch_result* out, double* chijj,
                                int ldc)
{
    if (!pts || !out || npts < 0 || workers < 0 || workers > 64 || (chijj && ldc <= 0)) return TPSC_E_INVALID;
    if (npts == 0) return TPSC_OK;
    if (workers == 0) {
        int rc = tpsc_batch_tune(pts, npts, &workers, nullptr);
        if (rc) return rc;
    }
    run(pts, by_cost(pts, npts), std::min(workers, npts), out, chijj, ldc);
    return TPSC_OK;
}

extern "C" int tpsc_batch_release(void)
{
    std::lock_guard<std::mutex> lock(g_queued_mu);
    delete g_queued;
    g_queued = nullptr;
    return TPSC_OK;
}

extern "C" int tpsc_batch_points(unsigned long long seed, int npts, tpsc_params* out)
{
    if (!out || npts < 0) return TPSC_E_INVALID;
    std::mt19937_64 gen(seed);
    std::uniform_int_distribution<int> pick(0, 4);
    std::uniform_real_distribution<double> uT(0.05, 0.5), un(0.8, 1.0), uU(1.0, 6.0), utp(-0.3, 0.0);
    const int Ns[5] = {32, 48, 64, 96, 128};
    for (int i = 0; i < npts; i++) {
        tpsc_params& p = out[i];
        std::memset(&p, 0, sizeof p);
        p.nk0 = Ns[pick(gen)];
        p.T = uT(gen);
        p.n = un(gen);
        p.U = uU(gen);
        p.tp = utp(gen);
        // smallest power-of-two Nw with (Nw - 1) pi T >= 20
        int Nw = 16;
        while ((Nw - 1) * M_PI * p.T < 20.0) Nw *= 2;
        p.Nw0 = Nw;
        p.eta = 0.1;
        p.coarse_pass = 0;
        p.precision = TPSC_PREC_FP64;
        p.t = 1.0;
    }
    return TPSC_OK;
}

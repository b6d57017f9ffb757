// Serving loop: continuous batching of SVD requests on one GPU.
//
// No counterpart in the reference (main.cpp runs one matrix, synchronously).  Requests
// (device-resident A and output buffers, owned by the caller) are queued by
// paco_server_submit from any thread.  A loop thread
//   * retires finished work (hipEventQuery on each lane), records per-request latency
//     from submit to observed completion, and frees the request's HBM footprint;
//   * admits queued requests while memory (the HBM budget for scratch) and lanes are
//     free: narrow matrices (N <= 512 and M N < 2^20) are packed, in arrival order,
//     into a batch on a free batch lane (one launch per phase over the batch); wide
//     ones go to a chip-wide lane (panel bidiagonalization + D&C entry points).
// Lanes are separate HIP streams, so batches and wide solves admitted while others
// run overlap on the device.  A request's result does not depend on what it was
// batched with: every kernel schedule is a function of the matrix shape only.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/paco_svd.h"
#include "lane.hpp"

namespace paco {

static double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    paco_svd_job job;
    uint64_t tag;
    double t_submit;
    size_t fp;
    bool wide;
    int err; // nonzero: its enqueue failed; reported with latency -1
    // a wide request whose chip-wide bidiagonalization timed out (info -2) beside the other
    // lanes' work runs once more, without kernels whose workgroups wait on each other
    bool retry = false;
};

// A finished request: tag, latency (-1: its enqueue failed) and its info word (0, or the
// LAPACK-style status of its SVD: < 0 bad input or a wait that timed out, > 0 roots at
// the iteration limit; 0 for a failed enqueue, which the latency reports)
struct Done {
    uint64_t tag;
    double ms;
    int info;
};

struct ServeLane {
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
    Lane *lane = nullptr;     // batch lane
    void *scratch = nullptr;  // wide lane
    size_t scratch_bytes = 0;
    size_t cap = 0;           // bytes of footprint this lane's buffers hold
    bool wide = false;
    bool busy = false;
    size_t fp = 0;
    std::vector<Req> reqs;
};

struct Server {
    size_t budget = 0, used = 0;
    int max_batch = 0;
    int wide_grid = 0; // tile grid of the wide lanes' chip-wide bidiagonalization
    std::vector<ServeLane> lanes;
    std::deque<Req> pending;
    std::vector<Done> finished; // not yet polled
    std::vector<double> lat_ms;
    std::mutex mu;
    std::condition_variable cv;
    std::thread th;
    bool stop = false;
    int error = 0;
    double t_first = 0.0, t_last = 0.0;
    long batches = 0, batched_jobs = 0, wide_jobs = 0, failed = 0, nonzero_info = 0, retried = 0, cancelled = 0;
    int max_inflight = 0;
    // PACO_SERVE_FAIL_EVERY=n: every n-th admission fails before anything is enqueued
    // (tests of the error path)
    long fail_every = 0, admissions = 0;

    void loop();
    bool admit_locked();
};

bool Server::admit_locked()
{
    bool any = false;
    for (auto &L : lanes) {
        if (L.busy || pending.empty()) continue;
        std::vector<Req> take;
        size_t fp = 0;
        if (L.wide) {
            for (auto it = pending.begin(); it != pending.end(); ++it) {
                if (!it->wide) continue;
                if (used + it->fp > budget || it->fp > L.cap) break;
                take.push_back(*it);
                fp = it->fp;
                pending.erase(it);
                break;
            }
        } else {
            for (auto it = pending.begin(); it != pending.end() && (int)take.size() < max_batch;) {
                if (it->wide) {
                    ++it;
                    continue;
                }
                if (used + fp + it->fp > budget || fp + it->fp > L.cap) break;
                fp += it->fp;
                take.push_back(*it);
                it = pending.erase(it);
            }
        }
        if (take.empty()) continue;
        int rc = 0;
        ++admissions;
        if (fail_every > 0 && admissions % fail_every == 0) {
            rc = PACO_EINJECTED;
        } else if (L.wide) {
            rc = wide_enqueue(take[0].job, L.scratch, L.s, wide_grid, take[0].retry);
            ++wide_jobs;
        } else {
            std::vector<paco_svd_job> jobs;
            for (auto &r : take) jobs.push_back(r.job);
            rc = lane_enqueue(L.lane, jobs.data(), (int)jobs.size(), L.s);
            ++batches;
            batched_jobs += (long)jobs.size();
        }
        if (rc != 0) {
            // the requests are reported as failed once the lane's stream has passed
            // whatever part of them was enqueued (their buffers stay in use until then)
            if (error == 0) error = rc;
            for (auto &r : take) r.err = rc;
        }
        const hipError_t er = hipEventRecord(L.done, L.s);
        if (er != hipSuccess) {
            if (error == 0) error = (int)er;
            (void)hipStreamSynchronize(L.s);
            const double t = now_s();
            for (auto &r : take) {
                finished.push_back({r.tag, -1.0, 0});
                ++failed;
            }
            t_last = t;
            cv.notify_all();
            continue;
        }
        L.busy = true;
        L.fp = fp;
        L.reqs.swap(take);
        used += fp;
        any = true;
    }
    int inflight = 0;
    for (auto &L : lanes) inflight += L.busy;
    max_inflight = std::max(max_inflight, inflight);
    return any;
}

void Server::loop()
{
    for (;;) {
        bool progress = false;
        // retire finished lanes (outside the lock: hipEventQuery, then the info words of
        // their requests; a busy lane's requests are touched by nothing else)
        std::vector<int> done_idx;
        std::vector<std::vector<int>> infos;
        for (size_t i = 0; i < lanes.size(); ++i) {
            ServeLane &L = lanes[i];
            if (!L.busy || hipEventQuery(L.done) != hipSuccess) continue;
            done_idx.push_back((int)i);
            std::vector<int> inf(L.reqs.size(), 0);
            bool enq_ok = true;
            for (auto &r : L.reqs) enq_ok &= r.err == 0;
            if (enq_ok) {
                const int rc = L.wide ? wide_info(L.reqs[0].job, L.scratch, inf.data())
                                      : lane_info(L.lane, inf.data(), (int)inf.size());
                if (rc != 0) std::fill(inf.begin(), inf.end(), PACO_EINVAL);
            }
            infos.push_back(std::move(inf));
        }
        std::unique_lock<std::mutex> lk(mu);
        if (!done_idx.empty()) {
            const double t = now_s();
            for (size_t d = 0; d < done_idx.size(); ++d) {
                ServeLane &L = lanes[done_idx[d]];
                for (size_t q = 0; q < L.reqs.size(); ++q) {
                    Req &r = L.reqs[q];
                    if (r.err) {
                        finished.push_back({r.tag, -1.0, 0});
                        ++failed;
                        continue;
                    }
                    if (L.wide && infos[d][q] == -2 && !r.retry) {
                        r.retry = true;
                        pending.push_front(r);
                        ++retried;
                        continue;
                    }
                    const double ms = 1000.0 * (t - r.t_submit);
                    lat_ms.push_back(ms);
                    finished.push_back({r.tag, ms, infos[d][q]});
                    nonzero_info += infos[d][q] != 0;
                }
                used -= L.fp;
                L.reqs.clear();
                L.busy = false;
                t_last = t;
            }
            progress = true;
            cv.notify_all();
        }
        if (admit_locked()) progress = true;
        bool busy = false;
        for (auto &L : lanes) busy |= L.busy;
        if (stop && pending.empty() && !busy) break;
        if (!progress) {
            if (busy)
                cv.wait_for(lk, std::chrono::microseconds(20));
            else
                cv.wait_for(lk, std::chrono::milliseconds(2));
        }
    }
}

} // namespace paco

using namespace paco;

extern "C" {

void *paco_server_create(size_t hbm_budget_bytes, int max_batch, int batch_lanes, int wide_lanes, int max_M,
                         int max_N)
{
    if (batch_lanes < 1 || max_batch < 1 || wide_lanes < 0) return nullptr;
    Server *S = new Server();
    S->budget = hbm_budget_bytes;
    S->max_batch = max_batch;
    // wide solves are bound by their per-column latency, not by bandwidth: several lanes
    // run at once, each on a share of the chip, round(16 / sqrt(lanes)) tiles per side (16
    // alone, 11 for two lanes, 8 for four, 6 for eight); a server's results depend on its
    // configuration only
    S->wide_grid = wide_lanes <= 1 ? 16 : std::max(4, (int)(16.0 / std::sqrt((double)wide_lanes) + 0.5));
    if (const char *fe = getenv("PACO_SERVE_FAIL_EVERY")) S->fail_every = atol(fe);
    const size_t cap = hbm_budget_bytes / (size_t)batch_lanes;
    for (int i = 0; i < batch_lanes + wide_lanes; ++i) {
        ServeLane L;
        L.wide = i >= batch_lanes;
        bool ok = hipStreamCreateWithFlags(&L.s, hipStreamNonBlocking) == hipSuccess &&
                  hipEventCreateWithFlags(&L.done, hipEventDisableTiming) == hipSuccess;
        if (ok && !L.wide) {
            L.cap = cap;
            L.lane = lane_create(cap, max_batch);
            ok = L.lane != nullptr;
        } else if (ok) {
            L.scratch_bytes = wide_scratch_bytes(max_M, max_N);
            L.cap = L.scratch_bytes;
            ok = hipMalloc(&L.scratch, L.scratch_bytes) == hipSuccess;
        }
        S->lanes.push_back(L);
        if (!ok) {
            paco_server_destroy(S);
            return nullptr;
        }
    }
    S->th = std::thread([S] { S->loop(); });
    return S;
}

int paco_server_submit(void *srv, const paco_svd_job *job, uint64_t tag)
{
    Server *S = static_cast<Server *>(srv);
    if (!S || !job || !job->A || !job->S || !job->U || !job->Vt || !job_supported(job->M, job->N))
        return PACO_EINVAL;
    Req r;
    r.err = 0;
    r.job = *job;
    r.tag = tag;
    r.t_submit = now_s();
    r.wide = job_is_wide(job->M, job->N);
    r.fp = r.wide ? wide_scratch_bytes(job->M, job->N) : job_footprint(job->M, job->N);
    size_t lane_cap = 0;
    for (auto &L : S->lanes)
        if (L.wide == r.wide) lane_cap = std::max(lane_cap, L.cap);
    if (r.fp > S->budget || r.fp > lane_cap) return PACO_EINVAL; // could never be admitted
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->t_first == 0.0) S->t_first = r.t_submit;
        S->pending.push_back(r);
    }
    S->cv.notify_all();
    return 0;
}

int paco_server_cancel(void *srv, uint64_t tag)
{
    Server *S = static_cast<Server *>(srv);
    if (!S) return PACO_EINVAL;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->pending.begin();
        while (it != S->pending.end() && (it->tag != tag || it->retry)) ++it;
        if (it == S->pending.end()) return 1; // admitted, finished or unknown
        S->pending.erase(it);
        ++S->cancelled;
    }
    S->cv.notify_all(); // drain may be waiting for the queue to empty
    return 0;
}

int paco_server_poll_info(void *srv, uint64_t *tags, double *latency_ms, int *info, int max)
{
    Server *S = static_cast<Server *>(srv);
    if (!S) return PACO_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    const int n = std::min<int>(max, (int)S->finished.size());
    for (int i = 0; i < n; ++i) {
        if (tags) tags[i] = S->finished[i].tag;
        if (latency_ms) latency_ms[i] = S->finished[i].ms;
        if (info) info[i] = S->finished[i].info;
    }
    S->finished.erase(S->finished.begin(), S->finished.begin() + n);
    return n;
}

int paco_server_poll(void *srv, uint64_t *tags, double *latency_ms, int max)
{
    return paco_server_poll_info(srv, tags, latency_ms, nullptr, max);
}

int paco_server_drain(void *srv)
{
    Server *S = static_cast<Server *>(srv);
    if (!S) return PACO_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv.wait(lk, [S] {
        bool busy = false;
        for (auto &L : S->lanes) busy |= L.busy;
        return S->pending.empty() && !busy;
    });
    return S->error;
}

int paco_server_stats(void *srv, double *out, int max)
{
    Server *S = static_cast<Server *>(srv);
    if (!S || !out) return PACO_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<double> v = S->lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
        if (v.empty()) return 0.0;
        const size_t k = std::min(v.size() - 1, (size_t)(p * (v.size() - 1) + 0.5));
        return v[k];
    };
    double mean = 0.0;
    for (double x : v) mean += x;
    if (!v.empty()) mean /= (double)v.size();
    const double span = S->t_last > S->t_first ? S->t_last - S->t_first : 0.0;
    const double vals[13] = {(double)v.size(),
                            pct(0.50),
                            pct(0.99),
                            mean,
                            span > 0 ? (double)v.size() / span : 0.0,
                            (double)S->batches,
                            S->batches ? (double)S->batched_jobs / (double)S->batches : 0.0,
                            (double)S->max_inflight,
                            (double)S->wide_jobs,
                            (double)S->failed,
                            (double)S->nonzero_info,
                            (double)S->retried,
                            (double)S->cancelled};
    const int n = std::min(max, 13);
    for (int i = 0; i < n; ++i) out[i] = vals[i];
    return n;
}

void paco_server_destroy(void *srv)
{
    Server *S = static_cast<Server *>(srv);
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->th.joinable()) S->th.join();
    for (auto &L : S->lanes) {
        if (L.s) (void)hipStreamSynchronize(L.s);
        if (L.lane) lane_destroy(L.lane);
        if (L.scratch) (void)hipFree(L.scratch);
        if (L.done) (void)hipEventDestroy(L.done);
        if (L.s) (void)hipStreamDestroy(L.s);
    }
    delete S;
}

} // extern "C"

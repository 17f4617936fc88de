// pann_server.cpp -- continuous-batching serving loop over a resident index.
//
// Submitters enqueue requests (host copies of their query rows and parameters).  One
// dispatcher thread owns the device work: it keeps `streams` lanes, each a batch of the
// persistent form with pinned host staging.  Whenever a lane is idle it packs pending
// queries (oldest request first, splitting requests across micro-batches as needed) up
// to max_batch (by default the wavefronts the LDS budget keeps resident) and to the
// device slot budget (HBM for query rows and results), and issues H2D copy, persistent launch and
// D2H copy on the lane's stream, followed by an event.  Finished lanes hand their results
// back to the requests.  Because each lane runs on its own stream, a micro-batch started
// for newly arrived requests runs on the device alongside the ones already there.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "pann_objects.h"

using namespace pann;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t id = 0;
    uint32_t nq = 0;
    std::vector<int8_t> rows;
    std::vector<pann_query_params> params;
    std::vector<uint32_t> ids;
    std::vector<int32_t> dists, flags;
    uint32_t next = 0, done = 0;
    bool complete = false;
    Clock::time_point arrival, completion;
};

struct Piece {
    std::shared_ptr<Request> req;
    uint32_t first, count, offset;
};

struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    pann_batch* batch = nullptr;
    int8_t* hRows = nullptr;
    pann_query_params* hParams = nullptr;
    uint32_t* hIds = nullptr;
    int32_t* hDists = nullptr;
    int32_t* hFlags = nullptr;
    std::vector<Piece> pieces;
    uint32_t nq = 0;
    bool busy = false;
    bool failed = false;
};

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const double pos = p * (v.size() - 1);
    const size_t lo = (size_t)std::floor(pos), hi = std::min(v.size() - 1, lo + 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - lo);
}

}  // namespace

struct pann_server {
    pann_server_config cfg{};
    uint32_t dim = 0;
    int device = 0;
    std::vector<Lane> lanes;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<std::shared_ptr<Request>> pending;
    std::map<uint64_t, std::shared_ptr<Request>> live;
    uint64_t nextId = 1;
    uint32_t inFlight = 0;
    std::atomic<bool> stop{false};
    std::thread worker;
    std::vector<double> latencies;
    uint64_t queries = 0, launches = 0, requests = 0;
    bool anySubmitted = false;
    Clock::time_point firstSubmit, lastDone;
    std::string workerError;

    void loop();
    bool completeLanes();
    bool dispatchLanes();
};

bool pann_server::completeLanes() {
    bool progressed = false;
    for (Lane& ln : lanes) {
        if (!ln.busy) continue;
        const hipError_t q = hipEventQuery(ln.done);
        if (q == hipErrorNotReady) continue;
        std::lock_guard<std::mutex> lk(mu);
        if (q != hipSuccess) workerError = hipGetErrorString(q);
        const uint32_t K = (uint32_t)cfg.max_k;
        const Clock::time_point now = Clock::now();
        for (const Piece& p : ln.pieces) {
            Request& r = *p.req;
            if (ln.failed) {
                std::fill(r.flags.begin() + p.first, r.flags.begin() + p.first + p.count, PANN_FLAG_CONFIG);
            } else {
                memcpy(r.ids.data() + (size_t)p.first * K, ln.hIds + (size_t)p.offset * K, 4ull * p.count * K);
                memcpy(r.dists.data() + (size_t)p.first * K, ln.hDists + (size_t)p.offset * K, 4ull * p.count * K);
                memcpy(r.flags.data() + p.first, ln.hFlags + p.offset, 4ull * p.count);
            }
            r.done += p.count;
            if (r.done == r.nq) {
                r.complete = true;
                r.completion = now;
                latencies.push_back(std::chrono::duration<double, std::milli>(now - r.arrival).count());
                requests++;
                lastDone = now;
            }
        }
        queries += ln.nq;
        inFlight -= ln.nq;
        ln.pieces.clear();
        ln.nq = 0;
        ln.busy = false;
        ln.failed = false;
        progressed = true;
        cvDone.notify_all();
    }
    return progressed;
}

bool pann_server::dispatchLanes() {
    bool progressed = false;
    for (Lane& ln : lanes) {
        if (ln.busy) continue;
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!pending.empty() && ln.nq < cfg.max_batch && inFlight < cfg.slots) {
                std::shared_ptr<Request> r = pending.front();
                const uint32_t take = std::min({r->nq - r->next, cfg.max_batch - ln.nq, cfg.slots - inFlight});
                if (take == 0) break;
                memcpy(ln.hRows + (size_t)ln.nq * dim, r->rows.data() + (size_t)r->next * dim, (size_t)take * dim);
                memcpy(ln.hParams + ln.nq, r->params.data() + r->next, take * sizeof(pann_query_params));
                ln.pieces.push_back({r, r->next, take, ln.nq});
                r->next += take;
                ln.nq += take;
                inFlight += take;
                if (r->next == r->nq) pending.pop_front();
            }
        }
        if (ln.nq == 0) continue;
        int rc = pann_batch_set_queries(ln.batch, ln.hRows, ln.hParams, ln.nq, ln.stream);
        if (rc == PANN_OK) rc = pann_batch_launch(ln.batch, ln.stream);
        if (rc == PANN_OK)
            rc = pann_batch_download(ln.batch, ln.hIds, ln.hDists, nullptr, ln.hFlags, nullptr, nullptr, ln.stream);
        if (rc != PANN_OK) {
            std::lock_guard<std::mutex> lk(mu);
            workerError = pann_last_error();
            ln.failed = true;
        }
        // the event also marks failed lanes, whose requests then complete with PANN_FLAG_CONFIG
        if (hipEventRecord(ln.done, ln.stream) != hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            workerError = "event record failed";
            ln.failed = true;
            hipStreamSynchronize(ln.stream);
        }
        ln.busy = true;
        launches++;
        progressed = true;
    }
    return progressed;
}

void pann_server::loop() {
    hipSetDevice(device);
    while (!stop.load()) {
        const bool a = completeLanes();
        const bool b = dispatchLanes();
        if (a || b) continue;
        bool anyBusy = false;
        for (const Lane& ln : lanes) anyBusy = anyBusy || ln.busy;
        std::unique_lock<std::mutex> lk(mu);
        if (anyBusy || !pending.empty()) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        } else {
            cvWork.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
    // drain what is still on the device before the lanes are torn down
    for (Lane& ln : lanes)
        if (ln.busy) hipEventSynchronize(ln.done);
    completeLanes();
}

extern "C" {

int pann_server_create(const pann_vectors* base, const pann_graph* g, const pann_hnsw* h,
                       const pann_server_config* cfg, pann_server** out) {
    if (!out) return setError(PANN_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!cfg || (!g && !h) || (g && !base)) return setError(PANN_ERR_ARG, "invalid arguments");
    if (cfg->slots == 0 || cfg->streams < 1 || cfg->streams > 8)
        return setError(PANN_ERR_ARG, "slots > 0 and 1 <= streams <= 8 required");
    auto s = std::make_unique<pann_server>();
    s->cfg = *cfg;
    // max_batch 0: the LDS budget decides (below); lanes are sized for all the slots
    s->cfg.max_batch = cfg->max_batch ? std::min(cfg->max_batch, cfg->slots) : cfg->slots;
    s->dim = h ? h->dim : base->dim;
    HIP_TRY(hipGetDevice(&s->device));
    const pann_batch_config bc{cfg->max_Ls, cfg->max_k, cfg->waves_per_block, cfg->lds_filter_bits,
                               cfg->visited_slots, 0};
    s->lanes.resize(cfg->streams);
    auto fail = [&](int rc) {
        for (Lane& ln : s->lanes) {
            pann_batch_free(ln.batch);
            if (ln.stream) hipStreamDestroy(ln.stream);
            if (ln.done) hipEventDestroy(ln.done);
            hipHostFree(ln.hRows);
            hipHostFree(ln.hParams);
            hipHostFree(ln.hIds);
            hipHostFree(ln.hDists);
            hipHostFree(ln.hFlags);
        }
        return rc;
    };
    const size_t mb = s->cfg.max_batch, K = (size_t)cfg->max_k;
    for (Lane& ln : s->lanes) {
        int rc = h ? pann_batch_prepare_hnsw(h, (uint32_t)mb, &bc, &ln.batch)
                   : pann_batch_prepare_flat(base, g, (uint32_t)mb, &bc, &ln.batch);
        if (rc != PANN_OK) return fail(rc);
        if (hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&ln.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void**)&ln.hRows, mb * s->dim, 0) != hipSuccess ||
            hipHostMalloc((void**)&ln.hParams, mb * sizeof(pann_query_params), 0) != hipSuccess ||
            hipHostMalloc((void**)&ln.hIds, mb * K * 4, 0) != hipSuccess ||
            hipHostMalloc((void**)&ln.hDists, mb * K * 4, 0) != hipSuccess ||
            hipHostMalloc((void**)&ln.hFlags, mb * 4, 0) != hipSuccess)
            return fail(setError(PANN_ERR_HIP, "allocating serving lanes"));
    }
    if (cfg->max_batch == 0) {
        // Admission by the LDS budget: a micro-batch takes at most the wavefronts that the
        // CUs' LDS (and registers) keep resident at once for the widest query class, so
        // every query of it starts when the launch does and it takes as long as its
        // slowest query, not several rounds of them.
        uint32_t resident = 0;
        const int rc = pann_batch_info(s->lanes[0].batch, nullptr, &resident);
        if (rc != PANN_OK) return fail(rc);
        s->cfg.max_batch = std::max<uint32_t>(1, std::min(s->cfg.max_batch, resident));
    }
    pann_server* raw = s.release();
    raw->worker = std::thread([raw] { raw->loop(); });
    *out = raw;
    return PANN_OK;
}

int pann_server_submit(pann_server* s, const int8_t* rows, const pann_query_params* params,
                       uint32_t nq, uint64_t* request_id) {
    if (!s || !request_id || nq == 0 || !rows || !params) return setError(PANN_ERR_ARG, "invalid arguments");
    for (uint32_t i = 0; i < nq; i++)
        if (params[i].Ls < 1 || params[i].Ls > s->cfg.max_Ls || params[i].k < 0 ||
            params[i].k > s->cfg.max_k || !(params[i].cut > 0))
            return setError(PANN_ERR_ARG, "query %u: parameters outside the server's limits", i);
    auto r = std::make_shared<Request>();
    r->nq = nq;
    r->rows.assign(rows, rows + (size_t)nq * s->dim);
    r->params.assign(params, params + nq);
    r->ids.assign((size_t)nq * s->cfg.max_k, PANN_NO_NEIGHBOR);
    r->dists.assign((size_t)nq * s->cfg.max_k, 0x7FFFFFFF);
    r->flags.assign(nq, 0);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->nextId++;
        r->arrival = Clock::now();
        if (!s->anySubmitted) {
            s->anySubmitted = true;
            s->firstSubmit = r->arrival;
        }
        s->pending.push_back(r);
        s->live[r->id] = r;
    }
    s->cvWork.notify_one();
    *request_id = r->id;
    return PANN_OK;
}

int pann_server_wait(pann_server* s, uint64_t request_id, uint32_t* ids, int32_t* dists,
                     int32_t* flags, double* latency_ms) {
    if (!s) return setError(PANN_ERR_ARG, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(request_id);
    if (it == s->live.end()) return setError(PANN_ERR_ARG, "unknown request %llu", (unsigned long long)request_id);
    std::shared_ptr<Request> r = it->second;
    s->cvDone.wait(lk, [&] { return r->complete; });
    s->live.erase(request_id);
    lk.unlock();
    if (ids) memcpy(ids, r->ids.data(), r->ids.size() * 4);
    if (dists) memcpy(dists, r->dists.data(), r->dists.size() * 4);
    if (flags) memcpy(flags, r->flags.data(), r->flags.size() * 4);
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->completion - r->arrival).count();
    return PANN_OK;
}

int pann_server_stats_get(pann_server* s, pann_server_stats* out) {
    if (!s || !out) return setError(PANN_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->requests = s->requests;
    out->queries = s->queries;
    out->launches = s->launches;
    out->p50_ms = pct(s->latencies, 0.50);
    out->p99_ms = pct(s->latencies, 0.99);
    double sum = 0, mx = 0;
    for (double v : s->latencies) {
        sum += v;
        mx = std::max(mx, v);
    }
    out->mean_ms = s->latencies.empty() ? 0 : sum / s->latencies.size();
    out->max_ms = mx;
    out->busy_s = s->anySubmitted && s->requests
                      ? std::chrono::duration<double>(s->lastDone - s->firstSubmit).count()
                      : 0;
    if (!s->workerError.empty()) return setError(PANN_ERR_HIP, "serving loop: %s", s->workerError.c_str());
    return PANN_OK;
}

void pann_server_free(pann_server* s) {
    if (!s) return;
    s->stop = true;
    s->cvWork.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Lane& ln : s->lanes) {
        pann_batch_free(ln.batch);
        if (ln.stream) hipStreamDestroy(ln.stream);
        if (ln.done) hipEventDestroy(ln.done);
        hipHostFree(ln.hRows);
        hipHostFree(ln.hParams);
        hipHostFree(ln.hIds);
        hipHostFree(ln.hDists);
        hipHostFree(ln.hFlags);
    }
    delete s;
}

int pann_server_run_poisson(pann_server* s, const int8_t* rows, uint32_t nrows, double rate_qps,
                            double duration_s, uint32_t max_request, int32_t Ls_lo, int32_t Ls_hi,
                            const int32_t* ks, uint32_t nks, uint64_t seed,
                            pann_server_stats* stats, uint32_t* record_ids, int32_t* record_dists,
                            int32_t* record_flags, uint32_t* record_rows,
                            pann_query_params* record_params, uint32_t record_cap,
                            uint32_t* recorded) {
    if (!s || !rows || nrows == 0 || !(rate_qps > 0) || !(duration_s > 0) || max_request == 0 ||
        Ls_lo < 1 || Ls_hi < Ls_lo || !ks || nks == 0 || !stats)
        return setError(PANN_ERR_ARG, "invalid load generator arguments");
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    const double logMax = std::log((double)max_request + 1.0);
    // mean of floor(exp(U[0, ln(max+1))))
    double meanSize = 0;
    for (uint32_t v = 1; v <= max_request; v++)
        meanSize += v * (std::log((double)v + 1) - std::log((double)v)) / logMax;
    const double reqRate = rate_qps / std::max(1.0, meanSize);
    struct Sent {
        uint64_t id;
        std::vector<uint32_t> rowIdx;
        std::vector<pann_query_params> params;
    };
    std::vector<Sent> sent;
    std::mutex sentMu;
    std::condition_variable sentCv;
    bool genDone = false;
    std::vector<double> lat;
    uint64_t servedQueries = 0;
    uint32_t rec = 0;
    int err = PANN_OK;
    const uint32_t K = (uint32_t)s->cfg.max_k;

    std::thread waiter([&] {
        size_t next = 0;
        std::vector<uint32_t> ids;
        std::vector<int32_t> dists, flags;
        while (true) {
            Sent cur;
            {
                std::unique_lock<std::mutex> lk(sentMu);
                sentCv.wait(lk, [&] { return next < sent.size() || genDone; });
                if (next >= sent.size()) break;
                cur = sent[next++];
            }
            ids.resize(cur.rowIdx.size() * (size_t)K);
            dists.resize(cur.rowIdx.size() * (size_t)K);
            flags.resize(cur.rowIdx.size());
            double ms = 0;
            if (pann_server_wait(s, cur.id, ids.data(), dists.data(), flags.data(), &ms) != PANN_OK) {
                err = PANN_ERR_HIP;
                continue;
            }
            lat.push_back(ms);
            servedQueries += cur.rowIdx.size();
            // the first min(k, 10) columns of each served query, in service order
            for (size_t i = 0; i < cur.rowIdx.size() && rec < record_cap; i++, rec++) {
                const uint32_t kk = (uint32_t)std::min<int32_t>(cur.params[i].k, 10);
                for (uint32_t j = 0; j < 10; j++) {
                    if (record_ids) record_ids[(size_t)rec * 10 + j] = j < kk ? ids[i * K + j] : PANN_NO_NEIGHBOR;
                    if (record_dists) record_dists[(size_t)rec * 10 + j] = j < kk ? dists[i * K + j] : 0x7FFFFFFF;
                }
                if (record_flags) record_flags[rec] = flags[i];
                if (record_rows) record_rows[rec] = cur.rowIdx[i];
                if (record_params) record_params[rec] = cur.params[i];
            }
        }
    });

    const Clock::time_point start = Clock::now();
    double t = 0;
    std::vector<int8_t> buf;
    std::vector<pann_query_params> par;
    while (true) {
        t += -std::log(1.0 - uni(rng)) / reqRate;
        if (t >= duration_s) break;
        const uint32_t nq = std::min<uint32_t>(max_request, std::max<uint32_t>(1, (uint32_t)std::exp(uni(rng) * logMax)));
        Sent sd;
        buf.resize((size_t)nq * s->dim);
        par.resize(nq);
        for (uint32_t i = 0; i < nq; i++) {
            const uint32_t row = (uint32_t)(rng() % nrows);
            memcpy(buf.data() + (size_t)i * s->dim, rows + (size_t)row * s->dim, s->dim);
            const int32_t Ls = (int32_t)std::lround(std::exp(std::log((double)Ls_lo) +
                                                             uni(rng) * (std::log((double)Ls_hi) - std::log((double)Ls_lo))));
            const int32_t k = ks[rng() % nks];
            par[i] = pann_query_params{row, std::max(Ls_lo, std::min(Ls_hi, Ls)), k, 0, 0, 1.35};
            sd.rowIdx.push_back(row);
            sd.params.push_back(par[i]);
        }
        std::this_thread::sleep_until(start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t)));
        if (pann_server_submit(s, buf.data(), par.data(), nq, &sd.id) != PANN_OK) {
            err = PANN_ERR_ARG;
            break;
        }
        {
            std::lock_guard<std::mutex> lk(sentMu);
            sent.push_back(std::move(sd));
        }
        sentCv.notify_one();
    }
    {
        std::lock_guard<std::mutex> lk(sentMu);
        genDone = true;
    }
    sentCv.notify_one();
    waiter.join();
    const double wall = std::chrono::duration<double>(Clock::now() - start).count();
    stats->requests = lat.size();
    stats->queries = servedQueries;
    stats->p50_ms = pct(lat, 0.50);
    stats->p99_ms = pct(lat, 0.99);
    double sum = 0, mx = 0;
    for (double v : lat) {
        sum += v;
        mx = std::max(mx, v);
    }
    stats->mean_ms = lat.empty() ? 0 : sum / lat.size();
    stats->max_ms = mx;
    stats->busy_s = wall;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        stats->launches = s->launches;
    }
    if (recorded) *recorded = rec;
    if (err != PANN_OK) return setError(err, "load generator: a request failed");
    return PANN_OK;
}

}  // extern "C"

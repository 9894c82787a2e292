// Serving form: a request loop over a resident database. Queries are queued in one FIFO; each
// lane (a scratch with its own stream and worker thread) admits the queries at the head of the
// FIFO, up to max_batch, as soon as it is free and runs them as one batched scan. With several
// lanes, batches admitted while others are on the GPU run concurrently on their own streams, so
// a new request waits at most for a lane, not for the batch in flight to finish.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"
#include "hy_state.h"

namespace {

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Request {
    uint64_t ticket;
    std::vector<uint8_t> query;
    uint64_t submit_ns;
};

struct Ticket {
    bool done = false;
    hy_server_result result{};
};

}  // namespace

struct hy_server {
    hy_db* db = nullptr;                  // retained for the server's lifetime
    uint32_t max_batch = 1;
    uint64_t lane_bytes = 0;
    std::vector<hy_scratch*> lanes;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<Request> fifo;
    std::unordered_map<uint64_t, Ticket> tickets;
    uint64_t next_ticket = 1;
    bool stop = false;
    uint64_t submitted = 0, completed = 0, batches = 0, query_residues = 0;

    void work(uint32_t lane) {
        hy_scratch* s = lanes[lane];
        std::vector<Request> batch;
        std::vector<const uint8_t*> ptrs;
        std::vector<size_t> lens;
        std::vector<int32_t> scores;
        std::vector<uint64_t> idx;
        std::vector<int64_t> qe, te;
        for (;;) {
            batch.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                work_cv.wait(lk, [&] { return stop || !fifo.empty(); });
                if (fifo.empty()) return;  // stopping, and nothing left to run
                while (!fifo.empty() && batch.size() < max_batch) {
                    batch.push_back(std::move(fifo.front()));
                    fifo.pop_front();
                }
            }
            const uint64_t admit = now_ns();
            const size_t n = batch.size();
            ptrs.resize(n);
            lens.resize(n);
            scores.assign(n, 0);
            idx.assign(n, 0);
            qe.assign(n, -1);
            te.assign(n, -1);
            for (size_t k = 0; k < n; ++k) {
                ptrs[k] = batch[k].query.data();
                lens[k] = batch[k].query.size();
            }
            const int rc = hy_scan_batch(db, s, ptrs.data(), lens.data(), n, scores.data(),
                                         idx.data(), qe.data(), te.data());
            const uint64_t done = now_ns();
            {
                std::lock_guard<std::mutex> lk(mu);
                for (size_t k = 0; k < n; ++k) {
                    Ticket& t = tickets[batch[k].ticket];
                    hy_server_result& r = t.result;
                    r.score = scores[k];
                    r.status = rc;
                    r.db_index = idx[k];
                    r.query_end = qe[k];
                    r.target_end = te[k];
                    r.submit_ns = batch[k].submit_ns;
                    r.admit_ns = admit;
                    r.done_ns = done;
                    r.batch_size = (uint32_t)n;
                    r.lane = lane;
                    t.done = true;
                    query_residues += lens[k];
                }
                completed += n;
                batches += 1;
            }
            done_cv.notify_all();
        }
    }

    ~hy_server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        work_cv.notify_all();
        for (auto& w : workers)
            if (w.joinable()) w.join();
        for (hy_scratch* s : lanes) hy_scratch_free(s);
        if (db) hy_db_free(db);
    }
};

extern "C" int hy_server_new(const hy_db* db, const hy_server_config* cfg, hy_server** out) {
    if (!db || !cfg || !out) return HY_E_NULL_ARGUMENT;
    *out = nullptr;
    for (uint32_t r : cfg->reserved)
        if (r != 0) return HY_E_INVALID_ARGUMENT;
    if (cfg->max_lanes < 1 || cfg->max_lanes > 64 || cfg->max_batch < 1 ||
        cfg->max_batch > 65536)
        return HY_E_INVALID_ARGUMENT;
    hy_server* srv = new (std::nothrow) hy_server();
    if (!srv) return HY_E_OUT_OF_MEMORY;
    hy_db* mdb = const_cast<hy_db*>(db);
    hy_db_retain(mdb);
    srv->db = mdb;
    srv->max_batch = cfg->max_batch;
    auto fail = [&](int code) {
        delete srv;
        return code;
    };
    // The first lane measures what one lane holds; the budget then fixes how many lanes exist.
    uint32_t nlanes = cfg->max_lanes;
    for (uint32_t k = 0; k < nlanes; ++k) {
        hy_scratch* s = nullptr;
        int rc = hy_scratch_new(db, &s);
        if (rc != HY_OK) return fail(rc);
        srv->lanes.push_back(s);
        if ((rc = hy_scratch_reserve_batch(s, cfg->max_batch)) != HY_OK) return fail(rc);
        if (k == 0) {
            srv->lane_bytes = s->dev_bytes + s->batch_bytes;
            if (cfg->scratch_budget_bytes) {
                const uint64_t fit = cfg->scratch_budget_bytes / std::max<uint64_t>(srv->lane_bytes, 1);
                if (fit == 0) return fail(HY_E_OUT_OF_MEMORY);
                nlanes = (uint32_t)std::min<uint64_t>(nlanes, fit);
            }
        }
    }
    for (uint32_t k = 0; k < nlanes; ++k) {
        try {
            srv->workers.emplace_back([srv, k] { srv->work(k); });
        } catch (...) {
            return fail(HY_E_OUT_OF_MEMORY);
        }
    }
    *out = srv;
    return HY_OK;
}

extern "C" int hy_server_submit(hy_server* srv, const uint8_t* query, size_t qlen,
                                uint64_t* ticket) {
    if (!srv || !ticket || (qlen && !query)) return HY_E_NULL_ARGUMENT;
    const hy_db* db = srv->db;
    if (qlen > db->mql) return HY_E_QUERY_TOO_LONG;
    for (size_t k = 0; k < qlen; ++k)
        if ((uint32_t)query[k] >= db->al) return HY_E_QUERY_SYMBOL;
    Request r;
    r.query.assign(query, query + qlen);
    r.submit_ns = now_ns();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) return HY_E_INVALID_ARGUMENT;
        r.ticket = srv->next_ticket++;
        srv->tickets.emplace(r.ticket, Ticket{});
        *ticket = r.ticket;
        srv->fifo.push_back(std::move(r));
        srv->submitted += 1;
    }
    srv->work_cv.notify_one();
    return HY_OK;
}

extern "C" int hy_server_wait(hy_server* srv, uint64_t ticket, int64_t timeout_us,
                              hy_server_result* out) {
    if (!srv || !out) return HY_E_NULL_ARGUMENT;
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return HY_E_INVALID_ARGUMENT;
    // Another thread may collect the same ticket meanwhile: look it up again after each wake.
    auto ready = [&] {
        auto f = srv->tickets.find(ticket);
        return f == srv->tickets.end() || f->second.done;
    };
    if (timeout_us < 0) {
        srv->done_cv.wait(lk, ready);
    } else if (!srv->done_cv.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) {
        return HY_E_TIMEOUT;
    }
    it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) return HY_E_INVALID_ARGUMENT;
    *out = it->second.result;
    srv->tickets.erase(it);
    return HY_OK;
}

extern "C" int hy_server_stats_read(hy_server* srv, hy_server_stats* out) {
    if (!srv || !out) return HY_E_NULL_ARGUMENT;
    std::lock_guard<std::mutex> lk(srv->mu);
    out->submitted = srv->submitted;
    out->completed = srv->completed;
    out->batches = srv->batches;
    out->query_residues = srv->query_residues;
    out->lanes = (uint32_t)srv->workers.size();
    out->max_batch = srv->max_batch;
    out->lane_bytes = srv->lane_bytes;
    return HY_OK;
}

extern "C" void hy_server_free(hy_server* srv) { delete srv; }

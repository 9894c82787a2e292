// Serving form: a request loop over a resident database. Queries wait in one FIFO; a lane (a
// scratch with its own stream and worker thread) admits the queries at the head of the FIFO, up
// to max_batch, and runs them as one batched scan, so batches admitted while others are on the
// GPU run concurrently on their own streams.
//
// Admission is by device bytes. The budget (scratch_budget_bytes, 0 = none) covers the lanes'
// buffers (a lane's boundary rows and batch reservation, measured on the first lane) and the
// requests in flight, each charged its own bytes: its query copy, its key and result slots, and
// for a query beyond 512 rows its winner pass's boundary row of max_target_len + 1 entries. A
// request is admitted when its bytes fit what the budget has left; lanes are not fixed at
// construction: the server starts with one and adds one (up to max_lanes) when a request is
// waiting, every lane is busy and the budget holds another lane plus that request.
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
    uint32_t max_lanes = 1;
    uint64_t budget = 0;                  // 0 = no limit
    uint64_t lane_bytes = 0;
    std::vector<hy_scratch*> lanes;       // lanes[k] is owned by worker k once it runs
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<Request> fifo;
    std::unordered_map<uint64_t, Ticket> tickets;
    uint64_t next_ticket = 1;
    bool stop = false;
    uint32_t idle = 0;                    // lanes waiting for work
    uint64_t held_bytes = 0;              // bytes the lanes hold
    uint64_t inflight_bytes = 0, inflight_queries = 0;
    uint64_t peak_bytes = 0, peak_queries = 0;
    uint64_t submitted = 0, completed = 0, batches = 0, query_residues = 0;

    // A request's own device bytes while it runs (see the file comment).
    uint64_t request_bytes(size_t m) const {
        uint64_t b = ((uint64_t)std::max<size_t>(m, 1) + 255) / 256 * 256 +
                     sizeof(unsigned long long) + sizeof(hy::Result);
        if (m > 512) b += (db->mtl + 1) * sizeof(int2);
        return b;
    }
    bool fits(uint64_t extra) const { return budget == 0 || held_bytes + extra <= budget; }

    // Called with mu held after the FIFO or the budget changed: add a lane when a request waits,
    // no lane is idle and the budget holds a lane plus that request. The new lane's worker
    // builds its scratch itself; its bytes are charged now.
    void maybe_grow() {
        if (stop || fifo.empty() || idle > 0 || lanes.size() >= max_lanes) return;
        if (!fits(inflight_bytes + lane_bytes + request_bytes(fifo.front().query.size()))) return;
        const uint32_t k = (uint32_t)lanes.size();
        lanes.push_back(nullptr);
        held_bytes += lane_bytes;
        idle += 1;  // counted from now, so the next arrival does not start another one
        try {
            workers.emplace_back([this, k] { run_lane(k, nullptr); });
        } catch (...) {
            lanes.pop_back();
            held_bytes -= lane_bytes;
            idle -= 1;
        }
    }

    // A lane's worker: its scratch (built here unless given), then batches until stopped. A
    // lane whose scratch cannot be built returns its bytes and ends.
    void run_lane(uint32_t lane, hy_scratch* s) {
        if (!s) {
            int rc = hy_scratch_new(db, &s);
            if (rc == HY_OK && (rc = hy_scratch_reserve_batch(s, max_batch)) != HY_OK) {
                hy_scratch_free(s);
                s = nullptr;
            }
            std::lock_guard<std::mutex> lk(mu);
            if (!s) {
                held_bytes -= lane_bytes;
                idle -= 1;
                return;
            }
            lanes[lane] = s;
        }
        work(lane, s);
    }

    void work(uint32_t lane, hy_scratch* s) {
        std::vector<Request> batch;
        std::vector<const uint8_t*> ptrs;
        std::vector<size_t> lens;
        std::vector<int32_t> scores;
        std::vector<uint64_t> idx;
        std::vector<int64_t> qe, te;
        uint64_t batch_bytes = 0;
        for (;;) {
            batch.clear();
            batch_bytes = 0;
            {
                std::unique_lock<std::mutex> lk(mu);
                // The head of the FIFO is admitted once its bytes fit what the budget has left
                // (a request that no longer fits even an otherwise idle device still runs alone).
                auto head_fits = [&] {
                    return !fifo.empty() &&
                           (inflight_queries == 0 ||
                            fits(inflight_bytes + request_bytes(fifo.front().query.size())));
                };
                work_cv.wait(lk, [&] { return stop || head_fits(); });
                if (fifo.empty()) return;  // stopping, and nothing left to run
                if (!head_fits()) {        // stopping: drain regardless of the budget
                    work_cv.wait(lk, [&] { return inflight_queries == 0 || head_fits(); });
                }
                while (!fifo.empty() && batch.size() < max_batch &&
                       (batch.empty() ||
                        fits(inflight_bytes + request_bytes(fifo.front().query.size())))) {
                    const uint64_t b = request_bytes(fifo.front().query.size());
                    inflight_bytes += b;
                    batch_bytes += b;
                    inflight_queries += 1;
                    batch.push_back(std::move(fifo.front()));
                    fifo.pop_front();
                }
                peak_bytes = std::max(peak_bytes, held_bytes + inflight_bytes);
                peak_queries = std::max(peak_queries, inflight_queries);
                idle -= 1;
                maybe_grow();
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
                inflight_bytes -= batch_bytes;
                inflight_queries -= n;
                idle += 1;
            }
            done_cv.notify_all();
            work_cv.notify_all();  // bytes came back: a waiting head may fit now
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
        cfg->max_batch > 65536 || cfg->prealloc > 1)
        return HY_E_INVALID_ARGUMENT;
    hy_server* srv = new (std::nothrow) hy_server();
    if (!srv) return HY_E_OUT_OF_MEMORY;
    hy_db* mdb = const_cast<hy_db*>(db);
    hy_db_retain(mdb);
    srv->db = mdb;
    srv->max_batch = cfg->max_batch;
    srv->max_lanes = cfg->max_lanes;
    srv->budget = cfg->scratch_budget_bytes;
    auto fail = [&](int code) {
        delete srv;
        return code;
    };
    // The first lane measures what a lane holds; the budget must hold it and one request of the
    // longest query.
    hy_scratch* first = nullptr;
    int rc = hy_scratch_new(db, &first);
    if (rc != HY_OK) return fail(rc);
    if ((rc = hy_scratch_reserve_batch(first, cfg->max_batch)) != HY_OK) {
        hy_scratch_free(first);
        return fail(rc);
    }
    srv->lane_bytes = first->dev_bytes + first->batch_bytes;
    if (!srv->fits(srv->lane_bytes + srv->request_bytes(db->mql))) {
        hy_scratch_free(first);
        return fail(HY_E_OUT_OF_MEMORY);
    }
    srv->lanes.push_back(first);
    srv->held_bytes = srv->lane_bytes;
    srv->idle = 1;
    try {
        srv->workers.emplace_back([srv, first] { srv->run_lane(0, first); });
    } catch (...) {
        return fail(HY_E_OUT_OF_MEMORY);
    }
    // prealloc: every lane the budget holds, built now (all lanes fixed at construction).
    if (cfg->prealloc) {
        std::lock_guard<std::mutex> lk(srv->mu);
        while (srv->lanes.size() < srv->max_lanes &&
               srv->fits(srv->lane_bytes + srv->request_bytes(db->mql))) {
            const uint32_t k = (uint32_t)srv->lanes.size();
            srv->lanes.push_back(nullptr);
            srv->held_bytes += srv->lane_bytes;
            srv->idle += 1;
            try {
                srv->workers.emplace_back([srv, k] { srv->run_lane(k, nullptr); });
            } catch (...) {
                srv->lanes.pop_back();
                srv->held_bytes -= srv->lane_bytes;
                srv->idle -= 1;
                break;
            }
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
        srv->maybe_grow();
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
    out->lanes = (uint32_t)srv->lanes.size();
    out->max_batch = srv->max_batch;
    out->lane_bytes = srv->lane_bytes;
    out->peak_bytes = srv->peak_bytes;
    out->peak_queries = srv->peak_queries;
    out->inflight_bytes = srv->inflight_bytes;
    return HY_OK;
}

extern "C" void hy_server_free(hy_server* srv) { delete srv; }

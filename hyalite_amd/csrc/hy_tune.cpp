// Autotuner of the packed kernels' launch shapes (hy_tuning): tile rows and workgroup shape of
// the one-query kernel, tile rows and resident workgroups of the persistent batch kernel, and the
// batched form's submission size. Candidates are timed with HIP events on the scratch's own
// stream over the whole database; shapes change speed only, never results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hyalite_hip.h"
#include "hy_internal.h"
#include "hy_state.h"

namespace {

struct Sample {
    std::vector<const uint8_t*> q;
    std::vector<size_t> len;
};

// Up to k queries spread evenly over the length-sorted input (shortest and longest included).
Sample spread(const uint8_t* const* qs, const size_t* ls, size_t nq, size_t k) {
    std::vector<size_t> idx(nq);
    std::iota(idx.begin(), idx.end(), size_t{0});
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return ls[a] < ls[b]; });
    Sample out;
    const size_t take = std::min(nq, k);
    for (size_t i = 0; i < take; ++i) {
        const size_t pos = take == 1 ? nq / 2 : i * (nq - 1) / (take - 1);
        out.q.push_back(qs[idx[pos]]);
        out.len.push_back(ls[idx[pos]]);
    }
    return out;
}

// Up to 2 * npairs queries as npairs adjacent pairs of the length-sorted input, the pairs placed
// at evenly spaced fractions of the input's residues (so of its cells: long queries, which carry
// most of a batch's work, are sampled as often as they weigh). The batch kernel pairs queries of
// neighbouring lengths in its lanes, so its efficiency on a batch is that of such pairs: queries
// spread over the whole range pair a 2,000-residue query with a 1,000-residue one, and queries
// spread by count over-weight the short ones, whose remainder tiles favour smaller tiles.
Sample adjacent_pairs(const uint8_t* const* qs, const size_t* ls, size_t nq, size_t npairs) {
    std::vector<size_t> idx(nq);
    std::iota(idx.begin(), idx.end(), size_t{0});
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return ls[a] < ls[b]; });
    Sample out;
    if (nq <= 2 * npairs) {
        for (size_t i : idx) {
            out.q.push_back(qs[i]);
            out.len.push_back(ls[i]);
        }
        return out;
    }
    std::vector<double> cum(nq + 1, 0.0);
    for (size_t i = 0; i < nq; ++i) cum[i + 1] = cum[i] + (double)ls[idx[i]];
    size_t prev_end = 0;
    for (size_t p = 0; p < npairs; ++p) {
        const double at = (p + 0.5) / (double)npairs * cum[nq];
        size_t pos = (size_t)(std::lower_bound(cum.begin() + 1, cum.end(), at) - cum.begin()) - 1;
        // Distinct pairs: after the previous one, and within the input.
        pos = std::min(std::max(pos, prev_end), nq - 2 * (npairs - p));
        for (size_t i = pos; i < pos + 2; ++i) {
            out.q.push_back(qs[idx[i]]);
            out.len.push_back(ls[idx[i]]);
        }
        prev_end = pos + 2;
    }
    return out;
}

// Device time of the one-query kernels over the sample: per query the best of `reps` timed
// scans after one untimed scan, summed over the sample.
int time_plain(const hy_db* db, hy_scratch* s, const Sample& smp, int reps, double* ms) {
    double total = 0.0;
    int32_t score = 0;
    uint64_t idx = 0;
    int64_t qe = 0, te = 0;
    for (size_t k = 0; k < smp.q.size(); ++k) {
        int rc = hy_scan(db, s, smp.q[k], smp.len[k], &score, &idx, &qe, &te);
        if (rc != HY_OK) return rc;
        double best = std::numeric_limits<double>::infinity();
        for (int r = 0; r < reps; ++r) {
            hy_scratch_profile_read(s, nullptr, nullptr, nullptr, 1);
            if ((rc = hy_scan(db, s, smp.q[k], smp.len[k], &score, &idx, &qe, &te)) != HY_OK)
                return rc;
            double inter = 0.0, tail = 0.0;
            uint64_t n = 0;
            hy_scratch_profile_read(s, &inter, &tail, &n, 1);
            if (n) best = std::min(best, inter);
        }
        total += best;
    }
    *ms = total;
    return HY_OK;
}

// Persistent-kernel device time and wall time of one batched submission of the sample, after
// one untimed submission.
int time_batch(const hy_db* db, hy_scratch* s, const Sample& smp, double* kernel_ms,
               double* wall_ms) {
    const size_t nq = smp.q.size();
    std::vector<int32_t> sc(nq);
    std::vector<uint64_t> ix(nq);
    std::vector<int64_t> qe(nq), te(nq);
    int rc = hy_scan_batch(db, s, smp.q.data(), smp.len.data(), nq, sc.data(), ix.data(),
                           qe.data(), te.data());
    if (rc != HY_OK) return rc;
    hy_scratch_batch_profile_read(s, nullptr, nullptr, 1);
    const auto t0 = std::chrono::steady_clock::now();
    rc = hy_scan_batch(db, s, smp.q.data(), smp.len.data(), nq, sc.data(), ix.data(), qe.data(),
                       te.data());
    const auto t1 = std::chrono::steady_clock::now();
    if (rc != HY_OK) return rc;
    uint64_t n = 0;
    hy_scratch_batch_profile_read(s, kernel_ms, &n, 1);
    *wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return n ? HY_OK : HY_E_INTERNAL;
}

std::string cache_key(const hy_db* db, const size_t* ls, size_t nq, uint32_t forms) {
    hipDeviceProp_t prop{};
    (void)hipGetDeviceProperties(&prop, db->device);
    std::vector<size_t> sorted(ls, ls + nq);
    std::sort(sorted.begin(), sorted.end());
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the sorted query lengths
    for (size_t v : sorted) {
        for (int b = 0; b < 8; ++b) {
            h ^= (v >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    }
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s/cu%d/m%u/a%u/n%llu/r%llu/t%llu/q%llu/f%u/L%016llx",
                  prop.gcnArchName, prop.multiProcessorCount, db->mode, db->al_eff,
                  (unsigned long long)db->n, (unsigned long long)db->total_residues,
                  (unsigned long long)db->mtl, (unsigned long long)db->mql, forms,
                  (unsigned long long)h);
    for (char* c = buf; *c; ++c)
        if (*c == ' ' || *c == ':') *c = '_';
    return buf;
}

bool cache_lookup(const char* path, const std::string& key, hy_tuning* t) {
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char k[512];
    hy_tuning v{};
    bool found = false;
    while (std::fscanf(f, "%511s %d %d %d %d %d %d", k, &v.tile_rows, &v.waves,
                       &v.batch_tile_rows, &v.batch_wgs_per_cu, &v.batch_chunk,
                       &v.batch_pairs) == 7) {
        if (key == k) {
            *t = v;
            found = true;  // the last entry for a key wins
        }
    }
    std::fclose(f);
    return found;
}

void cache_store(const char* path, const std::string& key, const hy_tuning& t) {
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d %d %d %d\n", key.c_str(), t.tile_rows, t.waves,
                 t.batch_tile_rows, t.batch_wgs_per_cu, t.batch_chunk, t.batch_pairs);
    std::fclose(f);
}

}  // namespace

extern "C" int hy_scratch_autotune(const hy_db* db, hy_scratch* s, const uint8_t* const* queries,
                                   const size_t* qlens, size_t nq, uint32_t forms,
                                   hy_tuning* out) {
    if (!db || !s || (nq && (!queries || !qlens))) return HY_E_NULL_ARGUMENT;
    if (s->db != db || (forms & ~3u) != 0) return HY_E_INVALID_ARGUMENT;
    for (size_t k = 0; k < nq; ++k) {
        if (qlens[k] > db->mql) return HY_E_QUERY_TOO_LONG;
        if (qlens[k] && !queries[k]) return HY_E_NULL_ARGUMENT;
        for (size_t x = 0; x < qlens[k]; ++x)
            if ((uint32_t)queries[k][x] >= db->al) return HY_E_QUERY_SYMBOL;
    }
    hy_tuning best{};
    int rc = hy_scratch_get_tuning(s, &best);
    if (rc != HY_OK) return rc;
    // Nothing runs in the packed kernels without packed targets.
    if (db->groups16.empty()) forms = 0;
    size_t longest = 0;
    for (size_t k = 0; k < nq; ++k) longest = std::max(longest, qlens[k]);
    if (nq == 0 || longest == 0 || forms == 0) {
        if (out) *out = best;
        return HY_OK;
    }
    const char* cache = std::getenv("HYALITE_HIP_TUNE_CACHE");
    const std::string key = cache_key(db, qlens, nq, forms);
    hy_tuning cached{};
    if (cache && *cache && cache_lookup(cache, key, &cached) &&
        hy_scratch_set_tuning(s, &cached) == HY_OK) {
        if (out) *out = cached;
        return HY_OK;
    }

    const bool was_profiling = s->profile;
    if ((rc = hy_scratch_profile(s, 1)) != HY_OK) return rc;
    auto finish = [&](int code) {
        hy_scratch_profile(s, was_profiling ? 1 : 0);
        return code;
    };
    static const int kRows[3] = {16, 24, 32};

    if (forms & 1u) {
        const Sample smp = spread(queries, qlens, nq, 3);
        double best_ms = std::numeric_limits<double>::infinity();
        hy_tuning pick = best;
        for (int q : kRows) {
            for (int w : {1, 4, 8}) {
                hy_tuning t = best;
                t.tile_rows = q;
                t.waves = w;
                if (hy_scratch_set_tuning(s, &t) != HY_OK) continue;  // profile does not fit
                double ms = 0.0;
                if ((rc = time_plain(db, s, smp, 2, &ms)) != HY_OK) return finish(rc);
                if (ms < best_ms) {
                    best_ms = ms;
                    pick = t;
                }
            }
        }
        best.tile_rows = pick.tile_rows;
        best.waves = pick.waves;
        if ((rc = hy_scratch_set_tuning(s, &best)) != HY_OK) return finish(rc);
    }

    if (forms & 2u) {
        const Sample smp = adjacent_pairs(queries, qlens, nq, 8);
        double best_ms = std::numeric_limits<double>::infinity();
        hy_tuning pick = best;
        std::vector<int> pair_opts{best.batch_pairs};
        if (db->mode == HY_MODE_SW) pair_opts = {1, 0};
        for (int pairs : pair_opts)
        for (int q : kRows) {
            const int occ =
                (int)(db->mode == HY_MODE_SW && pairs
                          ? hy::inter16q_batch_occupancy(q, (int)db->al_eff)
                          : hy::inter16_batch_occupancy((int)db->mode, q, (int)db->al_eff));
            std::vector<int> wpcs{occ};
            if (occ / 2 >= 1) wpcs.push_back(occ / 2);
            for (int wpc : wpcs) {
                hy_tuning t = best;
                t.batch_tile_rows = q;
                t.batch_wgs_per_cu = wpc;
                t.batch_pairs = pairs;
                t.batch_chunk = (int32_t)smp.q.size();
                if ((rc = hy_scratch_set_tuning(s, &t)) != HY_OK) return finish(rc);
                double kms = 0.0, wall = 0.0;
                if ((rc = time_batch(db, s, smp, &kms, &wall)) != HY_OK) return finish(rc);
                (void)wall;
                if (kms < best_ms) {
                    best_ms = kms;
                    pick = t;
                }
            }
        }
        // Submission size: the whole call in one persistent launch, up to 1,024 queries (and no
        // fewer than the scratch's default). Within a launch the queue keeps every workgroup
        // busy to the end, longest items first; each further submission pays the tail of its
        // longest item again (256 protein queries x 1M targets: 16-query submissions ran 13%
        // slower than one).
        uint32_t chunk = 1;
        while (chunk < nq && chunk < 1024) chunk *= 2;
        chunk = std::max<uint32_t>(chunk, (uint32_t)best.batch_chunk);
        best.batch_tile_rows = pick.batch_tile_rows;
        best.batch_wgs_per_cu = pick.batch_wgs_per_cu;
        best.batch_pairs = pick.batch_pairs;
        best.batch_chunk = (int32_t)chunk;
        if ((rc = hy_scratch_set_tuning(s, &best)) != HY_OK) return finish(rc);
    }
    if (cache && *cache) cache_store(cache, key, best);
    if (out) *out = best;
    return finish(HY_OK);
}

// server.cpp -- continuous-batching assembly service on one GPU.
//
// The reference has no serving layer; its drivers run one mesh at a time on
// the legacy stream (Acc:750-893).  Here n worker threads each own a HIP
// stream and a workspace.  A request is admitted when a worker is idle and its
// HBM footprint fits the remaining budget, so requests that arrive while
// others run start as soon as a slot and memory free up, and their kernels
// overlap on the device.  The general workers take the queue in arrival
// order, with at most one heavy request (a plain assembly of more than about
// 8 ms) on the device at a time: a second general worker takes a lighter
// request from behind a heavy head.  With two or more workers the last one
// is an express lane for small requests, which may pass a larger request at
// the head of the queue while that head has waited less than an aging
// window (so the head is never starved).  Requests that carry a mesh key can keep that mesh's pattern on
// the device, so a repeat request with a new E field runs only the
// value-only path.
//
// Building a pattern costs more than a plain assembly and a cached repeat
// saves only part of one (A-LLS itself is close to the value-only path), so
// by default a pattern is built only when the key's expected future repeats
// pay for it.  Those are read from the server's history: of the keys (per
// dimension) seen as often as this one, how many more times they came back.
// They are scaled by the share of predicted hits that earlier patterns
// delivered before they were evicted, and the costs are per-element running
// averages of each path the server measures itself.  Patterns are evicted
// lowest expected saving per byte first.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host_internal.h"

using namespace alls;
using Clock = std::chrono::steady_clock;

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    int64_t ticket = 0;
    alls_mesh mesh{};
    uint64_t key = 0;
    alls_opts opts{};
    int64_t footprint = 0;          // charged while running (output + pattern, see estimate_footprint)
    int64_t pattern_bytes = 0;      // part of footprint when this request builds a cached pattern
    int64_t scratch = 0;            // list pool + scratch, held by the worker's workspace
    Clock::time_point t_submit, t_admit, t_done;
    alls_status status = ALLS_OK;
    bool done = false;
    bool cache_hit = false;
    bool built = false;
    int batch_size = 1;
    int running_at_admit = 0;       // requests already on the device when it was admitted
    bool heavy = false;             // counted in heavy_running while it runs
    uint64_t admit_index = 0;       // the server's admission count including it
    alls_csr out{};
    std::string err;
};

struct CacheEntry {
    alls_pattern* pat = nullptr;
    int64_t bytes = 0;
    int refs = 0;
    uint64_t last_use = 0;
    int64_t hits = 0;               // value-only assemblies served
    double predicted = 0.0;         // expected repeats when it was built
    int precision = 0;
    int64_t n_elems = 0, n_nodes = 0;
    // What a value-only repeat takes from the cached pattern besides the
    // geometry (which the key promises): element kind layout and material.
    int32_t dim = 0, npe = 0, uniform_type = 0;
    bool mixed = false;
    double nu = 0.0, thickness = 0.0;

    void describe(const alls_mesh& m, const alls_opts& o) {
        precision = o.precision;
        n_elems = m.n_elems;
        n_nodes = m.n_nodes;
        dim = m.dim;
        npe = m.nodes_per_elem;
        uniform_type = m.uniform_type;
        mixed = m.elem_type != nullptr;
        nu = m.nu;
        thickness = m.thickness;
    }
    // A request may reuse this pattern: same key (checked by the caller),
    // precision, sizes, element kinds and material constants; only E differs.
    bool serves(const alls_mesh& m, const alls_opts& o) const {
        return precision == o.precision && n_elems == m.n_elems && n_nodes == m.n_nodes && dim == m.dim &&
               npe == m.nodes_per_elem && uniform_type == m.uniform_type && mixed == (m.elem_type != nullptr) &&
               nu == m.nu && (dim == 3 || thickness == m.thickness);
    }
};

// Upper-bound device bytes of one assembly: returns output (+ pattern) bytes;
// *scratch gets the list pool and workspace bytes.
int64_t estimate_footprint(const alls_mesh* m, const MeshInfo& mi, bool with_pattern, int64_t* pattern_bytes,
                           int64_t* scratch_out) {
    const int64_t B = mi.dpn * mi.dpn;
    const int64_t vsz = mi.f32 ? 4 : 8;
    const int64_t ent = mi.entries_bound;
    const int64_t C = default_chunk(mi.kind);
    const int64_t chunks = pool_chunks(m->n_nodes, ent, (int)C, mi.kind, false);
    const int64_t n_rows = m->n_nodes * mi.dpn;
    const int64_t nnz_bound = ent * B;
    int64_t scratch = chunks * C * (8 + 2 + B * vsz) + m->n_nodes * (32 + 8) + 1024;   // keys, group ids, values
    int64_t inputs = m->n_nodes * m->dim * 8 + m->n_elems * (m->nodes_per_elem * 4 + 9);
    int64_t out = (n_rows + 1) * 8 + nnz_bound * 12;
    int64_t pat = 0;
    if (with_pattern) {
        const int64_t npm = mi.kind == 3 ? 20 : 4;
        pat = inputs + (n_rows + 1) * 8 + nnz_bound * 4 + m->n_elems * npm * npm * 4 + m->n_nodes * 24 +
              m->n_elems * mi.npe_max * 4;
    }
    if (pattern_bytes) *pattern_bytes = pat;
    if (scratch_out) *scratch_out = scratch + inputs;
    return out + pat;
}

// Runs reqs (plain, one dimension) as one packed batch on w/s and gives each
// request its own CSR: the block-diagonal result is split by one kernel.
alls_status run_batch(Workspace& w, hipStream_t s, const std::vector<std::shared_ptr<Request>>& reqs,
                      alls_stats* st, std::string* err) {
    const int n = (int)reqs.size();
    std::vector<alls_mesh> ms((size_t)n);
    for (int k = 0; k < n; ++k) ms[k] = reqs[k]->mesh;
    std::vector<int64_t> ro((size_t)n + 1), no((size_t)n + 1);
    alls_csr bout{};
    // Every batch has its own shape, so no graph capture (it would not replay).
    alls_status rc = batch_assemble(w, s, n, ms.data(), reqs[0]->opts, &bout, ro.data(), no.data(), st, err, nullptr,
                                    false);
    std::vector<SplitPart> parts((size_t)n);
    for (int k = 0; k < n && rc == ALLS_OK; ++k) {
        rc = ensure_csr(&reqs[k]->out, ro[k + 1] - ro[k], no[k + 1] - no[k], 0);
        parts[k] = SplitPart{ro[k], no[k], ro[k + 1] - ro[k], no[k + 1] - no[k],
                             static_cast<int32_t*>(reqs[k]->out.row_ptr), reqs[k]->out.col_ind, reqs[k]->out.val};
    }
    if (rc == ALLS_OK) {
        const size_t tb = (size_t)n * sizeof(SplitPart);
        char* h = nullptr;
        hipError_t e = w.staging(tb, &h);
        if (e == hipSuccess) e = w.bsplit.ensure(tb);
        if (e == hipSuccess) {
            std::memcpy(h, parts.data(), tb);
            e = hipMemcpyAsync(w.bsplit.p, h, tb, hipMemcpyHostToDevice, s);
        }
        if (e == hipSuccess) {
            launch_split(w.bsplit.as<SplitPart>(), n, ro[n], no[n], bout.row_ptr, bout.row_ptr_is_i64, bout.col_ind,
                         bout.val, s);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = hip_fail(e, "batch split", err);
    }
    free_csr(&bout);
    return rc;
}

}  // namespace

struct alls_server {
    alls_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Request>> reqs;
    std::unordered_map<uint64_t, CacheEntry> cache;
    std::unordered_map<uint64_t, int> building;   // keys whose pattern is being built
    int64_t budget = 0, in_use = 0;   // in_use: running requests + unreleased outputs
    int64_t ws_total = 0;             // bytes held by the workers' grow-only workspaces
    std::vector<int64_t> ws_bytes;    // per worker, updated by that worker under the lock
    int64_t next_ticket = 1;
    uint64_t use_clock = 0;
    std::unordered_map<uint64_t, int> seen;   // admitted requests per mesh key
    // ge[d][j]: keys of dimension d (0: 2D, 1: Hex20) admitted at least j times.
    std::vector<int64_t> ge[2];
    // Hits delivered and hits predicted-but-lost by patterns already evicted.
    double kept[2] = {0.0, 0.0}, lost[2] = {0.0, 0.0};
    alls_server_counts counts{};
    // Per-element cost (ns, admission to completion, host inputs staged) of
    // the plain, build and hit paths, per dimension (0: 2D, 1: Hex20), as
    // running averages of requests that ran alone on the device: times taken
    // while other requests share it measure the contention, not the path, and
    // once made the Hex20 builds look worth it on a saturated device.  Seeded
    // with the levels measured on the served mix.
    double cost[2][3] = {{5.0, 6.5, 2.5}, {40.0, 150.0, 36.0}};
    int running = 0;                 // requests on the device
    int heavy_running = 0;           // heavy requests on the device (see pick)
    uint64_t admits = 0;             // admissions so far

    enum { PATH_PLAIN = 0, PATH_BUILD = 1, PATH_HIT = 2 };
    static int dim_of(const Request& q) { return q.mesh.dim == 3 ? 1 : 0; }

    // Further requests expected for a key admitted `k` times so far: of the
    // keys of this dimension that reached k requests, the mean number of
    // requests they had beyond k (this undercounts early in a run, when keys
    // have had little time to come back, which only makes the policy
    // cautious).  Until kHistMin keys have reached k, the key's own history
    // stands in: its k - 1 earlier requests predict as many more.
    static constexpr int64_t kHistMin = 8;
    double expected_repeats(int d, int k) const {
        if (k < 1) return 0.0;
        const std::vector<int64_t>& g = ge[d];
        if (k < (int)g.size() && g[k] >= kHistMin) {
            double more = 0.0;
            for (size_t j = (size_t)k + 1; j < g.size(); ++j) more += (double)g[j];
            return more / (double)g[k];
        }
        return (double)(k - 1);
    }
    // Share of predicted hits that evicted patterns delivered (1 until some
    // pattern has been evicted).
    double survival(int d) const { return (kept[d] + 1.0) / (kept[d] + lost[d] + 1.0); }
    // Whether the request (the key's k-th, counting itself) should build its
    // pattern; *pred receives the repeats it expects.
    bool worth_building(const Request& q, int k, double* pred) const {
        const int d = dim_of(q);
        const double e = expected_repeats(d, k);
        if (pred) *pred = e;
        if (cfg.cache_policy == ALLS_CACHE_EAGER) return true;
        const double* c = cost[d];
        return e * survival(d) * (c[PATH_PLAIN] - c[PATH_HIT]) > c[PATH_BUILD] - c[PATH_PLAIN];
    }
    // Records an admission of key `key` (dimension d); returns the key's
    // request count including this one.
    int note_seen(uint64_t key, int d) {
        const int k = ++seen[key];
        std::vector<int64_t>& g = ge[d];
        if ((int)g.size() <= k) g.resize((size_t)k + 1, 0);
        g[k]++;
        return k;
    }
    // Small plain requests queued right behind a plain head join it in one
    // packed batch (one pipeline, persistent insertion, split afterwards).
    static constexpr int kBatchMax = 256;
    static int64_t batch_elems_max(int dim) { return dim == 3 ? 65536 : 1 << 20; }
    static bool batchable(const Request& q) {
        return q.mesh.n_elems > 0 && q.mesh.n_elems <= (q.mesh.dim == 3 ? 16384 : 131072);
    }
    static bool compatible(const Request& a, const Request& b) {
        return a.mesh.dim == b.mesh.dim && a.mesh.mem == b.mesh.mem && a.mesh.nu == b.mesh.nu &&
               (a.mesh.dim == 3 || a.mesh.thickness == b.mesh.thickness) &&
               (a.mesh.dim == 2 || a.opts.precision == b.opts.precision);
    }
    // Express lane: requests whose plain assembly takes at most kSmallMs at
    // the measured cost; they may pass a head that has waited < kAgeMs.
    static constexpr double kSmallMs = 2.0;
    static constexpr double kAgeMs = 40.0;
    bool small(const Request& q) const {
        return (double)q.mesh.n_elems * cost[dim_of(q)][PATH_PLAIN] * 1e-6 <= kSmallMs;
    }
    // Heavy: a plain assembly of more than kHeavyMs at the measured cost
    // (Hex20 from about 2e5 elements; no 2D mesh of up to 1e6).  One at a
    // time on the device (see pick).
    static constexpr double kHeavyMs = 8.0;
    bool heavy(const Request& q) const {
        return (double)q.mesh.n_elems * cost[dim_of(q)][PATH_PLAIN] * 1e-6 > kHeavyMs;
    }
    // Whether queued request q would run the plain path if it were the head.
    bool runs_plain(const Request& q) const {
        if (!q.key) return true;
        auto c = cache.find(q.key);
        if (c != cache.end() && c->second.serves(q.mesh, q.opts)) return false;
        auto sc = seen.find(q.key);
        const int k = (sc == seen.end() ? 0 : sc->second) + 1;
        return building.count(q.key) || !worth_building(q, k, nullptr);
    }

    // Folds one served request into the running per-element cost.  Small
    // requests are dominated by fixed costs and are left out, and so is any
    // request that shared the device with another.
    void observe(const Request& r, int path) {
        if (r.mesh.n_elems < 50000 || r.status != ALLS_OK) return;
        if (r.running_at_admit != 0 || admits != r.admit_index || running != 1) return;
        const double ns = ms_between(r.t_admit, r.t_done) * 1e6 / (double)r.mesh.n_elems;
        double& c = cost[dim_of(r)][path];
        c = 0.8 * c + 0.2 * ns;
    }
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    std::vector<std::unique_ptr<Workspace>> ws;
    Arena* arena = nullptr;           // the budget, reserved once; workers allocate from it

    int64_t cache_bytes() const {
        int64_t b = 0;
        for (auto& kv : cache) b += kv.second.bytes;
        return b;
    }
    // Drops a cached pattern (idle), booking its delivered and lost hits.
    void drop(std::unordered_map<uint64_t, CacheEntry>::iterator it) {
        const int d = it->second.dim == 3 ? 1 : 0;
        counts.evictions[d]++;
        if (it->second.hits == 0) counts.evicted_unused[d]++;
        kept[d] += (double)it->second.hits;
        lost[d] += std::max(0.0, it->second.predicted - (double)it->second.hits);
        pattern_destroy(it->second.pat);
        cache.erase(it);
    }
    // Expected saving (ns) of keeping a pattern, per byte it holds.
    double keep_value(uint64_t key, const CacheEntry& c) const {
        const int d = c.dim == 3 ? 1 : 0;
        auto sc = seen.find(key);
        const double e = expected_repeats(d, sc == seen.end() ? 0 : sc->second);
        return e * (cost[d][PATH_PLAIN] - cost[d][PATH_HIT]) * (double)c.n_elems / (double)std::max<int64_t>(c.bytes, 1);
    }
    // Frees idle patterns, lowest expected saving per byte first (the least
    // recently used among equals), until `need` bytes fit.
    void evict_for(int64_t need) {
        while (in_use + ws_total + cache_bytes() + need > budget) {
            auto victim = cache.end();
            double vv = 0.0;
            for (auto it = cache.begin(); it != cache.end(); ++it) {
                if (it->second.refs != 0) continue;
                const double v = keep_value(it->first, it->second);
                if (victim == cache.end() || v < vv ||
                    (v == vv && it->second.last_use < victim->second.last_use)) {
                    victim = it;
                    vv = v;
                }
            }
            if (victim == cache.end()) return;
            drop(victim);
        }
    }

    // Index of the queued request worker `slot` may start, or -1.  General
    // workers take the head, unless it is heavy and another heavy request is
    // on the device (see below).  The express worker takes only small requests:
    // the head if it is small, else the first small one among the next 64
    // while the head has waited less than kAgeMs (after that it waits, so
    // that memory frees up for the head).  Letting it take large heads as
    // well measured much worse on the served mix (two large Hex20 requests at
    // once hold two list pools and push the cached patterns out).
    int pick(int slot) const {
        if (queue.empty()) return -1;
        const bool express = cfg.n_streams >= 2 && slot == cfg.n_streams - 1;
        if (!express) {
            // General workers: a heavy head waits while another heavy request
            // runs; meanwhile a lighter request behind it may go, while the
            // head has waited less than kAgeMs.  (It matters with 3 or more
            // workers, two of them general.)  On the served mix (one MI355X,
            // same arrival rate) 3 workers without this rule had half the p99
            // of 2 (33-37 against 73-80 ms) in most runs, but in 2 of 12 the
            // queue diverged (p99 0.66 s and 5.4 s with 4 workers: two large
            // Hex20 requests at once hold two list pools of tens of GB, push the
            // cached patterns out and share the device); with it, 41-57 ms in
            // all 17 runs.
            if (heavy_running == 0 || !heavy(*queue.front())) return 0;
            if (ms_between(queue.front()->t_submit, Clock::now()) >= kAgeMs) return -1;
            const int lim = (int)std::min<size_t>(queue.size(), 64);
            for (int i = 1; i < lim; ++i)
                if (!heavy(*queue[i])) return i;
            return -1;
        }
        if (small(*queue.front())) return 0;
        if (ms_between(queue.front()->t_submit, Clock::now()) >= kAgeMs) return -1;   // the head's turn
        const int lim = (int)std::min<size_t>(queue.size(), 64);
        for (int i = 1; i < lim; ++i)
            if (small(*queue[i])) return i;
        return -1;
    }

    void worker(int slot) {
        (void)hipSetDevice(cfg.device);
        arena_bind(arena);
        hipStream_t s = streams[slot];
        Workspace& w = *ws[slot];
        for (;;) {
            std::shared_ptr<Request> r;
            std::vector<std::shared_ptr<Request>> batch;   // plain requests batched behind r
            bool hit = false, build = false;
            CacheEntry* ce = nullptr;
            double predicted = 0.0;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (stop) return;
                    // Two tries: the second after this idle worker gives its
                    // workspace back, so held scratch can never starve a request.
                    for (int attempt = 0; attempt < 2 && !r; ++attempt) {
                        const int qi = pick(slot);
                        if (qi < 0) break;
                        std::shared_ptr<Request> q = queue[(size_t)qi];
                        auto c = q->key ? cache.find(q->key) : cache.end();
                        const bool can_hit = c != cache.end() && c->second.serves(q->mesh, q->opts);
                        const auto sc = q->key ? seen.find(q->key) : seen.end();
                        const int k = (sc == seen.end() ? 0 : sc->second) + 1;
                        double pred = 0.0;
                        const bool can_build = q->key && !can_hit && !building.count(q->key) &&
                                               (c == cache.end() || c->second.refs == 0) &&
                                               q->pattern_bytes <= budget / 4 && worth_building(*q, k, &pred);
                        // This worker's workspace is reused; only its growth is new memory.
                        const int64_t grow = std::max<int64_t>(0, q->scratch - (int64_t)w.device_bytes());
                        const int64_t need = (can_build ? q->footprint : q->footprint - q->pattern_bytes) + grow;
                        if (can_hit) c->second.refs++;   // pinned while we make room
                        int64_t evictable = 0;
                        for (auto& kv : cache)
                            if (kv.second.refs == 0) evictable += kv.second.bytes;
                        if (in_use + ws_total + cache_bytes() - evictable + need <= budget) {
                            evict_for(need);
                            auto stale = can_build ? cache.find(q->key) : cache.end();
                            if (stale != cache.end()) drop(stale);   // idle entry of another mesh under this key
                            r = q;
                            queue.erase(queue.begin() + qi);
                            r->running_at_admit = running++;
                            r->heavy = heavy(*r);
                            heavy_running += r->heavy ? 1 : 0;
                            r->admit_index = ++admits;
                            if (qi > 0) counts.passed++;
                            if (r->key) note_seen(r->key, dim_of(*r));
                            hit = can_hit;
                            build = can_build;
                            predicted = pred;
                            if (hit) {
                                ce = &cache[r->key];
                                ce->last_use = ++use_clock;
                            }
                            if (build) building[r->key] = 1;
                            r->footprint = need - grow;
                            in_use += need - grow;
                            ws_total += grow;   // trued up to the real size when the request ends
                            break;
                        }
                        if (can_hit) c->second.refs--;
                        if (w.device_bytes() == 0) break;
                        w.release_device();
                        ws_bytes[slot] = 0;
                        ws_total = 0;
                        for (int64_t b : ws_bytes) ws_total += b;
                    }
                    if (r) break;
                    cv_work.wait_for(lk, std::chrono::milliseconds(5));   // the express lane re-checks ages
                }
                if (!hit && !build && batchable(*r)) {
                    // Compatible small plain requests queued right behind r join it.
                    int64_t elems = r->mesh.n_elems;
                    size_t i = 0;
                    while (i < queue.size() && (int)batch.size() + 1 < kBatchMax) {
                        std::shared_ptr<Request> q = queue[i];
                        if (!batchable(*q) || !compatible(*r, *q) || !runs_plain(*q) ||
                            elems + q->mesh.n_elems > batch_elems_max(r->mesh.dim))
                            break;
                        // Its output, the block-diagonal copy and its share of the
                        // batch's scratch are charged while it runs.
                        const int64_t out_b = q->footprint - q->pattern_bytes;
                        const int64_t need = 2 * out_b + q->scratch;
                        if (in_use + ws_total + cache_bytes() + need > budget) break;
                        queue.erase(queue.begin() + (std::ptrdiff_t)i);
                        ++admits;   // batched along: r no longer ran alone
                        if (q->key) note_seen(q->key, dim_of(*q));
                        q->footprint = need;
                        in_use += need;
                        elems += q->mesh.n_elems;
                        batch.push_back(q);
                    }
                }
            }
            r->t_admit = Clock::now();
            for (auto& q : batch) q->t_admit = r->t_admit;
            alls_stats st{};
            alls_status rc;
            if (hit) {
                // New material field on the cached pattern: value-only path.
                const double* E = r->mesh.E_per_elem;
                rc = ALLS_OK;
                if (E && r->mesh.mem == ALLS_MEM_HOST) {
                    const size_t eb = (size_t)r->mesh.n_elems * sizeof(double);
                    hipError_t e = w.E.ensure(std::max<size_t>(eb, 8));
                    void* dst[1] = {w.E.p};
                    const void* src[1] = {E};
                    if (e == hipSuccess) e = w.upload(dst, src, &eb, 1, s);
                    if (e != hipSuccess) rc = hip_fail(e, "stage E", &r->err);
                    E = w.E.as<double>();
                }
                if (rc == ALLS_OK) rc = pattern_values(w, ce->pat, E, r->mesh.E, &r->out, s, &st, &r->err);
            } else if (build) {
                alls_pattern* p = nullptr;
                rc = pattern_build(w, &r->mesh, r->opts, s, cfg.device, &p, &r->out, &st, &r->err);
                std::lock_guard<std::mutex> lk(mu);
                building.erase(r->key);
                if (rc == ALLS_OK) {
                    CacheEntry e;
                    e.pat = p;
                    e.bytes = pattern_bytes(p);
                    e.last_use = ++use_clock;
                    e.predicted = predicted;
                    e.describe(r->mesh, r->opts);
                    if ((int)cache.size() >= std::max(1, cfg.pattern_cache_entries)) {
                        auto victim = cache.end();
                        for (auto it = cache.begin(); it != cache.end(); ++it)
                            if (it->second.refs == 0 &&
                                (victim == cache.end() || it->second.last_use < victim->second.last_use))
                                victim = it;
                        if (victim != cache.end()) drop(victim);
                    }
                    cache[r->key] = e;
                    counts.builds[dim_of(*r)]++;
                }
            } else if (!batch.empty()) {
                batch.insert(batch.begin(), r);
                rc = run_batch(w, s, batch, &st, &r->err);
                batch.erase(batch.begin());
            } else {
                MeshInfo mi;
                DevMesh dm;
                rc = classify_mesh(&r->mesh, r->opts, &mi, &r->err);
                if (rc == ALLS_OK) rc = stage_mesh(w, &r->mesh, mi, &dm, s, &r->err);
                if (rc == ALLS_OK) rc = run_alls(w, dm, mi, r->opts, &r->out, s, &st, &r->err);
            }
            // A workspace grown past an eighth of the budget (the list pool
            // of a large Hex20 request) is given back: held, it would push
            // cached patterns out, and the arena carves it again quickly.
            // With two general workers (3 or more in all) two such pools can
            // be held, so the limit is a sixteenth.
            if ((int64_t)w.device_bytes() > budget / (cfg.n_streams >= 3 ? 16 : 8)) w.release_device();
            {
                std::lock_guard<std::mutex> lk(mu);
                ws_bytes[slot] = (int64_t)w.device_bytes();
                ws_total = 0;
                for (int64_t b : ws_bytes) ws_total += b;
                r->t_done = Clock::now();
                r->status = rc;
                if (batch.empty()) observe(*r, hit ? PATH_HIT : build ? PATH_BUILD : PATH_PLAIN);
                --running;
                heavy_running -= r->heavy ? 1 : 0;
                r->cache_hit = hit;
                r->built = build && rc == ALLS_OK;
                r->done = true;
                if (hit) {
                    ce->refs--;
                    if (rc == ALLS_OK) {
                        ce->hits++;
                        counts.hits[dim_of(*r)]++;
                    }
                }
                counts.completed += 1 + (int64_t)batch.size();
                if (!batch.empty()) counts.batched += 1 + (int64_t)batch.size();
                // The output stays charged until the caller releases it.
                const int64_t outb = r->out.cap_nnz * 12 + r->out.cap_rows * 8;
                in_use -= r->footprint;
                r->footprint = rc == ALLS_OK ? outb : 0;
                in_use += r->footprint;
                if (!batch.empty()) r->batch_size = (int)batch.size() + 1;
                for (auto& q : batch) {
                    q->batch_size = r->batch_size;
                    q->t_done = r->t_done;
                    q->status = rc;
                    if (rc != ALLS_OK) q->err = r->err;
                    q->done = true;
                    in_use -= q->footprint;
                    q->footprint = rc == ALLS_OK ? q->out.cap_nnz * 12 + q->out.cap_rows * 8 : 0;
                    in_use += q->footprint;
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

alls_status alls_server_create(const alls_server_config* cfg, alls_server** out) {
    if (!cfg || !out) return ALLS_E_INVALID;
    *out = nullptr;
    if (alls_device_count() <= cfg->device || cfg->device < 0) return ALLS_E_NO_DEVICE;
    if (hipSetDevice(cfg->device) != hipSuccess) return ALLS_E_HIP;
    auto* s = new alls_server();
    s->cfg = *cfg;
    const int n = cfg->n_streams <= 0 ? 3 : std::min(cfg->n_streams, 16);
    s->cfg.n_streams = n;
    if (s->cfg.pattern_cache_entries <= 0) s->cfg.pattern_cache_entries = 1 << 20;   // bounded by the budget
    // Idle blocks in the process's allocation cache (buffers of destroyed
    // contexts and freed outputs, up to a quarter of HBM) go back to the
    // device first, so the default budget is taken from memory that is really
    // free rather than shrunk by blocks nothing uses.
    dev_trim();
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        delete s;
        return ALLS_E_HIP;
    }
    // Default budget: 55% of free HBM; the allocator may keep up to a quarter
    // of HBM in freed blocks for reuse, so live + cached stays below capacity.
    s->budget = cfg->hbm_budget_bytes > 0 ? cfg->hbm_budget_bytes : (int64_t)(0.55 * (double)free_b);
    // Reserve the budget up front (workspaces, outputs and patterns are carved
    // out of it); without the reservation the caching allocator serves them.
    s->arena = arena_create((size_t)std::min<int64_t>(s->budget, (int64_t)free_b));
    for (int i = 0; i < n; ++i) {
        hipStream_t st;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            alls_server_destroy(s);
            return ALLS_E_HIP;
        }
        s->streams.push_back(st);
        s->ws.emplace_back(new Workspace());
        if (s->ws.back()->init() != hipSuccess) {
            alls_server_destroy(s);
            return ALLS_E_HIP;
        }
    }
    s->ws_bytes.assign(n, 0);
    for (int i = 0; i < n; ++i) s->workers.emplace_back(&alls_server::worker, s, i);
    *out = s;
    return ALLS_OK;
}

alls_status alls_server_submit(alls_server* s, const alls_mesh* mesh, uint64_t key, const alls_opts* opts,
                               int64_t* ticket) {
    if (!s || !mesh || !ticket) return ALLS_E_INVALID;
    auto r = std::make_shared<Request>();
    r->mesh = *mesh;
    r->key = key;
    if (opts) r->opts = *opts;
    if (r->opts.val_f32) return ALLS_E_UNSUPPORTED;   // FP32 output values: alls_assemble only
    MeshInfo mi;
    std::string err;
    alls_status rc = classify_mesh(mesh, r->opts, &mi, &err);
    if (rc != ALLS_OK) return rc;
    r->footprint = estimate_footprint(mesh, mi, key != 0, &r->pattern_bytes, &r->scratch);
    if (r->footprint - r->pattern_bytes + r->scratch > s->budget) return ALLS_E_BUSY;
    if (r->footprint + r->scratch > s->budget) r->key = 0;   // fits only without caching its pattern
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->ticket = s->next_ticket++;
        s->reqs[r->ticket] = r;
        s->queue.push_back(r);
    }
    s->cv_work.notify_all();
    *ticket = r->ticket;
    return ALLS_OK;
}

alls_status alls_server_wait(alls_server* s, int64_t ticket, alls_csr* out, alls_request_stats* st) {
    if (!s) return ALLS_E_INVALID;
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return ALLS_E_INVALID;
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->reqs.erase(ticket);
    }
    if (st) {
        st->queue_ms = ms_between(r->t_submit, r->t_admit);
        st->run_ms = ms_between(r->t_admit, r->t_done);
        st->latency_ms = ms_between(r->t_submit, r->t_done);
        st->nnz = r->out.nnz;
        st->cache_hit = r->cache_hit ? 1 : r->built ? 2 : 0;
        st->status = r->status;
        st->batch_size = r->batch_size;
        st->reserved = 0;
    }
    if (out) {
        *out = r->out;
        std::memset(&r->out, 0, sizeof(r->out));
    } else {
        free_csr(&r->out);
        std::lock_guard<std::mutex> lk(s->mu);
        s->in_use -= r->footprint;
        s->cv_work.notify_all();
    }
    return r->status;
}

alls_status alls_server_counters(alls_server* s, alls_server_counts* out) {
    if (!s || !out) return ALLS_E_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->counts;
    return ALLS_OK;
}

void alls_server_release(alls_server* s, alls_csr* csr) {
    if (!s || !csr) return;
    const int64_t b = csr->cap_nnz * 12 + csr->cap_rows * 8;
    (void)hipSetDevice(s->cfg.device);
    free_csr(csr);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->in_use -= b;
        if (s->in_use < 0) s->in_use = 0;
    }
    s->cv_work.notify_all();
}

void alls_server_destroy(alls_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers)
        if (t.joinable()) t.join();
    (void)hipSetDevice(s->cfg.device);
    for (auto& kv : s->reqs) free_csr(&kv.second->out);
    for (auto& kv : s->cache) pattern_destroy(kv.second.pat);
    for (auto& w : s->ws) w->destroy();
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    arena_destroy(s->arena);
    delete s;
}

}  // extern "C"

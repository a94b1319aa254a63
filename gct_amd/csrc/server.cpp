// Serving: continuous admission of reconstruction requests onto one device.
//
// Requests of any geometry and size are queued by gct_server_submit.  Each of
// the server's worker threads owns one non-blocking HIP stream; a worker takes
// the first queued request whose device footprint fits the remaining HBM
// budget (and the in-flight limit), runs its whole pipeline on its stream --
// plan set-up, upload, fused weighting + filter, backprojection, optional
// Joseph reprojection and residual, download -- and returns the memory to the
// budget.  Requests that arrive while others are on the device are admitted
// as soon as a stream and memory free up, so their work overlaps on the GPU.
// Device buffers come from the stream-ordered pool, whose release threshold
// is raised so that memory stays resident between requests.
//
// Plans outlive their requests in a small cache keyed by the request's whole
// geometry, image grid and options, so a repeated protocol skips the set-up
// (per-view tables, redundancy table, filter spectrum, kernel shape).  A cached
// plan's device memory stays counted against the budget; idle cached plans
// are released, least recently used first, when a request needs the room.
// Admission is first fit in arrival order, with aging: a request passed over
// max_bypass times blocks every later request until it fits, so a large
// request cannot be overtaken indefinitely.
//
// No reference counterpart: the reference reconstructs one scan per call
// (src/cuda/recon/fdk_axial.cuh:69-89).
#include "plan_internal.hpp"

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace gct {

void launch_residual(const float *a, const float *b, size_t n, double *partials, uint32_t blocks, hipStream_t s);

namespace {

using Clock = std::chrono::steady_clock;

struct CacheEntry;

struct Request {
    uint64_t ticket = 0;
    gct_ct_geom geom{};
    std::vector<gct_proj_geom> views;
    gct_image_geom img{};
    gct_recon_opts opts{};
    const float *proj = nullptr;
    float *vol = nullptr;
    int reproject = 0;
    uint64_t footprint = 0;     // device bytes while in flight, plan included
    uint64_t plan_bytes = 0;    // the plan's part of it
    uint64_t charged = 0;       // what admission added to the budget's use
    std::string key;            // plan cache key
    CacheEntry *cached = nullptr;
    uint32_t bypassed = 0;
    uint64_t updates = 0;
    Clock::time_point t_submit, t_done;
    bool done = false;
    gct_status status = GCT_ERR_NOT_READY;
    std::string error;
    double reproj_err = -1.0;
};

struct CacheEntry {
    std::string key;
    std::unique_ptr<gct_plan> plan;
    uint64_t bytes = 0;
    bool busy = false;
    uint64_t stamp = 0;
};

constexpr uint32_t kResidualBlocks = 256;

uint64_t request_footprint(const gct_ct_geom &g, const gct_image_geom &img, const gct_recon_opts &o, int reproject)
{
    const uint64_t nv = is_1d(g.kind) ? 1u : g.det_col.num_samples;
    const uint64_t proj = uint64_t(g.num_proj) * nv * g.det_row.num_samples * sizeof(float);
    uint64_t b = proj + volume_bytes(img) + plan_device_bytes(g, img, o, 0);
    if (reproject)
        b += proj + joseph_pad_elems(g.kind, img.x.num_samples, img.y.num_samples, img.dim == 3 ? img.z.num_samples : 1u) *
                        sizeof(float) +
             2u * kResidualBlocks * sizeof(double);
    return b;
}

template <class T>
void put(std::string &k, const T &v)
{
    k.append(reinterpret_cast<const char *>(&v), sizeof(T));
}

// Everything make_plan reads: the geometry with every view, the image grid
// and the options (all fixed-size PODs without padding).
std::string plan_key(const gct_ct_geom &g, const gct_image_geom &img, const gct_recon_opts &o)
{
    std::string k;
    k.reserve(64 + size_t(g.num_proj) * sizeof(gct_proj_geom));
    put(k, g.kind);
    put(k, g.det_row);
    put(k, g.det_col);
    put(k, g.num_proj);
    k.append(reinterpret_cast<const char *>(g.proj_geom), size_t(g.num_proj) * sizeof(gct_proj_geom));
    put(k, img);
    put(k, o);
    return k;
}

}  // namespace
}  // namespace gct

using namespace gct;

struct gct_server {
    int device = 0;
    uint64_t budget = 0;
    uint32_t max_inflight = 8;
    uint32_t cache_cap = 16;
    uint32_t max_bypass = 4;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    std::list<CacheEntry> cache;    // a list: entries keep their addresses while others come and go
    uint64_t stamp = 0;
    uint64_t next_ticket = 1;
    uint64_t used = 0;
    uint32_t inflight = 0;
    bool stopping = false;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> streams;
    gct_server_stats stats{};
    bool started = false;
    Clock::time_point t_first, t_last;

    void run(hipStream_t s);
    std::unique_ptr<gct_plan> process(Request &r, hipStream_t s);
    std::shared_ptr<Request> admit();
    bool make_room(uint64_t need, const CacheEntry *keep);
    void evict(std::list<CacheEntry>::iterator it);
};

void gct_server::evict(std::list<CacheEntry>::iterator it)
{
    used -= it->bytes;
    cache.erase(it);   // frees the plan's buffers (stream-ordered, after its last use)
}

// Releases idle cached plans, least recently used first (never `keep`), until
// `need` more bytes fit the budget; false, releasing nothing, when even all
// of them would not make the room.
bool gct_server::make_room(uint64_t need, const CacheEntry *keep)
{
    if (used + need <= budget) return true;
    uint64_t idle = 0;
    for (const CacheEntry &e : cache)
        if (!e.busy && &e != keep) idle += e.bytes;
    if (used - idle + need > budget) return false;
    while (used + need > budget) {
        auto lru = cache.end();
        for (auto it = cache.begin(); it != cache.end(); ++it)
            if (!it->busy && &*it != keep && (lru == cache.end() || it->stamp < lru->stamp)) lru = it;
        evict(lru);
    }
    return true;
}

// First fit in arrival order, with aging (called with mu held).
std::shared_ptr<Request> gct_server::admit()
{
    for (auto it = queue.begin(); it != queue.end(); ++it) {
        Request &q = **it;
        CacheEntry *hit = nullptr;
        for (CacheEntry &e : cache)
            if (!e.busy && e.key == q.key) { hit = &e; break; }
        const uint64_t need = hit ? q.footprint - q.plan_bytes : q.footprint;
        if (make_room(need, hit)) {
            for (auto e = queue.begin(); e != it; ++e) ++(*e)->bypassed;
            std::shared_ptr<Request> r = *it;
            queue.erase(it);
            used += need;
            r->charged = need;
            r->cached = hit;
            if (hit) {
                hit->busy = true;
                ++stats.plan_cache_hits;
            } else {
                ++stats.plan_cache_misses;
            }
            return r;
        }
        if (q.bypassed >= max_bypass) break;   // no later request may overtake it
    }
    return nullptr;
}

// Runs one request; returns the plan it created (nullptr when it used a cached one).
std::unique_ptr<gct_plan> gct_server::process(Request &r, hipStream_t s)
{
    std::unique_ptr<gct_plan> made;
    r.status = guarded([&] {
        DeviceGuard dg(device);
        gct_plan *plan = r.cached ? r.cached->plan.get() : nullptr;
        if (!plan) {
            made = make_plan(r.geom, r.img, r.opts, 0, s, true);
            plan = made.get();
        }
        const uint32_t nv = is_1d(r.geom.kind) ? 1u : r.geom.det_col.num_samples;
        const size_t npix = size_t(r.geom.num_proj) * nv * r.geom.det_row.num_samples;
        DevBuf dproj(npix * sizeof(float), s, true);
        DevBuf dvol(volume_bytes(r.img), s, true);
        hip_check(hipMemcpyAsync(dproj.p, r.proj, dproj.bytes, hipMemcpyHostToDevice, s), "upload projections");
        launch_filter(filter_args(*plan, dproj.as<float>(), 0, 0, plan->V, 0), s);
        launch_backproject(bp_args(*plan, dvol.as<float>(), 0, plan->V, false, true, 0), s);
        std::vector<double> part;
        DevBuf dre, dvt, dpart;
        if (r.reproject) {
            dre = DevBuf(npix * sizeof(float), s, true);
            const uint64_t pad = joseph_pad_elems(r.geom.kind, r.img.x.num_samples, r.img.y.num_samples,
                                                  r.img.dim == 3 ? r.img.z.num_samples : 1u);
            if (pad) dvt = DevBuf(pad * sizeof(float), s, true);
            dpart = DevBuf(2u * kResidualBlocks * sizeof(double), s, true);
            launch_forward_project(fp_args(*plan, dvol.as<float>(), dre.as<float>(), dvt.as<float>()), s);
            launch_residual(dre.as<float>(), dproj.as<float>(), npix, dpart.as<double>(), kResidualBlocks, s);
            part.resize(2u * kResidualBlocks);
            hip_check(hipMemcpyAsync(part.data(), dpart.p, dpart.bytes, hipMemcpyDeviceToHost, s), "download residual");
        }
        hip_check(hipMemcpyAsync(r.vol, dvol.p, dvol.bytes, hipMemcpyDeviceToHost, s), "download volume");
        hip_check(hipStreamSynchronize(s), "request");
        if (r.reproject) {
            double num = 0.0, den = 0.0;
            for (uint32_t i = 0; i < kResidualBlocks; ++i) { num += part[2 * i]; den += part[2 * i + 1]; }
            r.reproj_err = den > 0.0 ? std::sqrt(num / den) : 0.0;
        }
    });
    if (r.status != GCT_OK) {
        r.error = gct_last_error();
        made.reset();   // a plan from a failed request is not kept
    }
    return made;
}

void gct_server::run(hipStream_t s)
{
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (;;) {
                if (inflight < max_inflight) r = admit();
                if (r) break;
                if (stopping && queue.empty()) return;
                cv_work.wait(lk);
            }
            ++inflight;
            stats.peak_device_bytes = std::max<uint64_t>(stats.peak_device_bytes, used);
            stats.peak_inflight = std::max(stats.peak_inflight, inflight);
        }
        std::unique_ptr<gct_plan> made = process(*r, s);
        {
            std::lock_guard<std::mutex> lk(mu);
            used -= r->charged;
            if (r->cached) {
                r->cached->busy = false;
                r->cached->stamp = ++stamp;
            } else if (made && cache_cap > 0) {
                // keep the new plan when the cache has a free entry (or an idle one to
                // replace) and the budget has room for it; otherwise it is released
                if (cache.size() >= cache_cap) {
                    auto lru = cache.end();
                    for (auto it = cache.begin(); it != cache.end(); ++it)
                        if (!it->busy && (lru == cache.end() || it->stamp < lru->stamp)) lru = it;
                    if (lru != cache.end()) evict(lru);
                }
                if (cache.size() < cache_cap && used + r->plan_bytes <= budget) {
                    used += r->plan_bytes;
                    cache.push_back(CacheEntry{r->key, std::move(made), r->plan_bytes, false, ++stamp});
                }
            }
            made.reset();
            --inflight;
            r->cached = nullptr;
            r->t_done = Clock::now();
            r->done = true;
            t_last = r->t_done;
            if (r->status == GCT_OK) {
                ++stats.completed;
                stats.updates += r->updates;
            } else {
                ++stats.failed;
            }
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" gct_status gct_server_create(const gct_server_opts *opts, gct_server **out)
{
    return guarded([&] {
        if (!out) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (gct_device_count() == 0) fail(GCT_ERR_NO_DEVICE, "no HIP device");
        gct_server_opts o{};
        if (opts) o = *opts;
        auto sv = std::make_unique<gct_server>();
        sv->device = o.device;
        DeviceGuard dg(o.device);
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        sv->budget = o.hbm_budget_bytes ? o.hbm_budget_bytes : uint64_t(0.8 * double(free_b));
        sv->max_inflight = o.max_inflight ? o.max_inflight : 8u;
        sv->cache_cap = o.plan_cache_entries == 0 ? 16u : (o.plan_cache_entries == UINT32_MAX ? 0u : o.plan_cache_entries);
        sv->max_bypass = o.max_bypass ? o.max_bypass : 4u;
        const uint32_t ns = o.num_streams ? o.num_streams : 4u;
        // keep pool memory resident between requests
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, o.device) == hipSuccess && pool) {
            uint64_t thr = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
        for (uint32_t i = 0; i < ns; ++i) {
            hipStream_t s = nullptr;
            hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            sv->streams.push_back(s);
        }
        gct_server *raw = sv.get();
        for (hipStream_t s : sv->streams) sv->workers.emplace_back([raw, s] { raw->run(s); });
        *out = sv.release();
    });
}

extern "C" void gct_server_destroy(gct_server *sv)
{
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stopping = true;
    }
    sv->cv_work.notify_all();
    for (auto &t : sv->workers) t.join();
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(sv->device);
    sv->cache.clear();   // cached plans free on their streams, which are still alive
    for (hipStream_t s : sv->streams) (void)hipStreamSynchronize(s);
    for (hipStream_t s : sv->streams) (void)hipStreamDestroy(s);
    (void)hipSetDevice(prev);
    delete sv;
}

extern "C" gct_status gct_server_submit(gct_server *sv, const gct_ct_geom *geom, const gct_image_geom *img,
                                        const gct_recon_opts *opts, const float *proj_host, float *vol_host,
                                        int reproject, uint64_t *ticket)
{
    return guarded([&] {
        if (!sv || !geom || !img || !proj_host || !vol_host || !ticket) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        validate_geom(*geom);
        validate_image(*geom, *img);
        auto r = std::make_shared<Request>();
        r->views.assign(geom->proj_geom, geom->proj_geom + geom->num_proj);
        r->geom = *geom;
        r->geom.proj_geom = r->views.data();
        r->img = *img;
        r->opts = opts_or_default(opts);
        r->opts.device = sv->device;
        r->proj = proj_host;
        r->vol = vol_host;
        r->reproject = reproject;
        r->footprint = request_footprint(r->geom, r->img, r->opts, reproject);
        r->plan_bytes = plan_device_bytes(r->geom, r->img, r->opts, 0);
        r->key = plan_key(r->geom, r->img, r->opts);
        r->updates = uint64_t(img->x.num_samples) * img->y.num_samples * (img->dim == 3 ? img->z.num_samples : 1u) *
                     geom->num_proj;
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->stopping) fail(GCT_ERR_RUNTIME, "server is shutting down");
        ++sv->stats.submitted;
        if (r->footprint > sv->budget) {
            ++sv->stats.rejected;
            fail(GCT_ERR_REJECTED, "request needs " + std::to_string(r->footprint) + " device bytes, budget " +
                                       std::to_string(sv->budget));
        }
        r->ticket = sv->next_ticket++;
        r->t_submit = Clock::now();
        if (!sv->started) {
            sv->started = true;
            sv->t_first = r->t_submit;
        }
        sv->all[r->ticket] = r;
        sv->queue.push_back(r);
        *ticket = r->ticket;
        sv->cv_work.notify_one();
    });
}

static gct_status finish_ticket(gct_server *sv, uint64_t ticket, bool block, double *latency_ms, double *reproj_err)
{
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->all.find(ticket);
        if (it == sv->all.end()) {
            set_last_error("unknown ticket");
            return GCT_ERR_INVALID_ARGUMENT;
        }
        r = it->second;
        if (!r->done) {
            if (!block) return GCT_ERR_NOT_READY;
            sv->cv_done.wait(lk, [&] { return r->done; });
        }
        sv->all.erase(ticket);
    }
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (reproj_err) *reproj_err = r->reproj_err;
    if (r->status != GCT_OK) set_last_error(r->error);
    return r->status;
}

extern "C" gct_status gct_server_wait(gct_server *sv, uint64_t ticket, double *latency_ms, double *reproj_err)
{
    if (!sv) return GCT_ERR_INVALID_ARGUMENT;
    return finish_ticket(sv, ticket, true, latency_ms, reproj_err);
}

extern "C" gct_status gct_server_poll(gct_server *sv, uint64_t ticket, double *latency_ms, double *reproj_err)
{
    if (!sv) return GCT_ERR_INVALID_ARGUMENT;
    return finish_ticket(sv, ticket, false, latency_ms, reproj_err);
}

extern "C" gct_status gct_server_get_stats(gct_server *sv, gct_server_stats *stats)
{
    if (!sv || !stats) return GCT_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(sv->mu);
    *stats = sv->stats;
    stats->busy_seconds = sv->started ? std::chrono::duration<double>(sv->t_last - sv->t_first).count() : 0.0;
    return GCT_OK;
}

// Batches: many uneven scans backprojected by persistent kernels, one per
// group of scans that share a specialised kernel, optionally replayed from a
// HIP graph.
//
// No reference counterpart: the reference reconstructs one scan per call and
// one view per launch (src/cuda/recon/fdk_axial.cuh:152-205,
// src/cuda/recon/back_proj.cuh:16-104).
#include "plan_internal.hpp"

#include <algorithm>
#include <initializer_list>
#include <memory>
#include <tuple>
#include <vector>

namespace gct {

namespace {

uint64_t scan_cost(const gct_plan &p)
{
    return uint64_t(p.img.x.num_samples) * p.img.y.num_samples * (p.img.dim == 3 ? p.img.z.num_samples : 1u) * p.V;
}

// Scans that share a persistent kernel: geometry kind, storage, FP64 column
// set-up and tile type.
std::tuple<int, int, int, uint32_t> group_key(const gct_plan &p)
{
    return {p.geom.kind, int(p.opts.storage), int(p.setup_f64), batch_tile_depth(p.geom.kind, p.circular)};
}

}  // namespace

void check_batch_plans(gct_plan *const *plans, uint32_t n)
{
    if (!plans || !n) fail(GCT_ERR_INVALID_ARGUMENT, "empty batch");
    for (uint32_t i = 0; i < n; ++i) {
        if (!plans[i]) fail(GCT_ERR_INVALID_ARGUMENT, "null plan");
        if (plans[i]->device != plans[0]->device) fail(GCT_ERR_INVALID_ARGUMENT, "plans on different devices");
        if (plans[i]->filt_views < plans[i]->V) fail(GCT_ERR_INVALID_ARGUMENT, "plan holds a partial filtered buffer");
    }
}

// Work order: grouped by key, larger scans first within a group (they start
// first and the small ones fill the tail), creation order among equals.
// std::sort with the index as the last key gives the order a stable sort by
// (key, cost) would, without a temporary buffer.
void batch_order(gct_plan *const *plans, uint32_t n, uint32_t *order)
{
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::sort(order, order + n, [&](uint32_t a, uint32_t c) {
        const auto ka = group_key(*plans[a]), kc = group_key(*plans[c]);
        if (ka != kc) return ka < kc;
        const uint64_t ca = scan_cost(*plans[a]), cc = scan_cost(*plans[c]);
        if (ca != cc) return ca > cc;
        return a < c;
    });
}

uint32_t batch_groups(gct_plan *const *plans, const uint32_t *order, uint32_t n, BatchGroup *groups,
                      uint32_t max_groups)
{
    uint32_t ng = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const gct_plan &p = *plans[order[k]];
        if (k == 0 || group_key(p) != group_key(*plans[order[k - 1]])) {
            if (ng == max_groups) fail(GCT_ERR_UNSUPPORTED, "too many kernel groups in one batch");
            const uint32_t depth = batch_tile_depth(p.geom.kind, p.circular);
            groups[ng++] = BatchGroup{p.geom.kind, int(p.opts.storage), !is_1d(p.geom.kind) && depth == kBatchLdsZU,
                                      p.setup_f64, k, 0, 0};
        }
        ++groups[ng - 1].num_scans;
    }
    return ng;
}

BatchScan batch_entry(gct_plan &p, float *vol, uint32_t tile_begin)
{
    if (!vol) fail(GCT_ERR_INVALID_ARGUMENT, "null volume");
    BatchScan sc{};
    sc.bp = bp_args(p, vol, 0, p.V, false, true, 0);
    const uint32_t depth = batch_tile_depth(p.geom.kind, sc.bp.circular);
    sc.bp.z_unroll = depth;
    sc.bp.variant = depth == kBatchLdsZU ? GCT_BP_LDS : GCT_BP_DIRECT;
    sc.tiles_x = (p.img.x.num_samples + 15u) / 16u;
    sc.tiles_y = (p.img.y.num_samples + 15u) / 16u;
    sc.tiles_z = is_1d(p.geom.kind) ? 1u : (p.img.z.num_samples + depth - 1u) / depth;
    sc.tile_begin = tile_begin;
    return sc;
}

// Fills table[k] for the scans in work order; tile numbers restart at each
// group, whose total_tiles are set.
void batch_table(gct_plan *const *plans, const uint32_t *order, uint32_t n, float *const *vol, BatchGroup *groups,
                 uint32_t ngroups, BatchScan *table)
{
    uint32_t tile = 0, gi = 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (gi < ngroups && groups[gi].first_scan == k) tile = 0;
        const uint32_t oi = order[k];
        table[k] = batch_entry(*plans[oi], vol[oi], tile);
        tile += table[k].tiles_x * table[k].tiles_y * table[k].tiles_z;
        groups[gi].total_tiles = tile;
        if (k + 1 == groups[gi].first_scan + groups[gi].num_scans) ++gi;
    }
}

void batch_filters(gct_plan *const *plans, const uint32_t *order, uint32_t n, const float *const *proj, hipStream_t s)
{
    for (uint32_t k = 0; k < n; ++k) {
        gct_plan &p = *plans[order[k]];
        if (!proj[order[k]]) fail(GCT_ERR_INVALID_ARGUMENT, "null projections");
        launch_filter(filter_args(p, proj[order[k]], 0, 0, p.V, 0), s);
    }
}

}  // namespace gct

using namespace gct;

struct gct_batch {
    static constexpr uint32_t kSideStreams = 4;
    int device = 0;
    std::vector<gct_plan *> plans;    // creation order
    std::vector<uint32_t> order;      // work order: by kernel group, largest scans first
    std::vector<BatchGroup> groups;   // contiguous ranges of the work order
    std::vector<BatchScan> table;     // host work table (graph path)
    // The plain path and the graph keep their own device tables and counters:
    // the graph's kernels hold the addresses of the ones it was captured with,
    // so a plain execute in between must not overwrite them.
    DevBuf d_plain, d_graph, c_plain, c_graph;
    BatchScan *h_stage = nullptr;     // pinned host copy of the plain path's table
    hipEvent_t staged = nullptr;      // recorded after the last upload from h_stage
    hipEvent_t plain_done = nullptr;  // recorded after the last plain execute's work
    hipEvent_t graph_done = nullptr;  // recorded after the last graph launch
    // Groups too small to fill the device run as concurrent branches: forked
    // from the launching stream to side streams by events and joined back the
    // same way (also under capture); grid_cap is the persistent grid's size,
    // 8 workgroups per CU (launch_backproject_batch_group).
    uint32_t grid_cap = 2048;
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork = nullptr, join[kSideStreams] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const float *> g_proj;
    std::vector<float *> g_vol;
    hipStream_t g_stream = nullptr;
    ~gct_batch()
    {
        if (graph_done) (void)hipEventSynchronize(graph_done);
        if (plain_done) (void)hipEventSynchronize(plain_done);
        if (staged) (void)hipEventSynchronize(staged);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (uint32_t i = 0; i < kSideStreams; ++i) {
            if (side[i]) (void)hipStreamDestroy(side[i]);
            if (join[i]) (void)hipEventDestroy(join[i]);
        }
        for (hipEvent_t e : {fork, staged, plain_done, graph_done})
            if (e) (void)hipEventDestroy(e);
        if (h_stage) (void)hipHostFree(h_stage);
    }
};

namespace gct {
namespace {

// Stream work only: filters of every scan, then one persistent backprojection
// per group on the table `scans` with counters `counters`.
void batch_record(gct_batch &b, const float *const *proj, const BatchScan *scans, uint32_t *counters, hipStream_t s)
{
    const uint32_t n = uint32_t(b.plans.size());
    batch_filters(b.plans.data(), b.order.data(), n, proj, s);
    launch_zero_counters(counters, uint32_t(b.groups.size()), s);
    // A group with at least as many tiles as its persistent grid has
    // workgroups fills the device by itself: those run one after another on
    // the launching stream.  Run beside each other, two such kernels only
    // share the CUs' workgroup slots, and the mix ran slower than the sum
    // (uneven batch: 0.212 s with the 2-D group on a side stream, 0.191 s in
    // sequence).  Groups too small to fill the device are forked to side
    // streams first, so they overlap the large ones and each other.
    std::vector<size_t> small;
    for (size_t gi = 0; gi < b.groups.size(); ++gi)
        if (b.groups[gi].total_tiles < b.grid_cap) small.push_back(gi);
    const uint32_t used = std::min<uint32_t>(gct_batch::kSideStreams, uint32_t(small.size()));
    if (used > 0) {
        hip_check(hipEventRecord(b.fork, s), "fork");
        for (uint32_t i = 0; i < used; ++i) hip_check(hipStreamWaitEvent(b.side[i], b.fork, 0), "fork wait");
        for (size_t i = 0; i < small.size(); ++i)
            launch_backproject_batch_group(scans, b.groups[small[i]], counters + small[i],
                                           b.side[i % gct_batch::kSideStreams]);
    }
    for (size_t gi = 0; gi < b.groups.size(); ++gi)
        if (b.groups[gi].total_tiles >= b.grid_cap) launch_backproject_batch_group(scans, b.groups[gi], counters + gi, s);
    for (uint32_t i = 0; i < used; ++i) {
        hip_check(hipEventRecord(b.join[i], b.side[i]), "join");
        hip_check(hipStreamWaitEvent(s, b.join[i], 0), "join wait");
    }
}

}  // namespace
}  // namespace gct

extern "C" gct_status gct_batch_create(gct_plan *const *plans, uint32_t n, gct_batch **out)
{
    return guarded([&] {
        if (!out) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        check_batch_plans(plans, n);
        auto b = std::make_unique<gct_batch>();
        b->device = plans[0]->device;
        b->plans.assign(plans, plans + n);
        b->order.resize(n);
        batch_order(plans, n, b->order.data());
        b->groups.resize(n);
        b->groups.resize(batch_groups(plans, b->order.data(), n, b->groups.data(), n));
        b->table.resize(n);
        DeviceGuard dg(b->device);
        int cus = 256;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess) cus = 256;
        b->grid_cap = uint32_t(cus) * 8u;
        const size_t tbytes = size_t(n) * sizeof(BatchScan), cbytes = b->groups.size() * sizeof(uint32_t);
        b->d_plain = DevBuf(tbytes);
        b->d_graph = DevBuf(tbytes);
        b->c_plain = DevBuf(cbytes);
        b->c_graph = DevBuf(cbytes);
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&b->h_stage), tbytes, hipHostMallocDefault), "hipHostMalloc");
        for (hipEvent_t *e : {&b->fork, &b->staged, &b->plain_done, &b->graph_done})
            hip_check(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate");
        for (uint32_t i = 0; i < gct_batch::kSideStreams; ++i) {
            hip_check(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming), "hipEventCreate");
        }
        *out = b.release();
    });
}

extern "C" void gct_batch_destroy(gct_batch *b)
{
    if (!b) return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    delete b;
    (void)hipSetDevice(prev);
}

extern "C" gct_status gct_batch_execute(gct_batch *b, const float *const *proj, float *const *vol, int use_graph,
                                        void *stream)
{
    return guarded([&] {
        if (!b || !proj || !vol) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        DeviceGuard dg(b->device);
        hipStream_t s = static_cast<hipStream_t>(stream);
        const uint32_t n = uint32_t(b->plans.size());
        if (!use_graph) {
            // The table is built into pinned memory once the previous upload
            // from it has completed, and uploaded on the caller's stream after
            // the previous plain execute's kernels (on whatever stream they ran)
            // have finished reading the device table.
            hip_check(hipEventSynchronize(b->staged), "wait for the previous table upload");
            batch_table(b->plans.data(), b->order.data(), n, vol, b->groups.data(), uint32_t(b->groups.size()),
                        b->h_stage);
            hip_check(hipStreamWaitEvent(s, b->plain_done, 0), "wait for the previous execute");
            hip_check(hipMemcpyAsync(b->d_plain.p, b->h_stage, size_t(n) * sizeof(BatchScan), hipMemcpyHostToDevice, s),
                      "upload batch table");
            hip_check(hipEventRecord(b->staged, s), "record");
            batch_record(*b, proj, b->d_plain.as<BatchScan>(), b->c_plain.as<uint32_t>(), s);
            hip_check(hipEventRecord(b->plain_done, s), "record");
            return;
        }
        std::vector<const float *> P(proj, proj + n);
        std::vector<float *> Vv(vol, vol + n);
        const bool same = b->exec && b->g_stream == s && b->g_proj == P && b->g_vol == Vv;
        if (!same) {
            // the previous replay must be done with the graph's table before it is rewritten
            hip_check(hipEventSynchronize(b->graph_done), "wait for the previous replay");
            if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
            if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
            batch_table(b->plans.data(), b->order.data(), n, vol, b->groups.data(), uint32_t(b->groups.size()),
                        b->table.data());
            hip_check(hipMemcpy(b->d_graph.p, b->table.data(), size_t(n) * sizeof(BatchScan), hipMemcpyHostToDevice),
                      "upload batch table");
            hipStream_t cap = nullptr;
            hip_check(hipStreamCreate(&cap), "hipStreamCreate");
            hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture");
            try {
                batch_record(*b, proj, b->d_graph.as<BatchScan>(), b->c_graph.as<uint32_t>(), cap);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(cap, &g);
                if (g) (void)hipGraphDestroy(g);
                (void)hipStreamDestroy(cap);
                throw;
            }
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(cap, &g);
            (void)hipStreamDestroy(cap);
            hip_check(e, "end capture");
            b->graph = g;
            hip_check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
            b->g_proj = P;
            b->g_vol = Vv;
            b->g_stream = s;
        }
        hip_check(hipGraphLaunch(b->exec, s), "graph launch");
        hip_check(hipEventRecord(b->graph_done, s), "record");
    });
}

// Internal state of the gfx950 backend shared by the C ABI (api.cpp) and the serving runtime
// (serve.cpp): the backend and guide objects behind the opaque handles of hussar_hip.h, and
// the error / device-memory helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/hussar_hip.h"
#include "guide.h"
#include "kernels.h"

namespace hsr {

// Records msg as the calling thread's hsr_last_error() and returns s.
hsr_status fail(hsr_status s, const std::string &msg);

template <class T> hsr_status dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) return fail(HSR_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return HSR_OK;
}

template <class T> void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace hsr

#define HSR_HIP(call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return hsr::fail(e_ == hipErrorOutOfMemory ? HSR_ERR_OOM : HSR_ERR_HIP,               \
                             std::string(#call) + ": " + hipGetErrorString(e_));                  \
    } while (0)

struct hsr_guide {
    hsr_backend *backend = nullptr;  // used only while the backend is alive (runs, steps)
    int device = 0;                  // own copy: destroy must not touch the backend
    hsr::GuideHost host;
    char *d_trees = nullptr;      // the sampling tree and the training tree's children, packed
    int32_t *d_sChild = nullptr;  // into d_trees at the last upload: child[ns] | p0[ns] | child[nt]
    float *d_sP0 = nullptr;
    int32_t *d_tChild = nullptr;
    double *d_tStat = nullptr;
    int32_t capacity = 1 << 16;
    bool dirty = true;  // host trees newer than the device copies
    // pinned host staging: uploads [sampling child | sampling p0 | training child | training
    // stat] (capacity entries each) and the downloaded training statistics; uploads are
    // asynchronous on the backend's stream, evUp marks when the staging area is free again
    char *h_stage = nullptr;
    double *h_stats = nullptr;
    hipEvent_t evUp = nullptr, evDl = nullptr;
    bool statsPending = false;  // a run has queued the statistics' download into h_stats (evDl)
};

struct hsr_backend {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu = 0;
    int ldsPerBlock = 65536;  // device limit of a workgroup's LDS (160 KiB on gfx950)
    int ldsPerCU = 65536;
    // mesh
    float4 *d_nodes = nullptr;   // quantised four-wide nodes
    float4 *d_nodesF = nullptr;  // float-box four-wide nodes
    float4 *d_tris = nullptr;
    int32_t nodeCount = 0, slotCount = 0, bvhDepth = 0;
    int64_t triangles = 0;
    double bvh_ms = 0.0;
    std::vector<int32_t> slotIds;  // triangle id of every slot (leaf order)
    float4 *d_vel = nullptr;       // per slot velocity, null for a static mesh
    int32_t *d_triSlot = nullptr;  // one slot of each triangle that has one (pvTriangles)
    int64_t pvTriangles = 0;
    // when set, runs trace this backend's mesh instead of their own (a serving slot running
    // a request on a resident scene; the server keeps the scene alive meanwhile)
    const hsr_backend *meshFrom = nullptr;
    // frame
    bool configured = false;
    hsr_frame_config fc{1, 0, 0};
    int64_t frameDoubles = 0;  // 2 * chirps * N * C
    int64_t frameCap = 0;      // doubles allocated (grow-only: no frees between frames)
    double *d_frame_sum = nullptr, *d_frame_comp = nullptr;
    double totalWeight = 0.0;
    // scratch
    double *d_partials = nullptr;
    size_t partialCap = 0;  // doubles
    double *d_stage_sum = nullptr, *d_stage_comp = nullptr;
    size_t stageCap = 0;
    unsigned long long *d_counters = nullptr;
    int32_t *d_cells = nullptr;  // coherence order of the wavefront generation (kCells)
    // cube frames: the event list of the LDS tile splat (evCap entries) and its count
    hsr::CubeEvent *d_events = nullptr;
    unsigned evCap = 0;
    unsigned *d_evCount = nullptr;
    double *d_cubeParts = nullptr;  // partial cubes of the tile splat (grow-only)
    double *d_statRows = nullptr;   // partial rows of the training statistics (kStatRows doubles)
    int64_t cubePartsCap = 0;       // doubles
    int32_t *d_recHits = nullptr;
    int8_t *d_recNee = nullptr;
    int64_t recCap = 0, recPathsLast = 0;
    int32_t recMaxDepth = 0, recC = 0;
    // options
    bool optCounters = false;
    int64_t optRecord = 0;
    int64_t optPathsPerLaunch = int64_t(1) << 24;
    int32_t optPhaseMode = 0;
    int32_t optChain = 0;  // HSR_OPT_CHAIN
    int32_t optTail = -1;  // HSR_OPT_TAIL: first depth of the fused tail, 0 = none, -1 = adaptive
    bool optMarkers = false;  // HSR_OPT_MARKERS: ROCTx ranges around passes, batches and stage launches
    bool optShadowSplat = true;  // HSR_OPT_SHADOW_SPLAT: decoupled batches splat in their shadow stages
    // the extension rays of each depth of the last wavefront batch (h_ext: pinned landing area
    // of their copy), for the adaptive tail depth
    static constexpr int kTrackDepths = 64;
    uint32_t *h_ext = nullptr;
    int64_t lastExt[kTrackDepths] = {};
    int lastBatchDepths = 0;
    int64_t lastBatchPaths = 0;
    int32_t optPipeline = 1;
    int32_t optTravWaves = 0;  // resident waves per CU of the traversal kernels, 0 = all that fit
    int64_t optCubeEvents = -1;  // cube event list capacity (HSR_OPT_CUBE_EVENTS), -1 = the default
    int32_t maxGrid = 0;
    // wavefront buffers
    void *d_wf = nullptr;
    size_t wfBytes = 0;
    int64_t wfCap = int64_t(1) << 24;  // paths per wavefront batch (HSR_OPT_BATCH_PATHS)
    // post-processing (grow-only)
    double *d_post = nullptr, *d_post2 = nullptr;
    int64_t postCap = 0;  // doubles
    void *d_postScratch = nullptr;
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // the last run returned before its frame reduction ended (its guide statistics were
    // ready first): reduce_ms and total_ms still to add from ev[1..3] (settle_timing)
    bool timingPending = false;
    std::vector<hipEvent_t> evTrav;  // pairs around traversal launches (wavefront)
    // the closest-hit traversal of a depth runs on stream2, beside that depth's shadow and
    // gather stages on stream (default priority both); evFork/evJoin order them per depth
    hipStream_t stream2 = nullptr;
    std::vector<hipEvent_t> evFork, evJoin;
    // off for the slots of a server with more than two: concurrent slots already share the
    // device, and a second stream per slot would make more streams than hardware queues
    bool useStream2 = true;
    // HSR_OPT_GRAPH: a batch's launch sequence (set-up to the last gather) built once per
    // launch shape into a graph and replayed, from the second batch of that shape among the
    // last kSeen shapes; at most kGraphs graphs are kept (least recently used out). A graph
    // that fails to build turns the option off for this backend. Off by default: replays
    // measured no faster than stage launches (their node updates and graph launches cost the
    // host what the launches they replace did).
    bool optGraph = false;
    static constexpr int kSeen = 32;
    std::vector<std::vector<int64_t>> graphSeen;
    struct GraphEntry {
        std::vector<int64_t> key;
        hsr::BatchPlan plan;
        hsr::BatchGraph g;
        uint64_t used = 0;
    };
    static constexpr int kGraphs = 8;
    std::vector<GraphEntry> graphs;
    uint64_t graphClock = 0;
    // primary visibility lists (HSR_OPT_PRIMARY): depth-0 closest hits from the lists of the
    // traced transmitter's position, built on the device at most once per position and frame
    // (epoch: advanced by every PathTracer::run) for runs of at least optPrimaryMin paths. Off:
    // switching them on is refused (their first GPU run faulted, cause not found)
    int32_t optPrimary = 0;
    int64_t optPrimaryMin = int64_t(1) << 20;
    int32_t optPrimaryRes = 64;
    void *d_pv = nullptr;
    size_t pvBytes = 0;
    int64_t pvCap = 0;      // list entries (grows when a build exceeds it)
    bool pvCapSet = false;  // HSR_OPT_PRIMARY_CAP set it
    hsr::PvBuild pv{};
    int32_t *h_pvMeta = nullptr;  // pinned copy of the last build's meta (near, over, entries)
    bool pvBuilt = false, pvCheck = false;
    const hsr_backend *pvMesh = nullptr;
    float pvPos[3] = {0.0f, 0.0f, 0.0f};
    uint64_t pvEpoch = 0, epoch = 0;
    hipEvent_t evPv[2] = {nullptr, nullptr};
    hsr_run_stats stats{};
};

namespace hsr {
// Allocates the wavefront buffers for batches of P paths with up to C receivers and maxDepth
// bounces, and a frame of frameDoubles doubles, ahead of time (serving slots allocate nothing
// while requests run).
hsr_status backend_reserve(hsr_backend *b, int64_t P, int C, int maxDepth, int64_t frameDoubles);
// Allocates a cube event list of at least `events` entries (its count starts at zero) and room
// for partial cubes of partDoubles doubles in all (both grow-only).
hsr_status backend_events(hsr_backend *b, int64_t partDoubles, unsigned events);
// Device bytes of a backend's mesh (BVH nodes, triangle slots, velocities).
int64_t backend_mesh_bytes(const hsr_backend *b);
// Device bytes of its work buffers, what they would be after backend_reserve (grow-only), and
// their release.
int64_t backend_device_bytes(const hsr_backend *b);
int64_t backend_reserve_bytes(const hsr_backend *b, int64_t P, int C, int maxDepth, int64_t frameDoubles,
                              bool training);
void backend_release(hsr_backend *b);
}  // namespace hsr


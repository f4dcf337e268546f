// C ABI of the gfx950 backend (include/hussar_hip.h).
//
// hsr_hip_create     ~ Backend(const TriangleMesh &, Integrator &)  (cpu.cpp:32-47, gpu.cpp:482-493)
// hsr_hip_run        ~ Backend::run(scene, budget, interruptFlag)   (cpu.h:44-46, gpu.cpp:499-532)
// hsr_path_tracer_run ~ PathTracer::run                             (path.h:171-224)
// hsr_hip_fetch_frame ~ PathTracer::fetchFrame                      (path.h:453-455)
// hsr_guide_*        ~ GuidingWrapper reset/step                    (path.h:48-51,119-126)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/hussar_hip.h"
#include "backend.h"
#include "bvh.h"

#ifndef HSR_GATHER_PER_CU
#define HSR_GATHER_PER_CU 4
#endif
// gather workgroups per CU (each holds a frame tile in LDS and adds a partial frame)
static constexpr int kGatherPerCU = HSR_GATHER_PER_CU;
// cube frames: at most this many doubles of partial cubes for the tile splat (1 GiB), and
// the partial cubes a serving slot reserves (frames' worth)
// the fused tail starts at the first depth with at most this many extension rays (adaptive)
static constexpr int64_t kTailPaths = int64_t(1) << 17;
static constexpr int64_t kCubePartDoubles = int64_t(1) << 27;
namespace hsr {
constexpr int64_t kCubeReserveParts = 4;
}

namespace hsr {

thread_local std::string g_last_error;

hsr_status fail(hsr_status s, const std::string &msg) {
    g_last_error = msg;
    return s;
}

}  // namespace hsr

using hsr::dalloc;
using hsr::dfree;
using hsr::fail;

// --------------------------------------------------------------------------------------
extern "C" {

const char *hsr_last_error(void) { return hsr::g_last_error.c_str(); }
int32_t hsr_abi_version(void) { return HSR_ABI_VERSION; }

void hsr_knobs_default(hsr_knobs *k) {
    if (!k) return;
    // path.h:153-169
    k->onlyIndirect = 1;
    k->maxDepth = 10;
    k->useGeometricalOptics = 1;
    k->poDiffraction = 0;
    k->doGuiding = 1;
    k->clearBeforeIteration = 1;
    k->correctPhase = 0;
    k->filterGuiding = 1;
    k->doFiltering = 1;
    k->filteringSphere = 1;
    k->filteringMin = 600.0f;
    k->filteringMax = 900.0f;
    k->filteringRadius = 160.0f;
    k->primaryOrder = 0;
    k->produceDebugImage = 0;
}

hsr_status hsr_hip_device_info(int32_t device, hsr_device_info *out) {
    if (!out) return fail(HSR_ERR_INVALID, "null output");
    hipDeviceProp_t p;
    HSR_HIP(hipGetDeviceProperties(&p, device));
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "%s", p.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
    out->compute_units = p.multiProcessorCount;
    out->xcds = 8;
    out->hbm_bytes = (int64_t)p.totalGlobalMem;
    out->lds_bytes_per_cu = (int32_t)p.maxSharedMemoryPerMultiProcessor;
    out->clock_mhz = p.clockRate / 1000;
    out->lds_bytes_per_block = (int32_t)p.sharedMemPerBlock;
    return HSR_OK;
}

hsr_status hsr_hip_create(const float *vertices, int64_t vertexCount, const int32_t *indices,
                          int64_t triangleCount, int32_t device, hsr_backend **out) {
    if (!out) return fail(HSR_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (vertexCount < 0 || triangleCount < 0 || (triangleCount > 0 && (!vertices || !indices)))
        return fail(HSR_ERR_INVALID, "invalid mesh buffers");
    int ndev = 0;
    HSR_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(HSR_ERR_INVALID, "no such HIP device");
    std::unique_ptr<hsr_backend> b(new hsr_backend());
    b->device = device;
    HSR_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    HSR_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(HSR_ERR_UNSUPPORTED, std::string("this library is built for gfx950, device is ") +
                                             prop.gcnArchName);
    b->cu = prop.multiProcessorCount;
    b->ldsPerBlock = (int)prop.sharedMemPerBlock;
    b->ldsPerCU = (int)prop.maxSharedMemoryPerMultiProcessor;
    HSR_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HSR_HIP(hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking));
    for (auto &e : b->ev) HSR_HIP(hipEventCreate(&e));
    HSR_HIP(hipHostMalloc((void **)&b->h_ext, hsr_backend::kTrackDepths * sizeof(uint32_t), hipHostMallocDefault));

    hsr::FlatBVH bvh;
    const auto t0 = std::chrono::steady_clock::now();
    if (!hsr::build_bvh(vertices, vertexCount, indices, triangleCount, bvh))
        return fail(HSR_ERR_INVALID, "BVH build failed (index out of range, too many triangles or depth)");
    b->bvh_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    b->nodeCount = bvh.nodeCount;
    b->slotCount = bvh.slotCount;
    b->bvhDepth = bvh.maxDepth;
    b->triangles = triangleCount;
    hsr_status s;
    if ((s = dalloc(&b->d_nodes, (size_t)bvh.nodeCount * 4)) != HSR_OK) return s;   // 4 float4 per node
    if ((s = dalloc(&b->d_nodesF, (size_t)bvh.nodeCount * 8)) != HSR_OK) return s;  // 8 float4 per node
    if ((s = dalloc(&b->d_tris, bvh.tris.size() / 4)) != HSR_OK) return s;  // the slots and the sentinel
    HSR_HIP(hipMemcpy(b->d_nodes, bvh.nodes.data(), bvh.nodes.size() * sizeof(float), hipMemcpyHostToDevice));
    HSR_HIP(hipMemcpy(b->d_nodesF, bvh.nodesF.data(), bvh.nodesF.size() * sizeof(float), hipMemcpyHostToDevice));
    HSR_HIP(hipMemcpy(b->d_tris, bvh.tris.data(), bvh.tris.size() * sizeof(float), hipMemcpyHostToDevice));
    b->slotIds.resize(bvh.slotCount);
    for (int32_t s2 = 0; s2 < bvh.slotCount; ++s2) std::memcpy(&b->slotIds[s2], &bvh.tris[12 * (size_t)s2 + 3], 4);
    {
        // one slot of every triangle that has one (its first in leaf order), for the primary
        // visibility lists: every slot of a triangle holds the same whole triangle
        std::vector<int32_t> first((size_t)triangleCount, -1), ts;
        for (int32_t s2 = 0; s2 < bvh.slotCount; ++s2) {
            const int32_t id = b->slotIds[s2];
            if (id >= 0 && id < triangleCount && first[id] < 0) first[id] = s2;
        }
        for (int32_t v : first)
            if (v >= 0) ts.push_back(v);
        b->pvTriangles = (int64_t)ts.size();
        if (!ts.empty()) {
            if ((s = dalloc(&b->d_triSlot, ts.size())) != HSR_OK) return s;
            HSR_HIP(hipMemcpy(b->d_triSlot, ts.data(), ts.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        }
    }
    HSR_HIP(hipHostMalloc((void **)&b->h_pvMeta, 4 * sizeof(int32_t), hipHostMallocDefault));
    for (auto &e : b->evPv) HSR_HIP(hipEventCreate(&e));
    if ((s = dalloc(&b->d_counters, 16)) != HSR_OK) return s;
    {
        static const std::vector<int32_t> cells = hsr::coherence_cells();
        if ((s = dalloc(&b->d_cells, cells.size())) != HSR_OK) return s;
        HSR_HIP(hipMemcpy(b->d_cells, cells.data(), cells.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    b->stats.bvh_build_ms = b->bvh_ms;
    *out = b.release();
    return HSR_OK;
}

void hsr_hip_destroy(hsr_backend *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->stream2) (void)hipStreamSynchronize(b->stream2);
    dfree(b->d_nodes);
    dfree(b->d_nodesF);
    dfree(b->d_tris);
    dfree(b->d_frame_sum);
    dfree(b->d_frame_comp);
    dfree(b->d_partials);
    dfree(b->d_stage_sum);
    dfree(b->d_stage_comp);
    dfree(b->d_counters);
    dfree(b->d_cells);
    dfree(b->d_events);
    dfree(b->d_evCount);
    dfree(b->d_cubeParts);
    dfree(b->d_statRows);
    dfree(b->d_recHits);
    dfree(b->d_recNee);
    dfree(b->d_vel);
    dfree(b->d_post);
    dfree(b->d_post2);
    dfree(b->d_triSlot);
    if (b->d_pv) (void)hipFree(b->d_pv);
    if (b->d_postScratch) (void)hipFree(b->d_postScratch);
    if (b->d_wf) (void)hipFree(b->d_wf);
    for (auto &g : b->graphs) hsr::destroy_graph(g.g);
    b->graphs.clear();
    if (b->h_ext) (void)hipHostFree(b->h_ext);
    if (b->h_pvMeta) (void)hipHostFree(b->h_pvMeta);
    for (auto &e : b->evPv)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : b->evTrav) (void)hipEventDestroy(e);
    for (auto &e : b->evFork) (void)hipEventDestroy(e);
    for (auto &e : b->evJoin) (void)hipEventDestroy(e);
    if (b->stream2) (void)hipStreamDestroy(b->stream2);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"

namespace hsr {
hsr_status backend_events(hsr_backend *b, int64_t partDoubles, unsigned events) {
    hsr_status s;
    if (!b->d_evCount) {
        if ((s = dalloc(&b->d_evCount, 64)) != HSR_OK) return s;
        HSR_HIP(hipMemset(b->d_evCount, 0, 64 * sizeof(unsigned)));
    }
    if (events > b->evCap) {
        dfree(b->d_events);
        b->evCap = 0;
        if ((s = dalloc(&b->d_events, events)) != HSR_OK) return s;
        b->evCap = events;
    }
    if (partDoubles > b->cubePartsCap) {
        dfree(b->d_cubeParts);
        b->cubePartsCap = 0;
        if ((s = dalloc(&b->d_cubeParts, (size_t)partDoubles)) != HSR_OK) return s;
        b->cubePartsCap = partDoubles;
    }
    return HSR_OK;
}

hsr_status backend_reserve(hsr_backend *b, int64_t P, int C, int maxDepth, int64_t frameDoubles) {
    HSR_HIP(hipSetDevice(b->device));
    const size_t need = wavefront_bytes(P, C, maxDepth);
    if (need > b->wfBytes) {
        if (b->d_wf) (void)hipFree(b->d_wf);
        b->d_wf = nullptr;
        b->wfBytes = 0;
        HSR_HIP(hipMalloc(&b->d_wf, need));
        b->wfBytes = need;
    }
    b->wfCap = P;
    // the largest one-chirp frame that still uses LDS tiles
    const size_t tileDoubles = (size_t)b->ldsPerBlock / sizeof(double);
    const size_t needP = (size_t)b->cu * 2 * tileDoubles, needS = (size_t)64 * tileDoubles;
    if (needP > b->partialCap) {
        dfree(b->d_partials);
        b->partialCap = 0;
        hsr_status s;
        if ((s = dalloc(&b->d_partials, needP)) != HSR_OK) return s;
        b->partialCap = needP;
    }
    if (needS > b->stageCap) {
        dfree(b->d_stage_sum);
        dfree(b->d_stage_comp);
        b->stageCap = 0;
        hsr_status s;
        if ((s = dalloc(&b->d_stage_sum, needS)) != HSR_OK) return s;
        if ((s = dalloc(&b->d_stage_comp, needS)) != HSR_OK) return s;
        b->stageCap = needS;
    }
    if (frameDoubles > b->frameCap) {
        dfree(b->d_frame_sum);
        dfree(b->d_frame_comp);
        b->frameCap = 0;
        hsr_status s;
        if ((s = dalloc(&b->d_frame_sum, frameDoubles)) != HSR_OK) return s;
        if ((s = dalloc(&b->d_frame_comp, frameDoubles)) != HSR_OK) return s;
        b->frameCap = frameDoubles;
    }
    // the tile splat: partial cubes of kCubeReserveParts frames, and an event list of four
    // events per path of a batch (more go straight to the frame with device atomics)
    return backend_events(b, kCubeReserveParts * frameDoubles,
                          (unsigned)std::min<int64_t>(kCubeEvents, 4 * P));
}

// Device bytes of a backend's work buffers: wavefront state, frame, tiles and stages, the cube
// event list and partial cubes, statistics rows, post-processing and hit records.
int64_t backend_device_bytes(const hsr_backend *b) {
    return (int64_t)b->wfBytes + (int64_t)b->partialCap * 8 + (int64_t)b->stageCap * 16 + b->frameCap * 16 +
           (int64_t)b->evCap * (int64_t)sizeof(CubeEvent) + b->cubePartsCap * 8 + (b->d_statRows ? kStatRows * 8 : 0) +
           b->postCap * 16 + b->recCap * 5;
}

// The bytes backend_reserve(b, P, C, maxDepth, frameDoubles) leaves the backend holding (its
// buffers only grow), plus the statistics rows of a guided training run.
int64_t backend_reserve_bytes(const hsr_backend *b, int64_t P, int C, int maxDepth, int64_t frameDoubles,
                              bool training) {
    const int64_t tileDoubles = b->ldsPerBlock / (int64_t)sizeof(double);
    const int64_t wf = std::max<int64_t>((int64_t)b->wfBytes, (int64_t)wavefront_bytes(P, C, maxDepth));
    const int64_t part = std::max<int64_t>((int64_t)b->partialCap, (int64_t)b->cu * 2 * tileDoubles) * 8;
    const int64_t stage = std::max<int64_t>((int64_t)b->stageCap, 64 * tileDoubles) * 16;
    const int64_t frame = std::max<int64_t>(b->frameCap, frameDoubles) * 16;
    const int64_t ev = std::max<int64_t>(b->evCap, std::min<int64_t>(kCubeEvents, 4 * P)) * (int64_t)sizeof(CubeEvent);
    const int64_t parts = std::max<int64_t>(b->cubePartsCap, kCubeReserveParts * frameDoubles) * 8;
    const int64_t rows = (b->d_statRows || training) ? kStatRows * 8 : 0;
    return wf + part + stage + frame + ev + parts + rows + b->postCap * 16 + b->recCap * 5;
}

// Frees a backend's work buffers (not its mesh, events or guide); the next run allocates again.
void backend_release(hsr_backend *b) {
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    if (b->d_wf) (void)hipFree(b->d_wf);
    b->d_wf = nullptr;
    b->wfBytes = 0;
    dfree(b->d_partials);
    b->partialCap = 0;
    dfree(b->d_stage_sum);
    dfree(b->d_stage_comp);
    b->stageCap = 0;
    dfree(b->d_frame_sum);
    dfree(b->d_frame_comp);
    b->frameCap = 0;
    b->configured = false;
    dfree(b->d_events);
    b->evCap = 0;
    dfree(b->d_cubeParts);
    b->cubePartsCap = 0;
    dfree(b->d_statRows);
    dfree(b->d_post);
    dfree(b->d_post2);
    b->postCap = 0;
    dfree(b->d_recHits);
    dfree(b->d_recNee);
    b->recCap = 0;
}

int64_t backend_mesh_bytes(const hsr_backend *b) {
    return (int64_t)b->nodeCount * (64 + 128) + ((int64_t)b->slotCount + 1) * 48 + (b->d_vel ? (int64_t)b->slotCount * 16 : 0);
}
}  // namespace hsr

extern "C" {

hsr_status hsr_hip_set_velocities(hsr_backend *b, const float *vel, int64_t triangleCount) {
    if (!b) return fail(HSR_ERR_INVALID, "null backend");
    HSR_HIP(hipSetDevice(b->device));
    HSR_HIP(hipStreamSynchronize(b->stream));
    if (!vel || triangleCount == 0) {
        dfree(b->d_vel);
        return HSR_OK;
    }
    if (triangleCount != b->triangles) return fail(HSR_ERR_INVALID, "one velocity per mesh triangle expected");
    std::vector<float4> v(b->slotIds.size());
    for (size_t s = 0; s < v.size(); ++s) {
        const int32_t id = b->slotIds[s];
        v[s] = (id >= 0 && id < b->triangles)
                   ? make_float4(vel[3 * (int64_t)id], vel[3 * (int64_t)id + 1], vel[3 * (int64_t)id + 2], 0.0f)
                   : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (!std::isfinite(v[s].x) || !std::isfinite(v[s].y) || !std::isfinite(v[s].z))
            return fail(HSR_ERR_INVALID, "velocities must be finite");
    }
    if (!b->d_vel) {
        hsr_status s;
        if ((s = dalloc(&b->d_vel, v.size())) != HSR_OK) return s;
    }
    HSR_HIP(hipMemcpy(b->d_vel, v.data(), v.size() * sizeof(float4), hipMemcpyHostToDevice));
    return HSR_OK;
}

hsr_status hsr_hip_configure_frame(hsr_backend *b, const hsr_frame_config *c) {
    if (!b || !c) return fail(HSR_ERR_INVALID, "null argument");
    if (c->chirpCount < 1 || c->samplesPerChirp < 1 || c->channelCount < 1 || c->channelCount > HSR_MAX_CHANNELS)
        return fail(HSR_ERR_INVALID, "frame dimensions out of range");
    if ((int64_t)c->chirpCount * c->samplesPerChirp * c->channelCount > (int64_t(1) << 28))
        return fail(HSR_ERR_INVALID, "frame too large");
    // frames whose one-chirp tile fits in LDS are accumulated in per-workgroup LDS tiles,
    // others directly in device memory (wavefront pipeline only)
    HSR_HIP(hipSetDevice(b->device));
    const int64_t nd = 2LL * c->chirpCount * c->samplesPerChirp * c->channelCount;
    if (nd > b->frameCap) {
        HSR_HIP(hipStreamSynchronize(b->stream));
        dfree(b->d_frame_sum);
        dfree(b->d_frame_comp);
        b->frameCap = 0;
        hsr_status s;
        if ((s = dalloc(&b->d_frame_sum, nd)) != HSR_OK) return s;
        if ((s = dalloc(&b->d_frame_comp, nd)) != HSR_OK) return s;
        b->frameCap = nd;
    }
    b->frameDoubles = nd;
    b->fc = *c;
    b->configured = true;
    b->maxGrid = 0;
    return hsr_hip_clear_frame(b);
}

hsr_status hsr_hip_clear_frame(hsr_backend *b) {
    if (!b) return fail(HSR_ERR_INVALID, "null backend");
    if (!b->configured) return fail(HSR_ERR_INVALID, "frame not configured");
    HSR_HIP(hipSetDevice(b->device));
    HSR_HIP(hsr::launch_zero2(b->d_frame_sum, b->d_frame_comp, b->frameDoubles, b->stream));
    b->totalWeight = 0.0;
    return HSR_OK;
}

hsr_status hsr_hip_set_option(hsr_backend *b, int32_t option, int64_t value) {
    if (!b) return fail(HSR_ERR_INVALID, "null backend");
    switch (option) {
        case HSR_OPT_COUNTERS:
            b->optCounters = value != 0;
            b->maxGrid = 0;
            return HSR_OK;
        case HSR_OPT_RECORD_HITS: b->optRecord = std::max<int64_t>(0, value); return HSR_OK;
        case HSR_OPT_PATHS_PER_LAUNCH:
            if (value < 1) return fail(HSR_ERR_INVALID, "paths per launch must be positive");
            b->optPathsPerLaunch = value;
            return HSR_OK;
        case HSR_OPT_PIPELINE:
            if (value < 0 || value > 1) return fail(HSR_ERR_INVALID, "pipeline must be 0 (fused) or 1 (wavefront)");
            b->optPipeline = (int32_t)value;
            return HSR_OK;
        case HSR_OPT_BATCH_PATHS:
            if (value < 1024 || value > (int64_t(1) << 24))
                return fail(HSR_ERR_INVALID, "paths per batch must lie in [1024, 2^24]");
            b->wfCap = value;
            return HSR_OK;
        case HSR_OPT_CUBE_EVENTS:
            if (value < 0 || value > (int64_t)hsr::kCubeEvents)
                return fail(HSR_ERR_INVALID, "cube event list capacity must lie in [0, 2^24]");
            b->optCubeEvents = value;
            return HSR_OK;
        case HSR_OPT_TRAV_WAVES:
            if (value < 0 || value > 64) return fail(HSR_ERR_INVALID, "traversal waves per CU out of range");
            b->optTravWaves = (int32_t)value;
            return HSR_OK;
        case HSR_OPT_PHASE_MODE:
            if (value < 0 || value > 2)
                return fail(HSR_ERR_UNSUPPORTED, "phase mode must be 0 (FP64), 1 (FP32) or 2 (double-float)");
            b->optPhaseMode = (int32_t)value;
            return HSR_OK;
        case HSR_OPT_TAIL:
            if (value < -1 || value > 64) return fail(HSR_ERR_INVALID, "tail depth must lie in [-1, 64]");
            b->optTail = (int32_t)value;
            return HSR_OK;
        case HSR_OPT_MARKERS:
            if (value < 0 || value > 1) return fail(HSR_ERR_INVALID, "markers must be 0 or 1");
            b->optMarkers = value != 0;
            return HSR_OK;
        case HSR_OPT_GRAPH:
            if (value < 0 || value > 1) return fail(HSR_ERR_INVALID, "graph must be 0 (stage launches) or 1 (replay)");
            b->optGraph = value != 0;
            return HSR_OK;
        case HSR_OPT_CHAIN:
            if (value < 0 || value > 1)
                return fail(HSR_ERR_UNSUPPORTED, "chain must be 0 (FP64) or 1 (the reference's FP32 chain)");
            b->optChain = (int32_t)value;
            return HSR_OK;
        case HSR_OPT_PRIMARY:
            if (value < 0 || value > 1) return fail(HSR_ERR_INVALID, "primary must be 0 (BVH) or 1 (visibility lists)");
            // the lists' first run on the device faulted and the cause is not found: disabled
            if (value == 1) return fail(HSR_ERR_UNSUPPORTED, "primary visibility lists are disabled");
            b->optPrimary = (int32_t)value;
            return HSR_OK;
        case HSR_OPT_PRIMARY_MIN:
            if (value < 0) return fail(HSR_ERR_INVALID, "primary minimum paths must not be negative");
            b->optPrimaryMin = value;
            return HSR_OK;
        case HSR_OPT_PRIMARY_RES:
            if (value < 4 || value > hsr::kPvMaxRes) return fail(HSR_ERR_INVALID, "primary resolution must lie in [4, 80]");
            if ((int32_t)value != b->optPrimaryRes) b->pvBuilt = false;
            b->optPrimaryRes = (int32_t)value;
            return HSR_OK;
        case HSR_OPT_SHADOW_SPLAT:
            if (value < 0 || value > 1) return fail(HSR_ERR_INVALID, "shadow splat must be 0 or 1");
            b->optShadowSplat = value != 0;
            return HSR_OK;
        case HSR_OPT_PRIMARY_CAP:
            if (value < 0 || value >= (int64_t(1) << 31))
                return fail(HSR_ERR_INVALID, "primary list capacity must lie in [0, 2^31)");
            b->pvCap = value;
            b->pvCapSet = value > 0;
            b->pvBuilt = false;
            return HSR_OK;
        default: return fail(HSR_ERR_INVALID, "unknown option");
    }
}

// A ROCTx range (HSR_OPT_MARKERS) held for the object's lifetime.
struct Marker {
    bool on;
    Marker(bool markers, const char *message) : on(markers) {
        if (on) roctxRangePushA(message);
    }
    ~Marker() {
        if (on) roctxRangePop();
    }
};

static size_t pv_align(size_t x) { return (x + 255) & ~(size_t)255; }

// The primary visibility lists of transmitter position pos over this run's mesh, built on the
// backend's stream unless the lists of the same mesh, position and frame epoch are in place;
// their pointers into a. A build's meta is copied to the host: a build over the lists'
// capacity is found after the batch (the depth-0 kernel then traversed the BVH) and the
// capacity grows for the next build.
static hsr_status pv_prepare(hsr_backend *b, const hsr_backend *mesh, const float pos[3], hsr::TraceArgs &a) {
    const int R = b->optPrimaryRes;
    const int64_t T = mesh->pvTriangles;
    const bool same = b->pvBuilt && b->pvMesh == mesh && b->pvEpoch == b->epoch && b->pv.R == R &&
                      std::memcmp(b->pvPos, pos, sizeof(b->pvPos)) == 0;
    if (!same) {
        const int64_t cells = 6 * (int64_t)R * R;
        const int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(256, T / 256));
        if (!b->pvCapSet) b->pvCap = std::max<int64_t>(b->pvCap, std::max<int64_t>(int64_t(1) << 23, 64 * T));
        b->pvCap = std::max<int64_t>(b->pvCap, 1);
        b->pvCap = std::min<int64_t>(b->pvCap, (int64_t(1) << 31) - 1);
        const size_t sortBytes = hsr::pv_sort_bytes(T);
        const size_t sz[] = {(size_t)T * 4,
                             (size_t)T * 4,
                             (size_t)T * 4,
                             (size_t)T * 4,
                             (size_t)hsr::pv_prep_blocks(T) * 25 * 4,
                             6 * 16,
                             16,
                             (size_t)chunks * cells * 4,
                             (size_t)(cells + 1) * 4,
                             (size_t)b->pvCap * 8,
                             sortBytes};
        size_t total = 0;
        for (size_t v : sz) total += pv_align(v);
        if (total > b->pvBytes) {
            if (b->d_pv) {
                HSR_HIP(hipStreamSynchronize(b->stream));
                (void)hipFree(b->d_pv);
            }
            b->d_pv = nullptr;
            b->pvBytes = 0;
            HSR_HIP(hipMalloc(&b->d_pv, total));
            b->pvBytes = total;
        }
        char *p = (char *)b->d_pv;
        void *ptr[11];
        for (int i = 0; i < 11; ++i) {
            ptr[i] = p;
            p += pv_align(sz[i]);
        }
        hsr::PvBuild &v = b->pv;
        v.tris = mesh->d_tris;
        v.triSlot = mesh->d_triSlot;
        v.T = T;
        std::memcpy(v.o, pos, sizeof(v.o));
        v.R = R;
        v.chunks = chunks;
        v.cap = b->pvCap;
        v.keys = (uint32_t *)ptr[0];
        v.keys2 = (uint32_t *)ptr[1];
        v.slots = (int32_t *)ptr[2];
        v.slots2 = (int32_t *)ptr[3];
        v.part = (float *)ptr[4];
        v.faces = (float4 *)ptr[5];
        v.meta = (int32_t *)ptr[6];
        v.matrix = (int32_t *)ptr[7];
        v.start = (int32_t *)ptr[8];
        v.list = (int2 *)ptr[9];
        v.sortTemp = ptr[10];
        v.sortBytes = sortBytes;
        const Marker mark(b->optMarkers, "hsr primary lists");
        HSR_HIP(hipEventRecord(b->evPv[0], b->stream));
        HSR_HIP(hsr::launch_pv_build(v, b->stream));
        HSR_HIP(hipEventRecord(b->evPv[1], b->stream));
        HSR_HIP(hipMemcpyAsync(b->h_pvMeta, v.meta, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, b->stream));
        b->pvBuilt = true;
        b->pvCheck = true;
        b->pvMesh = mesh;
        b->pvEpoch = b->epoch;
        std::memcpy(b->pvPos, pos, sizeof(b->pvPos));
        ++b->stats.primary_builds;
    }
    a.pvMeta = b->pv.meta;
    a.pvFaces = b->pv.faces;
    a.pvKeys = b->pv.keys2;
    a.pvSlots = b->pv.slots2;
    a.pvStart = b->pv.start;
    a.pvList = b->pv.list;
    a.pvR = R;
    return HSR_OK;
}

// After a batch that followed a build: its device time, and the capacity check.
static hsr_status pv_settle(hsr_backend *b) {
    if (!b->pvCheck) return HSR_OK;
    b->pvCheck = false;
    float ms = 0.0f;
    HSR_HIP(hipEventElapsedTime(&ms, b->evPv[0], b->evPv[1]));
    b->stats.primary_ms += ms;
    if (b->h_pvMeta[1] != 0) {
        const int64_t need = (int64_t)(uint32_t)b->h_pvMeta[2];
        b->pvCap = std::max<int64_t>(2 * b->pvCap, need + need / 4);
        b->pvBuilt = false;
    }
    return HSR_OK;
}

static hipError_t event_wait(hipEvent_t e);

// Adds the reduction and total device time of a run that returned before its frame reduction
// ended, once that has ended.
static hsr_status settle_timing(hsr_backend *b) {
    if (!b->timingPending) return HSR_OK;
    b->timingPending = false;
    HSR_HIP(event_wait(b->ev[3]));
    float reduce = 0.0f, total = 0.0f;
    HSR_HIP(hipEventElapsedTime(&reduce, b->ev[1], b->ev[3]));
    HSR_HIP(hipEventElapsedTime(&total, b->ev[2], b->ev[3]));
    b->stats.reduce_ms += reduce;
    b->stats.total_ms += total;
    return HSR_OK;
}

// Starts the statistics of a new top-level call. A guided training run of an earlier call
// that failed before its next run may still hold unsettled timings; they are settled (and
// so dropped) here, never added to the new call's figures.
static hsr_status reset_stats(hsr_backend *b) {
    hsr_status s = settle_timing(b);
    b->stats = hsr_run_stats{};
    return s;
}

hsr_status hsr_hip_last_stats(hsr_backend *b, hsr_run_stats *out) {
    if (!b || !out) return fail(HSR_ERR_INVALID, "null argument");
    HSR_HIP(hipSetDevice(b->device));
    hsr_status s = settle_timing(b);
    if (s != HSR_OK) return s;
    *out = b->stats;
    return HSR_OK;
}

static hsr_status guide_upload(hsr_guide *g);
static hsr_status guide_queue_stats(hsr_guide *g);

// Waits for an event by polling it: a blocking wait sleeps and wakes tens of microseconds late,
// and the host's next step (the guide, the next batch) waits on the answer.
static hipError_t event_wait(hipEvent_t e) {
    hipError_t r;
    while ((r = hipEventQuery(e)) == hipErrorNotReady) {
    }
    return r;
}

// The graph entry of a batch's launch shape `key`: a cached one, or, when the shape was seen
// before among the last kSeen shapes, one built now from `plan` (with the batch's arguments h
// as the nodes' first parameters). A shape met for the first time returns null (its batch
// launches stage by stage), so shapes that never recur cost no graph. Returns null too when
// building the graph fails; graphs are then off for this backend.
extern "C++" template <class Plan>
static hsr_backend::GraphEntry *batch_graph(hsr_backend *b, const std::vector<int64_t> &key, Plan &&plan,
                                            const hsr::BatchArgs &h) {
    for (auto &g : b->graphs)
        if (g.key == key) {
            g.used = ++b->graphClock;
            return &g;
        }
    bool seen = false;
    for (const auto &k : b->graphSeen) seen = seen || k == key;
    if (!seen) {
        if ((int)b->graphSeen.size() >= hsr_backend::kSeen) b->graphSeen.erase(b->graphSeen.begin());
        b->graphSeen.push_back(key);
        return nullptr;
    }
    hsr_backend::GraphEntry e;
    e.key = key;
    e.plan = plan();
    if (hsr::graph_plan(&e.g, e.plan, h) != hipSuccess) {
        (void)hipGetLastError();
        b->optGraph = false;
        return nullptr;
    }
    if ((int)b->graphs.size() >= hsr_backend::kGraphs) {
        auto lru = std::min_element(b->graphs.begin(), b->graphs.end(),
                                    [](const hsr_backend::GraphEntry &x, const hsr_backend::GraphEntry &y) {
                                        return x.used < y.used;
                                    });
        hsr::destroy_graph(lru->g);
        b->graphs.erase(lru);
    }
    e.used = ++b->graphClock;
    b->graphs.push_back(std::move(e));
    ++b->stats.graph_captures;
    return &b->graphs.back();
}

static void fill_scene(hsr::TraceArgs &a, const hsr_scene *sc) {
    const hsr_rf_config &rf = sc->rfConfig;
    a.startFreq = rf.startFreq;
    a.freqSlope = rf.freqSlope;
    a.adcRate = rf.adcRate;
    a.rampTime = rf.rampTime;
    a.antennaDelay = rf.antennaDelay;
    a.rxCount = sc->rxCount;
    std::memcpy(a.tx.pos, sc->tx.position, sizeof(a.tx.pos));
    std::memcpy(a.tx.rot, sc->tx.rotation, sizeof(a.tx.rot));
    for (int k = 0; k < 3; ++k) a.ego[k] = sc->velocity[k];
    a.chirpFreq = 1.0 / ((double)rf.idleTime + (double)rf.rampTime);
    for (int i = 0; i < HSR_MAX_CHANNELS; ++i) {
        std::memcpy(a.rx[i].pos, sc->rx[i].position, sizeof(a.rx[i].pos));
        std::memcpy(a.rx[i].rot, sc->rx[i].rotation, sizeof(a.rx[i].rot));
    }
}

// One Backend::run; with `accumulate` the statistics add to those of the enclosing
// schedule call instead of replacing them. discardFrame: the caller clears the frame before
// anything can read it (a training iteration of a guiding schedule that clears before each
// iteration), so the wavefront pipeline splats nothing and reduces no frame.
static hsr_status run_impl(hsr_backend *b, const hsr_scene *scene, const hsr_knobs *knobs, int64_t sampleIndexOffset,
                           int64_t budget, float sampleWeight, int32_t isFinalIteration, hsr_guide *guide,
                           const volatile uint8_t *interruptFlag, bool accumulate, bool queueStats = false,
                           bool discardFrame = false) {
    if (!b || !scene || !knobs) return fail(HSR_ERR_INVALID, "null argument");
    if (!b->configured) return fail(HSR_ERR_INVALID, "frame not configured (hsr_hip_configure_frame)");
    if (budget < 0 || sampleIndexOffset < 0) return fail(HSR_ERR_INVALID, "negative budget or offset");
    char passName[96];
    if (b->optMarkers)
        std::snprintf(passName, sizeof(passName), "hsr pass: %lld paths from %lld%s", (long long)budget,
                      (long long)sampleIndexOffset, isFinalIteration ? ", final" : "");
    const Marker passMark(b->optMarkers, b->optMarkers ? passName : "");
    if (knobs->correctPhase && !knobs->filteringSphere)  // the reference's Assert, path.h:318
        return fail(HSR_ERR_INVALID, "correctPhase needs filteringSphere");
    // geometrical optics off: the reference Asserts it for correctPhase (path.h:317) and for the
    // sphere filter (path.h:344), and its Halton sampler has dimensions 0-9 (halton.h:58-67,
    // 1082-1084), which a path at depth 3 uses up (two per bounce from dimension 3)
    if (!knobs->useGeometricalOptics && knobs->correctPhase)
        return fail(HSR_ERR_INVALID, "correctPhase needs geometrical optics");
    if (!knobs->useGeometricalOptics && knobs->filteringSphere)
        return fail(HSR_ERR_INVALID, "the sphere filter needs geometrical optics (set filteringSphere = 0)");
    if (!knobs->useGeometricalOptics && knobs->maxDepth > 3)
        return fail(HSR_ERR_INVALID, "without geometrical optics maxDepth is at most 3 (Halton dimensions 0-9)");
    if (knobs->maxDepth < 0) return fail(HSR_ERR_INVALID, "negative maxDepth");
    if (b->optChain && b->optPipeline != 1)
        return fail(HSR_ERR_UNSUPPORTED, "the reference FP32 chain needs the wavefront pipeline");
    if (scene->rxCount < 1 || scene->rxCount > HSR_MAX_CHANNELS) return fail(HSR_ERR_INVALID, "rxCount out of range");
    if (scene->rxCount > b->fc.channelCount)
        return fail(HSR_ERR_INVALID, "more receivers than frame channels");
    if (knobs->doGuiding && !guide) return fail(HSR_ERR_INVALID, "doGuiding requires a guide");
    if (guide && guide->backend != b) return fail(HSR_ERR_INVALID, "guide belongs to another backend");
    if (!(scene->rfConfig.adcRate > 0.0f)) return fail(HSR_ERR_INVALID, "adcRate must be positive");
    const int T = scene->txCount < 1 ? 1 : scene->txCount;
    if (T > HSR_MAX_TX) return fail(HSR_ERR_INVALID, "txCount out of range");
    if (T * scene->rxCount > b->fc.channelCount)
        return fail(HSR_ERR_INVALID, "txCount * rxCount exceeds the frame's channels");
    const hsr_backend *mesh = b->meshFrom ? b->meshFrom : b;
    const bool motion = mesh->d_vel != nullptr || scene->velocity[0] != 0.0f || scene->velocity[1] != 0.0f ||
                        scene->velocity[2] != 0.0f;
    const int chirps = b->fc.chirpCount;
    const size_t tileBytes = hsr::gather_lds_bytes(b->fc.samplesPerChirp, b->fc.channelCount);
    const bool cube = chirps > 1 || tileBytes > (size_t)b->ldsPerBlock;
    if ((motion || chirps > 1) && b->optPipeline == 0)
        return fail(HSR_ERR_UNSUPPORTED, "motion and multi-chirp frames need the wavefront pipeline");
    if (b->optPipeline == 0 && hsr::trace_lds_bytes(b->fc.samplesPerChirp, b->fc.channelCount) > 65536)
        return fail(HSR_ERR_UNSUPPORTED, "the fused kernel's frame tile exceeds 64 KiB of LDS (samples x channels)");
    if (motion && !(scene->rfConfig.idleTime + scene->rfConfig.rampTime > 0.0f))
        return fail(HSR_ERR_INVALID, "motion needs a positive chirp period (idleTime + rampTime)");
    HSR_HIP(hipSetDevice(b->device));
    if (guide && guide->dirty) {
        hsr_status s = guide_upload(guide);
        if (s != HSR_OK) return s;
    }
    if (guide) guide->statsPending = false;  // this run adds records: a queued download is stale

    const int N = b->fc.samplesPerChirp, C = b->fc.channelCount;
    const int MD = 2 * N * C;  // doubles of one chirp's frame (one LDS tile / partial)
    hsr::TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    fill_scene(a, scene);
    a.onlyIndirect = knobs->onlyIndirect;
    a.maxDepth = knobs->maxDepth;
    a.poDiffraction = knobs->poDiffraction;
    a.doGuiding = knobs->doGuiding;
    a.doFiltering = knobs->doFiltering;
    a.filteringSphere = knobs->filteringSphere;
    a.filterGuiding = knobs->filterGuiding;
    a.primaryOrder = knobs->primaryOrder;
    a.produceDebugImage = knobs->produceDebugImage;
    a.correctPhase = knobs->correctPhase;
    a.geometricalOptics = knobs->useGeometricalOptics != 0;
    a.filteringMin = knobs->filteringMin;
    a.filteringMax = knobs->filteringMax;
    a.filteringRadius = knobs->filteringRadius;
    a.phaseMode = b->optPhaseMode;
    a.refChain = b->optChain;
    a.N = N;
    a.C = C;
    a.sampleWeight = sampleWeight;
    a.isFinal = isFinalIteration;
    const bool noSplat = discardFrame && b->optPipeline == 1;
    a.noSplat = noSplat;
    a.nodes = mesh->d_nodes;
    a.nodesF = mesh->d_nodesF;
    a.tris = mesh->d_tris;
    if (guide) {
        a.gsChild = guide->d_sChild;
        a.gsP0 = guide->d_sP0;
        a.gsSize = guide->host.sampling.size();
        a.gtChild = guide->d_tChild;
        a.gtStat = guide->d_tStat;
        a.gtStride = guide->capacity;
        a.uniformProb = guide->host.uniformProb;
        a.secondMoment = guide->host.secondMoment;
        a.guideOn = 1;
    }
    a.counters = b->d_counters;
    a.M = chirps;
    a.motion = motion;
    a.vel = mesh->d_vel;
    a.frameAcc = cube ? b->d_frame_sum : nullptr;
    a.cube = cube;
    // cube: events go through the LDS tiles of cubePlan when one fits, cubeG workgroups (and
    // partial cubes) per tile, at most kCubePartDoubles of partials
    const int lists = std::min(scene->rxCount, C);
    const hsr::CubeSplat cubePlan =
        cube ? hsr::cube_splat_plan(chirps, N, (size_t)b->ldsPerBlock, (size_t)b->ldsPerCU) : hsr::CubeSplat{};
    const int cubeTiles = cubePlan.lds > 0 ? lists * cubePlan.tilesPerList : 1;
    const int64_t cubeCells = (int64_t)chirps * N * C;
    // doubles of one partial cube and its share of the guard rows
    const int64_t perPart = 2 * cubeCells + hsr::cube_guard_doubles(cubePlan, chirps, lists, 1);
    int cubeG = (int)std::min<int64_t>(std::max(1, (2 * b->cu + cubeTiles - 1) / cubeTiles),
                                       kCubePartDoubles / perPart);
    const unsigned evWant = b->optCubeEvents < 0 ? (b->evCap ? b->evCap : hsr::kCubeEvents)
                                                 : (unsigned)b->optCubeEvents;
    if (cubePlan.lds > 0 && cubeG >= 1 && b->optPipeline == 1 && evWant > 0) {
        // a slot that reserved less room for partial cubes uses fewer of them
        if (b->cubePartsCap >= perPart) cubeG = (int)std::min<int64_t>(cubeG, b->cubePartsCap / perPart);
        hsr_status s = hsr::backend_events(b, (int64_t)cubeG * perPart, evWant);
        if (s != HSR_OK) return s;
        a.events = b->d_events;
        a.evCount = b->d_evCount;
        a.evCap = std::min(b->evCap, evWant) / (unsigned)lists;  // per receiver
    }

    // grid: as many workgroups as fit at once, fewer for small budgets (>= 8 paths per lane)
    if (b->maxGrid == 0) {
        const int per_cu = hsr::trace_blocks_per_cu(N, C, b->optCounters);
        b->maxGrid = b->cu * std::max(1, per_cu);
    }

    // hit record
    const int64_t recPaths = std::min(b->optRecord, budget);
    b->recPathsLast = recPaths;
    const int md = std::max(1, knobs->maxDepth);
    if (recPaths > 0) {
        const int64_t need = recPaths * md * C;
        if (need > b->recCap) {
            dfree(b->d_recHits);
            dfree(b->d_recNee);
            hsr_status s;
            if ((s = dalloc(&b->d_recHits, (size_t)recPaths * md)) != HSR_OK) return s;
            if ((s = dalloc(&b->d_recNee, (size_t)need)) != HSR_OK) return s;
            b->recCap = need;
        }
        HSR_HIP(hipMemsetAsync(b->d_recHits, 0xFF, (size_t)recPaths * md * sizeof(int32_t), b->stream));
        HSR_HIP(hipMemsetAsync(b->d_recNee, 0, (size_t)need, b->stream));
        b->recMaxDepth = md;
        b->recC = C;
    }
    if (b->optCounters) HSR_HIP(hipMemsetAsync(b->d_counters, 0, 16 * sizeof(unsigned long long), b->stream));

    {
        hsr_status s = settle_timing(b);  // before ev[] is recorded again
        if (s != HSR_OK) return s;
    }
    if (!accumulate) b->stats = hsr_run_stats{};
    b->stats.bvh_build_ms = b->bvh_ms;
    double trace_ms = 0.0, reduce_ms = 0.0;
    int64_t launches = 0;
    bool first = true;
    for (int tx = 0; tx < T; ++tx) {
    // MIMO: transmitter tx emits the same paths into virtual channels tx*rxCount + r
    const hsr_antenna &txa = tx == 0 ? scene->tx : scene->txMore[tx - 1];
    std::memcpy(a.tx.pos, txa.position, sizeof(a.tx.pos));
    std::memcpy(a.tx.rot, txa.rotation, sizeof(a.tx.rot));
    a.chBase = tx * scene->rxCount;
    a.pvMeta = nullptr;
    a.pvFaces = nullptr;
    a.pvKeys = nullptr;
    a.pvSlots = nullptr;
    a.pvStart = nullptr;
    a.pvList = nullptr;
    a.pvR = 0;
    if (b->optPipeline == 1 && b->optPrimary && budget >= b->optPrimaryMin && budget > 0 && mesh->d_triSlot) {
        hsr_status s = pv_prepare(b, mesh, txa.position, a);
        if (s != HSR_OK) return s;
    }
    for (int64_t done = 0; done < budget && b->optPipeline == 1;) {
        // wavefront batches: path state in HBM, all stage sizes on the device
        if (interruptFlag && *interruptFlag) return fail(HSR_ERR_INTERRUPTED, "interrupted");
        const int64_t n = std::min({budget - done, b->optPathsPerLaunch, b->wfCap});
        char batchName[96];
        if (b->optMarkers)
            std::snprintf(batchName, sizeof(batchName), "hsr batch: tx %d, %lld paths", tx, (long long)n);
        const Marker batchMark(b->optMarkers, b->optMarkers ? batchName : "");
        const int64_t P = std::min(budget, std::min(b->optPathsPerLaunch, b->wfCap));
        const size_t need = hsr::wavefront_bytes(P, C, knobs->maxDepth);
        if (need > b->wfBytes) {
            if (b->d_wf) (void)hipFree(b->d_wf);
            b->d_wf = nullptr;
            b->wfBytes = 0;
            HSR_HIP(hipMalloc(&b->d_wf, need));
            b->wfBytes = need;
        }
        hsr::WfArgs w;
        std::memset(&w, 0, sizeof(w));
        hsr::WfSet sets[2];
        hsr::wavefront_layout(w, sets, b->d_wf, P, C, knobs->maxDepth);
        w.cells = b->d_cells;
        w.base = done;
        w.n = (int32_t)n;
        w.b0mod = (int32_t)((sampleIndexOffset + done) % hsr::kCells);
        a.offset = sampleIndexOffset;
        if (recPaths > 0) {
            a.recHits = b->d_recHits;
            a.recNee = b->d_recNee;
            a.recPaths = recPaths;
        }
        const int travBlocks = hsr::wavefront_trav_blocks_per_cu(b->optCounters);
        const int gridTrav =
            b->cu * (b->optTravWaves > 0 ? std::max(1, std::min(travBlocks, (b->optTravWaves + 3) / 4)) : travBlocks);
        const int gridWork = b->cu * 4;
        // gather: up to kGatherPerCU resident workgroups per CU (their LDS tiles permitting),
        // and no more workgroups than the batch has paths for (each one adds a tile to reduce)
        const int tilesPerCU =
            cube ? kGatherPerCU
                 : (int)std::max<size_t>(1, std::min<size_t>(kGatherPerCU, (size_t)b->ldsPerCU / tileBytes));
        const int gridGather = (int)std::max<int64_t>(
            1, std::min<int64_t>((int64_t)b->cu * tilesPerCU, (n + hsr::kTraceBlock - 1) / hsr::kTraceBlock));
        const bool training = guide && knobs->doGuiding && !isFinalIteration;
        if (training && !b->d_statRows) {
            hsr_status s = dalloc(&b->d_statRows, (size_t)hsr::kStatRows);
            if (s != HSR_OK) return s;
        }
        HSR_HIP(hipEventRecord(b->ev[0], b->stream));
        if (first) HSR_HIP(hipEventRecord(b->ev[2], b->stream));
        first = false;
        const size_t nev = 4 * (size_t)(knobs->maxDepth + 1);
        while (b->evTrav.size() < nev) {
            hipEvent_t e;
            HSR_HIP(hipEventCreate(&e));
            b->evTrav.push_back(e);
        }
        int nClosest = 0, nShadow = 0;
        hipEvent_t *evC = b->evTrav.data(), *evS = b->evTrav.data() + nev / 2;
        while ((int)b->evFork.size() < knobs->maxDepth + 2) {
            hipEvent_t f, j;
            HSR_HIP(hipEventCreateWithFlags(&f, hipEventDisableTiming));
            HSR_HIP(hipEventCreateWithFlags(&j, hipEventDisableTiming));
            b->evFork.push_back(f);
            b->evJoin.push_back(j);
        }
        // the fused tail from depth K on, when the batch has bounces past K and its LDS fits: K =
        // optTail, or (-1) the first depth k >= 2 whose extension rays in this backend's last
        // batch, scaled to this batch's paths, were at most kTailPaths (depth 2 at the first)
        int tailDepth = b->optTail;
        if (tailDepth < 0) {
            tailDepth = 2;
            if (b->lastBatchPaths > 0)
                while (tailDepth < knobs->maxDepth && tailDepth < b->lastBatchDepths &&
                       (double)b->lastExt[tailDepth] * (double)n / (double)b->lastBatchPaths > (double)kTailPaths)
                    ++tailDepth;
        }
        if (tailDepth < 1 || tailDepth >= knobs->maxDepth ||
            hsr::tail_lds_bytes(N, C, cube || noSplat) > (size_t)b->ldsPerBlock)
            tailDepth = 0;
        ++b->stats.tail_batches[std::min(tailDepth, 7)];
        // a batch without guiding records runs its gather stages beside the path chain (the
        // decoupled plan), and its fused tail adds its frame tiles into partial rows of its own
        const bool decouple = !training;
        // its shadow stages splat (no gather stage): a cube frame's events go to its event
        // lists, a one-chirp frame's into a frame tile when the tile, the event lists and the
        // traversal stack fit a workgroup's LDS (HSR_OPT_SHADOW_SPLAT)
        const bool fuseSplat = b->optShadowSplat &&
                               (cube || hsr::shadow_splat_lds_bytes(N, C) + hsr::kStackShort * 256 * sizeof(int32_t) <=
                                            (size_t)b->ldsPerBlock);
        const int tiles = gridGather * (decouple && tailDepth >= 1 ? 2 : 1);
        const size_t needP = (size_t)tiles * MD;
        if (!cube && needP > b->partialCap) {
            dfree(b->d_partials);
            hsr_status s;
            if ((s = dalloc(&b->d_partials, needP)) != HSR_OK) return s;
            b->partialCap = needP;
        }
        const int S = std::min(64, tiles);
        if (!cube && (size_t)S * MD > b->stageCap) {
            dfree(b->d_stage_sum);
            dfree(b->d_stage_comp);
            hsr_status s;
            if ((s = dalloc(&b->d_stage_sum, (size_t)S * MD)) != HSR_OK) return s;
            if ((s = dalloc(&b->d_stage_comp, (size_t)S * MD)) != HSR_OK) return s;
            b->stageCap = (size_t)S * MD;
        }
        // the batch's set-up launch clears the partial frames (unless nothing is splatted) and
        // the per-slot statistics of a training batch, summed per leaf in a fixed order after it
        const int64_t partialDoubles = cube || noSplat ? 0 : (int64_t)tiles * MD;
        // the batch's launch sequence: replayed from the graph of its launch shape with the
        // batch's arguments, or stage by stage
        HSR_HIP(hsr::prepare_wavefront(N, C, cube, cube || noSplat));
        hsr::BatchArgs host;
        host.a = a;
        host.w = w;
        host.set[0] = sets[0];
        host.set[1] = sets[1];
        host.partials = (double2 *)b->d_partials;
        auto plan = [&] {
            return hsr::plan_wavefront(a, w, gridTrav, gridWork, gridGather, partialDoubles, training,
                                       b->optCounters, evC, evS, tailDepth, decouple, fuseSplat);
        };
        hsr_backend::GraphEntry *graph = nullptr;
        if (b->optGraph) {
            const std::vector<int64_t> key = {n, gridTrav, gridWork, gridGather, partialDoubles, training,
                                              b->optCounters, a.maxDepth, a.onlyIndirect, a.poDiffraction,
                                              a.refChain, a.cube, a.rxCount, a.C, a.N, (int64_t)(intptr_t)evC,
                                              tailDepth, noSplat, a.pvList != nullptr, fuseSplat};
            graph = batch_graph(b, key, plan, host);
        }
        if (graph) {
            HSR_HIP(hsr::replay_plan(graph->g, graph->plan, host, b->stream));
            nClosest = graph->plan.nClosest;
            nShadow = graph->plan.nShadow;
            ++b->stats.graph_launches;
        } else {
            const hsr::BatchPlan p = plan();
            HSR_HIP(hsr::launch_plan(p, host, b->stream, b->useStream2 ? b->stream2 : b->stream, b->evFork.data(),
                                     b->evJoin.data(), b->optMarkers));
            nClosest = p.nClosest;
            nShadow = p.nShadow;
        }
        // the batch's queue sizes to the host (read once the batch is over): the extension rays
        // of each depth
        const int depths = std::min(knobs->maxDepth + 1, hsr_backend::kTrackDepths);
        HSR_HIP(hipMemcpy2DAsync(b->h_ext, sizeof(uint32_t), &w.cnt[0].n_ext.v, sizeof(hsr::WfCounts),
                                 sizeof(uint32_t), depths, hipMemcpyDeviceToHost, b->stream));
        if (a.events && !noSplat)
            HSR_HIP(hsr::launch_cube_splat(cubePlan, a.events, a.evCount, a.evCap, a.chBase, lists, chirps, N, C, cubeG,
                                           (double2 *)b->d_cubeParts, b->d_frame_sum, b->stream));
        if (training)
            HSR_HIP(hsr::launch_stat_reduce(w.statV, w.statL, n, guide->host.training.size(), b->d_statRows,
                                            hsr::kStatRows, guide->d_tStat, b->stream));
        HSR_HIP(hipEventRecord(b->ev[1], b->stream));
        // the caller steps the guide next: its statistics' download goes with the last batch,
        // ahead of the frame reduction, which then runs while the host refines the trees
        const bool statsFirst = queueStats && guide && tx == T - 1 && done + n >= budget;
        if (statsFirst) {
            hsr_status s = guide_queue_stats(guide);
            if (s != HSR_OK) return s;
        }
        // LDS tiles: the workgroup tiles are reduced into the frame; cube: the gather kernel
        // has added into the frame directly
        if (!cube && !noSplat) {
            const Marker reduceMark(b->optMarkers, "hsr reduce");
            HSR_HIP(hsr::launch_reduce(b->d_partials, tiles, MD, b->d_stage_sum, b->d_stage_comp, S,
                                       b->d_frame_sum, b->d_frame_comp, b->stream));
        }
        HSR_HIP(hipEventRecord(b->ev[3], b->stream));
        // the traversal events all precede ev[1] on the stream, and so the statistics' copy
        HSR_HIP(event_wait(statsFirst ? guide->evDl : b->ev[3]));
        {
            hsr_status s = pv_settle(b);
            if (s != HSR_OK) return s;
        }
        for (int k = 0; k < depths; ++k) b->lastExt[k] = b->h_ext[k];
        b->lastBatchDepths = depths;
        b->lastBatchPaths = n;
        float ms1 = 0.0f;
        HSR_HIP(hipEventElapsedTime(&ms1, b->ev[0], b->ev[1]));
        trace_ms += ms1;
        if (statsFirst) {
            b->timingPending = true;
        } else {
            float ms2 = 0.0f;
            HSR_HIP(hipEventElapsedTime(&ms2, b->ev[1], b->ev[3]));
            reduce_ms += ms2;
        }
        launches += 1;
        for (int q = 0; q < nClosest; ++q) {
            float m = 0.0f;
            HSR_HIP(hipEventElapsedTime(&m, evC[2 * q], evC[2 * q + 1]));
            b->stats.closest_ms += m;
        }
        for (int q = 0; q < nShadow; ++q) {
            float m = 0.0f;
            HSR_HIP(hipEventElapsedTime(&m, evS[2 * q], evS[2 * q + 1]));
            b->stats.shadow_ms += m;
        }
        b->stats.closest_launches += nClosest;
        b->stats.shadow_launches += nShadow;
        done += n;
    }
    for (int64_t done = 0; done < budget && b->optPipeline == 0;) {
        if (interruptFlag && *interruptFlag) return fail(HSR_ERR_INTERRUPTED, "interrupted");
        const int64_t n = std::min(budget - done, b->optPathsPerLaunch);
        int grid = (int)std::min<int64_t>(b->maxGrid, (n + 8 * hsr::kTraceBlock - 1) / (8 * hsr::kTraceBlock));
        grid = std::max(grid, 1);
        const int64_t waves = (int64_t)grid * (hsr::kTraceBlock / 64);
        a.offset = sampleIndexOffset + done;
        a.count = n;
        a.chunk = (n + waves - 1) / waves;
        // record: only the first recPaths paths of the whole run, i.e. of the first launch
        if (recPaths > 0 && done == 0) {
            a.recHits = b->d_recHits;
            a.recNee = b->d_recNee;
            a.recPaths = std::min(recPaths, n);
        } else {
            a.recHits = nullptr;
            a.recNee = nullptr;
            a.recPaths = 0;
        }
        const size_t needP = (size_t)grid * MD;
        if (needP > b->partialCap) {
            dfree(b->d_partials);
            hsr_status s;
            if ((s = dalloc(&b->d_partials, needP)) != HSR_OK) return s;
            b->partialCap = needP;
        }
        const int S = std::min(64, grid);
        if ((size_t)S * MD > b->stageCap) {
            dfree(b->d_stage_sum);
            dfree(b->d_stage_comp);
            hsr_status s;
            if ((s = dalloc(&b->d_stage_sum, (size_t)S * MD)) != HSR_OK) return s;
            if ((s = dalloc(&b->d_stage_comp, (size_t)S * MD)) != HSR_OK) return s;
            b->stageCap = (size_t)S * MD;
        }
        a.partials = (double2 *)b->d_partials;
        HSR_HIP(hipEventRecord(b->ev[0], b->stream));
        if (first) HSR_HIP(hipEventRecord(b->ev[2], b->stream));
        first = false;
        HSR_HIP(hsr::launch_trace(a, grid, b->optCounters, b->stream));
        HSR_HIP(hipEventRecord(b->ev[1], b->stream));
        HSR_HIP(hsr::launch_reduce(b->d_partials, grid, MD, b->d_stage_sum, b->d_stage_comp, S, b->d_frame_sum,
                                   b->d_frame_comp, b->stream));
        HSR_HIP(hipEventRecord(b->ev[3], b->stream));
        HSR_HIP(hipEventSynchronize(b->ev[3]));
        float ms1 = 0.0f, ms2 = 0.0f;
        HSR_HIP(hipEventElapsedTime(&ms1, b->ev[0], b->ev[1]));
        HSR_HIP(hipEventElapsedTime(&ms2, b->ev[1], b->ev[3]));
        trace_ms += ms1;
        reduce_ms += ms2;
        launches += 1;
        done += n;
    }
    }  // transmitters
    float total = 0.0f;
    if (budget > 0 && !b->timingPending) HSR_HIP(hipEventElapsedTime(&total, b->ev[2], b->ev[3]));
    b->totalWeight += (double)budget * (double)sampleWeight;
    b->stats.trace_ms += trace_ms;
    b->stats.reduce_ms += reduce_ms;
    b->stats.total_ms += total;
    b->stats.paths += budget;
    b->stats.launches += launches;
    if (b->optCounters) {
        unsigned long long c[16];
        HSR_HIP(hipMemcpyAsync(c, b->d_counters, sizeof(c), hipMemcpyDeviceToHost, b->stream));
        HSR_HIP(hipStreamSynchronize(b->stream));
        b->stats.nodes_visited += (int64_t)(c[0] + c[5] + c[7]);
        b->stats.tris_tested += (int64_t)(c[1] + c[6] + c[8]);
        b->stats.nodes_tail += (int64_t)c[7];
        b->stats.tris_tail += (int64_t)c[8];
        b->stats.nodes_shadow += (int64_t)c[5];
        b->stats.tris_shadow += (int64_t)c[6];
        b->stats.rays_closest += (int64_t)c[2];
        b->stats.rays_shadow += (int64_t)c[3];
        b->stats.events += (int64_t)c[4];
    }
    return HSR_OK;
}

hsr_status hsr_hip_run(hsr_backend *b, const hsr_scene *scene, const hsr_knobs *knobs, int64_t sampleIndexOffset,
                       int64_t budget, float sampleWeight, int32_t isFinalIteration, hsr_guide *guide,
                       const volatile uint8_t *interruptFlag) {
    return run_impl(b, scene, knobs, sampleIndexOffset, budget, sampleWeight, isFinalIteration, guide,
                    interruptFlag, false);
}

hsr_status hsr_hip_fetch_frame(hsr_backend *b, float *out_c64, double *out_c128, double *totalWeight) {
    if (!b) return fail(HSR_ERR_INVALID, "null backend");
    if (!b->configured) return fail(HSR_ERR_INVALID, "frame not configured");
    HSR_HIP(hipSetDevice(b->device));
    HSR_HIP(hipStreamSynchronize(b->stream));
    std::vector<double> s(b->frameDoubles), c(b->frameDoubles);
    HSR_HIP(hipMemcpy(s.data(), b->d_frame_sum, s.size() * sizeof(double), hipMemcpyDeviceToHost));
    HSR_HIP(hipMemcpy(c.data(), b->d_frame_comp, c.size() * sizeof(double), hipMemcpyDeviceToHost));
    const double tw = b->totalWeight;
    const double inv = tw != 0.0 ? 1.0 / tw : 0.0;
    for (int64_t i = 0; i < b->frameDoubles; ++i) {
        const double v = (s[i] - c[i]) * inv;
        if (out_c128) out_c128[i] = v;
        if (out_c64) out_c64[i] = (float)v;
    }
    if (totalWeight) *totalWeight = tw;
    return HSR_OK;
}

// The normalised frame (and, with fft, its forward 3-D DFT) in post, with post2 as the DFT's
// second buffer (both of b->frameDoubles doubles).
static hsr_status post_fill(hsr_backend *b, bool fft, double *post, double *post2) {
    const hsr_frame_config &f = b->fc;
    const double inv = b->totalWeight != 0.0 ? 1.0 / b->totalWeight : 0.0;
    HSR_HIP(hsr::launch_post_normalize(b->d_frame_sum, b->d_frame_comp, inv, b->frameDoubles, post, b->stream));
    if (fft)
        HSR_HIP(hsr::launch_post_fft3((double2 *)post, (double2 *)post2, f.chirpCount, f.samplesPerChirp,
                                      f.channelCount, b->stream));
    return HSR_OK;
}

static hsr_status post_check(hsr_backend *b, bool fft) {
    if (!b->configured) return fail(HSR_ERR_INVALID, "frame not configured");
    const hsr_frame_config &f = b->fc;
    if (fft && (f.chirpCount > hsr::kPostMaxAxis || f.samplesPerChirp > hsr::kPostMaxAxis ||
                f.channelCount > hsr::kPostMaxAxis))
        return fail(HSR_ERR_UNSUPPORTED, "frame axis longer than the device DFT supports");
    return HSR_OK;
}

// The normalised frame (and, with fft, its forward 3-D DFT) in b->d_post, the backend's own
// post-processing buffers (allocated on first use).
static hsr_status post_prepare(hsr_backend *b, bool fft) {
    hsr_status s = post_check(b, fft);
    if (s != HSR_OK) return s;
    HSR_HIP(hipSetDevice(b->device));
    if (b->frameDoubles > b->postCap) {
        HSR_HIP(hipStreamSynchronize(b->stream));
        dfree(b->d_post);
        dfree(b->d_post2);
        b->postCap = 0;
        if ((s = dalloc(&b->d_post, b->frameDoubles)) != HSR_OK) return s;
        if ((s = dalloc(&b->d_post2, b->frameDoubles)) != HSR_OK) return s;
        b->postCap = b->frameDoubles;
    }
    if (!b->d_postScratch) HSR_HIP(hipMalloc(&b->d_postScratch, hsr::post_scratch_bytes()));
    return post_fill(b, fft, b->d_post, b->d_post2);
}

hsr_status hsr_hip_fetch_frame_fft(hsr_backend *b, float *out_c64, double *out_c128) {
    if (!b) return fail(HSR_ERR_INVALID, "null backend");
    hsr_status s = post_prepare(b, true);
    if (s != HSR_OK) return s;
    std::vector<double> v(b->frameDoubles);
    HSR_HIP(hipMemcpyAsync(v.data(), b->d_post, v.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HSR_HIP(hipStreamSynchronize(b->stream));
    for (int64_t i = 0; i < b->frameDoubles; ++i) {
        if (out_c128) out_c128[i] = v[i];
        if (out_c64) out_c64[i] = (float)v[i];
    }
    return HSR_OK;
}

// Frame::argmax and frequencyEstimation over the prepared frame `post` (complex, b->frameDoubles
// doubles) with the argmax scratch `scratch` (post_scratch_bytes()).
static hsr_status peak_search(hsr_backend *b, const double *post, void *scratch, int32_t index[3],
                              double refined[3], double *magnitude) {
    const hsr_frame_config &f = b->fc;
    const int64_t n = b->frameDoubles / 2;
    int64_t *d_out = (int64_t *)((char *)scratch + hsr::post_scratch_bytes() - sizeof(int64_t));
    HSR_HIP(hsr::launch_post_argmax((const double2 *)post, n, scratch, d_out, b->stream));
    int64_t best = 0;
    HSR_HIP(hipMemcpyAsync(&best, d_out, sizeof(best), hipMemcpyDeviceToHost, b->stream));
    HSR_HIP(hipStreamSynchronize(b->stream));
    // Frame::makeIndex (radar.h:497-504) and frequencyEstimation (radar.h:517-531, 631-643):
    // per axis, from the last to the first, the neighbours at -off and +off (wrapping) give
    // b - l/(l+m) when l > r, else b + r/(r+m)
    const int dims[3] = {f.chirpCount, f.samplesPerChirp, f.channelCount};
    int64_t rem = best;
    for (int i = 2; i >= 0; --i) {
        index[i] = (int32_t)(rem % dims[i]);
        rem /= dims[i];
    }
    auto mag = [&](int64_t i, double *out) -> hsr_status {
        double v[2];
        HSR_HIP(hipMemcpy(v, post + 2 * i, sizeof(v), hipMemcpyDeviceToHost));
        *out = std::hypot(v[0], v[1]);
        return HSR_OK;
    };
    double m = 0.0;
    hsr_status s;
    if ((s = mag(best, &m)) != HSR_OK) return s;
    if (magnitude) *magnitude = m;
    int64_t off = 1;
    for (int i = 2; i >= 0; --i) {
        const int bi = index[i], mx = dims[i];
        double l = 0.0, r = 0.0;
        if ((s = mag(bi > 0 ? best - off : best + (int64_t)(mx - 1) * off, &l)) != HSR_OK) return s;
        if ((s = mag(bi < mx - 1 ? best + off : best - (int64_t)(mx - 1) * off, &r)) != HSR_OK) return s;
        if (l + m + r == 0.0) refined[i] = bi;
        else if (l > r) refined[i] = bi - l / (l + m);
        else refined[i] = bi + r / (r + m);
        off *= mx;
    }
    return HSR_OK;
}

hsr_status hsr_hip_frame_peak(hsr_backend *b, int32_t fourier, int32_t index[3], double refined[3],
                              double *magnitude) {
    if (!b || !index || !refined) return fail(HSR_ERR_INVALID, "null argument");
    hsr_status s = post_prepare(b, fourier != 0);
    if (s != HSR_OK) return s;
    return peak_search(b, b->d_post, b->d_postScratch, index, refined, magnitude);
}

hsr_status hsr_hip_frame_peak_noalloc(hsr_backend *b, int32_t fourier, int32_t index[3], double refined[3],
                                      double *magnitude) {
    if (!b || !index || !refined) return fail(HSR_ERR_INVALID, "null argument");
    hsr_status s = post_check(b, fourier != 0);
    if (s != HSR_OK) return s;
    if (b->frameDoubles / 2 > hsr::kNoAllocCells)
        return fail(HSR_ERR_INVALID, "frame of " + std::to_string(b->frameDoubles / 2) +
                                         " cells exceeds the allocation-free peak search's limit of " +
                                         std::to_string(hsr::kNoAllocCells) + " cells");
    HSR_HIP(hipSetDevice(b->device));
    double2 *scr = nullptr;
    HSR_HIP(hsr::noalloc_scratch(b->device, &scr));
    double *post = (double *)scr, *post2 = (double *)(scr + hsr::kNoAllocCells);
    if ((s = post_fill(b, fourier != 0, post, post2)) != HSR_OK) return s;
    return peak_search(b, post, scr + 2 * hsr::kNoAllocCells, index, refined, magnitude);
}

hsr_status hsr_hip_get_hit_record(hsr_backend *b, int32_t *hit_ids, int8_t *nee, int64_t paths, int32_t maxDepth,
                                  int32_t rxCount) {
    if (!b) return fail(HSR_ERR_INVALID, "null backend");
    if (paths > b->recPathsLast || maxDepth != b->recMaxDepth || rxCount != b->recC)
        return fail(HSR_ERR_INVALID, "record shape does not match the last run");
    HSR_HIP(hipSetDevice(b->device));
    HSR_HIP(hipStreamSynchronize(b->stream));
    if (hit_ids)
        HSR_HIP(hipMemcpy(hit_ids, b->d_recHits, (size_t)paths * maxDepth * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (nee) HSR_HIP(hipMemcpy(nee, b->d_recNee, (size_t)paths * maxDepth * rxCount, hipMemcpyDeviceToHost));
    return HSR_OK;
}

// ---------------------------------------------------------------- guiding
// The trees go to the device asynchronously on the backend's stream, after every kernel
// queued there before (which may still read the previous trees): packed into the pinned
// staging area (once the previous upload has left it) as child[ns] | p0[ns] | child[nt] and
// copied in one piece, the statistics' copies cleared by a kernel (and the host's statistics
// copied only when one is nonzero). The host returns at once.
static hsr_status guide_upload(hsr_guide *g) {
    const hsr::GuideTree &s = g->host.sampling, &t = g->host.training;
    if (s.size() > g->capacity || t.size() > g->capacity) return fail(HSR_ERR_INVALID, "guide tree too large");
    hipStream_t st = g->backend->stream;
    const size_t cap = (size_t)g->capacity, ns = (size_t)s.size(), nt = (size_t)t.size();
    double *tStat = (double *)(g->h_stage + 12 * cap);
    HSR_HIP(hipEventSynchronize(g->evUp));
    std::memcpy(g->h_stage, s.child.data(), ns * sizeof(int32_t));
    std::memcpy(g->h_stage + 4 * ns, s.p0.data(), ns * sizeof(float));
    std::memcpy(g->h_stage + 8 * ns, t.child.data(), nt * sizeof(int32_t));
    bool statsZero = true;
    for (size_t i = 0; i < nt && statsZero; ++i) statsZero = t.stat[i] == 0.0;
    HSR_HIP(hsr::launch_zero(g->d_tStat, (int64_t)hsr::kStatCopies * g->capacity, st));
    HSR_HIP(hipMemcpyAsync(g->d_trees, g->h_stage, 8 * ns + 4 * nt, hipMemcpyHostToDevice, st));
    if (!statsZero) {
        std::memcpy(tStat, t.stat.data(), nt * sizeof(double));
        HSR_HIP(hipMemcpyAsync(g->d_tStat, tStat, nt * sizeof(double), hipMemcpyHostToDevice, st));
    }
    HSR_HIP(hipEventRecord(g->evUp, st));
    g->d_sChild = (int32_t *)g->d_trees;
    g->d_sP0 = (float *)(g->d_trees + 4 * ns);
    g->d_tChild = (int32_t *)(g->d_trees + 8 * ns);
    g->dirty = false;
    g->statsPending = false;
    return HSR_OK;
}

// Queues the training statistics' download: the kStatCopies device copies folded into the
// first, in order, then copied into the pinned h_stats (complete at evDl).
static hsr_status guide_queue_stats(hsr_guide *g) {
    const hsr::GuideTree &t = g->host.training;
    hipStream_t s = g->backend->stream;
    HSR_HIP(hsr::launch_stat_fold(g->d_tStat, g->capacity, (int64_t)t.size(), s));
    HSR_HIP(hipMemcpyAsync(g->h_stats, g->d_tStat, t.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HSR_HIP(hipEventRecord(g->evDl, s));
    g->statsPending = true;
    return HSR_OK;
}

// The training statistics into the host tree: the download a run has queued at its end, or a
// new one.
static hsr_status guide_download_stats(hsr_guide *g) {
    hsr::GuideTree &t = g->host.training;
    if (!g->statsPending) {
        hsr_status s = guide_queue_stats(g);
        if (s != HSR_OK) return s;
    }
    HSR_HIP(event_wait(g->evDl));
    std::memcpy(t.stat.data(), g->h_stats, t.size() * sizeof(double));
    g->statsPending = false;
    return HSR_OK;
}

hsr_status hsr_guide_create(hsr_backend *b, float uniformProb, float splitThreshold, int32_t secondMoment,
                            hsr_guide **out) {
    if (!b || !out) return fail(HSR_ERR_INVALID, "null argument");
    if (!(uniformProb >= 0.0f && uniformProb <= 1.0f)) return fail(HSR_ERR_INVALID, "uniformProb outside [0, 1]");
    *out = nullptr;
    HSR_HIP(hipSetDevice(b->device));
    std::unique_ptr<hsr_guide> g(new hsr_guide());
    g->backend = b;
    g->device = b->device;
    g->host.uniformProb = uniformProb;
    g->host.splitThreshold = splitThreshold;
    g->host.secondMoment = secondMoment != 0;
    hsr_status s;
    if ((s = dalloc(&g->d_trees, (size_t)12 * g->capacity)) != HSR_OK) return s;
    if ((s = dalloc(&g->d_tStat, (size_t)hsr::kStatCopies * g->capacity)) != HSR_OK) return s;
    const size_t cap = (size_t)g->capacity;
    HSR_HIP(hipHostMalloc((void **)&g->h_stage, 20 * cap, hipHostMallocDefault));
    HSR_HIP(hipHostMalloc((void **)&g->h_stats, 8 * cap, hipHostMallocDefault));
    HSR_HIP(hipEventCreateWithFlags(&g->evUp, hipEventDisableTiming));
    HSR_HIP(hipEventCreateWithFlags(&g->evDl, hipEventDisableTiming));
    g->host.reset();
    if ((s = guide_upload(g.get())) != HSR_OK) return s;
    *out = g.release();
    return HSR_OK;
}

void hsr_guide_destroy(hsr_guide *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    // the last upload or download may still use the buffers
    if (g->evUp) (void)hipEventSynchronize(g->evUp);
    if (g->evDl) (void)hipEventSynchronize(g->evDl);
    if (g->evUp) (void)hipEventDestroy(g->evUp);
    if (g->evDl) (void)hipEventDestroy(g->evDl);
    if (g->h_stage) (void)hipHostFree(g->h_stage);
    if (g->h_stats) (void)hipHostFree(g->h_stats);
    dfree(g->d_trees);
    dfree(g->d_tStat);
    delete g;
}

hsr_status hsr_guide_reset(hsr_guide *g) {
    if (!g) return fail(HSR_ERR_INVALID, "null guide");
    HSR_HIP(hipSetDevice(g->backend->device));
    g->host.reset();
    return guide_upload(g);
}

hsr_status hsr_guide_step(hsr_guide *g) {
    if (!g) return fail(HSR_ERR_INVALID, "null guide");
    HSR_HIP(hipSetDevice(g->backend->device));
    hsr_status s = guide_download_stats(g);
    if (s != HSR_OK) return s;
    g->host.step();
    return guide_upload(g);
}

hsr_status hsr_guide_sampling_size(hsr_guide *g, int32_t *n) {
    if (!g || !n) return fail(HSR_ERR_INVALID, "null argument");
    *n = g->host.sampling.size();
    return HSR_OK;
}

hsr_status hsr_guide_get_sampling(hsr_guide *g, int32_t *child, float *p0, int32_t capacity) {
    if (!g) return fail(HSR_ERR_INVALID, "null guide");
    const hsr::GuideTree &s = g->host.sampling;
    if (capacity < s.size()) return fail(HSR_ERR_INVALID, "capacity too small");
    if (child) std::memcpy(child, s.child.data(), s.size() * sizeof(int32_t));
    if (p0) std::memcpy(p0, s.p0.data(), s.size() * sizeof(float));
    return HSR_OK;
}

hsr_status hsr_guide_set_sampling(hsr_guide *g, const int32_t *child, const float *p0, int32_t n) {
    if (!g || !child || !p0 || n < 1 || n > g->capacity) return fail(HSR_ERR_INVALID, "invalid sampling tree");
    hsr::GuideTree &s = g->host.sampling;
    s.child.assign(child, child + n);
    s.p0.assign(p0, p0 + n);
    s.stat.assign(n, 0.0);
    s.depth.assign(n, 0);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t c = s.child[i];
        if (c != 0) {
            if (c <= i || c + 1 >= n) return fail(HSR_ERR_INVALID, "child index out of order");
            s.depth[c] = s.depth[c + 1] = s.depth[i] + 1;
        }
    }
    HSR_HIP(hipSetDevice(g->backend->device));
    return guide_upload(g);
}

hsr_status hsr_guide_training_size(hsr_guide *g, int32_t *n) {
    if (!g || !n) return fail(HSR_ERR_INVALID, "null argument");
    *n = g->host.training.size();
    return HSR_OK;
}

hsr_status hsr_guide_get_training(hsr_guide *g, int32_t *child, double *stat, int32_t capacity) {
    if (!g) return fail(HSR_ERR_INVALID, "null guide");
    HSR_HIP(hipSetDevice(g->backend->device));
    hsr_status s = guide_download_stats(g);
    if (s != HSR_OK) return s;
    const hsr::GuideTree &t = g->host.training;
    if (capacity < t.size()) return fail(HSR_ERR_INVALID, "capacity too small");
    if (child) std::memcpy(child, t.child.data(), t.size() * sizeof(int32_t));
    if (stat) std::memcpy(stat, t.stat.data(), t.size() * sizeof(double));
    return HSR_OK;
}

// ---------------------------------------------------------------- PathTracer::run
hsr_status hsr_path_tracer_run(hsr_backend *b, const hsr_scene *scene, const hsr_knobs *knobs, int64_t samples,
                               hsr_guide *guide, const volatile uint8_t *interruptFlag) {
    if (!b || !scene || !knobs) return fail(HSR_ERR_INVALID, "null argument");
    ++b->epoch;  // a new frame: its primary visibility lists are built again
    hsr_status s = hsr_hip_clear_frame(b);  // setup() + clearFrame(), path.h:173-174
    if (s != HSR_OK) return s;
    if ((s = reset_stats(b)) != HSR_OK) return s;
    int64_t offset = 0;
    float weight = 1.0f;
    if (!knobs->doGuiding) return run_impl(b, scene, knobs, 0, samples, 1.0f, 0, nullptr, interruptFlag, true);
    if (!guide) return fail(HSR_ERR_INVALID, "doGuiding requires a guide");
    if ((s = hsr_guide_reset(guide)) != HSR_OK) return s;
    bool isFinal = false;
    int64_t milestone = 16384, remaining = samples;
    while (true) {
        milestone = std::min(milestone, remaining);
        // an iteration followed by clearFrame() leaves nothing readable in the frame unless an
        // interrupt ends the run after it
        const bool more = remaining > milestone;
        s = run_impl(b, scene, knobs, offset, milestone, weight, isFinal, guide, interruptFlag, true, more,
                     more && knobs->clearBeforeIteration && !interruptFlag);
        if (s != HSR_OK) return s;
        offset += milestone;
        remaining -= milestone;
        if (remaining == 0) break;
        milestone *= 2;
        if (remaining < milestone * 2) {
            isFinal = true;
            milestone = remaining;
        }
        if (knobs->clearBeforeIteration) {
            if ((s = hsr_hip_clear_frame(b)) != HSR_OK) return s;
            offset = 0;
        } else {
            weight = 10.0f * weight;
        }
        if ((s = hsr_guide_step(guide)) != HSR_OK) return s;
    }
    return HSR_OK;
}

hsr_status hsr_path_tracer_run_schedule(hsr_backend *b, const hsr_scene *scene, const hsr_knobs *knobs,
                                        int32_t trainingPasses, int64_t trainingPaths, int64_t finalPaths,
                                        hsr_guide *guide, const volatile uint8_t *interruptFlag) {
    if (!b || !scene || !knobs) return fail(HSR_ERR_INVALID, "null argument");
    if (trainingPasses < 0 || trainingPaths < 0 || finalPaths < 0) return fail(HSR_ERR_INVALID, "negative schedule");
    ++b->epoch;  // a new frame: its primary visibility lists are built again
    hsr_status s;
    const bool guiding = knobs->doGuiding && guide;
    if (knobs->doGuiding && !guide) return fail(HSR_ERR_INVALID, "doGuiding requires a guide");
    HSR_HIP(hipSetDevice(b->device));
    if ((s = reset_stats(b)) != HSR_OK) return s;
    if (guiding) {
        if ((s = hsr_guide_reset(guide)) != HSR_OK) return s;
        for (int p = 0; p < trainingPasses; ++p) {
            // a training pass splats nothing unless an interrupt may end the schedule after it
            if (interruptFlag && (s = hsr_hip_clear_frame(b)) != HSR_OK) return s;
            if ((s = run_impl(b, scene, knobs, 0, trainingPaths, 1.0f, 0, guide, interruptFlag, true, true,
                              !interruptFlag)) != HSR_OK)
                return s;
            if ((s = hsr_guide_step(guide)) != HSR_OK) return s;
        }
    }
    if ((s = hsr_hip_clear_frame(b)) != HSR_OK) return s;
    return run_impl(b, scene, knobs, 0, finalPaths, 1.0f, 1, guiding ? guide : nullptr, interruptFlag, true);
}

}  // extern "C"

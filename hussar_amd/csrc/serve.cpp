// Serving runtime of the gfx950 backend (hsr_server_* of include/hussar_hip.h).
//
// Resident scenes are backends that own a mesh (BVH nodes, triangle slots, velocities).
// Slots are backends without a mesh of their own: each owns a HIP stream, wavefront buffers,
// a frame and a guide, reserved when the server is created, and runs a request on a resident
// scene by borrowing that scene's mesh (hsr_backend::meshFrom). One worker thread per slot
// takes requests from a FIFO queue, so requests of different scenes and frame shapes overlap
// on the device. Only the reference's visualizer re-render loop is comparable
// (visualizer/src/Simulator.h:128-183): it restarts one blocking run per frame.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "backend.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    int64_t ticket = 0;
    hsr_request req{};
    std::vector<float> frame;
    hsr_status status = HSR_OK;
    std::string error;
    bool done = false;
    int slot = -1;
    double totalWeight = 0.0;
    Clock::time_point tSubmit, tStart, tDone;
};

struct Scene {
    hsr_backend *owner = nullptr;
    int64_t bytes = 0;
    int pending = 0;  // queued + running requests
    bool closing = false;
};

struct Slot {
    hsr_backend *b = nullptr;
    hsr_guide *g = nullptr;
    std::thread worker;
};

}  // namespace

struct hsr_server {
    hsr_server_config cfg{};
    int64_t slotBytes = 0;
    int64_t frameReserveDoubles = 0;
    std::vector<Slot> slots;
    std::map<int32_t, Scene> scenes;
    int32_t nextScene = 1;
    std::deque<std::shared_ptr<Request>> queue;
    std::map<int64_t, std::shared_ptr<Request>> tickets;
    int64_t nextTicket = 1;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    bool stop = false;
    volatile uint8_t interrupt = 0;
    int64_t hbmUsed = 0;
    // statistics
    int64_t submitted = 0, completed = 0, failed = 0, running = 0;
    std::vector<double> latencies;
    bool haveFirst = false;
    Clock::time_point firstSubmit, lastDone;
};

namespace {

void worker_loop(hsr_server *s, int slotIndex) {
    Slot &slot = s->slots[slotIndex];
    for (;;) {
        std::shared_ptr<Request> r;
        hsr_backend *mesh = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cvWork.wait(lk, [&] { return s->stop || !s->queue.empty(); });
            if (s->stop) return;
            r = s->queue.front();
            s->queue.pop_front();
            mesh = s->scenes[r->req.sceneId].owner;
            r->slot = slotIndex;
            r->tStart = Clock::now();
            ++s->running;
        }
        hsr_backend *b = slot.b;
        b->meshFrom = mesh;
        hsr_status st = hsr_hip_configure_frame(b, &r->req.frame);
        if (st == HSR_OK)
            st = hsr_path_tracer_run(b, &r->req.scene, &r->req.knobs, r->req.paths, slot.g, &s->interrupt);
        double tw = 0.0;
        if (st == HSR_OK) {
            const hsr_frame_config &f = r->req.frame;
            r->frame.resize(2 * (size_t)f.chirpCount * f.samplesPerChirp * f.channelCount);
            st = hsr_hip_fetch_frame(b, r->frame.data(), nullptr, &tw);
        }
        const std::string err = st == HSR_OK ? std::string() : std::string(hsr_last_error());
        b->meshFrom = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            r->status = st;
            r->error = err;
            r->totalWeight = tw;
            r->tDone = Clock::now();
            r->done = true;
            --s->running;
            Scene &sc = s->scenes[r->req.sceneId];
            --sc.pending;
            if (st == HSR_OK) {
                ++s->completed;
                s->latencies.push_back(ms_between(r->tSubmit, r->tDone));
                s->lastDone = r->tDone;
            } else {
                ++s->failed;
            }
            if (sc.closing && sc.pending == 0) {
                s->hbmUsed -= sc.bytes;
                hsr_hip_destroy(sc.owner);
                s->scenes.erase(r->req.sceneId);
            }
        }
        s->cvDone.notify_all();
    }
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

void hsr_server_destroy(hsr_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->interrupt = 1;
        for (auto &r : s->queue) {
            r->status = HSR_ERR_INTERRUPTED;
            r->error = "server destroyed";
            r->done = true;
        }
        s->queue.clear();
    }
    s->cvWork.notify_all();
    for (auto &sl : s->slots)
        if (sl.worker.joinable()) sl.worker.join();
    s->cvDone.notify_all();
    for (auto &sl : s->slots) {
        if (sl.g) hsr_guide_destroy(sl.g);
        if (sl.b) hsr_hip_destroy(sl.b);
    }
    for (auto &kv : s->scenes) hsr_hip_destroy(kv.second.owner);
    delete s;
}

hsr_status hsr_server_create(const hsr_server_config *config, hsr_server **out) {
    if (!config || !out) return hsr::fail(HSR_ERR_INVALID, "null argument");
    *out = nullptr;
    hsr_server_config c = *config;
    if (c.slots == 0) c.slots = 4;
    if (c.queueCapacity == 0) c.queueCapacity = 4096;
    if (c.maxDepth == 0) c.maxDepth = 16;
    if (c.pathsPerBatch == 0) c.pathsPerBatch = int64_t(1) << 22;
    if (c.frameReserveBytes == 0) c.frameReserveBytes = int64_t(64) << 20;
    if (c.slots < 1 || c.slots > 64 || c.queueCapacity < 1 || c.maxDepth < 1 || c.maxDepth > 64 ||
        c.pathsPerBatch < 1024 || c.pathsPerBatch > (int64_t(1) << 26) || c.frameReserveBytes < 4096 ||
        c.hbmBudget < 0)
        return hsr::fail(HSR_ERR_INVALID, "server configuration out of range");
    int ndev = 0;
    HSR_HIP(hipGetDeviceCount(&ndev));
    if (c.device < 0 || c.device >= ndev) return hsr::fail(HSR_ERR_INVALID, "no such HIP device");
    HSR_HIP(hipSetDevice(c.device));
    if (c.hbmBudget == 0) {
        size_t freeB = 0, totalB = 0;
        HSR_HIP(hipMemGetInfo(&freeB, &totalB));
        c.hbmBudget = (int64_t)((double)freeB * 0.9);
    }
    std::unique_ptr<hsr_server> s(new hsr_server());
    s->cfg = c;
    s->frameReserveDoubles = c.frameReserveBytes / 16;  // sum + compensation
    const int64_t wf = (int64_t)hsr::wavefront_bytes(c.pathsPerBatch, HSR_MAX_CHANNELS, c.maxDepth);
    const int64_t events = std::min<int64_t>(hsr::kCubeEvents, 4 * c.pathsPerBatch);
    s->slotBytes = wf + (2 + 4) * s->frameReserveDoubles * 8 + events * (int64_t)sizeof(hsr::CubeEvent) +
                   (int64_t(1) << 26);  // + partial cubes, the cube event list, tiles and small buffers
    if ((int64_t)c.slots * s->slotBytes > c.hbmBudget)
        return hsr::fail(HSR_ERR_OOM, "the slots' reservations exceed the HBM budget");
    s->slots.resize(c.slots);
    for (int i = 0; i < c.slots; ++i) {
        Slot &sl = s->slots[i];
        hsr_status st = hsr_hip_create(nullptr, 0, nullptr, 0, c.device, &sl.b);
        if (st != HSR_OK) {
            hsr_server_destroy(s.release());
            return st;
        }
        // a slot's second stream only while every stream can have a hardware queue of its own
        // (HIP gives a process 4 by default); beyond that the slots' streams would share queues
        sl.b->useStream2 = 2 * c.slots <= 4;
        if ((st = hsr::backend_reserve(sl.b, c.pathsPerBatch, HSR_MAX_CHANNELS, c.maxDepth,
                                       s->frameReserveDoubles)) != HSR_OK ||
            (st = hsr_guide_create(sl.b, 0.1f, 0.005f, 1, &sl.g)) != HSR_OK) {
            const std::string e = hsr_last_error();
            hsr_server_destroy(s.release());
            return hsr::fail(st, e);
        }
    }
    s->hbmUsed = (int64_t)c.slots * s->slotBytes;
    for (int i = 0; i < c.slots; ++i) s->slots[i].worker = std::thread(worker_loop, s.get(), i);
    *out = s.release();
    return HSR_OK;
}

hsr_status hsr_server_open_scene(hsr_server *s, const float *vertices, int64_t vertexCount, const int32_t *indices,
                                 int64_t triangleCount, int32_t *sceneId) {
    if (!s || !sceneId) return hsr::fail(HSR_ERR_INVALID, "null argument");
    *sceneId = 0;
    // admission on the mesh's size before building: nodes (< 2 per triangle) and slots
    const int64_t estimate = std::max<int64_t>(1, triangleCount) * (2 * 64 + 48);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->hbmUsed + estimate > s->cfg.hbmBudget)
            return hsr::fail(HSR_ERR_OOM, "admission refused: the scene does not fit in the HBM budget");
        s->hbmUsed += estimate;  // held while building
    }
    hsr_backend *owner = nullptr;
    hsr_status st = hsr_hip_create(vertices, vertexCount, indices, triangleCount, s->cfg.device, &owner);
    std::lock_guard<std::mutex> lk(s->mu);
    s->hbmUsed -= estimate;
    if (st != HSR_OK) return st;
    Scene sc;
    sc.owner = owner;
    sc.bytes = hsr::backend_mesh_bytes(owner);
    s->hbmUsed += sc.bytes;
    const int32_t id = s->nextScene++;
    s->scenes[id] = sc;
    *sceneId = id;
    return HSR_OK;
}

hsr_status hsr_server_set_velocities(hsr_server *s, int32_t sceneId, const float *velocities, int64_t triangleCount) {
    if (!s) return hsr::fail(HSR_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->scenes.find(sceneId);
    if (it == s->scenes.end() || it->second.closing) return hsr::fail(HSR_ERR_INVALID, "no such scene");
    if (it->second.pending) return hsr::fail(HSR_ERR_BUSY, "the scene has queued or running requests");
    Scene &sc = it->second;
    const hsr_status st = hsr_hip_set_velocities(sc.owner, velocities, triangleCount);
    s->hbmUsed -= sc.bytes;
    sc.bytes = hsr::backend_mesh_bytes(sc.owner);
    s->hbmUsed += sc.bytes;
    return st;
}

hsr_status hsr_server_close_scene(hsr_server *s, int32_t sceneId) {
    if (!s) return hsr::fail(HSR_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->scenes.find(sceneId);
    if (it == s->scenes.end() || it->second.closing) return hsr::fail(HSR_ERR_INVALID, "no such scene");
    if (it->second.pending == 0) {
        s->hbmUsed -= it->second.bytes;
        hsr_hip_destroy(it->second.owner);
        s->scenes.erase(it);
    } else {
        it->second.closing = true;
    }
    return HSR_OK;
}

hsr_status hsr_server_submit(hsr_server *s, const hsr_request *req, int64_t *ticket) {
    if (!s || !req || !ticket) return hsr::fail(HSR_ERR_INVALID, "null argument");
    *ticket = 0;
    const hsr_frame_config &f = req->frame;
    if (req->paths < 0) return hsr::fail(HSR_ERR_INVALID, "negative path count");
    if (f.chirpCount < 1 || f.samplesPerChirp < 1 || f.channelCount < 1 || f.channelCount > HSR_MAX_CHANNELS)
        return hsr::fail(HSR_ERR_INVALID, "frame dimensions out of range");
    if (2 * (int64_t)f.chirpCount * f.samplesPerChirp * f.channelCount > s->frameReserveDoubles)
        return hsr::fail(HSR_ERR_OOM, "admission refused: the frame exceeds the slot's frame reservation");
    if (req->knobs.maxDepth > s->cfg.maxDepth)
        return hsr::fail(HSR_ERR_UNSUPPORTED, "maxDepth exceeds the server's reservation");
    auto r = std::make_shared<Request>();
    r->req = *req;
    r->tSubmit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return hsr::fail(HSR_ERR_INVALID, "server is shutting down");
        auto it = s->scenes.find(req->sceneId);
        if (it == s->scenes.end() || it->second.closing) return hsr::fail(HSR_ERR_INVALID, "no such scene");
        if ((int64_t)s->queue.size() >= s->cfg.queueCapacity) return hsr::fail(HSR_ERR_BUSY, "queue full");
        ++it->second.pending;
        r->ticket = s->nextTicket++;
        s->tickets[r->ticket] = r;
        s->queue.push_back(r);
        ++s->submitted;
        if (!s->haveFirst) {
            s->haveFirst = true;
            s->firstSubmit = r->tSubmit;
        }
        *ticket = r->ticket;
    }
    s->cvWork.notify_one();
    return HSR_OK;
}

hsr_status hsr_server_poll(hsr_server *s, int64_t ticket, int32_t *done) {
    if (!s || !done) return hsr::fail(HSR_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return hsr::fail(HSR_ERR_INVALID, "unknown ticket");
    *done = it->second->done ? 1 : 0;
    return HSR_OK;
}

hsr_status hsr_server_wait(hsr_server *s, int64_t ticket, float *frame_c64, int64_t capacity,
                           hsr_request_info *info) {
    if (!s) return hsr::fail(HSR_ERR_INVALID, "null server");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->tickets.find(ticket);
        if (it == s->tickets.end()) return hsr::fail(HSR_ERR_INVALID, "unknown ticket");
        r = it->second;
        s->cvDone.wait(lk, [&] { return r->done; });
        s->tickets.erase(ticket);
    }
    if (info) {
        info->status = r->status;
        info->slot = r->slot;
        info->queue_ms = r->slot >= 0 ? ms_between(r->tSubmit, r->tStart) : 0.0;
        info->run_ms = r->slot >= 0 ? ms_between(r->tStart, r->tDone) : 0.0;
        info->latency_ms = r->slot >= 0 ? ms_between(r->tSubmit, r->tDone) : 0.0;
        info->totalWeight = r->totalWeight;
    }
    if (r->status != HSR_OK) return hsr::fail(r->status, r->error);
    if (frame_c64) {
        const int64_t n = (int64_t)r->frame.size() / 2;
        if (capacity < n) return hsr::fail(HSR_ERR_INVALID, "frame buffer too small");
        std::memcpy(frame_c64, r->frame.data(), r->frame.size() * sizeof(float));
    }
    return HSR_OK;
}

hsr_status hsr_server_get_stats(hsr_server *s, hsr_server_stats *out) {
    if (!s || !out) return hsr::fail(HSR_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->queued = (int64_t)s->queue.size();
    out->running = s->running;
    out->p50_ms = percentile(s->latencies, 0.50);
    out->p99_ms = percentile(s->latencies, 0.99);
    double sum = 0.0, mx = 0.0;
    for (double v : s->latencies) {
        sum += v;
        mx = std::max(mx, v);
    }
    out->mean_ms = s->latencies.empty() ? 0.0 : sum / (double)s->latencies.size();
    out->max_ms = mx;
    if (s->completed > 0 && s->haveFirst) {
        const double span = ms_between(s->firstSubmit, s->lastDone);
        out->frames_per_s = span > 0.0 ? (double)s->completed / (span * 1e-3) : 0.0;
    }
    out->hbm_used = s->hbmUsed;
    out->hbm_budget = s->cfg.hbmBudget;
    out->scenes = (int32_t)s->scenes.size();
    out->slots = (int32_t)s->slots.size();
    return HSR_OK;
}

hsr_status hsr_server_reset_stats(hsr_server *s) {
    if (!s) return hsr::fail(HSR_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted = s->completed = s->failed = 0;
    s->latencies.clear();
    s->haveFirst = false;
    return HSR_OK;
}

}  // extern "C"

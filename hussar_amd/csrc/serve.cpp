// Serving runtime of the gfx950 backend (hsr_server_* of include/hussar_hip.h).
//
// Resident scenes are backends that own a mesh (BVH nodes, triangle slots, velocities) and,
// once trained, a resident guiding tree. Slots are backends without a mesh of their own: each
// owns a HIP stream, wavefront buffers, a frame and a guide, and runs a request on a resident
// scene by borrowing that scene's mesh (hsr_backend::meshFrom). One worker thread per slot
// takes requests from a FIFO queue, so requests of different scenes and frame shapes overlap
// on the device.
//
// Admission is by footprint: a request needs the device bytes of its own wavefront state,
// frame and tiles (hsr::backend_reserve_bytes for its batch, channels, depth and frame), and
// starts only when a slot is free and those bytes fit in the HBM budget beside the resident
// scenes and the other slots' buffers; otherwise it waits at the head of the queue while
// memory frees up (idle slots give their buffers back first). A request that could never fit
// is refused at submit.
//
// Guiding (hsr_request.guideMode): HSR_GUIDE_PER_REQUEST runs the reference's PathTracer::run
// schedule per request with the slot's own guide; HSR_GUIDE_TRAIN does the same and then keeps
// the trained sampling tree as the scene's resident tree; HSR_GUIDE_SCENE spends the whole
// budget as one final pass with the scene's resident tree (a slot keeps the tree it last
// loaded), or, while the scene has none, trains it like HSR_GUIDE_TRAIN. Only the reference's
// visualizer re-render loop is comparable (visualizer/src/Simulator.h:128-183): it restarts
// one blocking run per frame.
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
    // the resident guiding tree (sampling tree: children, p0); version 0 = none yet
    std::vector<int32_t> gChild;
    std::vector<float> gP0;
    int64_t gVersion = 0;
};

struct Slot {
    hsr_backend *b = nullptr;
    hsr_guide *g = nullptr;
    std::thread worker;
    int64_t bytes = 0;  // device bytes of the slot's work buffers (accounted in hbmUsed)
    bool busy = false;
    // the scene tree the slot's guide holds as its sampling tree (0: none)
    int32_t cacheScene = 0;
    int64_t cacheVersion = 0;
};

}  // namespace

struct hsr_server {
    hsr_server_config cfg{};
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
    int64_t hbmUsed = 0;     // scenes + slots' work buffers
    int64_t treeVersion = 0;  // versions of resident guiding trees
    // statistics
    int64_t submitted = 0, completed = 0, failed = 0, running = 0, rejected = 0, memoryWaits = 0;
    std::vector<double> latencies;
    bool haveFirst = false;
    Clock::time_point firstSubmit, lastDone;
};

namespace {

// Batch, channels, depth and frame of a request, as the slot reserves them.
struct Shape {
    int64_t P;
    int C, maxDepth;
    int64_t frameDoubles;
    bool training;
};

Shape shape_of(const hsr_server *s, const hsr_request &q) {
    const hsr_frame_config &f = q.frame;
    Shape sh;
    sh.P = std::min<int64_t>(std::max<int64_t>(q.paths, 1), s->cfg.pathsPerBatch);
    sh.C = f.channelCount;
    sh.maxDepth = std::max(1, q.knobs.maxDepth);
    sh.frameDoubles = 2LL * f.chirpCount * f.samplesPerChirp * f.channelCount;
    sh.training = q.knobs.doGuiding != 0;
    return sh;
}

int64_t footprint(const hsr_backend *b, const Shape &sh) {
    return hsr::backend_reserve_bytes(b, sh.P, sh.C, sh.maxDepth, sh.frameDoubles, sh.training);
}

// The footprint of a request on a slot that holds no buffers yet.
int64_t fresh_footprint(const hsr_backend *like, const Shape &sh) {
    hsr_backend empty;
    empty.cu = like->cu;
    empty.ldsPerBlock = like->ldsPerBlock;
    return footprint(&empty, sh);
}

// Under s->mu: whether the head of the queue may start on slot `me` now, giving back idle
// slots' buffers if that makes it fit; `grow` receives the bytes the slot's buffers will gain.
bool admit_head(hsr_server *s, int me, const Shape &sh, int64_t &grow) {
    Slot &slot = s->slots[me];
    grow = std::max<int64_t>(0, footprint(slot.b, sh) - slot.bytes);
    if (s->hbmUsed + grow <= s->cfg.hbmBudget) return true;
    for (int i = 0; i < (int)s->slots.size() && s->hbmUsed + grow > s->cfg.hbmBudget; ++i) {
        Slot &o = s->slots[i];
        if (i == me || o.busy || o.bytes == 0) continue;
        hsr::backend_release(o.b);
        s->hbmUsed -= o.bytes;
        o.bytes = 0;
    }
    if (s->hbmUsed + grow <= s->cfg.hbmBudget) return true;
    // this slot's own buffers, if they are not what the request needs anyway
    if (slot.bytes > 0) {
        hsr::backend_release(slot.b);
        s->hbmUsed -= slot.bytes;
        slot.bytes = 0;
        grow = footprint(slot.b, sh);
    }
    return s->hbmUsed + grow <= s->cfg.hbmBudget;
}

// The request on the slot: the reference's schedule with the slot's guide, or one final pass
// with the scene's resident tree. Returns whether the slot's guide now holds a tree to keep as
// the scene's (trained), and the status.
hsr_status run_request(hsr_server *s, Slot &slot, const Request &r, const std::vector<int32_t> *child,
                       const std::vector<float> *p0, int64_t version, bool &trained) {
    hsr_backend *b = slot.b;
    const hsr_request &q = r.req;
    trained = false;
    hsr_status st = hsr_hip_configure_frame(b, &q.frame);
    if (st != HSR_OK) return st;
    if (q.knobs.doGuiding && child) {
        if (slot.cacheScene != q.sceneId || slot.cacheVersion != version) {
            slot.cacheScene = 0;
            st = hsr_guide_set_sampling(slot.g, child->data(), p0->data(), (int32_t)child->size());
            if (st != HSR_OK) return st;
            slot.cacheScene = q.sceneId;
            slot.cacheVersion = version;
        }
        return hsr_hip_run(b, &q.scene, &q.knobs, 0, q.paths, 1.0f, 1, slot.g, &s->interrupt);
    }
    slot.cacheScene = 0;  // the schedule resets and retrains the slot's guide
    st = hsr_path_tracer_run(b, &q.scene, &q.knobs, q.paths, slot.g, &s->interrupt);
    trained = st == HSR_OK && q.knobs.doGuiding;
    return st;
}

void worker_loop(hsr_server *s, int slotIndex) {
    Slot &slot = s->slots[slotIndex];
    for (;;) {
        std::shared_ptr<Request> r;
        hsr_backend *mesh = nullptr;
        std::vector<int32_t> child;
        std::vector<float> p0;
        int64_t version = 0;
        bool useTree = false, keepTree = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            bool waited = false;
            for (;;) {
                if (s->stop) return;
                if (!s->queue.empty()) {
                    int64_t grow = 0;
                    const Shape sh = shape_of(s, s->queue.front()->req);
                    if (admit_head(s, slotIndex, sh, grow)) {
                        s->hbmUsed += grow;  // held from now; settled after the reservation
                        slot.bytes += grow;
                        break;
                    }
                    if (s->running == 0) {
                        // nothing will free memory: the scenes opened since its submit leave no
                        // room for this request
                        std::shared_ptr<Request> q = s->queue.front();
                        s->queue.pop_front();
                        q->status = HSR_ERR_OOM;
                        q->error = "admission refused: the request's footprint exceeds the HBM budget";
                        q->tDone = Clock::now();
                        q->done = true;
                        ++s->failed;
                        ++s->rejected;
                        Scene &qs = s->scenes[q->req.sceneId];
                        if (--qs.pending == 0 && qs.closing) {
                            s->hbmUsed -= qs.bytes;
                            hsr_hip_destroy(qs.owner);
                            s->scenes.erase(q->req.sceneId);
                        }
                        s->cvDone.notify_all();
                        continue;
                    }
                    if (!waited) {
                        waited = true;
                        ++s->memoryWaits;
                    }
                }
                s->cvWork.wait(lk);
            }
            r = s->queue.front();
            s->queue.pop_front();
            Scene &sc = s->scenes[r->req.sceneId];
            mesh = sc.owner;
            const int mode = r->req.guideMode;
            if (r->req.knobs.doGuiding && mode == HSR_GUIDE_SCENE && sc.gVersion > 0) {
                useTree = true;
                version = sc.gVersion;
                if (slot.cacheScene != r->req.sceneId || slot.cacheVersion != version) {
                    child = sc.gChild;
                    p0 = sc.gP0;
                }
            }
            keepTree = r->req.knobs.doGuiding &&
                       (mode == HSR_GUIDE_TRAIN || (mode == HSR_GUIDE_SCENE && sc.gVersion == 0));
            r->slot = slotIndex;
            r->tStart = Clock::now();
            slot.busy = true;
            ++s->running;
        }
        hsr_backend *b = slot.b;
        b->meshFrom = mesh;
        const Shape sh = shape_of(s, r->req);
        hsr_status st = hsr::backend_reserve(b, sh.P, sh.C, sh.maxDepth, sh.frameDoubles);
        bool trained = false;
        if (st == HSR_OK) {
            // a cached tree needs no copy: run_request finds it in the slot's guide
            const bool cached = useTree && child.empty();
            st = run_request(s, slot, *r, useTree ? (cached ? &slot.g->host.sampling.child : &child) : nullptr,
                             useTree ? (cached ? &slot.g->host.sampling.p0 : &p0) : nullptr, version, trained);
        }
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
            // the slot's buffers as they are now
            const int64_t now = hsr::backend_device_bytes(b);
            s->hbmUsed += now - slot.bytes;
            slot.bytes = now;
            slot.busy = false;
            r->status = st;
            r->error = err;
            r->totalWeight = tw;
            r->tDone = Clock::now();
            r->done = true;
            --s->running;
            Scene &sc = s->scenes[r->req.sceneId];
            --sc.pending;
            if (st == HSR_OK && trained && keepTree &&
                (r->req.guideMode == HSR_GUIDE_TRAIN || sc.gVersion == 0)) {
                sc.gChild = slot.g->host.sampling.child;
                sc.gP0 = slot.g->host.sampling.p0;
                sc.gVersion = ++s->treeVersion;
            }
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
        s->cvWork.notify_all();  // memory may have freed up for a waiting head
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
    if (c.slots < 1 || c.slots > 64 || c.queueCapacity < 1 || c.maxDepth < 1 || c.maxDepth > 64 ||
        c.pathsPerBatch < 1024 || c.pathsPerBatch > (int64_t(1) << 26) || c.frameReserveBytes < 0 ||
        c.hbmBudget < 0 || c.graphs < 0 || c.graphs > 1)
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
        sl.b->optGraph = c.graphs != 0;
        // a slot's work buffers are reserved by the requests it runs (admission by footprint)
        if ((st = hsr_guide_create(sl.b, 0.1f, 0.005f, 1, &sl.g)) != HSR_OK) {
            const std::string e = hsr_last_error();
            hsr_server_destroy(s.release());
            return hsr::fail(st, e);
        }
    }
    s->hbmUsed = 0;
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
    if (req->guideMode < HSR_GUIDE_PER_REQUEST || req->guideMode > HSR_GUIDE_TRAIN)
        return hsr::fail(HSR_ERR_INVALID, "unknown guideMode");
    if (s->frameReserveDoubles > 0 &&
        2 * (int64_t)f.chirpCount * f.samplesPerChirp * f.channelCount > s->frameReserveDoubles) {
        std::lock_guard<std::mutex> lk(s->mu);
        ++s->rejected;
        return hsr::fail(HSR_ERR_OOM, "admission refused: the frame exceeds the server's largest frame");
    }
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
        // a request whose footprint does not fit beside the resident scenes even with every
        // slot's buffers given back can never start
        int64_t scenesBytes = 0;
        for (const auto &kv : s->scenes) scenesBytes += kv.second.bytes;
        if (scenesBytes + fresh_footprint(s->slots[0].b, shape_of(s, *req)) > s->cfg.hbmBudget) {
            ++s->rejected;
            return hsr::fail(HSR_ERR_OOM, "admission refused: the request's footprint exceeds the HBM budget");
        }
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

hsr_status hsr_server_request_bytes(hsr_server *s, const hsr_request *req, int64_t *bytes) {
    if (!s || !req || !bytes) return hsr::fail(HSR_ERR_INVALID, "null argument");
    *bytes = fresh_footprint(s->slots[0].b, shape_of(s, *req));
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
    out->rejected = s->rejected;
    out->memory_waits = s->memoryWaits;
    return HSR_OK;
}

hsr_status hsr_server_reset_stats(hsr_server *s) {
    if (!s) return hsr::fail(HSR_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted = s->completed = s->failed = s->rejected = s->memoryWaits = 0;
    s->latencies.clear();
    s->haveFirst = false;
    return HSR_OK;
}

}  // extern "C"

/*
 * hussar_hip.hpp — header-only C++ front end of the gfx950 backend.
 *
 * It keeps the shape of libhussar's host API so that code written against the reference
 * switches backends by changing one type:
 *
 *   hussar::hip::TriangleMesh mesh;  mesh.addBox(...);            (core/mesh.h:30-54)
 *   hussar::hip::Scene scene;        scene.rx = NFAntenna{...};   (core/scene.h:10-16)
 *   hussar::hip::PathTracer integrator;                           (integrators/path.h:133-224)
 *   integrator.configureFrame(frameConfig);                       (core/integrator.h:60-62)
 *   hussar::hip::Backend backend{mesh, integrator};               (arch/gpu.h:18-20)
 *   integrator.run(backend, scene, sampleCount);                  (path.h:171-224)
 *   auto frame = integrator.fetchFrame();                         (path.h:453-455)
 *
 * Backend and PathTracer also accept the reference's own TriangleMesh / Scene / NFAntenna
 * types (Eigen vectors and matrices): the templates read `vertexBuffer`, `indexBuffer`,
 * `rfConfig`, `rx`, `tx`, `position()[i]` and `rotation()(r, c)` only.
 *
 * Errors: every failed call throws hussar::hip::Error carrying the C status and
 * hsr_last_error()'s message. An interrupt (the caller's flag raised) is not an error: as
 * in the reference, run() returns with a partial frame.
 *
 * Ownership (as in the reference, gpu.h:21, cpu.cpp:36-40): the backend copies the mesh and
 * keeps a reference to the integrator, which must outlive it. The frame and the guiding
 * trees live in device memory next to the backend the integrator last ran on.
 */
#ifndef HUSSAR_HIP_HPP
#define HUSSAR_HIP_HPP

#include "hussar_hip.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace hussar {
namespace hip {

class Error : public std::runtime_error {
public:
    Error(hsr_status s, const std::string &what) : std::runtime_error(what), status(s) {}
    hsr_status status;
};

inline void check(hsr_status s) {
    if (s != HSR_OK) throw Error(s, hsr_last_error());
}

// ------------------------------------------------------------------ value types (H17)
struct Vector3f {
    float v[3] = {0.0f, 0.0f, 0.0f};
    Vector3f() = default;
    Vector3f(float x, float y, float z) : v{x, y, z} {}
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    float operator[](int i) const { return v[i]; }
    float &operator[](int i) { return v[i]; }
    friend Vector3f operator+(const Vector3f &a, const Vector3f &b) {
        return {a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]};
    }
    friend Vector3f operator-(const Vector3f &a, const Vector3f &b) {
        return {a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]};
    }
    friend Vector3f operator*(float s, const Vector3f &a) { return {s * a.v[0], s * a.v[1], s * a.v[2]}; }
};

// Row-major 3x3 matrix; m(r, c) as Eigen's Matrix33f.
struct Matrix33f {
    float m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float operator()(int r, int c) const { return m[3 * r + c]; }
    float &operator()(int r, int c) { return m[3 * r + c]; }
    static Matrix33f rows(float a, float b, float c, float d, float e, float f, float g, float h, float i) {
        Matrix33f r;
        const float x[9] = {a, b, c, d, e, f, g, h, i};
        std::copy(x, x + 9, r.m);
        return r;
    }
    // Rotation by `radians` about +y (Eigen::AngleAxisf(radians, Vector3f::UnitY())).
    static Matrix33f rotationY(float radians) {
        const float c = std::cos(radians), s = std::sin(radians);
        return rows(c, 0, s, 0, 1, 0, -s, 0, c);
    }
    friend Matrix33f operator*(const Matrix33f &a, const Matrix33f &b) {
        Matrix33f r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
        return r;
    }
    friend Vector3f operator*(const Matrix33f &a, const Vector3f &v) {
        return {a(0, 0) * v[0] + a(0, 1) * v[1] + a(0, 2) * v[2], a(1, 0) * v[0] + a(1, 1) * v[1] + a(1, 2) * v[2],
                a(2, 0) * v[0] + a(2, 1) * v[1] + a(2, 2) * v[2]};
    }
    float determinant() const {
        return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
               m[2] * (m[3] * m[7] - m[4] * m[6]);
    }
};

// core/mesh.h:13-54: vertex and index buffers; a quad (a, b, c) adds the vertices a, a+b,
// a+b+c, a+c and the triangles (0,1,2), (0,2,3); a box adds its six faces.
struct TriangleMesh {
    struct IndexTriplet {
        int v0, v1, v2;
    };
    std::vector<Vector3f> vertexBuffer;
    std::vector<IndexTriplet> indexBuffer;

    void addQuad(const Vector3f &a, const Vector3f &b, const Vector3f &c) {
        const int i = (int)vertexBuffer.size();
        vertexBuffer.insert(vertexBuffer.end(), {a, a + b, a + b + c, a + c});
        indexBuffer.push_back({i, i + 1, i + 2});
        indexBuffer.push_back({i, i + 2, i + 3});
    }
    void addBox(const Vector3f &lo, const Vector3f &hi) {
        Vector3f d = hi - lo;
        addQuad(lo, {0, d.y(), 0}, {0, 0, d.z()});
        addQuad(lo, {0, 0, d.z()}, {d.x(), 0, 0});
        addQuad(lo, {d.x(), 0, 0}, {0, d.y(), 0});
        d = lo - hi;
        addQuad(hi, {0, d.y(), 0}, {d.x(), 0, 0});
        addQuad(hi, {0, 0, d.z()}, {0, d.y(), 0});
        addQuad(hi, {d.x(), 0, 0}, {0, 0, d.z()});
    }
};

// radar.h:113-136
struct RFConfig {
    float startFreq = 0, freqSlope = 0, adcRate = 0, idleTime = 0, rampTime = 0, antennaDelay = 0;
    float bandwidth() const { return freqSlope * rampTime; }
    float chirpFrequency() const { return 1.f / (idleTime + rampTime); }
};

// radar.h:141-164
struct FrameConfig {
    int chirpCount = 1, samplesPerChirp = 0, channelCount = 1;
    size_t sampleCount() const { return (size_t)chirpCount * samplesPerChirp * channelCount; }
};

// The AWR1443 pattern (emitter.h:109-137) is the only angular distribution of the path.
struct AWRAngularDistribution {};

// emitter.h:159-214 (position and rotation; the pattern is fixed).
class NFAntenna {
public:
    NFAntenna() = default;
    NFAntenna(const Vector3f &position, const Matrix33f &rotation, const AWRAngularDistribution & = {})
        : m_position(position), m_rotation(rotation) {
        if (!(std::abs(std::abs(rotation.determinant()) - 1) < 1e-2f))
            throw Error(HSR_ERR_INVALID, "supplied rotation matrix is not a rotation matrix");
    }
    const Vector3f &position() const { return m_position; }
    const Matrix33f &rotation() const { return m_rotation; }
    Vector3f &position() { return m_position; }
    Matrix33f &rotation() { return m_rotation; }

private:
    Vector3f m_position;
    Matrix33f m_rotation;
};

// scene.h:10-16, with the extensions of hsr_scene: when rxArray is not empty its antennas
// are the receivers (otherwise `rx`); when txArray is not empty its antennas are the MIMO
// transmitters (otherwise `tx`), the pair (t, r) feeding channel t * receivers + r;
// velocity is the platform velocity of all antennas [m/s].
struct Scene {
    RFConfig rfConfig;
    NFAntenna rx;
    NFAntenna tx;
    std::vector<NFAntenna> rxArray;
    std::vector<NFAntenna> txArray;
    Vector3f velocity;
};

// radar.h:169-... : [chirp][sample][channel] row-major complex64 (makeIndex, radar.h:487-494).
struct Frame {
    FrameConfig config;
    std::vector<std::complex<float>> data;
    size_t sampleCount() const { return data.size(); }
    std::complex<float> operator()(size_t i) const { return data[i]; }
    size_t makeIndex(int chirp, int sample, int channel) const {
        return ((size_t)chirp * config.samplesPerChirp + sample) * config.channelCount + channel;
    }
    // radar.h:534-546: index of the largest |value|; the first of equal maxima wins.
    size_t argmax() const {
        size_t best = 0;
        float bv = data.empty() ? 0.0f : std::abs(data[0]);
        for (size_t s = 1; s < data.size(); ++s) {
            const float c = std::abs(data[s]);
            if (c > bv) {
                bv = c;
                best = s;
            }
        }
        return best;
    }
};

// The raw .SIM layout of examples/src/utils.h:29-40: the frame's complex64 values in index
// order, (re, im) float pairs, frames appended one after another.
inline void writeFrameToFile(std::ostream &file, const Frame &frame) {
    file.write(reinterpret_cast<const char *>(frame.data.data()),
               (std::streamsize)(frame.data.size() * sizeof(std::complex<float>)));
    file.flush();
}

inline void saveFrame(const Frame &frame, const std::string &path) {
    std::ofstream file(path + ".SIM", std::ios::binary);
    writeFrameToFile(file, frame);
}

// hussar::WavefrontFile (io/wavefront.h:12-64): `v x y z` and `f a b c` lines only; face
// indices are 1-based and relative to this file's vertices; a face uses its first three
// indices; `#` comment lines, `s` lines and empty lines are skipped, other commands are
// reported and skipped. A face token that is not a plain integer (v/vt/vn syntax) throws
// instead of stalling the reader. Works with any mesh whose vertexBuffer holds 3 floats
// (operator[]) and whose indexBuffer holds {v0, v1, v2}.
class WavefrontFile {
public:
    explicit WavefrontFile(const std::string &path) : m_path(path) {}

    template <class Mesh>
    void read(Mesh &mesh) {
        std::ifstream in(m_path);
        if (!in) throw Error(HSR_ERR_INVALID, "cannot open wavefront file " + m_path);
        const int offset = (int)mesh.vertexBuffer.size();
        std::string line;
        for (int lineno = 1; std::getline(in, line); ++lineno) {
            if (!line.empty() && line[0] == '#') continue;
            std::istringstream ls(line);
            std::string cmd;
            if (!(ls >> cmd)) continue;
            if (cmd == "v") {
                typename std::decay<decltype(mesh.vertexBuffer[0])>::type v{};
                for (int i = 0; i < 3; ++i) {
                    float x = 0.0f;
                    if (!(ls >> x)) throw Error(HSR_ERR_INVALID, where(lineno) + "a vertex needs three numbers");
                    v[i] = x;
                }
                mesh.vertexBuffer.push_back(v);
            } else if (cmd == "f") {
                int idx[3];
                for (int i = 0; i < 3; ++i) {
                    std::string tok;
                    if (!(ls >> tok)) throw Error(HSR_ERR_INVALID, where(lineno) + "a face needs three indices");
                    size_t used = 0;
                    try {
                        idx[i] = std::stoi(tok, &used);
                    } catch (const std::exception &) {
                        used = 0;
                    }
                    if (used == 0 || used != tok.size())
                        throw Error(HSR_ERR_INVALID,
                                    where(lineno) + "face indices must be plain integers (no v/vt/vn syntax)");
                    idx[i] += offset - 1;
                }
                mesh.indexBuffer.push_back({idx[0], idx[1], idx[2]});
            } else if (cmd == "s") {
            } else {
                std::cout << "unsupported wavefront command: " << cmd << std::endl;
            }
        }
    }

private:
    std::string where(int lineno) const { return m_path + ":" + std::to_string(lineno) + ": "; }
    std::string m_path;
};

namespace detail {
template <class V>
inline void put3(float *dst, const V &v) {
    for (int i = 0; i < 3; ++i) dst[i] = static_cast<float>(v[i]);
}
template <class A>
inline hsr_antenna antenna(const A &a) {
    hsr_antenna r{};
    put3(r.position, a.position());
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.rotation[3 * i + j] = static_cast<float>(a.rotation()(i, j));
    return r;
}
template <class S, class = void>
struct has_rx_array : std::false_type {};
template <class S>
struct has_rx_array<S, std::void_t<decltype(std::declval<const S &>().rxArray.size())>> : std::true_type {};
template <class S, class = void>
struct has_tx_array : std::false_type {};
template <class S>
struct has_tx_array<S, std::void_t<decltype(std::declval<const S &>().txArray.size())>> : std::true_type {};
template <class S, class = void>
struct has_velocity : std::false_type {};
template <class S>
struct has_velocity<S, std::void_t<decltype(std::declval<const S &>().velocity[0])>> : std::true_type {};

template <class S>
inline hsr_scene scene(const S &s) {
    hsr_scene r{};
    const auto &rf = s.rfConfig;
    r.rfConfig = {static_cast<float>(rf.startFreq), static_cast<float>(rf.freqSlope), static_cast<float>(rf.adcRate),
                  static_cast<float>(rf.idleTime),  static_cast<float>(rf.rampTime),  static_cast<float>(rf.antennaDelay)};
    r.tx = antenna(s.tx);
    r.rx[0] = antenna(s.rx);
    r.rxCount = 1;
    if constexpr (has_rx_array<S>::value) {
        if (!s.rxArray.empty()) {
            if (s.rxArray.size() > HSR_MAX_CHANNELS) throw Error(HSR_ERR_INVALID, "more receivers than channels");
            for (size_t i = 0; i < s.rxArray.size(); ++i) r.rx[i] = antenna(s.rxArray[i]);
            r.rxCount = (int32_t)s.rxArray.size();
        }
    }
    r.txCount = 1;
    if constexpr (has_tx_array<S>::value) {
        if (!s.txArray.empty()) {
            if (s.txArray.size() > HSR_MAX_TX) throw Error(HSR_ERR_INVALID, "more transmitters than supported");
            r.tx = antenna(s.txArray[0]);
            for (size_t i = 1; i < s.txArray.size(); ++i) r.txMore[i - 1] = antenna(s.txArray[i]);
            r.txCount = (int32_t)s.txArray.size();
        }
    }
    if constexpr (has_velocity<S>::value) put3(r.velocity, s.velocity);
    return r;
}
}  // namespace detail

class Backend;

// ------------------------------------------------------------------ PathTracer (path.h:133-475)
class PathTracer {
public:
    // path.h:153-169, same names and defaults
    bool onlyIndirect = true;
    int maxDepth = 10;
    bool useGeometricalOptics = true;
    bool poDiffraction = false;
    bool doGuiding = true;
    bool clearBeforeIteration = true;
    bool correctPhase = false;
    bool filterGuiding = true;
    bool doFiltering = true;
    bool filteringSphere = true;
    float filteringMin = 600;
    float filteringMax = 900;
    float filteringRadius = 160;
    // integrator.h:58 (the debug image is not produced; see hsr_knobs)
    bool produceDebugImage = false;
    // primary-sample dimension order: 0 as g++ builds halton.h:24-27, 1 as clang does
    int primaryOrder = 0;

    PathTracer() = default;
    PathTracer(const PathTracer &) = delete;
    PathTracer &operator=(const PathTracer &) = delete;
    ~PathTracer() { releaseGuide(); }

    void configureFrame(const FrameConfig &config) {
        if (config.chirpCount < 1 || config.samplesPerChirp < 1 || config.channelCount < 1 ||
            config.channelCount > HSR_MAX_CHANNELS)
            throw Error(HSR_ERR_INVALID, "invalid frame configuration");
        m_frameConfig = config;
        ++m_frameGeneration;
    }
    const FrameConfig &frameConfig() const { return m_frameConfig; }

    // path.h:171-224: the doubling guiding schedule, issued as backend runs
    template <class SceneT>
    void run(Backend &backend, const SceneT &scene, long samples, bool *interruptFlag = nullptr);

    // path.h:453-455: frame / totalWeight, complex64
    Frame fetchFrame() const;
    // the same values before rounding to float, complex128
    std::vector<std::complex<double>> fetchFrameDouble() const;
    double totalWeight() const;

    hsr_knobs knobs() const {
        hsr_knobs k;
        hsr_knobs_default(&k);
        k.onlyIndirect = onlyIndirect;
        k.maxDepth = maxDepth;
        k.useGeometricalOptics = useGeometricalOptics;
        k.poDiffraction = poDiffraction;
        k.doGuiding = doGuiding;
        k.clearBeforeIteration = clearBeforeIteration;
        k.correctPhase = correctPhase;
        k.filterGuiding = filterGuiding;
        k.doFiltering = doFiltering;
        k.filteringSphere = filteringSphere;
        k.filteringMin = filteringMin;
        k.filteringMax = filteringMax;
        k.filteringRadius = filteringRadius;
        k.primaryOrder = primaryOrder;
        k.produceDebugImage = produceDebugImage;
        return k;
    }

private:
    friend class Backend;

    FrameConfig m_frameConfig{};
    long m_frameGeneration = 0;
    // state the reference keeps in PathTracer and its backends read (path.h:138-141,458-464)
    long sampleIndexOffset = 0;
    float currentSampleWeight = 1.f;
    bool isFinalIteration = false;
    Backend *m_backend = nullptr;   // the backend holding the frame and the guide
    long m_syncedGeneration = -1;   // frame generation configured on m_backend
    hsr_guide *m_guide = nullptr;

    hsr_backend *handle() const;
    void bind(Backend &b);
    void unbind(Backend &b) {
        if (m_backend == &b) {
            releaseGuide();
            m_backend = nullptr;
            m_syncedGeneration = -1;
        }
    }
    void releaseGuide() {
        if (m_guide) hsr_guide_destroy(m_guide);
        m_guide = nullptr;
    }
    void clearFrame() { check(hsr_hip_clear_frame(handle())); }
};

// ------------------------------------------------------------------ Backend (arch/gpu.h:17-47)
class Backend {
public:
    template <class Mesh>
    Backend(const Mesh &mesh, PathTracer &integrator, int device = 0) : m_integrator(integrator) {
        static_assert(sizeof(mesh.vertexBuffer[0]) == 3 * sizeof(float), "vertices must be 3 packed floats");
        static_assert(sizeof(mesh.indexBuffer[0]) == 3 * sizeof(int32_t), "indices must be 3 packed ints");
        check(hsr_hip_create(reinterpret_cast<const float *>(mesh.vertexBuffer.data()),
                             (int64_t)mesh.vertexBuffer.size(),
                             reinterpret_cast<const int32_t *>(mesh.indexBuffer.data()),
                             (int64_t)mesh.indexBuffer.size(), device, &m_handle));
        integrator.bind(*this);
    }
    Backend(const Backend &) = delete;
    Backend &operator=(const Backend &) = delete;
    ~Backend() {
        m_integrator.unbind(*this);
        hsr_hip_destroy(m_handle);
    }

    // Backend::run (gpu.h:26-32): the per-path work for Halton indices
    // sampleIndexOffset .. +budget-1 of the integrator's current iteration.
    template <class SceneT>
    void run(const SceneT &scene, long budget, bool *interruptFlag = nullptr) {
        static_assert(sizeof(bool) == sizeof(uint8_t), "the interrupt flag is read as one byte");
        PathTracer &pt = m_integrator;
        pt.bind(*this);
        const hsr_scene s = detail::scene(scene);
        const hsr_knobs k = pt.knobs();
        const hsr_status st = hsr_hip_run(m_handle, &s, &k, pt.sampleIndexOffset, budget, pt.currentSampleWeight,
                                          pt.isFinalIteration, k.doGuiding ? pt.m_guide : nullptr,
                                          reinterpret_cast<const volatile uint8_t *>(interruptFlag));
        if (st != HSR_ERR_INTERRUPTED) check(st);
    }

    hsr_run_stats lastStats() const {
        hsr_run_stats s;
        check(hsr_hip_last_stats(m_handle, &s));
        return s;
    }
    void setOption(int32_t option, int64_t value) { check(hsr_hip_set_option(m_handle, option, value)); }
    // per-triangle velocities [m/s] in the mesh's triangle order; an empty vector: static mesh
    void setVelocities(const std::vector<Vector3f> &velocities) {
        check(hsr_hip_set_velocities(m_handle,
                                     velocities.empty() ? nullptr : reinterpret_cast<const float *>(velocities.data()),
                                     (int64_t)velocities.size()));
    }
    hsr_backend *handle() const { return m_handle; }

private:
    hsr_backend *m_handle = nullptr;
    PathTracer &m_integrator;
};

// ------------------------------------------------------------------ PathTracer members
inline hsr_backend *PathTracer::handle() const {
    if (!m_backend) throw Error(HSR_ERR_INVALID, "the integrator has no backend");
    return m_backend->handle();
}

// Makes b the backend holding this integrator's frame and guide, configuring its frame
// when the configuration changed since.
inline void PathTracer::bind(Backend &b) {
    if (m_backend != &b) {
        releaseGuide();
        m_backend = &b;
        m_syncedGeneration = -1;
    }
    if (m_syncedGeneration != m_frameGeneration && m_frameGeneration > 0) {
        const hsr_frame_config fc{m_frameConfig.chirpCount, m_frameConfig.samplesPerChirp,
                                  m_frameConfig.channelCount};
        check(hsr_hip_configure_frame(b.handle(), &fc));
        m_syncedGeneration = m_frameGeneration;
    }
    if (!m_guide) check(hsr_guide_create(b.handle(), 0.1f, 0.005f, 1, &m_guide));  // path.h:229-232
}

template <class SceneT>
void PathTracer::run(Backend &backend, const SceneT &scene, long samples, bool *interruptFlag) {
    bind(backend);
    clearFrame();
    sampleIndexOffset = 0;
    currentSampleWeight = 1.f;
    isFinalIteration = false;
    if (!doGuiding) {
        backend.run(scene, samples, interruptFlag);
        return;
    }
    check(hsr_guide_reset(m_guide));
    long milestone = 16384;
    long remaining = samples;
    for (;;) {
        milestone = std::min(milestone, remaining);
        backend.run(scene, milestone, interruptFlag);
        if (interruptFlag && *interruptFlag) return;
        sampleIndexOffset += milestone;
        remaining -= milestone;
        if (remaining == 0) break;
        milestone *= 2;
        if (remaining < 2 * milestone) {  // the rest becomes the final iteration
            isFinalIteration = true;
            milestone = remaining;
        }
        if (clearBeforeIteration) {
            clearFrame();
            sampleIndexOffset = 0;
        } else {
            currentSampleWeight = 10 * currentSampleWeight;
        }
        check(hsr_guide_step(m_guide));
    }
}

inline Frame PathTracer::fetchFrame() const {
    if (m_syncedGeneration != m_frameGeneration)
        throw Error(HSR_ERR_INVALID, "the frame configuration changed since the last run");
    Frame f;
    f.config = m_frameConfig;
    f.data.resize(m_frameConfig.sampleCount());
    check(hsr_hip_fetch_frame(handle(), reinterpret_cast<float *>(f.data.data()), nullptr, nullptr));
    return f;
}

inline std::vector<std::complex<double>> PathTracer::fetchFrameDouble() const {
    if (m_syncedGeneration != m_frameGeneration)
        throw Error(HSR_ERR_INVALID, "the frame configuration changed since the last run");
    std::vector<std::complex<double>> z(m_frameConfig.sampleCount());
    check(hsr_hip_fetch_frame(handle(), nullptr, reinterpret_cast<double *>(z.data()), nullptr));
    return z;
}

inline double PathTracer::totalWeight() const {
    double w = 0.0;
    check(hsr_hip_fetch_frame(handle(), nullptr, nullptr, &w));
    return w;
}

}  // namespace hip
}  // namespace hussar

#endif /* HUSSAR_HIP_HPP */

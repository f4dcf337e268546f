// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of libhussar's Monte Carlo FMCW radar integrator, written from the
// reference sources (cited file:line below) so that the HIP backend can be checked against
// it. It is never linked into, loaded by or called from the product library; only tests,
// the smoke check and the CPU-baseline leg of the benchmark use it, and only as a checker
// or as the host-core baseline.
//
// Precision instances (template parameters G = geometry/field scalar, P = time/phase/splat
// scalar):
//   prec 0 "ref32"  G=float,  P=float   — the reference's own arithmetic (hussar.h:37,
//                                         radar.h:36): FP32 everywhere.
//   prec 1 "mixed"  G=float,  P=double  — the product's precision model: FP32 traversal
//                                         and fields (antenna lobes included); every segment
//                                         length is
//                                         recomputed in FP64 against the hit triangle's
//                                         plane and chained through FP64 hit points, so the
//                                         path time, phase and frame sums are FP64.
//                                         Hit triangle ids are bit-exact against the GPU.
//   prec 2 "fp64"   G=double, P=double  — the accuracy oracle ("truth").
//   prec 3 "chain32" G=float, P=double  — the reference's FP32 geometry chain with FP64 time,
//                                         phase and sums (shows that an FP32 chain alone
//                                         misses the 1e-3 bar).
//   prec 4 "dual"   the mixed instance plus the reference's FP32 chain, which gives the
//                   traversal and shadow rays and every discrete decision (hit ids bit-exact
//                   with ref32) while lengths, time, phase and weights stay FP64: the product
//                   with HSR_OPT_CHAIN = 1.
// All instances evaluate the same sample points: Halton values are the reference's FP32
// values (halton.h:38-54) in every instance, and mesh/scene inputs are the same floats.
//
// Modelling conventions shared with the product (each one a documented choice where the
// reference leaves room):
//  * speed of light = the reference's float constant 299792458f, i.e. 299792448 m/s after
//    float rounding (radar.h:107), in every precision;
//  * Embree (cpu.cpp:63-87) is replaced by a Möller–Trumbore test with tnear < t < tfar
//    (tnear = Epsilon = 5e-5, hussar.h:52), closest hit = smallest t, ties broken by the
//    smaller triangle id; geometric normal normalize(cross(v1-v0, v2-v0)) face-forwarded
//    against the ray (kernel.cu:86-97, cpu.cpp:73-86);
//  * get2D's unspecified argument order (halton.h:24-27) is a knob (primaryOrder);
//  * FP32 sin/cos of the sphere warp (sampling.h:38-44) use one fixed double-precision
//    Cody–Waite + Taylor evaluation rounded to float, so CPU and GPU agree bit for bit;
//  * scalar division of a complex vector is component-wise.
//
// Parity pins (tests/test_oracle_*.py): Halton known values and the radical-inverse
// definition; sphere warp closed form; AWR pattern symmetry and H ⟂ d; BVH == brute force on
// random and edge-aimed rays; splat taps == v·e^{jπd}·sinc(d) and the tap sum against a
// numpy DFT of a tone; single-bounce range bin and phase against the closed form; the
// reference-held dihedral fixtures (evaluation/simulation/*) for peak range bins; guiding
// invariants (∫pdf = 1, pdf(sample(x)) == returned pdf, untrained == uniform); the NEE weight
// (nee_weight: GO density correction, sphere and plane phase errors, soft-sphere and plane
// filters) against closed forms of a plate vertex (tests/test_oracle_pins.py); the guide
// step against hand-worked trees and the density it must build (∝ sqrt(stat/area) with the
// second moment, ∝ stat/area without) and the refinement it must make (a leaf of share s > t
// becomes a complete subtree of depth ceil(log2(s/t))); geometrical optics off against an FP64
// brute force of a two-plate corner written from path.h and emitter.h in numpy (plane hits in
// closed form, contributions splatted with the pinned splat) and against the default branch at
// one bounce, where both weight the same quantity (tests/test_oracle_go_off.py).
// libguiding's source is absent, so that rule is the published SD-tree rule (Müller et al.
// 2017) in binary form, not a reading of libguiding; Eigen's ulp-level order of mixed-type
// products (Matrix33f·Vector3c, Vector3f.cross(Vector3c)) is parity unpinned (Eigen absent; the
// ref32 instance sums (a0 + a1) + a2).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

namespace orc {

static constexpr int kMaxCh = 16;
static constexpr int kMaxTx = 4;

// ---------------------------------------------------------------------------------------
// POD mirrors of the C-ABI structs (same layout as include/hussar_hip.h, restated here so
// that the oracle does not depend on the product).
struct Rf { float startFreq, freqSlope, adcRate, idleTime, rampTime, antennaDelay; };
struct Antenna { float pos[3]; float rot[9]; };
struct SceneC {
    Rf rf;
    Antenna tx;
    Antenna rx[kMaxCh];
    int32_t rxCount;
    int32_t txCount;             // MIMO: 0 or 1 = tx only; tx t = txMore[t-1]
    Antenna txMore[kMaxTx - 1];
    float velocity[3];           // platform velocity of all antennas
    int32_t reserved0;
};
struct Knobs {
    int32_t onlyIndirect, maxDepth, useGeometricalOptics, poDiffraction, doGuiding,
        clearBeforeIteration, correctPhase, filterGuiding, doFiltering, filteringSphere;
    float filteringMin, filteringMax, filteringRadius;
    int32_t primaryOrder;
    int32_t produceDebugImage;
};
struct FrameCfg { int32_t chirpCount, samplesPerChirp, channelCount; };

// ---------------------------------------------------------------------------------------
// Scalar constants per precision (hussar.h:37-52, radar.h:107).
template <class T> struct K;
template <> struct K<float> {
    static float pi() { return 3.14159265358979323846f; }
    static float inv4pi() { return 0.07957747154594766788f; }
};
template <> struct K<double> {
    static double pi() { return 3.14159265358979323846; }
    static double inv4pi() { return 0.07957747154594766788; }
};
// radar.h:107 `constexpr float SPEED_OF_LIGHT = 299792458;` rounds to 299792448.
static const double kSpeedOfLight = (double)299792458.0f;
static const float kEpsilon = 5e-5f;  // hussar.h:52

// ---------------------------------------------------------------------------------------
// Small vector / complex types. Sums of three terms are (a0 + a1) + a2, products of a real
// matrix and a vector are row-wise dot products in that order (Eigen's unrolled redux).
template <class T> struct V3 {
    T x, y, z;
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T norm(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalized(V3<T> a) {
    T n2 = dot(a, a);
    return n2 > T(0) ? a / std::sqrt(n2) : a;
}

// radar::complex (libradar/include/radar/complex.h:31-265).
template <class T> struct Cx { T re, im; };
template <class T> inline Cx<T> operator+(Cx<T> a, Cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T> inline Cx<T> operator-(Cx<T> a, Cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T> inline Cx<T> operator*(Cx<T> a, Cx<T> b) {  // complex.h:56-61
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T> inline Cx<T> operator*(T s, Cx<T> a) { return {s * a.re, s * a.im}; }
template <class T> inline Cx<T> operator*(Cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <class T> inline Cx<T> operator/(Cx<T> a, T s) { return {a.re / s, a.im / s}; }
template <class T> inline Cx<T> conj(Cx<T> a) { return {a.re, -a.im}; }
template <class T> inline T cabs(Cx<T> z) {  // complex.h:225-237 (scaled hypot)
    T x = std::abs(z.re), y = std::abs(z.im);
    const T s = x < y ? y : x;
    if (s == T(0)) return s;
    x /= s;
    y /= s;
    return s * std::sqrt(x * x + y * y);
}

template <class T> struct CV3 { Cx<T> x, y, z; };
template <class T> inline CV3<T> operator+(CV3<T> a, CV3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline CV3<T> operator-(CV3<T> a, CV3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline CV3<T> operator*(T s, CV3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline CV3<T> operator*(CV3<T> a, Cx<T> s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline CV3<T> operator/(CV3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
// real vector × complex vector (Vector3f.cross(Vector3c)), component-wise real scaling
template <class T> inline CV3<T> cross(V3<T> a, CV3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// real vector · complex vector, Eigen dot = Σ conj(a_i) b_i with a real
template <class T> inline Cx<T> rdot(V3<T> a, CV3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// complex · complex, Σ conj(a_i) b_i (geometry.h:138 measureH, conjugate-linear in a)
template <class T> inline Cx<T> cdot(CV3<T> a, CV3<T> b) {
    return (conj(a.x) * b.x + conj(a.y) * b.y) + conj(a.z) * b.z;
}
// real vector scaled component-wise into a complex vector: (s * v_k) * c
template <class T> inline CV3<T> rscale(V3<T> v, Cx<T> c) { return {v.x * c, v.y * c, v.z * c}; }
template <class T> inline bool is_zero(CV3<T> h) {  // Eigen isZero(1e-20) for complex entries
    const T p = T(1e-20);
    const T thr = p * p;
    auto z = [&](Cx<T> c) { return c.re * c.re + c.im * c.im <= thr; };
    return z(h.x) && z(h.y) && z(h.z);
}

template <class T> struct M33 {
    T m[9];  // row-major
    V3<T> mul(V3<T> v) const {
        return {(m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[3] * v.x + m[4] * v.y) + m[5] * v.z,
                (m[6] * v.x + m[7] * v.y) + m[8] * v.z};
    }
    V3<T> mulT(V3<T> v) const {
        return {(m[0] * v.x + m[3] * v.y) + m[6] * v.z, (m[1] * v.x + m[4] * v.y) + m[7] * v.z,
                (m[2] * v.x + m[5] * v.y) + m[8] * v.z};
    }
    CV3<T> mul(CV3<T> v) const {
        return {(m[0] * v.x + m[1] * v.y) + m[2] * v.z, (m[3] * v.x + m[4] * v.y) + m[5] * v.z,
                (m[6] * v.x + m[7] * v.y) + m[8] * v.z};
    }
};

// ---------------------------------------------------------------------------------------
// Halton (libhussar/include/hussar/samplers/halton.h:11-87). FP32 exactly as the reference:
// digits reversed into a uint64, invBase = 1/float(base), invBaseN by repeated FP32
// multiply, float(reversed) * invBaseN, clamped to 1 - ulp.
static float radical_inverse(uint64_t v, int base) {
    const float invBase = 1.0f / float(base);
    uint64_t reversed = 0;
    float invBaseN = 1.0f;
    while (v) {
        uint64_t next = v / (uint64_t)base;
        uint64_t digit = v - next * (uint64_t)base;
        reversed = reversed * (uint64_t)base + digit;
        invBaseN *= invBase;
        v = next;
    }
    const float oneMinusEps = std::nextafter(1.0f, 0.0f);
    return std::min(float(reversed) * invBaseN, oneMinusEps);
}
static const int kPrimes[10] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29};  // halton.h:56-67

// HaltonSampler::get1D of dimension dim (halton.h:20-22, 56-67): the reference's switch has
// dimensions 0-9; a higher one is its Assert (halton.h:1082-1084), which the library refuses
// up front (useGeometricalOptics = 0 with maxDepth > 3), so it never comes here.
static float halton_dim(uint64_t index, int dim) { return dim < 10 ? radical_inverse(index, kPrimes[dim]) : 0.0f; }

// ---------------------------------------------------------------------------------------
// sin/cos of the sphere warp, reproducible bit for bit on CPU and GPU: Cody–Waite reduction
// by pi/2 with the two-part constant of fdlibm's __ieee754_rem_pio2 (pio2_1 = first 33 bits
// of pi/2, pio2_1t = the remainder; valid for |x| < 2^20), then the Taylor series of sin and
// cos to degree 17/18 on |r| <= pi/4 (truncation < 1e-19), evaluated in double without fused
// multiply-add. The FP32 variant rounds the double results to float.
static void sincos_d(double x, double *s_out, double *c_out) {
    const double k = std::rint(x * 0.63661977236758134308);
    const double r = (x - k * 1.57079632673412561417e+00) - k * 6.07710050650619224932e-11;
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;             //  1/17!  (Horner from the top)
    ps = ps * r2 - 1.0 / 1307674368000.0;            // -1/15!
    ps = ps * r2 + 1.0 / 6227020800.0;               //  1/13!
    ps = ps * r2 - 1.0 / 39916800.0;                 // -1/11!
    ps = ps * r2 + 1.0 / 362880.0;                   //  1/9!
    ps = ps * r2 - 1.0 / 5040.0;                     // -1/7!
    ps = ps * r2 + 1.0 / 120.0;                      //  1/5!
    ps = ps * r2 - 1.0 / 6.0;                        // -1/3!
    const double sr = r + (r * r2) * ps;
    double pc = -1.0 / 6402373705728000.0;           // -1/18!
    pc = pc * r2 + 1.0 / 20922789888000.0;           //  1/16!
    pc = pc * r2 - 1.0 / 87178291200.0;              // -1/14!
    pc = pc * r2 + 1.0 / 479001600.0;                //  1/12!
    pc = pc * r2 - 1.0 / 3628800.0;                  // -1/10!
    pc = pc * r2 + 1.0 / 40320.0;                    //  1/8!
    pc = pc * r2 - 1.0 / 720.0;                      // -1/6!
    pc = pc * r2 + 1.0 / 24.0;                       //  1/4!
    pc = pc * r2 - 0.5;                              // -1/2!
    const double cr = 1.0 + r2 * pc;
    const int q = (int)((int64_t)k & 3);
    double s, c;
    switch (q) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
    *s_out = s;
    *c_out = c;
}
static void sincos_f32(float xf, float *s_out, float *c_out) {
    double s, c;
    sincos_d((double)xf, &s, &c);
    *s_out = (float)s;
    *c_out = (float)c;
}
template <class G> inline void warp_sincos(G phi, G *s, G *c);
template <> inline void warp_sincos<float>(float phi, float *s, float *c) { sincos_f32(phi, s, c); }
template <> inline void warp_sincos<double>(double phi, double *s, double *c) { sincos_d(phi, s, c); }

// UniformSampleSphere (libhussar/include/hussar/core/sampling.h:38-44).
template <class G> inline V3<G> uniform_sample_sphere(G u0, G u1) {
    const G phi = (G(2) * K<G>::pi()) * u0;
    const G y = G(2) * u1 - G(1);
    const G r = std::sqrt(std::max(G(0), G(1) - y * y));
    G s, c;
    warp_sincos<G>(phi, &s, &c);
    return v3<G>(-r * s, y, -r * c);
}

// AWRAngularDistribution::evaluate (emitter.h:126-136). H = e_y × d, scaled by the two
// polynomial lobes. kPromote: the lobes in double as the reference's double literals promote
// them, then applied as a G factor (reference arithmetic); otherwise in G throughout, in the
// order ((a2·a2)·a2) with a2 = a·a (the product's FP32 field arithmetic).
template <class G, bool kPromote = true> inline CV3<G> awr_evaluate(V3<G> d) {
    CV3<G> H{{d.z, G(0)}, {G(0), G(0)}, {-d.x, G(0)}};  // (0,1,0) × d
    const G cos_h = std::sqrt(G(1) - d.x * d.x);
    const G cos_e = std::sqrt(G(1) - d.y * d.y);
    G f1, f2;
    if (kPromote) {
        f1 = G(2.622 / std::pow((double)cos_h - 1.8, 6));
        f2 = G(0.625 / std::pow((double)cos_e - 1.5, 4));
    } else {
        const G a = cos_h - G(1.8), b = cos_e - G(1.5);
        const G a2 = a * a, b2 = b * b;
        f1 = G(2.622) / ((a2 * a2) * a2);
        f2 = G(0.625) / (b2 * b2);
    }
    H = {H.x * f1, H.y * f1, H.z * f1};
    H = {H.x * f2, H.y * f2, H.z * f2};
    return H;
}

template <class G> struct AntennaG {
    V3<G> pos;
    M33<G> R;
};
template <class G> AntennaG<G> to_g(const Antenna &a) {
    AntennaG<G> r;
    r.pos = v3<G>(G(a.pos[0]), G(a.pos[1]), G(a.pos[2]));
    for (int i = 0; i < 9; ++i) r.R.m[i] = G(a.rot[i]);
    return r;
}

// ---------------------------------------------------------------------------------------
// Mesh and intersection.
template <class G> struct TriG { V3<G> v0, e1, e2; };

// Möller–Trumbore in the order both implementations use; returns t or +inf.
template <class G>
inline G tri_intersect(const TriG<G> &tr, V3<G> o, V3<G> d, G tnear, G tfar) {
    const G inf = std::numeric_limits<G>::infinity();
    const V3<G> pv = cross(d, tr.e2);
    const G det = dot(tr.e1, pv);
    if (det == G(0)) return inf;
    const G inv = G(1) / det;
    const V3<G> tv = o - tr.v0;
    const G u = dot(tv, pv) * inv;
    if (u < G(0) || u > G(1)) return inf;
    const V3<G> qv = cross(tv, tr.e1);
    const G v = dot(d, qv) * inv;
    if (v < G(0) || u + v > G(1)) return inf;
    const G t = dot(tr.e2, qv) * inv;
    if (!(t > tnear && t < tfar)) return inf;
    return t;
}

template <class G> struct Node {
    G lo[3], hi[3];
    int32_t left;   // inner: index of left child (right = left + 1); leaf: first triangle slot
    int32_t count;  // 0 for inner nodes
};

struct Counters {
    int64_t nodes = 0, tris = 0, rays_closest = 0, rays_shadow = 0, events = 0;
};

// The oracle's own BVH2: binned SAH (16 bins) over triangle centroids, leaves of <= 4
// triangles; node boxes are enlarged by a relative margin so that a ray the triangle test
// accepts is never culled — traversal then equals brute force (tested).
template <class G> struct Accel {
    std::vector<TriG<G>> tris;      // by original id
    std::vector<int32_t> order;     // leaf slot -> original id
    std::vector<Node<G>> nodes;

    void build(const float *V, const int32_t *I, int64_t nt) {
        tris.resize(nt);
        std::vector<G> blo(nt * 3), bhi(nt * 3), cen(nt * 3);
        for (int64_t t = 0; t < nt; ++t) {
            const float *a = V + 3 * (int64_t)I[3 * t + 0];
            const float *b = V + 3 * (int64_t)I[3 * t + 1];
            const float *c = V + 3 * (int64_t)I[3 * t + 2];
            V3<G> v0 = v3<G>(a[0], a[1], a[2]), v1 = v3<G>(b[0], b[1], b[2]),
                  v2 = v3<G>(c[0], c[1], c[2]);
            tris[t] = {v0, v1 - v0, v2 - v0};
            for (int k = 0; k < 3; ++k) {
                blo[3 * t + k] = std::min(v0[k], std::min(v1[k], v2[k]));
                bhi[3 * t + k] = std::max(v0[k], std::max(v1[k], v2[k]));
                cen[3 * t + k] = (blo[3 * t + k] + bhi[3 * t + k]) * G(0.5);
            }
        }
        order.resize(nt);
        for (int64_t i = 0; i < nt; ++i) order[i] = (int32_t)i;
        nodes.clear();
        nodes.reserve(2 * nt + 1);
        nodes.push_back(Node<G>{});
        if (nt == 0) {
            for (int k = 0; k < 3; ++k) { nodes[0].lo[k] = G(1); nodes[0].hi[k] = G(-1); }
            nodes[0].left = 0;
            nodes[0].count = 0;
            return;
        }
        struct Job { int32_t node, begin, end; };
        std::vector<Job> stack{{0, 0, (int32_t)nt}};
        while (!stack.empty()) {
            Job j = stack.back();
            stack.pop_back();
            G lo[3], hi[3], clo[3], chi[3];
            for (int k = 0; k < 3; ++k) {
                lo[k] = clo[k] = std::numeric_limits<G>::infinity();
                hi[k] = chi[k] = -std::numeric_limits<G>::infinity();
            }
            for (int32_t i = j.begin; i < j.end; ++i) {
                int32_t t = order[i];
                for (int k = 0; k < 3; ++k) {
                    lo[k] = std::min(lo[k], blo[3 * t + k]);
                    hi[k] = std::max(hi[k], bhi[3 * t + k]);
                    clo[k] = std::min(clo[k], cen[3 * t + k]);
                    chi[k] = std::max(chi[k], cen[3 * t + k]);
                }
            }
            Node<G> &n = nodes[j.node];
            for (int k = 0; k < 3; ++k) {
                const G m = (std::abs(lo[k]) + std::abs(hi[k]) + (hi[k] - lo[k])) * G(1.0 / 1048576.0) +
                            G(1e-6);
                n.lo[k] = lo[k] - m;
                n.hi[k] = hi[k] + m;
            }
            const int32_t cnt = j.end - j.begin;
            int best_axis = -1;
            int32_t mid = j.begin;
            if (cnt > 4) {
                // binned SAH
                const int B = 16;
                G best_cost = std::numeric_limits<G>::infinity();
                int best_bin = -1;
                for (int k = 0; k < 3; ++k) {
                    const G ext = chi[k] - clo[k];
                    if (!(ext > G(0))) continue;
                    int32_t bc[B] = {0};
                    G bl[B][3], bh[B][3];
                    for (int b = 0; b < B; ++b)
                        for (int q = 0; q < 3; ++q) {
                            bl[b][q] = std::numeric_limits<G>::infinity();
                            bh[b][q] = -std::numeric_limits<G>::infinity();
                        }
                    for (int32_t i = j.begin; i < j.end; ++i) {
                        int32_t t = order[i];
                        int b = (int)((cen[3 * t + k] - clo[k]) / ext * G(B));
                        b = std::min(std::max(b, 0), B - 1);
                        bc[b]++;
                        for (int q = 0; q < 3; ++q) {
                            bl[b][q] = std::min(bl[b][q], blo[3 * t + q]);
                            bh[b][q] = std::max(bh[b][q], bhi[3 * t + q]);
                        }
                    }
                    auto area = [](const G *l, const G *h) {
                        G dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
                        if (!(dx >= 0)) return G(0);
                        return dx * dy + dy * dz + dz * dx;
                    };
                    G rl[B][3], rh[B][3];
                    int32_t rc[B];
                    {
                        G al[3], ah[3];
                        for (int q = 0; q < 3; ++q) { al[q] = std::numeric_limits<G>::infinity(); ah[q] = -al[q]; }
                        int32_t c = 0;
                        for (int b = B - 1; b >= 0; --b) {
                            for (int q = 0; q < 3; ++q) { al[q] = std::min(al[q], bl[b][q]); ah[q] = std::max(ah[q], bh[b][q]); }
                            c += bc[b];
                            for (int q = 0; q < 3; ++q) { rl[b][q] = al[q]; rh[b][q] = ah[q]; }
                            rc[b] = c;
                        }
                    }
                    G al[3], ah[3];
                    for (int q = 0; q < 3; ++q) { al[q] = std::numeric_limits<G>::infinity(); ah[q] = -al[q]; }
                    int32_t c = 0;
                    for (int b = 0; b < B - 1; ++b) {
                        for (int q = 0; q < 3; ++q) { al[q] = std::min(al[q], bl[b][q]); ah[q] = std::max(ah[q], bh[b][q]); }
                        c += bc[b];
                        if (c == 0 || rc[b + 1] == 0) continue;
                        G cost = area(al, ah) * G(c) + area(rl[b + 1], rh[b + 1]) * G(rc[b + 1]);
                        if (cost < best_cost) { best_cost = cost; best_axis = k; best_bin = b; }
                    }
                }
                if (best_axis >= 0) {
                    const int k = best_axis;
                    const G ext = chi[k] - clo[k];
                    auto it = std::partition(order.begin() + j.begin, order.begin() + j.end, [&](int32_t t) {
                        int b = (int)((cen[3 * t + k] - clo[k]) / ext * G(B));
                        b = std::min(std::max(b, 0), B - 1);
                        return b <= best_bin;
                    });
                    mid = (int32_t)(it - order.begin());
                    if (mid == j.begin || mid == j.end) best_axis = -1;
                }
                if (best_axis < 0) {  // all centroids coincide: split by count
                    mid = j.begin + cnt / 2;
                    best_axis = 0;
                }
            }
            if (cnt <= 4) {
                n.left = j.begin;
                n.count = cnt;
            } else {
                const int32_t l = (int32_t)nodes.size();
                nodes[j.node].left = l;
                nodes[j.node].count = 0;
                nodes.push_back(Node<G>{});
                nodes.push_back(Node<G>{});
                stack.push_back({l + 1, mid, j.end});
                stack.push_back({l, j.begin, mid});
            }
        }
    }

    // Slab test with Ize's robust factor on tmax.
    static bool box_hit(const Node<G> &n, V3<G> o, V3<G> invd, G tnear, G tbest, G *tentry) {
        G t0 = tnear, t1 = std::numeric_limits<G>::infinity();
        for (int k = 0; k < 3; ++k) {
            G a = (n.lo[k] - o[k]) * invd[k];
            G b = (n.hi[k] - o[k]) * invd[k];
            if (std::isnan(a) || std::isnan(b)) continue;  // o on a slab plane with d_k = 0
            if (a > b) std::swap(a, b);
            t0 = std::max(t0, a);
            t1 = std::min(t1, b);
        }
        const G g3 = G(3) * std::numeric_limits<G>::epsilon() * G(0.5);
        t1 = t1 * (G(1) + G(2) * g3 / (G(1) - g3));
        *tentry = t0;
        return t0 <= t1 && t0 <= tbest;
    }

    // closest hit; returns id or -1. Children are visited near-first; a popped entry is
    // skipped when its box entry lies beyond the current best hit.
    int32_t closest(V3<G> o, V3<G> d, G tnear, G tfar, G *tout, Counters *cnt) const {
        const G inf = std::numeric_limits<G>::infinity();
        *tout = inf;
        if (tris.empty()) return -1;
        V3<G> invd = v3<G>(G(1) / d.x, G(1) / d.y, G(1) / d.z);
        G best = inf;
        int32_t best_id = -1;
        struct Entry { int32_t node; G t; };
        Entry stack[128];
        int sp = 0;
        G te;
        if (cnt) cnt->nodes++;
        if (!box_hit(nodes[0], o, invd, tnear, inf, &te)) return -1;
        stack[sp++] = {0, te};
        while (sp) {
            const Entry e = stack[--sp];
            if (e.t > best) continue;
            const Node<G> &n = nodes[e.node];
            if (n.count > 0) {
                for (int32_t i = n.left; i < n.left + n.count; ++i) {
                    const int32_t id = order[i];
                    if (cnt) cnt->tris++;
                    const G t = tri_intersect(tris[id], o, d, tnear, tfar);
                    if (t < inf && (best_id < 0 || t < best || (t == best && id < best_id))) {
                        best = t;
                        best_id = id;
                    }
                }
            } else {
                G t0, t1;
                if (cnt) cnt->nodes += 2;
                const bool h0 = box_hit(nodes[n.left], o, invd, tnear, best, &t0);
                const bool h1 = box_hit(nodes[n.left + 1], o, invd, tnear, best, &t1);
                if (h0 && h1) {
                    if (t0 <= t1) { stack[sp++] = {n.left + 1, t1}; stack[sp++] = {n.left, t0}; }
                    else { stack[sp++] = {n.left, t0}; stack[sp++] = {n.left + 1, t1}; }
                } else if (h0) {
                    stack[sp++] = {n.left, t0};
                } else if (h1) {
                    stack[sp++] = {n.left + 1, t1};
                }
            }
        }
        *tout = best;
        return best_id;
    }

    bool occluded(V3<G> o, V3<G> d, G tnear, G tfar, Counters *cnt) const {
        const G inf = std::numeric_limits<G>::infinity();
        if (tris.empty()) return false;
        V3<G> invd = v3<G>(G(1) / d.x, G(1) / d.y, G(1) / d.z);
        int32_t stack[128];
        int sp = 0;
        G te;
        if (cnt) cnt->nodes++;
        if (!box_hit(nodes[0], o, invd, tnear, tfar, &te)) return false;
        stack[sp++] = 0;
        while (sp) {
            const Node<G> &n = nodes[stack[--sp]];
            if (n.count > 0) {
                for (int32_t i = n.left; i < n.left + n.count; ++i) {
                    if (cnt) cnt->tris++;
                    if (tri_intersect(tris[order[i]], o, d, tnear, tfar) < inf) return true;
                }
            } else {
                G t0, t1;
                if (cnt) cnt->nodes += 2;
                if (box_hit(nodes[n.left + 1], o, invd, tnear, tfar, &t1)) stack[sp++] = n.left + 1;
                if (box_hit(nodes[n.left], o, invd, tnear, tfar, &t0)) stack[sp++] = n.left;
            }
        }
        return false;
    }

    int32_t closest_brute(V3<G> o, V3<G> d, G tnear, G tfar, G *tout) const {
        const G inf = std::numeric_limits<G>::infinity();
        G best = inf;
        int32_t best_id = -1;
        for (size_t id = 0; id < tris.size(); ++id) {
            G t = tri_intersect(tris[id], o, d, tnear, tfar);
            if (t < best) { best = t; best_id = (int32_t)id; }  // ascending ids: ties keep the smaller
        }
        *tout = best;
        return best_id;
    }
    bool occluded_brute(V3<G> o, V3<G> d, G tnear, G tfar) const {
        const G inf = std::numeric_limits<G>::infinity();
        for (size_t id = 0; id < tris.size(); ++id)
            if (tri_intersect(tris[id], o, d, tnear, tfar) < inf) return true;
        return false;
    }
    V3<G> normal(int32_t id) const {  // kernel.cu:86, normalize(cross(e1, e2))
        return normalized(cross(tris[id].e1, tris[id].e2));
    }
};

struct Mesh {
    Accel<float> f;
    Accel<double> d;
    int64_t nt = 0;
    int brute = 0;           // 1: closest/occluded by brute force instead of the BVH
    std::vector<float> vel;  // motion: 3 floats per triangle id, empty for a static mesh
};
template <class G> const Accel<G> &accel_of(const Mesh &m);
template <> const Accel<float> &accel_of<float>(const Mesh &m) { return m.f; }
template <> const Accel<double> &accel_of<double>(const Mesh &m) { return m.d; }

// ---------------------------------------------------------------------------------------
// Guiding: the build's own 2-D binary tree over primary sample space with the semantics of
// GuidingWrapper (path.h:17-131). libguiding's BTree is absent; this tree: node 0 is the
// root covering [0,1)^2; an inner node splits its cell in half along axis (depth % 2); its
// two children are stored consecutively (child, child + 1), child 0 the lower half.
struct Tree {
    std::vector<int32_t> child{0};
    std::vector<float> p0{0.5f};
    std::vector<double> stat{0.0};
    std::vector<double> mass{0.0};
    std::vector<int32_t> depth{0};
    static constexpr int kMaxDepth = 24;
    static constexpr int kMaxNodes = 1 << 16;

    void reset() {
        child = {0};
        p0 = {0.5f};
        stat = {0.0};
        mass = {0.0};
        depth = {0};
    }
    // tree pdf at a point (descent by cell midpoint)
    float pdf(const float x[2]) const {
        float pdfv = 1.0f, lo[2] = {0.0f, 0.0f}, sz[2] = {1.0f, 1.0f};
        int32_t node = 0, dep = 0;
        while (child[node] != 0) {
            const int a = dep & 1;
            const float p = p0[node];
            const float mid = lo[a] + sz[a] * 0.5f;
            if (x[a] < mid) {
                pdfv = pdfv * (2.0f * p);
                node = child[node];
            } else {
                pdfv = pdfv * (2.0f * (1.0f - p));
                lo[a] = mid;
                node = child[node] + 1;
            }
            sz[a] = sz[a] * 0.5f;
            dep++;
        }
        return pdfv;
    }
    // hierarchical warp of x in place; returns the tree pdf of the warped point
    float sample(float x[2]) const {
        const float oneMinusEps = std::nextafter(1.0f, 0.0f);
        float pdfv = 1.0f, lo[2] = {0.0f, 0.0f}, sz[2] = {1.0f, 1.0f};
        int32_t node = 0, dep = 0;
        while (child[node] != 0) {
            const int a = dep & 1;
            const float p = p0[node];
            if (x[a] < p) {
                x[a] = std::min(x[a] / p, oneMinusEps);
                pdfv = pdfv * (2.0f * p);
                node = child[node];
            } else {
                x[a] = std::min((x[a] - p) / (1.0f - p), oneMinusEps);
                pdfv = pdfv * (2.0f * (1.0f - p));
                lo[a] = lo[a] + sz[a] * 0.5f;
                node = child[node] + 1;
            }
            sz[a] = sz[a] * 0.5f;
            dep++;
        }
        x[0] = lo[0] + x[0] * sz[0];
        x[1] = lo[1] + x[1] * sz[1];
        return pdfv;
    }
    int32_t leaf_of(const float x[2]) const {
        float lo[2] = {0.0f, 0.0f}, sz[2] = {1.0f, 1.0f};
        int32_t node = 0, dep = 0;
        while (child[node] != 0) {
            const int a = dep & 1;
            const float mid = lo[a] + sz[a] * 0.5f;
            if (x[a] < mid) {
                node = child[node];
            } else {
                lo[a] = mid;
                node = child[node] + 1;
            }
            sz[a] = sz[a] * 0.5f;
            dep++;
        }
        return node;
    }
    // training statistic of one path (path.h:448-450 → GuidingWrapper::splat path.h:96-111)
    static double stat_value(float density, float weight, bool secondMoment) {
        const double dd = (double)density;
        return (secondMoment ? dd * dd : dd) * (double)weight;
    }
    // build: leaf mass = secondMoment ? sqrt(stat * area) : stat; p0 = m(child0)/m(node)
    void build(bool secondMoment) {
        const int32_t n = (int32_t)child.size();
        mass.assign(n, 0.0);
        for (int32_t i = n - 1; i >= 0; --i) {
            if (child[i] == 0) {
                const double area = std::ldexp(1.0, -depth[i]);
                mass[i] = secondMoment ? std::sqrt(std::max(stat[i], 0.0) * area) : std::max(stat[i], 0.0);
            }
        }
        // children always have larger indices than their parent
        for (int32_t i = n - 1; i >= 0; --i) {
            if (child[i] != 0) {
                const double m0 = mass[child[i]], m1 = mass[child[i] + 1];
                mass[i] = m0 + m1;
                p0[i] = mass[i] > 0.0 ? (float)(m0 / mass[i]) : 0.5f;
            }
        }
    }
    // refine: in node order (new nodes included), split every leaf whose share of the root
    // mass exceeds the threshold; each child is assumed to carry half of its parent's share,
    // so one refine subdivides as deep as the statistics warrant
    void refine(float splitThreshold) {
        const double total = mass.empty() ? 0.0 : mass[0];
        if (total > 0.0) {
            std::vector<double> share(child.size());
            for (size_t i = 0; i < child.size(); ++i) share[i] = mass[i] / total;
            for (int32_t i = 0; i < (int32_t)child.size(); ++i) {
                if (child[i] != 0) continue;
                if (depth[i] >= kMaxDepth) continue;
                if ((int32_t)child.size() + 2 > kMaxNodes) break;
                if (share[i] > (double)splitThreshold) {
                    const int32_t c = (int32_t)child.size();
                    child[i] = c;
                    p0[i] = 0.5f;
                    for (int q = 0; q < 2; ++q) {
                        child.push_back(0);
                        p0.push_back(0.5f);
                        stat.push_back(0.0);
                        mass.push_back(0.0);
                        depth.push_back(depth[i] + 1);
                        share.push_back(share[i] * 0.5);
                    }
                }
            }
        }
        std::fill(stat.begin(), stat.end(), 0.0);
    }
};

struct Guide {
    float uniformProb = 0.1f;
    float splitThreshold = 0.005f;
    bool secondMoment = true;
    Tree sampling, training;
    void reset() {  // path.h:48-51
        sampling.reset();
        training.reset();
    }
    // GuidingWrapper::sample (path.h:53-82)
    float sample(float x[2]) const {
        const float u = uniformProb;
        if (u == 1.0f) return 1.0f;
        float pdf = 1.0f - u;
        if (x[0] < u) {
            x[0] /= u;
            pdf *= sampling.pdf(x);
        } else {
            x[0] -= u;
            x[0] /= 1.0f - u;
            float gpdf = sampling.sample(x);
            pdf *= gpdf;
        }
        pdf += u;
        return pdf;
    }
    // GuidingWrapper::step (path.h:119-126)
    void step() {
        training.build(secondMoment);
        sampling = training;
        training.refine(splitThreshold);
    }
};

// ---------------------------------------------------------------------------------------
// Frame splat (radar::Frame::splat<16>, radar.h:564-627) of one NEE event through
// Integrator::splat (integrator.h:113-135): fractional sample index via setTime
// (radar.h:232-241), IF phase via measureRay (integrator.h:107-110), fractional chirp index
// via setVelocity (radar.h:248-251) from the contribution's radial velocity delta_v (the
// reference passes 0, integrator.h:130; `motion` off keeps chirp 0 exactly), and the
// separable kernel: per spread axis value·e^{jπs}, weight·sin(πs)/π and taps
// weight/(s_chirp-l)/(s_sample-k) over ±16 bins, an axis with |s| < 1e-4 being a delta.
// Bins wrap modulo the axis length (the reference would index one past the end when
// round(x) equals the length). chirpFreq = RFConfig::chirpFrequency in double.
template <class P>
inline void splat_event(P *frame, int M, int N, int C, int ch, P delta_t, Cx<P> v, P weight, const Rf &rf,
                        Counters *cnt, bool motion = false, P delta_v = P(0)) {
    delta_t = delta_t + P(rf.antennaDelay);
    // chirp axis (setVelocity)
    int cc = 0;
    P sc = P(0);
    bool deltaC = true;
    if (motion) {
        const P chirpFreq = P(1.0 / ((double)rf.idleTime + (double)rf.rampTime));
        const P dp = ((P(2) * P(rf.startFreq)) * delta_v) / P(kSpeedOfLight);
        const P q = dp / chirpFreq;
        const P xc = P(M) * (q - std::floor(q));
        cc = (int)std::round(xc);
        sc = xc - P(cc);
        deltaC = std::abs(sc) < P(1e-4);
    }
    P x;
    if (rf.freqSlope == 0.0f) {
        x = P(0);
    } else {
        const P delta_f = delta_t * P(rf.freqSlope);
        const P q = delta_f / P(rf.adcRate);
        x = P(N) * (q - std::floor(q));
    }
    const P arg = ((P(2) * K<P>::pi()) * (P(rf.startFreq) - delta_t * P(rf.freqSlope) / P(2))) * delta_t;
    const Cx<P> phase{std::cos(arg), std::sin(arg)};
    Cx<P> value = v * phase;
    value = weight * value;
    const int center = (int)std::round(x);
    P s = x - P(center);
    if (cnt) cnt->events++;
    const bool deltaS = std::abs(s) < P(1e-4);
    P w = P(1);
    if (!deltaC) {
        const P shiftPi = P(M_PI * (double)sc);
        value = value * Cx<P>{std::cos(shiftPi), std::sin(shiftPi)};
        w = w * P(std::sin(shiftPi) / M_PI);
    }
    if (!deltaS) {
        const P shiftPi = P(M_PI * (double)s);
        value = value * Cx<P>{std::cos(shiftPi), std::sin(shiftPi)};
        w = w * P(std::sin(shiftPi) / M_PI);
    }
    const int lc = deltaC ? 0 : 16, ls = deltaS ? 0 : 16;
    for (int l = -lc; l <= lc; ++l) {
        const P wl = deltaC ? w : w / (sc - P(l));
        const int bc = (((cc + l) % M) + M) % M;
        for (int k = -ls; k <= ls; ++k) {
            const P wk = deltaS ? wl : wl / (s - P(k));
            const int b = (((center + k) % N) + N) % N;
            P *dst = frame + 2 * (((int64_t)bc * N + b) * C + ch);
            dst[0] += wk * value.re;
            dst[1] += wk * value.im;
        }
    }
}

// ---------------------------------------------------------------------------------------
// PathTracer::sample (path.h:236-451), both branches of useGeometricalOptics: specular
// reflection (the reference's default), or with it off a uniform sphere sample per bounce
// (SurfaceEmitter::sample, emitter.h:58-70, from Halton dimensions 3 + 2(k-1) and 4 + 2(k-1)
// at depth k, path.h:277-279) weighted after each hit by the emitter's connect and the
// hemisphere-to-surface density t²/cosθ (path.h:420-427); correctPhase as the knobs say.
//
// Four precision roles: T — BVH traversal and ray/triangle tests; Gd — the geometry chain
// that fixes path lengths (sampled and reflected directions, hit points, NEE legs,
// filtering distances); F — fields and weights (H, PO phasor, antenna pattern); P — path
// time, IF phase, splat and frame sums. ref32: all float, in the reference's order.
// mixed: T = F = float, Gd = P = double — traversal rays are the chain's double rays rounded
// to float, and a hit's distance is re-evaluated in double against the found triangle's
// plane (normal from the float edges promoted to double). fp64: all double.
struct PathRecord {
    int32_t *hits;  // [maxDepth] triangle id per depth, -1 none
    int8_t *nee;    // [maxDepth * C] 0 none, 1 occluded, 2 visible, 3 visible + splatted
};

template <class A, class B> inline V3<A> vcast(V3<B> v) { return v3<A>(A(v.x), A(v.y), A(v.z)); }
template <class A> inline V3<A> vcast3(const float *p) { return v3<A>(A(p[0]), A(p[1]), A(p[2])); }
template <class A> inline M33<A> mcast(const float *m) {
    M33<A> r;
    for (int i = 0; i < 9; ++i) r.m[i] = A(m[i]);
    return r;
}

// The weight of an NEE contribution v from a bounce vertex (depth > 0, path.h:330-394): the GO
// density correction v·r/cosθ/4π (path.h:330-341; false = skipped when cosθ < 1e-3), the phase
// error dphase in wavelengths (path.h:343-355: for the sphere, the distance from the receiver
// to the reflected ray o + λd, λ >= 0; for the plane, r·tan of the angle between the half
// vector normalize(dn - d_in) and the normal n_in) and the filter (path.h:357-387: the soft
// sphere max(1 - 0.2·(dphase/R - 1), 0) beyond R; the plane's linear ramp from filteringMin to
// filteringMax and zero beyond, whose cut value may feed the guiding weight), and the guiding
// weight update gw += v (path.h:392-394). o, d: the vertex and the reflected direction; dn, rr:
// unit direction and distance to the receiver; d_in, n_in: incoming direction and normal at the
// vertex; r: path length up to the vertex. dphase_out (may be null) receives dphase.
// Phase error dphase in wavelengths of an NEE contribution from a bounce vertex (path.h:343-355):
// for the sphere, the distance from the receiver to the reflected ray o + λd, λ >= 0; for the
// plane, r·tan of the angle between the half vector normalize(dn - d_in) and the normal n_in.
template <class Gd>
Gd phase_error(const Knobs &knobs, Gd r, V3<Gd> o, V3<Gd> d, V3<Gd> dn, Gd rr, V3<Gd> d_in, V3<Gd> n_in,
               Gd lambda) {
    Gd dphase;
    if (knobs.filteringSphere) {  // path.h:343-349
        const V3<Gd> rxPos = o + rr * dn;
        const Gd lam = std::max(Gd(0), dot(d, rxPos - o));
        dphase = norm((o + lam * d) - rxPos);
    } else {  // path.h:350-353
        const Gd c = dot(normalized(dn - d_in), n_in);
        dphase = r * std::sqrt(Gd(1) - c * c) / c;
    }
    return dphase / lambda;  // path.h:355
}

// The filter of path.h:357-387: the soft sphere max(1 - 0.2·(dphase/R - 1), 0) beyond R (with
// correctPhase a hard sphere, path.h:360-368, whose cut value may feed the guiding weight
// through 1/((dphase/R)² + 1)); the plane's linear ramp from filteringMin to filteringMax and
// zero beyond, whose cut value may feed the guiding weight.
template <class Gd, class F>
void filter_weight(const Knobs &knobs, Gd dphase, Cx<F> &v, Cx<F> &guidingWeight) {
    if (!knobs.doFiltering) return;
    if (knobs.filteringSphere) {
        const Gd R = Gd(knobs.filteringRadius);
        if (dphase > R && knobs.correctPhase) {
            if (knobs.filterGuiding) {  // std::pow(Float, int) is evaluated in double
                const double q = (double)(dphase / R);
                guidingWeight = guidingWeight + v / F(q * q + 1.0);
            }
            v = Cx<F>{F(0), F(0)};
        } else if (dphase > R) {
            const Gd f = Gd(1.0 - 0.20 * (double)(dphase / R - Gd(1)));
            v = v * F(std::max(f, Gd(0)));
        }
    } else {
        const Gd fmin = Gd(knobs.filteringMin), fmax = Gd(knobs.filteringMax);
        if (dphase > fmax) {
            if (knobs.filterGuiding) {
                const double q = (double)(dphase / fmax);
                guidingWeight = guidingWeight + v / F(q * q + 1.0);
            }
            v = Cx<F>{F(0), F(0)};
        }
        if (dphase > fmin) v = v * F((fmax - dphase) / (fmax - fmin));
    }
}

// The weight of an NEE contribution v from a bounce vertex (depth > 0, path.h:330-394): with
// geometrical optics the density correction v·r/cosθ/4π (path.h:330-341; false = skipped when
// cosθ < 1e-3), then the phase error and the filter, and the guiding weight update gw += v
// (path.h:392-394). o, d: the vertex and the ray's direction (reflected, or sampled with
// geometrical optics off); dn, rr: unit direction and distance to the receiver; d_in, n_in:
// incoming direction and normal at the vertex; r: path length up to the vertex. dphase_out
// (may be null) receives dphase.
template <class Gd, class F>
bool nee_weight(const Knobs &knobs, Gd r, Gd cosTheta, V3<Gd> o, V3<Gd> d, V3<Gd> dn, Gd rr, V3<Gd> d_in,
                V3<Gd> n_in, Gd lambda, Cx<F> &v, Cx<F> &guidingWeight, Gd *dphase_out) {
    const F piF = K<F>::pi();
    if (knobs.useGeometricalOptics) {
        v = v * F(r / cosTheta);
        v = v / (F(4) * piF);
        if (cosTheta < Gd(1e-3)) return false;
    }
    const Gd dphase = phase_error<Gd>(knobs, r, o, d, dn, rr, d_in, n_in, lambda);
    if (dphase_out) *dphase_out = dphase;
    filter_weight<Gd, F>(knobs, dphase, v, guidingWeight);
    guidingWeight = guidingWeight + v;  // path.h:392-394
    return true;
}

// kDual: beside the Gd chain, the reference's FP32 chain (the ref32 instance's arithmetic) is
// carried along; it gives the traversal and shadow rays and every discrete decision (hit ids,
// visibility, grazing and range cut-offs, back-face cull, filter thresholds), while path
// lengths, time, phase and the GO weight come from the Gd chain refined on the same triangles.
template <class T, class Gd, class F, class P, bool kPromote, bool kDual = false> struct Tracer {
    static constexpr bool refine = !std::is_same<T, Gd>::value;
    const Mesh *mesh;
    const Accel<T> *acc;
    const SceneC *scene;
    const Knobs *knobs;
    int N, C, chirps;
    P *frame;               // 2 * chirps * N * C
    double *guide_stat;     // training statistics (per node) or null
    const Guide *guide;     // null if no guiding
    bool isFinal;
    float sampleWeight;
    Counters *cnt;
    // MIMO and motion: the transmitter of this pass, the channel of its receiver 0, and
    // whether contributions carry a radial velocity (mesh velocities or platform motion)
    const Antenna *txa = nullptr;
    int chBase = 0;
    bool motion = false;

    int32_t trace_closest(V3<T> o, V3<T> d, T *t) const {
        if (cnt) cnt->rays_closest++;
        const T inf = std::numeric_limits<T>::infinity();
        if (mesh->brute) return acc->closest_brute(o, d, T(kEpsilon), inf, t);
        return acc->closest(o, d, T(kEpsilon), inf, t, cnt);
    }
    bool trace_visible(V3<T> o, V3<T> d, T tmax) const {
        if (cnt) cnt->rays_shadow++;
        if (mesh->brute) return !acc->occluded_brute(o, d, T(kEpsilon), tmax);
        return !acc->occluded(o, d, T(kEpsilon), tmax, cnt);
    }
    // geometric normal of a hit triangle in the chain precision (kernel.cu:86): for the
    // refined chain from the float edges promoted to double
    V3<Gd> normal(int32_t id) const {
        const TriG<float> &tf = mesh->f.tris[id];
        if (refine) return normalized(cross(vcast<Gd>(tf.e1), vcast<Gd>(tf.e2)));
        return vcast<Gd>(acc->normal(id));
    }
    // distance along d from o to the plane of triangle id (refined chain only)
    Gd plane_distance(int32_t id, V3<Gd> o, V3<Gd> d) const {
        const TriG<float> &tf = mesh->f.tris[id];
        const V3<Gd> n = cross(vcast<Gd>(tf.e1), vcast<Gd>(tf.e2));
        return dot(n, vcast<Gd>(tf.v0) - o) / dot(n, d);
    }

    void sample(uint64_t index, PathRecord *rec) {
        const Rf &rf = scene->rf;
        const F piF = K<F>::pi();
        const int maxDepth = knobs->maxDepth;
        const int nrx = std::min<int>(scene->rxCount, C);

        // path.h:237-238, 257, 268: Halton dims 0 (frequency), 1 and 2 (primary)
        const float u_f = radical_inverse(index, kPrimes[0]);
        const float u_a = radical_inverse(index, kPrimes[1]);
        const float u_b = radical_inverse(index, kPrimes[2]);
        float primary[2];
        if (knobs->primaryOrder == 0) { primary[0] = u_b; primary[1] = u_a; }
        else { primary[0] = u_a; primary[1] = u_b; }

        const Gd cG = Gd(kSpeedOfLight);
        const Gd maxDist = (Gd(rf.adcRate) / Gd(rf.freqSlope)) * cG;   // path.h:240
        const F freq = F(rf.startFreq) + F(u_f) * (F(rf.freqSlope) * F(rf.rampTime));  // path.h:257
        const F k0 = ((F(2) * piF) * freq) / F(kSpeedOfLight);        // geometry.h:97-99
        const Gd lambda = cG / Gd(freq);                                // geometry.h:102-104
        const Antenna &txAnt = txa ? *txa : scene->tx;
        const M33<Gd> txRg = mcast<Gd>(txAnt.rot);
        const M33<F> txRf = mcast<F>(txAnt.rot);
        // motion: d/dt of the path length up to the current vertex, and that vertex's velocity
        const V3<Gd> ego = vcast3<Gd>(scene->velocity);
        Gd rate = Gd(0);
        V3<Gd> vprev = ego;

        float primaryPdf = 1.0f;
        Cx<F> guidingWeight{F(0), F(0)};

        // depth 0: tx.sample (emitter.h:174-180) after the guiding warp (path.h:270-275)
        if (knobs->doGuiding && guide) primaryPdf = guide->sample(primary);
        const V3<Gd> dl = uniform_sample_sphere<Gd>(Gd(primary[0]), Gd(primary[1]));
        const CV3<F> Hl = awr_evaluate<F, kPromote>(vcast<F>(dl)) / K<F>::inv4pi();  // evaluate(d) / pdf(d)
        V3<Gd> o = vcast3<Gd>(txAnt.pos);
        V3<Gd> d = txRg.mul(dl);
        CV3<F> H = txRf.mul(Hl);
        // kDual: the FP32 chain (vertex, direction, incoming direction and normal, length, cosθ)
        V3<float> o32{}, d32{}, d32_in{}, n32_in{};
        float r32 = 0.0f, cos32 = 1.0f;
        const float maxDist32 = (float(rf.adcRate) / float(rf.freqSlope)) * float(kSpeedOfLight);
        if constexpr (kDual) {
            const V3<float> dl32 = uniform_sample_sphere<float>(primary[0], primary[1]);
            o32 = vcast3<float>(txAnt.pos);
            d32 = mcast<float>(txAnt.rot).mul(dl32);
        }
        H = (F(1) / F(primaryPdf)) * H;  // ray.weightBy(1 / primaryPdf)
        P time = P(0);
        Gd r = Gd(0), cosTheta = Gd(1);
        int depth = 0;
        // surface.incoming (path.h:439): normal, field and direction at the last hit
        V3<Gd> n_in{}, d_in{};
        CV3<F> H_in{};

        while (true) {
            // NEE (path.h:290-397); at depth 0 only when onlyIndirect is off
            if (!(depth == 0 && knobs->onlyIndirect)) {
                for (int ch = 0; ch < nrx; ++ch) {
                    const M33<Gd> rxRg = mcast<Gd>(scene->rx[ch].rot);
                    const M33<F> rxRf = mcast<F>(scene->rx[ch].rot);
                    // rx.nee (emitter.h:195-203)
                    V3<Gd> dn = vcast3<Gd>(scene->rx[ch].pos) - o;
                    const Gd rr = norm(dn);
                    dn = dn / rr;
                    const CV3<F> Hrx = rxRf.mul(awr_evaluate<F, kPromote>(vcast<F>(rxRg.mulT(-dn))));
                    V3<float> dn32{};
                    float rr32 = 0.0f;
                    if constexpr (kDual) {
                        dn32 = vcast3<float>(scene->rx[ch].pos) - o32;
                        rr32 = norm(dn32);
                        dn32 = dn32 / rr32;
                    }
                    if (!knobs->poDiffraction && !(kDual ? trace_visible(vcast<T>(o32), vcast<T>(dn32), T(rr32))
                                                         : trace_visible(vcast<T>(o), vcast<T>(dn), T(rr)))) {
                        if (rec && depth >= 1) rec->nee[(depth - 1) * C + chBase + ch] = 1;
                        continue;
                    }
                    // the leg to the receiver: d|rx - p|/dt = dn . (v_rx - v_p)
                    const Gd dvel = motion ? -((rate + dot(dn, ego - vprev)) / Gd(2)) : Gd(0);
                    const F rrF = F(rr);
                    CV3<F> Hn;
                    if (depth == 0) {
                        // tx.evaluate + tx.connect (emitter.h:182-193)
                        Hn = txRf.mul(awr_evaluate<F, kPromote>(vcast<F>(txRg.mulT(dn))));  // tx.evaluate
                        Hn = (F(1) / ((F(4) * piF) * rrF)) * Hn;
                    } else {
                        // SurfaceEmitter::evaluate (emitter.h:64-70)
                        const CV3<F> J = F(2) * cross(vcast<F>(n_in), H_in);
                        Hn = cross(vcast<F>(dn), J);
                        if (kDual ? dot(dn32, n32_in) < 0.0f : dot(dn, n_in) < Gd(0)) Hn = CV3<F>{};
                        // SurfaceEmitter::connect (emitter.h:72-81)
                        const Cx<F> green = Cx<F>{F(1) / std::max(rrF, F(1e-3)), k0} / ((F(4) * piF) * rrF);
                        Hn = Hn * green;
                    }
                    P tn = time + P(rr) / P(kSpeedOfLight);  // addDistance in connect
                    Cx<F> v = cdot(Hn, Hrx);                 // measureH, geometry.h:138
                    if (depth > 0 && knobs->correctPhase) {
                        // phase correction (path.h:316-329): the reflected ray itself measured at
                        // the receiver, its time the distance from the transmitter's mirror image
                        // o - r·d to the receiver (setDistance, geometry.h:113), its value
                        // measureH(H, Hrx)·π·dist / (π·R·λ)² / 4π; then the phase error and the
                        // hard sphere (the FP32 chain's in the dual instance), and gw += v
                        const V3<Gd> rxPos = o + rr * dn;
                        const Gd dist = norm((o - r * d) - rxPos);
                        tn = P(dist) / P(kSpeedOfLight);
                        const Gd piG = K<Gd>::pi();
                        const double den = (double)(piG * Gd(knobs->filteringRadius) * lambda);
                        v = cdot(H, Hrx);
                        v = v * F(piG * dist);
                        v = v / F(den * den);
                        v = v / (F(4) * piF);
                        if constexpr (kDual) {
                            const float lambda32 = float(kSpeedOfLight) / float(freq);
                            const float dp = phase_error<float>(*knobs, r32, o32, d32, dn32, rr32, d32_in, n32_in,
                                                                lambda32);
                            filter_weight<float, F>(*knobs, dp, v, guidingWeight);
                        } else {
                            const Gd dphase = phase_error<Gd>(*knobs, r, o, d, dn, rr, d_in, n_in, lambda);
                            filter_weight<Gd, F>(*knobs, dphase, v, guidingWeight);
                        }
                        guidingWeight = guidingWeight + v;  // path.h:392-394
                    } else if (depth > 0) {
                        if constexpr (kDual) {
                            // the GO weight from the Gd chain, the grazing skip, phase error and
                            // filter from the FP32 chain (as the ref32 instance)
                            if (knobs->useGeometricalOptics) {
                                v = v * F(r / cosTheta);
                                v = v / (F(4) * piF);
                                if (cos32 < 1e-3f) continue;
                            }
                            const float lambda32 = float(kSpeedOfLight) / float(freq);
                            const float dp = phase_error<float>(*knobs, r32, o32, d32, dn32, rr32, d32_in, n32_in,
                                                                lambda32);
                            filter_weight<float, F>(*knobs, dp, v, guidingWeight);
                            guidingWeight = guidingWeight + v;
                        } else if (!nee_weight<Gd, F>(*knobs, r, cosTheta, o, d, dn, rr, d_in, n_in, lambda, v,
                                                      guidingWeight, nullptr)) {
                            continue;
                        }
                    }
                    // Integrator::splat skip rule (integrator.h:121-124), off with the debug image
                    if (!knobs->produceDebugImage && (sampleWeight == 0.0f || v.re == F(0) || v.im == F(0))) {
                        if (rec && depth >= 1) rec->nee[(depth - 1) * C + chBase + ch] = 2;
                        continue;
                    }
                    splat_event<P>(frame, chirps, N, C, chBase + ch, tn, Cx<P>{P(v.re), P(v.im)}, P(sampleWeight),
                                   rf, cnt, motion, P(dvel));
                    if (rec && depth >= 1) rec->nee[(depth - 1) * C + chBase + ch] = 3;
                }
            }
            // random walk (path.h:399-442)
            if (depth >= maxDepth || (kDual ? r32 >= maxDist32 : r >= maxDist)) break;
            if (is_zero(H)) break;
            T t;
            const int32_t id = kDual ? trace_closest(vcast<T>(o32), vcast<T>(d32), &t)
                                     : trace_closest(vcast<T>(o), vcast<T>(d), &t);
            if (rec && depth < maxDepth) rec->hits[depth] = id;
            if (id < 0) break;
            V3<Gd> n = normal(id);
            if (dot(n, d) > Gd(0)) n = -n;  // face-forward (cpu.cpp:82-85)
            cosTheta = std::abs(dot(n, -d));
            V3<float> n32{};
            if constexpr (kDual) {
                n32 = mesh->f.normal(id);
                if (dot(n32, d32) > 0.0f) n32 = -n32;
                cos32 = std::abs(dot(n32, -d32));
                if (cos32 < 1e-3f) break;
            } else if (cosTheta < Gd(1e-3)) {
                break;
            }
            if (motion) {
                // the incoming leg: d|p - o|/dt = d . (v_p - v_o), the vertex moving with its triangle
                const V3<Gd> vh = mesh->vel.empty() ? V3<Gd>{Gd(0), Gd(0), Gd(0)} : vcast3<Gd>(&mesh->vel[3 * (size_t)id]);
                rate = rate + dot(d, vh - vprev);
                vprev = vh;
            }
            const Gd tg = refine ? plane_distance(id, o, d) : Gd(t);
            r = r + tg;
            time = time + P(tg) / P(kSpeedOfLight);  // addDistance, geometry.h:110
            if (!knobs->useGeometricalOptics) {
                // path.h:420-427: the emitter's connect (NFAntenna::connect 1/(4πt) from the
                // transmitter, emitter.h:188-193; SurfaceEmitter::connect (jk0 + 1/max(t, 1e-3))
                // / (4πt) from a surface, emitter.h:72-81), then the hemisphere-to-surface
                // density t²/cosθ
                const F tF = F(tg);
                if (depth == 0) {
                    H = (F(1) / ((F(4) * piF) * tF)) * H;
                } else {
                    const Cx<F> green = Cx<F>{F(1) / std::max(tF, F(1e-3)), k0} / ((F(4) * piF) * tF);
                    H = H * green;
                }
                H = F(tg * tg / cosTheta) * H;
            }
            if (is_zero(H)) break;
            n_in = n;
            H_in = H;
            d_in = d;
            o = o + tg * d;
            depth++;
            if constexpr (kDual) {
                const float t32 = float(t);
                r32 = r32 + t32;
                n32_in = n32;
                d32_in = d32;
                o32 = o32 + t32 * d32;
                const V3<float> wi32 = -d32;
                d32 = (2.0f * dot(n32, wi32)) * n32 - wi32;
            }
            if (!knobs->useGeometricalOptics) {
                // SurfaceEmitter::sample (emitter.h:58-62) with get2D (path.h:277-278): the
                // direction from Halton dimensions 3 + 2(k-1), 4 + 2(k-1) in get2D's order, H
                // from SurfaceEmitter::evaluate (emitter.h:64-70), J = 2 n x H_in, H = d x J (the
                // 4π weight of the sphere pdf is overwritten by setH), zero below the surface
                const int dimA = 3 + 2 * (depth - 1);
                const float ua = halton_dim(index, dimA), ub = halton_dim(index, dimA + 1);
                const float s0 = knobs->primaryOrder == 0 ? ub : ua, s1 = knobs->primaryOrder == 0 ? ua : ub;
                d = uniform_sample_sphere<Gd>(Gd(s0), Gd(s1));
                const CV3<F> J = F(2) * cross(vcast<F>(n_in), H_in);
                H = cross(vcast<F>(d), J);
                if constexpr (kDual) {
                    d32 = uniform_sample_sphere<float>(s0, s1);
                    if (dot(d32, n32_in) < 0.0f) H = CV3<F>{};
                } else if (dot(d, n_in) < Gd(0)) {
                    H = CV3<F>{};
                }
                continue;
            }
            // reflectRay (integrator.h:164-168) with Intersection::R (intersection.h:38)
            const V3<Gd> wi = -d;
            d = (Gd(2) * dot(n, wi)) * n - wi;
            const V3<F> nF = vcast<F>(n);
            H = rscale(F(2) * nF, rdot(nF, H)) - H;
        }
        // guiding record (path.h:448-450)
        if (knobs->doGuiding && guide && guide_stat && !isFinal && primaryPdf > 0.0f) {
            const float density = (float)cabs(guidingWeight) * primaryPdf;
            const float weight = 1.0f / primaryPdf;
            const int32_t leaf = guide->training.leaf_of(primary);
            guide_stat[leaf] += Tree::stat_value(density, weight, guide->secondMoment);
        }
    }
};

// ---------------------------------------------------------------------------------------
// Backend::run (cpu.h:21-41): a std::thread pool grabbing batches of 256 consecutive
// indices under a mutex; one accumulation frame per worker, summed after the join.
template <class T, class Gd, class F, class P, bool kPromote, bool kDual = false>
void render(const Mesh &mesh, const SceneC &scene, const Knobs &knobs, const FrameCfg &fc, int64_t offset,
            int64_t count, float sampleWeight, bool isFinal, Guide *guide, int nthreads, double *frame_out,
            double *totalWeight, int32_t *hit_ids, int8_t *nee, int64_t record_paths, Counters *counters) {
    const int N = fc.samplesPerChirp, C = fc.channelCount, chirps = std::max(1, fc.chirpCount);
    const int64_t fsize = 2LL * chirps * N * C;
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::vector<P>> frames(nthreads, std::vector<P>(fsize, P(0)));
    std::vector<std::vector<double>> gstats(nthreads);
    std::vector<Counters> cnts(nthreads);
    if (guide) for (auto &g : gstats) g.assign(guide->training.child.size(), 0.0);
    std::mutex mtx;
    int64_t next = 0;
    auto worker = [&](int tid) {
        Tracer<T, Gd, F, P, kPromote, kDual> tr;
        tr.mesh = &mesh;
        tr.acc = &accel_of<T>(mesh);
        tr.scene = &scene;
        tr.knobs = &knobs;
        tr.N = N;
        tr.C = C;
        tr.chirps = chirps;
        tr.frame = frames[tid].data();
        tr.guide = guide;
        tr.guide_stat = guide ? gstats[tid].data() : nullptr;
        tr.isFinal = isFinal;
        tr.sampleWeight = sampleWeight;
        tr.cnt = counters ? &cnts[tid] : nullptr;
        tr.motion = !mesh.vel.empty() || scene.velocity[0] != 0.0f || scene.velocity[1] != 0.0f ||
                    scene.velocity[2] != 0.0f;
        const int ntx = std::max(1, std::min<int>(scene.txCount, kMaxTx));
        const int nrx = std::min<int>(scene.rxCount, C);
        const int maxDepth = std::max(1, knobs.maxDepth);
        while (true) {
            int64_t begin, batch;
            {
                std::lock_guard<std::mutex> lock(mtx);
                batch = std::min<int64_t>(count - next, 256);
                if (batch <= 0) break;
                begin = next;
                next += batch;
            }
            for (int64_t j = 0; j < batch; ++j) {
                const int64_t i = begin + j;
                PathRecord rec{nullptr, nullptr};
                PathRecord *recp = nullptr;
                if (i < record_paths && hit_ids && nee) {
                    rec.hits = hit_ids + i * maxDepth;
                    rec.nee = nee + i * (int64_t)maxDepth * C;
                    for (int q = 0; q < maxDepth; ++q) rec.hits[q] = -1;
                    std::memset(rec.nee, 0, (size_t)maxDepth * C);
                    recp = &rec;
                }
                // MIMO: every transmitter emits the same path index into its block of channels
                for (int t = 0; t < ntx; ++t) {
                    tr.txa = t == 0 ? &scene.tx : &scene.txMore[t - 1];
                    tr.chBase = t * nrx;
                    tr.sample((uint64_t)(offset + i), recp);
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t) pool.emplace_back(worker, t);
    for (auto &th : pool) th.join();
    for (int t = 0; t < nthreads; ++t)
        for (int64_t k = 0; k < fsize; ++k) frame_out[k] += (double)frames[t][k];
    if (guide)
        for (int t = 0; t < nthreads; ++t)
            for (size_t k = 0; k < gstats[t].size(); ++k) guide->training.stat[k] += gstats[t][k];
    if (totalWeight) *totalWeight += (double)count * (double)sampleWeight;
    if (counters)
        for (auto &c : cnts) {
            counters->nodes += c.nodes;
            counters->tris += c.tris;
            counters->rays_closest += c.rays_closest;
            counters->rays_shadow += c.rays_shadow;
            counters->events += c.events;
        }
}

typedef void (*RenderFn)(const Mesh &, const SceneC &, const Knobs &, const FrameCfg &, int64_t, int64_t, float,
                         bool, Guide *, int, double *, double *, int32_t *, int8_t *, int64_t, Counters *);
static RenderFn render_for(int prec) {
    switch (prec) {
        case 0: return &render<float, float, float, float, true>;
        case 1: return &render<float, double, float, double, false>;
        case 3: return &render<float, float, float, double, false>;
        case 4: return &render<float, double, float, double, false, true>;
        default: return &render<double, double, double, double, true>;
    }
}

}  // namespace orc

// =======================================================================================
// C entry points for the test harness (ctypes).
using namespace orc;

extern "C" {

int orc_abi_version() { return 1; }

float orc_radical_inverse(uint64_t v, int base) { return radical_inverse(v, base); }

void orc_halton_batch(uint64_t start, int64_t n, int dim, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = radical_inverse(start + (uint64_t)i, kPrimes[dim]);
}

void orc_sincos_f32(const float *x, int64_t n, float *s, float *c) {
    for (int64_t i = 0; i < n; ++i) sincos_f32(x[i], s + i, c + i);
}

// prec: 0 float, else double. u: [n*2]; out: [n*3]
void orc_sphere_sample(int prec, const float *u, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        if (prec == 0) {
            V3<float> d = uniform_sample_sphere<float>(u[2 * i], u[2 * i + 1]);
            out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
        } else {
            V3<double> d = uniform_sample_sphere<double>(u[2 * i], u[2 * i + 1]);
            out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
        }
    }
}

// d: [n*3] (double), out: [n*6] (re, im per component)
void orc_awr_evaluate(int prec, const double *d, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) {
        if (prec == 0) {
            CV3<float> H = awr_evaluate<float>(v3<float>((float)d[3 * i], (float)d[3 * i + 1], (float)d[3 * i + 2]));
            const Cx<float> c[3] = {H.x, H.y, H.z};
            for (int k = 0; k < 3; ++k) { out[6 * i + 2 * k] = c[k].re; out[6 * i + 2 * k + 1] = c[k].im; }
        } else {
            CV3<double> H = awr_evaluate<double>(v3<double>(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
            const Cx<double> c[3] = {H.x, H.y, H.z};
            for (int k = 0; k < 3; ++k) { out[6 * i + 2 * k] = c[k].re; out[6 * i + 2 * k + 1] = c[k].im; }
        }
    }
}

void *orc_mesh_create(const float *V, int64_t nv, const int32_t *I, int64_t nt) {
    (void)nv;
    Mesh *m = new Mesh();
    m->nt = nt;
    m->f.build(V, I, nt);
    m->d.build(V, I, nt);
    return m;
}
void orc_mesh_destroy(void *m) { delete (Mesh *)m; }
void orc_mesh_set_brute(void *m, int brute) { ((Mesh *)m)->brute = brute; }
// per-triangle velocities (3 floats each, triangle order); vel NULL or n 0: static mesh
int orc_mesh_set_velocities(void *mh, const float *vel, int64_t n) {
    Mesh *m = (Mesh *)mh;
    if (!vel || n == 0) {
        m->vel.clear();
        return 0;
    }
    if (n != m->nt) return -1;
    m->vel.assign(vel, vel + 3 * n);
    return 0;
}
int64_t orc_mesh_node_count(void *m) { return (int64_t)((Mesh *)m)->f.nodes.size(); }

// rays: [n*6] o,d (float). mode: 0 BVH, 1 brute force. any_hit: 0 closest, 1 occlusion.
// out_ids: closest id (-1 miss) or 1/0 occluded; out_t: t (closest only). counters: [2] nodes, tris
void orc_intersect(void *mh, int prec, int mode, int any_hit, const float *rays, int64_t n, float tnear, float tfar,
                   int32_t *out_ids, double *out_t, int64_t *counters) {
    const Mesh &m = *(Mesh *)mh;
    Counters c;
    for (int64_t i = 0; i < n; ++i) {
        const float *r = rays + 6 * i;
        if (prec == 0) {
            V3<float> o = v3<float>(r[0], r[1], r[2]), d = v3<float>(r[3], r[4], r[5]);
            if (any_hit) {
                bool occ = mode ? m.f.occluded_brute(o, d, tnear, tfar) : m.f.occluded(o, d, tnear, tfar, &c);
                out_ids[i] = occ ? 1 : 0;
            } else {
                float t;
                out_ids[i] = mode ? m.f.closest_brute(o, d, tnear, tfar, &t) : m.f.closest(o, d, tnear, tfar, &t, &c);
                if (out_t) out_t[i] = t;
            }
        } else {
            V3<double> o = v3<double>(r[0], r[1], r[2]), d = v3<double>(r[3], r[4], r[5]);
            if (any_hit) {
                bool occ = mode ? m.d.occluded_brute(o, d, tnear, tfar) : m.d.occluded(o, d, tnear, tfar, &c);
                out_ids[i] = occ ? 1 : 0;
            } else {
                double t;
                out_ids[i] = mode ? m.d.closest_brute(o, d, tnear, tfar, &t) : m.d.closest(o, d, tnear, tfar, &t, &c);
                if (out_t) out_t[i] = t;
            }
        }
    }
    if (counters) { counters[0] += c.nodes; counters[1] += c.tris; }
}

// One splat event into frame[2*N*C] (double). prec 0: FP32 arithmetic, else FP64.
// one splat event into frame[2*M*N*C]; motion != 0 places it on the chirp axis by delta_v
void orc_splat(int prec, const Rf *rf, int M, int N, int C, int ch, double delta_t, double vre, double vim,
               double weight, int motion, double delta_v, double *frame) {
    if (prec == 0) {
        std::vector<float> f(2 * (size_t)M * N * C, 0.0f);
        splat_event<float>(f.data(), M, N, C, ch, (float)delta_t, Cx<float>{(float)vre, (float)vim}, (float)weight,
                           *rf, nullptr, motion != 0, (float)delta_v);
        for (size_t k = 0; k < f.size(); ++k) frame[k] += f[k];
    } else {
        splat_event<double>(frame, M, N, C, ch, delta_t, Cx<double>{vre, vim}, weight, *rf, nullptr, motion != 0,
                            delta_v);
    }
}

// Frame::argmax (radar.h:534-546: the first largest |value|) and Frame::frequencyEstimation
// (radar.h:517-531, 631-643: per axis from the last, neighbours at -off/+off wrapping, b -
// l/(l+m) when l > r, else b + r/(r+m), b when all three are zero) of a complex128 frame
// [M][N][C]. index[3] and refined[3] in (chirp, sample, channel) order.
void orc_frame_peak(const double *frame, int M, int N, int C, int32_t *index, double *refined) {
    const int dims[3] = {M, N, C};
    const int64_t n = (int64_t)M * N * C;
    auto mag = [&](int64_t i) { return std::hypot(frame[2 * i], frame[2 * i + 1]); };
    int64_t best = 0;
    double bm = mag(0);
    for (int64_t i = 1; i < n; ++i) {
        const double c = mag(i);
        if (c > bm) {
            bm = c;
            best = i;
        }
    }
    int64_t rem = best;
    for (int i = 2; i >= 0; --i) {
        index[i] = (int32_t)(rem % dims[i]);
        rem /= dims[i];
    }
    int64_t off = 1;
    for (int i = 2; i >= 0; --i) {
        const int b = index[i], mx = dims[i];
        const double l = mag(b > 0 ? best - off : best + (int64_t)(mx - 1) * off);
        const double m = bm;
        const double r = mag(b < mx - 1 ? best + off : best - (int64_t)(mx - 1) * off);
        if (l + m + r == 0.0) refined[i] = b;
        else if (l > r) refined[i] = b - l / (l + m);
        else refined[i] = b + r / (r + m);
        off *= mx;
    }
}

// ---- guide
void *orc_guide_create(float uniformProb, float splitThreshold, int secondMoment) {
    Guide *g = new Guide();
    g->uniformProb = uniformProb;
    g->splitThreshold = splitThreshold;
    g->secondMoment = secondMoment != 0;
    return g;
}
void orc_guide_destroy(void *g) { delete (Guide *)g; }
void orc_guide_reset(void *g) { ((Guide *)g)->reset(); }
void orc_guide_step(void *g) { ((Guide *)g)->step(); }
// mixture sample: x [n*2] warped in place, pdf [n]
void orc_guide_sample(void *gh, float *x, int64_t n, float *pdf) {
    Guide *g = (Guide *)gh;
    for (int64_t i = 0; i < n; ++i) pdf[i] = g->sample(x + 2 * i);
}
// tree-only sample / pdf on the sampling tree
void orc_tree_sample(void *gh, float *x, int64_t n, float *pdf) {
    Guide *g = (Guide *)gh;
    for (int64_t i = 0; i < n; ++i) pdf[i] = g->sampling.sample(x + 2 * i);
}
void orc_tree_pdf(void *gh, const float *x, int64_t n, float *pdf) {
    Guide *g = (Guide *)gh;
    for (int64_t i = 0; i < n; ++i) pdf[i] = g->sampling.pdf(x + 2 * i);
}
void orc_guide_record(void *gh, const float *x, const float *density, const float *weight, int64_t n) {
    Guide *g = (Guide *)gh;
    for (int64_t i = 0; i < n; ++i)
        g->training.stat[g->training.leaf_of(x + 2 * i)] += Tree::stat_value(density[i], weight[i], g->secondMoment);
}
int32_t orc_guide_size(void *gh, int which) {
    Guide *g = (Guide *)gh;
    return (int32_t)(which == 0 ? g->sampling.child.size() : g->training.child.size());
}
// which 0: sampling (child, p0); 1: training (child, stat)
void orc_guide_get(void *gh, int which, int32_t *child, float *p0, double *stat) {
    Guide *g = (Guide *)gh;
    const Tree &t = which == 0 ? g->sampling : g->training;
    for (size_t i = 0; i < t.child.size(); ++i) {
        if (child) child[i] = t.child[i];
        if (p0) p0[i] = t.p0[i];
        if (stat) stat[i] = t.stat[i];
    }
}
// replace the training tree and its statistics (depth recomputed from the structure)
void orc_guide_set_training(void *gh, const int32_t *child, const double *stat, int32_t n) {
    Guide *g = (Guide *)gh;
    Tree &t = g->training;
    t.child.assign(child, child + n);
    t.p0.assign(n, 0.5f);
    t.stat.assign(stat, stat + n);
    t.mass.assign(n, 0.0);
    t.depth.assign(n, 0);
    for (int32_t i = 0; i < n; ++i)
        if (t.child[i]) { t.depth[t.child[i]] = t.depth[i] + 1; t.depth[t.child[i] + 1] = t.depth[i] + 1; }
}
// NEE weight of one contribution v at a bounce vertex (nee_weight), in double (prec 2) or in
// float (prec 0): vec = o[3], d[3], dn[3], d_in[3], n_in[3]; sc = r, cosTheta, rr, lambda;
// v [2] in/out, gw [2] in/out; returns 1 when the contribution is kept, 0 when skipped.
int orc_nee_weight(int prec, const Knobs *k, const double *vec, const double *sc, double *v, double *gw,
                   double *dphase) {
    auto run = [&](auto zero) -> int {
        using G = decltype(zero);
        auto V = [&](int i) { return v3<G>(G(vec[3 * i]), G(vec[3 * i + 1]), G(vec[3 * i + 2])); };
        Cx<G> vv{G(v[0]), G(v[1])}, gg{G(gw[0]), G(gw[1])};
        G dp = G(0);
        const bool kept = nee_weight<G, G>(*k, G(sc[0]), G(sc[1]), V(0), V(1), V(2), G(sc[2]), V(3), V(4), G(sc[3]),
                                           vv, gg, &dp);
        v[0] = vv.re; v[1] = vv.im;
        gw[0] = gg.re; gw[1] = gg.im;
        if (dphase) *dphase = dp;
        return kept ? 1 : 0;
    };
    return prec == 0 ? run(0.0f) : run(0.0);
}
// replace the sampling tree (depth recomputed from the structure)
void orc_guide_set_sampling(void *gh, const int32_t *child, const float *p0, int32_t n) {
    Guide *g = (Guide *)gh;
    Tree &t = g->sampling;
    t.child.assign(child, child + n);
    t.p0.assign(p0, p0 + n);
    t.stat.assign(n, 0.0);
    t.mass.assign(n, 0.0);
    t.depth.assign(n, 0);
    for (int32_t i = 0; i < n; ++i)
        if (t.child[i]) { t.depth[t.child[i]] = t.depth[i] + 1; t.depth[t.child[i] + 1] = t.depth[i] + 1; }
}

// ---- integrator
// Backend::run equivalent: paths offset .. offset+count-1 accumulated into frame_out
// (double[2*chirps*N*C], +=). counters: [5] nodes, tris, rays_closest, rays_shadow, events.
int orc_render(void *mh, const SceneC *scene, const Knobs *knobs, const FrameCfg *fc, int prec, int64_t offset,
               int64_t count, float sampleWeight, int isFinal, void *guide, int nthreads, double *frame_out,
               double *totalWeight, int32_t *hit_ids, int8_t *nee, int64_t record_paths, int64_t *counters) {
    if (!mh || !scene || !knobs || !fc || !frame_out) return 1;
    if (knobs->correctPhase && !knobs->filteringSphere) return 1;  // path.h:318
    // geometrical optics off: the reference Asserts it for correctPhase (path.h:317) and the
    // sphere filter (path.h:344), and its sampler has Halton dimensions 0-9 (halton.h:58-67),
    // which depth 3 uses up
    if (!knobs->useGeometricalOptics && (knobs->correctPhase || knobs->filteringSphere || knobs->maxDepth > 3))
        return 1;
    if (fc->samplesPerChirp <= 0 || fc->channelCount <= 0) return 1;
    Counters c;
    render_for(prec)(*(Mesh *)mh, *scene, *knobs, *fc, offset, count, sampleWeight, isFinal != 0, (Guide *)guide,
                     nthreads, frame_out, totalWeight, hit_ids, nee, record_paths, counters ? &c : nullptr);
    if (counters) {
        counters[0] += c.nodes;
        counters[1] += c.tris;
        counters[2] += c.rays_closest;
        counters[3] += c.rays_shadow;
        counters[4] += c.events;
    }
    return 0;
}

// PathTracer::run (path.h:171-224): frame_out receives the final (un-normalised) frame,
// totalWeight the final totalWeight.
int orc_path_tracer_run(void *mh, const SceneC *scene, const Knobs *knobs, const FrameCfg *fc, int prec,
                        int64_t samples, void *guide_h, int nthreads, double *frame_out, double *totalWeight) {
    if (!mh || !scene || !knobs || !fc || !frame_out || !totalWeight) return 1;
    const int64_t fsize = 2LL * std::max(1, fc->chirpCount) * fc->samplesPerChirp * fc->channelCount;
    std::fill(frame_out, frame_out + fsize, 0.0);
    *totalWeight = 0.0;
    int64_t offset = 0;
    float weight = 1.0f;
    Guide *guide = (Guide *)guide_h;
    if (!knobs->doGuiding || !guide) {
        return orc_render(mh, scene, knobs, fc, prec, 0, samples, 1.0f, 0, nullptr, nthreads, frame_out, totalWeight,
                          nullptr, nullptr, 0, nullptr);
    }
    guide->reset();
    bool isFinal = false;
    int64_t milestone = 16384, remaining = samples;
    while (true) {
        milestone = std::min(milestone, remaining);
        int rc = orc_render(mh, scene, knobs, fc, prec, offset, milestone, weight, isFinal, guide, nthreads, frame_out,
                            totalWeight, nullptr, nullptr, 0, nullptr);
        if (rc) return rc;
        offset += milestone;
        remaining -= milestone;
        if (remaining == 0) break;
        milestone *= 2;
        if (remaining < milestone * 2) {
            isFinal = true;
            milestone = remaining;
        }
        if (knobs->clearBeforeIteration) {
            std::fill(frame_out, frame_out + fsize, 0.0);
            *totalWeight = 0.0;
            offset = 0;
        } else {
            weight = 10.0f * weight;
        }
        guide->step();
    }
    return 0;
}

// Explicit schedule: `passes` training passes of `train` paths (each from index 0, frame
// cleared, followed by a guide step), then `final_paths` paths into a cleared frame.
int orc_path_tracer_run_schedule(void *mh, const SceneC *scene, const Knobs *knobs, const FrameCfg *fc, int prec,
                                 int passes, int64_t train, int64_t final_paths, void *guide_h, int nthreads,
                                 double *frame_out, double *totalWeight) {
    if (!mh || !scene || !knobs || !fc || !frame_out || !totalWeight) return 1;
    const int64_t fsize = 2LL * std::max(1, fc->chirpCount) * fc->samplesPerChirp * fc->channelCount;
    Guide *guide = (Guide *)guide_h;
    const bool guiding = knobs->doGuiding && guide;
    if (guiding) {
        guide->reset();
        for (int p = 0; p < passes; ++p) {
            std::fill(frame_out, frame_out + fsize, 0.0);
            *totalWeight = 0.0;
            int rc = orc_render(mh, scene, knobs, fc, prec, 0, train, 1.0f, 0, guide, nthreads, frame_out, totalWeight,
                                nullptr, nullptr, 0, nullptr);
            if (rc) return rc;
            guide->step();
        }
    }
    std::fill(frame_out, frame_out + fsize, 0.0);
    *totalWeight = 0.0;
    return orc_render(mh, scene, knobs, fc, prec, 0, final_paths, 1.0f, 1, guiding ? guide : nullptr, nthreads,
                      frame_out, totalWeight, nullptr, nullptr, 0, nullptr);
}

}  // extern "C"

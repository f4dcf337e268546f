// Device code shared by the fused path kernel (kernels.hip) and the wavefront pipeline
// (wavefront.hip): constants, BVH traversal with the LDS-staged short stack, the guiding tree,
// path generation, NEE set-up, the hit update, and the splat helpers.
//
// Precision model: FP32 traversal, triangle tests, fields and weights; FP64 geometry chain
// (directions, hit points, segment lengths re-evaluated against the hit triangle's plane),
// FP64 path time, IF phase and frame sums. Compiled without FP contraction.
#pragma once

#include "device_math.h"
#include "kernels.h"

namespace hsr {

static constexpr double kC = 299792448.0;  // radar.h:107, 299792458f after float rounding
static constexpr float kPiF = 3.14159265358979323846f;
static constexpr float kInv4PiF = 0.07957747154594766788f;
static constexpr double kPiD = 3.14159265358979323846;
static constexpr float kEps = 5e-5f;  // hussar.h:52
static constexpr float kOneMinusEps = 0x1.fffffep-1f;

__device__ inline mat3<float> rotf(const DevAntenna &a) {
    mat3<float> m;
#pragma unroll
    for (int i = 0; i < 9; ++i) m.m[i] = a.rot[i];
    return m;
}
__device__ inline mat3<double> rotd(const DevAntenna &a) {
    mat3<double> m;
#pragma unroll
    for (int i = 0; i < 9; ++i) m.m[i] = (double)a.rot[i];
    return m;
}
__device__ inline d3 posd(const DevAntenna &a) { return {(double)a.pos[0], (double)a.pos[1], (double)a.pos[2]}; }

// ------------------------------------------------------------------------ traversal stack
// The spill entries live in a private array of the kernel, reached through a pointer, so that
// the stack pointer itself stays in a register (kept in one aggregate with the array, it would
// be placed in private memory with it and every push and pop would load it from there).
// The LDS column is typed as LDS (address space 3) so that its accesses are ds_read/ds_write
// and are not merged with the spill's into one flat access.
typedef __attribute__((address_space(3))) int32_t lds_int32;
typedef __attribute__((address_space(3))) double lds_double;
// FP64 add to an LDS word without a return value (ds_add_f64).
__device__ inline void lds_add(lds_double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
struct Stack {
    lds_int32 *lds;  // column of this thread: lds[i * kTraceBlock]
    int32_t *spill;  // [kStackSpill] private
    int sp;
    __device__ inline void set_column(int32_t *column) { lds = (lds_int32 *)column; }
    __device__ inline void push(int32_t v) {
        if (__builtin_expect(sp < kStackShort, 1)) lds[sp * kTraceBlock] = v;
        else spill[sp - kStackShort] = v;
        ++sp;
    }
    __device__ inline int32_t pop() {
        --sp;
        int32_t v;
        if (__builtin_expect(sp < kStackShort, 1)) v = lds[sp * kTraceBlock];
        else v = spill[sp - kStackShort];
        return v;
    }
    // Pushes the children v1..v(n-1) of a visited node, sorted nearest first (n = 1..4 entered
    // children, v0 taken next), farthest first, so that v1 ends on top. While three entries fit
    // in the LDS part it writes sp, sp + 1 and sp + 2 without branches (entries past the new top
    // are dead) and advances sp by n - 1.
    __device__ inline void push_sorted(int n, int32_t v1, int32_t v2, int32_t v3) {
        if (__builtin_expect(sp + 3 <= kStackShort, 1)) {
            lds[sp * kTraceBlock] = n == 4 ? v3 : (n == 3 ? v2 : v1);
            lds[(sp + 1) * kTraceBlock] = n == 4 ? v2 : v1;
            lds[(sp + 2) * kTraceBlock] = v1;
            sp += n - 1;
        } else {
            if (n == 4) push(v3);
            if (n >= 3) push(v2);
            if (n >= 2) push(v1);
        }
    }
};

// Four-wide BVH node (64 B = 4 float4) with 8-bit child boxes: origin.xyz, scale.x | qlo.x,
// qhi.x, qlo.y, qhi.y | qlo.z, qhi.z, scale.y, scale.z | child[4]; byte i of a q word belongs
// to child i and decodes to origin + q * scale (scale a power of two: product exact, sum
// rounded; the builder chose q so that the decoded box contains the child's box). A child >= 0
// is an inner node; a negative value a leaf: bit 31 set, bits 27-30 count-1, bits 0-26 first
// triangle slot. An unused child slot holds the inverted box (qlo 255, qhi 0) and the sentinel
// leaf, a degenerate triangle that no ray hits, so the box test needs no check of the slot.
// The float nodes (Node4, any-hit) mark unused slots with kNoChild.
constexpr int32_t kNoChild = 0x7fffffff;

__device__ inline void leaf_decode(int32_t c, int &first, int &count) {
    first = c & 0x07FFFFFF;
    count = ((c >> 27) & 0xF) + 1;
}

struct SlabRay {
    f3 o, invd;
    float tnear;
};

__device__ inline bool slab(float lx, float hx, float ly, float hy, float lz, float hz, const SlabRay &r,
                            float tbest, float &tentry) {
    const float ax = (lx - r.o.x) * r.invd.x, bx = (hx - r.o.x) * r.invd.x;
    const float ay = (ly - r.o.y) * r.invd.y, by = (hy - r.o.y) * r.invd.y;
    const float az = (lz - r.o.z) * r.invd.z, bz = (hz - r.o.z) * r.invd.z;
    const float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), r.tnear));
    float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    t1 = t1 * 1.0000004f;  // 1 + 2*gamma(3): robust traversal (Ize 2013)
    tentry = t0;
    return t0 <= t1 && t0 <= tbest;
}

struct Node4 {
    float4 lx, hx, ly, hy, lz, hz;
    int4 ch;
};

// The float layout of the same node (128 B = 8 float4: lo.x[4] | hi.x[4] | lo.y[4] | hi.y[4] |
// lo.z[4] | hi.z[4] | child[4] | -), used by the any-hit traversals: their packets test every
// child against several rays, where exact boxes save more visits than the smaller fetch.
__device__ inline Node4 load_node4f(const float4 *__restrict__ nodes, int32_t n) {
    const float4 *p = nodes + 8 * (int64_t)n;
    Node4 q;
    q.lx = p[0];
    q.hx = p[1];
    q.ly = p[2];
    q.hy = p[3];
    q.lz = p[4];
    q.hz = p[5];
    q.ch = __builtin_bit_cast(int4, p[6]);
    return q;
}

template <int i> __device__ inline float lane4(const float4 &v) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
template <int i> __device__ inline int32_t lane4(const int4 &v) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

template <int i> __device__ inline bool slab_child(const Node4 &q, const SlabRay &r, float tb, float &te) {
    return slab(lane4<i>(q.lx), lane4<i>(q.hx), lane4<i>(q.ly), lane4<i>(q.hy), lane4<i>(q.lz), lane4<i>(q.hz), r,
                tb, te) &&
           lane4<i>(q.ch) != kNoChild;
}

// Boxes of a node's four children against one ray: bit i of the mask = child i is hit before tb.
__device__ inline uint32_t slab4(const Node4 &q, const SlabRay &r, float tb, float te[4]) {
    uint32_t m = 0;
    if (slab_child<0>(q, r, tb, te[0])) m |= 1u;
    if (slab_child<1>(q, r, tb, te[1])) m |= 2u;
    if (slab_child<2>(q, r, tb, te[2])) m |= 4u;
    if (slab_child<3>(q, r, tb, te[3])) m |= 8u;
    return m;
}

// The closest-hit box test of a quantised node's four children, in the node's frame: a child
// bound origin + q*2^e is at ray distance t = q*a + b, with a = 2^e * invd (exact: a power of
// two) and b = (origin - o) * invd, so each plane costs one fused multiply-add instead of a
// decode, a subtraction and a product, and the ray's direction signs pick each axis' near and
// far planes (no min/max). b carries two roundings, so t is exact to within one rounding of
// itself plus 2.00001 u |b| (u = 2^-24): the near planes use b - err and the far planes
// b + err, err = 2^-20 |b| per axis, and the exit keeps the robust factor 1 + 2 gamma_3 (Ize
// 2013), which makes the test a superset of the exact one. The bound must stay per axis: an
// axis whose direction component is tiny has a huge |b|, and its error would open every box
// along the other axes. The real boxes origin + q*2^e lie
// within one rounding of the decoded boxes the builder checked, and it widened every box by
// far more. invd is finite (safe_rcp), so no product is NaN.
struct QuadHit {
    uint32_t m;     // bit i: child i is entered
    float k[4];     // sort keys: the entry distance of an entered child, +inf otherwise
    int32_t c[4];   // child refs
};

__device__ inline QuadHit slab4q(const float4 *__restrict__ nodes, int32_t n, const SlabRay &r, float tb) {
    const float4 *p = nodes + 4 * (int64_t)n;
    const float4 h = p[0];
    const uint4 qa = __builtin_bit_cast(uint4, p[1]);
    const float4 pb = p[2];
    const int4 ch = __builtin_bit_cast(int4, p[3]);
    QuadHit qh;
    qh.c[0] = ch.x;
    qh.c[1] = ch.y;
    qh.c[2] = ch.z;
    qh.c[3] = ch.w;
    const float ax = h.w * r.invd.x, ay = pb.z * r.invd.y, az = pb.w * r.invd.z;
    const float bx = (h.x - r.o.x) * r.invd.x, by = (h.y - r.o.y) * r.invd.y, bz = (h.z - r.o.z) * r.invd.z;
    constexpr float kErr = 9.5367431640625e-07f;  // 2^-20
    const float ex_ = fabsf(bx) * kErr, ey_ = fabsf(by) * kErr, ez_ = fabsf(bz) * kErr;
    const float bnx = bx - ex_, bfx = bx + ex_, bny = by - ey_, bfy = by + ey_, bnz = bz - ez_, bfz = bz + ez_;
    // near and far plane bytes per axis: the lower bounds are near for a positive direction
    const bool sx = r.invd.x >= 0.0f, sy = r.invd.y >= 0.0f, sz = r.invd.z >= 0.0f;
    const uint32_t qz0 = __float_as_uint(pb.x), qz1 = __float_as_uint(pb.y);
    const uint32_t nxw = sx ? qa.x : qa.y, fxw = sx ? qa.y : qa.x;
    const uint32_t nyw = sy ? qa.z : qa.w, fyw = sy ? qa.w : qa.z;
    const uint32_t nzw = sz ? qz0 : qz1, fzw = sz ? qz1 : qz0;
    qh.m = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float nx = __builtin_fmaf((float)((nxw >> (8 * i)) & 0xFFu), ax, bnx);
        const float fx = __builtin_fmaf((float)((fxw >> (8 * i)) & 0xFFu), ax, bfx);
        const float ny = __builtin_fmaf((float)((nyw >> (8 * i)) & 0xFFu), ay, bny);
        const float fy = __builtin_fmaf((float)((fyw >> (8 * i)) & 0xFFu), ay, bfy);
        const float nz = __builtin_fmaf((float)((nzw >> (8 * i)) & 0xFFu), az, bnz);
        const float fz = __builtin_fmaf((float)((fzw >> (8 * i)) & 0xFFu), az, bfz);
        const float t0 = fmaxf(fmaxf(nx, ny), fmaxf(nz, r.tnear));
        const float t1 = fminf(fminf(fx, fy), fz) * 1.0000004f;
        const bool hit = t0 <= t1 && t0 <= tb;
        qh.k[i] = hit ? t0 : __builtin_inff();
        if (hit) qh.m |= 1u << i;
    }
    return qh;
}

// Reciprocal direction for the slab tests, finite: a zero (or subnormal) component becomes
// +-2^100, which keeps every box test conservative (the products stay finite).
__device__ inline float safe_rcp(float v) {
    return fabsf(v) > 7.888609052210118e-31f ? 1.0f / v : copysignf(1.2676506002282294e30f, v);
}

__device__ inline int32_t child_ref(const int4 &c, int i) {
    return i == 0 ? c.x : (i == 1 ? c.y : (i == 2 ? c.z : c.w));
}

// Compare-exchange of (key, child) pairs as selects (no branch): the smaller key first.
__device__ inline void cswap(float &ka, int32_t &va, float &kb, int32_t &vb) {
    const bool s = kb < ka;
    const float k0 = s ? kb : ka, k1 = s ? ka : kb;
    const int32_t v0 = s ? vb : va, v1 = s ? va : vb;
    ka = k0;
    kb = k1;
    va = v0;
    vb = v1;
}

// Closest-hit order from a visited inner node with hit mask m != 0: the hit children sorted by
// entry distance (a five-comparator network over four keys, misses keyed +inf; entry distances
// of entered children are finite), the nearest returned as the next node, the others pushed
// farthest first so that they pop nearest first. The same instructions run for any number of
// entered children, so the lanes of a wave do not diverge here.
__device__ inline int32_t order_children(const QuadHit &q, Stack &st) {
    float k0 = q.k[0], k1 = q.k[1], k2 = q.k[2], k3 = q.k[3];
    int32_t v0 = q.c[0], v1 = q.c[1], v2 = q.c[2], v3 = q.c[3];
    cswap(k0, v0, k1, v1);
    cswap(k2, v2, k3, v3);
    cswap(k0, v0, k2, v2);
    cswap(k1, v1, k3, v3);
    cswap(k1, v1, k2, v2);
    st.push_sorted(__popc(q.m), v1, v2, v3);
    return v0;
}

// Any-hit order from a visited inner node with hit mask m != 0: the lowest hit child next,
// the others pushed.
__device__ inline int32_t any_children(const Node4 &q, uint32_t m, Stack &st) {
    const int first = __builtin_ctz(m);
    m &= m - 1;
    while (m) {
        st.push(child_ref(q.ch, __builtin_ctz(m)));
        m &= m - 1;
    }
    return child_ref(q.ch, first);
}

template <bool kCount>
__device__ inline int32_t closest_hit(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, f3 o,
                                      f3 d, Stack &st, float &tout, int32_t &slot, uint32_t &cn, uint32_t &ct) {
    SlabRay r{o, {safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)}, kEps};
    float best = __builtin_inff();
    int32_t best_id = -1, best_slot = -1;
    st.sp = 0;
    int32_t node = 0;
    while (true) {
        if (node >= 0) {
            if (kCount) ++cn;
            const QuadHit qh = slab4q(nodes, node, r, best);
            if (qh.m) {
                node = order_children(qh, st);
                continue;
            }
        } else {
            int first, count;
            leaf_decode(node, first, count);
            for (int i = first; i < first + count; ++i) {
                if (kCount) ++ct;
                const float4 p = tris[3 * i + 0];
                const float4 e1 = tris[3 * i + 1];
                const float4 e2 = tris[3 * i + 2];
                const float t = tri_intersect({p.x, p.y, p.z}, {e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, o, d, kEps,
                                              __builtin_inff());
                const int32_t id = __float_as_int(p.w);
                if (t < __builtin_inff() && (best_id < 0 || t < best || (t == best && id < best_id))) {
                    best = t;
                    best_id = id;
                    best_slot = i;
                }
            }
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
    tout = best;
    slot = best_slot;
    return best_id;
}

template <bool kCount>
__device__ inline bool occluded(const float4 *__restrict__ nodesF, const float4 *__restrict__ tris, f3 o, f3 d,
                                float tfar, Stack &st, uint32_t &cn, uint32_t &ct) {
    SlabRay r{o, {1.0f / d.x, 1.0f / d.y, 1.0f / d.z}, kEps};
    st.sp = 0;
    int32_t node = 0;
    while (true) {
        if (node >= 0) {
            if (kCount) ++cn;
            const Node4 q = load_node4f(nodesF, node);
            float te[4];
            const uint32_t m = slab4(q, r, tfar, te);
            if (m) {
                node = any_children(q, m, st);
                continue;
            }
        } else {
            int first, count;
            leaf_decode(node, first, count);
            for (int i = first; i < first + count; ++i) {
                if (kCount) ++ct;
                const float4 p = tris[3 * i + 0];
                const float4 e1 = tris[3 * i + 1];
                const float4 e2 = tris[3 * i + 2];
                if (tri_intersect({p.x, p.y, p.z}, {e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, o, d, kEps, tfar) <
                    __builtin_inff())
                    return true;
            }
        }
        if (st.sp == 0) return false;
        node = st.pop();
    }
}

// ------------------------------------------------------------------------ guiding tree
// Node 0 covers [0,1)^2; an inner node halves its cell along axis depth%2; children are
// stored consecutively (child, child+1), child 0 the lower half.
__device__ inline float tree_pdf(const int32_t *__restrict__ child, const float *__restrict__ p0, float x0,
                                 float x1) {
    float pdf = 1.0f, lo0 = 0.0f, lo1 = 0.0f, sz0 = 1.0f, sz1 = 1.0f;
    int32_t node = 0;
    int dep = 0;
    int32_t c;
    while ((c = child[node]) != 0) {
        const float p = p0[node];
        if ((dep & 1) == 0) {
            const float mid = lo0 + sz0 * 0.5f;
            if (x0 < mid) { pdf = pdf * (2.0f * p); node = c; }
            else { pdf = pdf * (2.0f * (1.0f - p)); lo0 = mid; node = c + 1; }
            sz0 = sz0 * 0.5f;
        } else {
            const float mid = lo1 + sz1 * 0.5f;
            if (x1 < mid) { pdf = pdf * (2.0f * p); node = c; }
            else { pdf = pdf * (2.0f * (1.0f - p)); lo1 = mid; node = c + 1; }
            sz1 = sz1 * 0.5f;
        }
        ++dep;
    }
    return pdf;
}

__device__ inline float tree_sample(const int32_t *__restrict__ child, const float *__restrict__ p0, float &x0,
                                    float &x1) {
    float pdf = 1.0f, lo0 = 0.0f, lo1 = 0.0f, sz0 = 1.0f, sz1 = 1.0f;
    int32_t node = 0;
    int dep = 0;
    int32_t c;
    while ((c = child[node]) != 0) {
        const float p = p0[node];
        if ((dep & 1) == 0) {
            if (x0 < p) { x0 = fminf(x0 / p, kOneMinusEps); pdf = pdf * (2.0f * p); node = c; }
            else {
                x0 = fminf((x0 - p) / (1.0f - p), kOneMinusEps);
                pdf = pdf * (2.0f * (1.0f - p));
                lo0 = lo0 + sz0 * 0.5f;
                node = c + 1;
            }
            sz0 = sz0 * 0.5f;
        } else {
            if (x1 < p) { x1 = fminf(x1 / p, kOneMinusEps); pdf = pdf * (2.0f * p); node = c; }
            else {
                x1 = fminf((x1 - p) / (1.0f - p), kOneMinusEps);
                pdf = pdf * (2.0f * (1.0f - p));
                lo1 = lo1 + sz1 * 0.5f;
                node = c + 1;
            }
            sz1 = sz1 * 0.5f;
        }
        ++dep;
    }
    x0 = lo0 + x0 * sz0;
    x1 = lo1 + x1 * sz1;
    return pdf;
}

__device__ inline int32_t tree_leaf(const int32_t *__restrict__ child, float x0, float x1) {
    float lo0 = 0.0f, lo1 = 0.0f, sz0 = 1.0f, sz1 = 1.0f;
    int32_t node = 0;
    int dep = 0;
    int32_t c;
    while ((c = child[node]) != 0) {
        if ((dep & 1) == 0) {
            const float mid = lo0 + sz0 * 0.5f;
            if (x0 < mid) node = c;
            else { lo0 = mid; node = c + 1; }
            sz0 = sz0 * 0.5f;
        } else {
            const float mid = lo1 + sz1 * 0.5f;
            if (x1 < mid) node = c;
            else { lo1 = mid; node = c + 1; }
            sz1 = sz1 * 0.5f;
        }
        ++dep;
    }
    return node;
}

// GuidingWrapper::sample (path.h:53-82), over the sampling tree (child, p0), which may be a
// copy in LDS
__device__ inline float guide_sample(const TraceArgs &a, const int32_t *child, const float *p0, float &x0,
                                     float &x1) {
    const float u = a.uniformProb;
    if (u == 1.0f) return 1.0f;
    float pdf = 1.0f - u;
    if (x0 < u) {
        x0 /= u;
        pdf *= tree_pdf(child, p0, x0, x1);
    } else {
        x0 -= u;
        x0 /= 1.0f - u;
        const float gpdf = tree_sample(child, p0, x0, x1);
        pdf *= gpdf;
    }
    pdf += u;
    return pdf;
}

// ------------------------------------------------------------------------ path state
// The reference's FP32 geometry chain (the arithmetic of path.h:236-442 and cpu.cpp:9-18,
// 73-86 in float), carried beside the FP64 chain when a run traces with it (kRef): it gives
// the traversal and shadow rays and every discrete decision (hit ids, visibility, the grazing
// and range cut-offs, the back-face cull, the filter thresholds), while path lengths, time,
// phase and the GO weight come from the FP64 chain refined on the same triangles.
struct Chain32 {
    f3 o, d;        // vertex and direction
    f3 n_in, d_in;  // face-forwarded normal and incoming direction at the last hit
    float r;        // path length
    float cosTheta;
};

struct PathState {
    d3 o, d;        // FP64 chain ray
    d3 n_in, d_in;  // normal and direction at the last hit (surface.incoming)
    cf3 H, H_in;    // field of the ray and at the last hit
    double time;    // path time [s]
    double r;       // path length [m]
    double cosTheta;
    cf gw;          // guidingWeight
    float x0, x1;   // primary sample after the guiding warp
    float pdf;      // primaryPdf
    float freq;
    int depth;
    int64_t local;  // index - offset
    double rate;    // motion: d/dt of the path length up to the current vertex [m/s]
    f3 vprev;       // motion: velocity of the current vertex (the Tx at depth 0)
    Chain32 c32;    // kRef: the reference's FP32 chain
};

__device__ inline f3 posf(const DevAntenna &a) { return {a.pos[0], a.pos[1], a.pos[2]}; }

// UniformSampleSphere (sampling.h:38-44) in FP32 as the reference evaluates it, sin and cos
// from the shared sincos_d rounded to float
__device__ inline f3 uniform_sample_sphere_f(float u0, float u1) {
    const float phi = (2.0f * kPiF) * u0;
    const float y = 2.0f * u1 - 1.0f;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - y * y));
    double sd, cd;
    sincos_d((double)phi, &sd, &cd);
    const float s = (float)sd, c = (float)cd;
    return {-r * s, y, -r * c};
}

// The primary sample of a Halton index before the guiding warp: dimensions 1 and 2 in the order
// get2D evaluates them (halton.h:24-27).
__device__ inline void primary_sample(const TraceArgs &a, uint64_t index, float &x0, float &x1) {
    const float u_a = radical_inverse_b<3>(index);
    const float u_b = radical_inverse_b<5>(index);
    x0 = a.primaryOrder == 0 ? u_b : u_a;
    x1 = a.primaryOrder == 0 ? u_a : u_b;
}

template <bool kRef = false>
__device__ inline void generate(const TraceArgs &a, uint64_t index, PathState &p, const int32_t *gsChild,
                                const float *gsP0) {
    const float u_f = radical_inverse_2(index);
    float x0, x1;
    primary_sample(a, index, x0, x1);
    float pdf = 1.0f;
    if (a.doGuiding && a.guideOn) pdf = guide_sample(a, gsChild, gsP0, x0, x1);
    const d3 dl = uniform_sample_sphere((double)x0, (double)x1);
    cf3 Hl = awr_evaluate(cast3<float>(dl));
    Hl = Hl / kInv4PiF;
    p.o = posd(a.tx);
    p.d = mulv(rotd(a.tx), dl);
    p.H = (1.0f / pdf) * mulv(rotf(a.tx), Hl);
    p.freq = a.startFreq + u_f * (a.freqSlope * a.rampTime);
    p.time = 0.0;
    p.r = 0.0;
    p.cosTheta = 1.0;
    p.gw = {0.0f, 0.0f};
    p.x0 = x0;
    p.x1 = x1;
    p.pdf = pdf;
    p.depth = 0;
    p.rate = 0.0;
    p.vprev = {a.ego[0], a.ego[1], a.ego[2]};
    if (kRef) {
        p.c32.o = posf(a.tx);
        p.c32.d = mulv(rotf(a.tx), uniform_sample_sphere_f(x0, x1));
        p.c32.r = 0.0f;
        p.c32.cosTheta = 1.0f;
    }
}

// NEE set-up for receiver ch at the path's current vertex: everything of path.h:296-396
// except the visibility query (rx.nee, the PO phasor of SurfaceEmitter::evaluate/connect or
// the direct Tx evaluate/connect at depth 0, measureH, the GO density correction and the
// filtering). Pure: the caller applies the guiding-weight updates only for visible rays.
struct NeeSetup {
    f3 o, d;       // shadow ray: origin and direction rounded from the FP64 chain
    float tfar;    // distance to the receiver
    double tn;     // path time at the receiver
    cf v;          // measurement
    cf gw1;        // plane-filter guiding term (filterGuiding, path.h:376-380), else 0
    bool gwAdd;    // guidingWeight += v when visible (depth > 0, path.h:392-394)
    bool splat;    // the splat skip rule lets v through (integrator.h:121-124)
    double dv;     // motion: radial velocity -1/2 d/dt(path length) [m/s], else 0
};

template <bool kRef = false>
__device__ inline NeeSetup nee_setup(const TraceArgs &a, const PathState &p, int ch) {
    const float fourPi = 4.0f * kPiF;
    NeeSetup ns;
    const DevAntenna &rxa = a.rx[ch];
    d3 dn = posd(rxa) - p.o;
    const double rr = norm(dn);
    dn = dn / rr;
    f3 dn32{};
    float rr32 = 0.0f;
    if (kRef) {  // the shadow ray of the FP32 chain
        dn32 = posf(rxa) - p.c32.o;
        rr32 = norm(dn32);
        dn32 = dn32 / rr32;
        ns.o = p.c32.o;
        ns.d = dn32;
        ns.tfar = rr32;
    } else {
        ns.o = cast3<float>(p.o);
        ns.d = cast3<float>(dn);
        ns.tfar = (float)rr;
    }
    ns.dv = 0.0;
    if (a.motion) {
        // the leg to the receiver: d|rx - p|/dt = dn . (v_rx - v_p)
        const d3 rel = {(double)a.ego[0] - (double)p.vprev.x, (double)a.ego[1] - (double)p.vprev.y,
                        (double)a.ego[2] - (double)p.vprev.z};
        const double rn = p.rate + dot(dn, rel);
        ns.dv = -(rn / 2.0);
    }
    ns.gw1 = cf{0.0f, 0.0f};
    ns.gwAdd = false;
    const cf3 Hrx = mulv(rotf(rxa), awr_evaluate(cast3<float>(mulTv(rotd(rxa), -dn))));
    const float rrF = (float)rr;
    cf3 Hn;
    if (p.depth == 0) {
        // tx.evaluate + tx.connect (emitter.h:182-193)
        Hn = mulv(rotf(a.tx), awr_evaluate(cast3<float>(mulTv(rotd(a.tx), dn))));
        Hn = (1.0f / (fourPi * rrF)) * Hn;
    } else {
        // SurfaceEmitter::evaluate (emitter.h:64-70) and ::connect (emitter.h:72-81)
        const cf3 J = 2.0f * cross(cast3<float>(p.n_in), p.H_in);
        Hn = cross(cast3<float>(dn), J);
        if (kRef ? dot(dn32, p.c32.n_in) < 0.0f : dot(dn, p.n_in) < 0.0) Hn = cf3{};
        const float k0 = ((2.0f * kPiF) * p.freq) / 299792448.0f;
        const cf green = cf{1.0f / fmaxf(rrF, 1e-3f), k0} / (fourPi * rrF);
        Hn = Hn * green;
    }
    ns.tn = p.time + rr / kC;
    cf v = cdot(Hn, Hrx);  // measureH (geometry.h:138)
    const bool hard = a.correctPhase && a.filteringSphere;  // the sphere filter cuts (path.h:360-368)
    if (p.depth > 0 && a.correctPhase) {
        // phase correction (path.h:316-329): the reflected ray itself measured at the receiver,
        // its time the distance from the transmitter's mirror image (the vertex moved back by
        // the path length along the reflected direction) to the receiver, its value
        // measureH(H, Hrx)·π·dist / (π·R·λ)² / 4π
        const d3 rxPos = p.o + rr * dn;
        const double dist = norm((p.o - p.r * p.d) - rxPos);
        ns.tn = dist / kC;  // setDistance (geometry.h:113)
        const double den = kPiD * (double)a.filteringRadius * (kC / (double)p.freq);
        v = cdot(p.H, Hrx);
        v = v * (float)(kPiD * dist);
        v = v / (float)(den * den);
        v = v / fourPi;
    } else if (p.depth > 0 && a.geometricalOptics) {
        // GO density correction (path.h:330-341)
        v = v * (float)(p.r / p.cosTheta);
        v = v / fourPi;
        if (kRef ? p.c32.cosTheta < 1e-3f : p.cosTheta < 1e-3) {
            ns.v = v;
            ns.splat = false;
            return ns;
        }
    }
    if (p.depth > 0) {
        if (kRef) {
            // phase error and filter of the FP32 chain (path.h:343-387 in float)
            float dp;
            if (a.filteringSphere) {
                const f3 rxPos = p.c32.o + rr32 * dn32;
                const float lam = fmaxf(0.0f, dot(p.c32.d, rxPos - p.c32.o));
                dp = norm((p.c32.o + lam * p.c32.d) - rxPos);
            } else {
                const float c = dot(normalized(dn32 - p.c32.d_in), p.c32.n_in);
                dp = p.c32.r * sqrtf(1.0f - c * c) / c;
            }
            dp = dp / (299792448.0f / p.freq);
            if (a.doFiltering) {
                if (a.filteringSphere) {
                    const float R = a.filteringRadius;
                    if (dp > R && hard) {
                        if (a.filterGuiding) {  // std::pow(float, int) is evaluated in double
                            const double q = (double)(dp / R);
                            ns.gw1 = v / (float)(q * q + 1.0);
                        }
                        v = cf{0.0f, 0.0f};
                    } else if (dp > R) {
                        const float f = (float)(1.0 - 0.20 * (double)(dp / R - 1.0f));
                        v = v * fmaxf(f, 0.0f);
                    }
                } else {
                    const float fmin_ = a.filteringMin, fmax_ = a.filteringMax;
                    if (dp > fmax_) {
                        if (a.filterGuiding) {
                            const double q = (double)(dp / fmax_);
                            ns.gw1 = v / (float)(q * q + 1.0);
                        }
                        v = cf{0.0f, 0.0f};
                    }
                    if (dp > fmin_) v = v * ((fmax_ - dp) / (fmax_ - fmin_));
                }
            }
            ns.gwAdd = true;
            ns.v = v;
            ns.splat = a.produceDebugImage || !(a.sampleWeight == 0.0f || v.re == 0.0f || v.im == 0.0f);
            return ns;
        }
        double dphase;
        if (a.filteringSphere) {  // path.h:343-349
            const d3 rxPos = p.o + rr * dn;
            const double lam = fmax(0.0, dot(p.d, rxPos - p.o));
            dphase = norm((p.o + lam * p.d) - rxPos);
        } else {  // path.h:350-353
            const double c = dot(normalized(dn - p.d_in), p.n_in);
            dphase = p.r * sqrt(1.0 - c * c) / c;
        }
        dphase = dphase / (kC / (double)p.freq);  // path.h:355
        if (a.doFiltering) {                      // path.h:357-387
            if (a.filteringSphere) {
                const double R = (double)a.filteringRadius;
                if (dphase > R && hard) {
                    if (a.filterGuiding) {
                        const double q = dphase / R;
                        ns.gw1 = v / (float)(q * q + 1.0);
                    }
                    v = cf{0.0f, 0.0f};
                } else if (dphase > R) {
                    const double f = 1.0 - 0.20 * (dphase / R - 1.0);
                    v = v * (float)fmax(f, 0.0);
                }
            } else {
                const double fmin_ = a.filteringMin, fmax_ = a.filteringMax;
                if (dphase > fmax_) {
                    if (a.filterGuiding) {
                        const double q = dphase / fmax_;
                        ns.gw1 = v / (float)(q * q + 1.0);
                    }
                    v = cf{0.0f, 0.0f};
                }
                if (dphase > fmin_) v = v * (float)((fmax_ - dphase) / (fmax_ - fmin_));
            }
        }
        ns.gwAdd = true;
    }
    ns.v = v;
    // splat skip rule (integrator.h:121-124); with the debug image on, every contribution is
    // splatted, a zero one too (it adds nothing to the frame)
    ns.splat = a.produceDebugImage || !(a.sampleWeight == 0.0f || v.re == 0.0f || v.im == 0.0f);
    return ns;
}

// Hit update and specular reflection (path.h:413-442, integrator.h:164-168,
// intersection.h:38): geometric normal (kernel.cu:86) face-forwarded (cpu.cpp:82-85), and
// the segment length re-evaluated in double against the hit triangle's plane, both from the
// float edges promoted to double. Returns false when the hit is grazing (path ends), or, with
// geometrical optics off, when the weighted field is zero (path.h:428-429).
// Geometrical optics off (path.h:277-279, 420-427): the field is weighted by the emitter's
// connect (1/(4πt) from the transmitter, emitter.h:188-193; (jk0 + 1/max(t, 1e-3))/(4πt) from
// a surface, emitter.h:72-81) and by the hemisphere-to-surface density t²/cosθ, and the next
// direction is SurfaceEmitter::sample (emitter.h:58-70): a uniform sphere sample of Halton
// dimensions 3 + 2(k-1) and 4 + 2(k-1) at the new depth k, in get2D's order, with H = d x J,
// J = 2 n x H_in, zero when d points below the surface.
// kRef: the hit of the FP32 chain's ray at distance t32 (the traversal's), whose normal,
// face-forwarding, grazing test and update are the reference's float arithmetic; the FP64
// chain is refined on the same triangle.
template <bool kRef = false>
__device__ inline bool apply_hit(const TraceArgs &a, PathState &p, int32_t slot, float t32 = 0.0f) {
    const float4 P = a.tris[3 * slot + 0];
    const float4 E1 = a.tris[3 * slot + 1];
    const float4 E2 = a.tris[3 * slot + 2];
    const d3 e1 = {(double)E1.x, (double)E1.y, (double)E1.z};
    const d3 e2 = {(double)E2.x, (double)E2.y, (double)E2.z};
    const d3 v0 = {(double)P.x, (double)P.y, (double)P.z};
    const d3 nc = cross(e1, e2);
    d3 n = normalized(nc);
    if (dot(n, p.d) > 0.0) n = -n;
    const double cosT = fabs(dot(n, -p.d));
    f3 n32{};
    if (kRef) {
        n32 = normalized(cross(f3{E1.x, E1.y, E1.z}, f3{E2.x, E2.y, E2.z}));
        if (dot(n32, p.c32.d) > 0.0f) n32 = -n32;
        const float cos32 = fabsf(dot(n32, -p.c32.d));
        if (cos32 < 1e-3f) return false;
        p.c32.cosTheta = cos32;
    } else if (cosT < 1e-3) {
        return false;
    }
    if (a.motion) {
        // the incoming leg: d|p - o|/dt = d . (v_p - v_o), the vertex moving with its triangle
        const float4 vh = a.vel ? a.vel[slot] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const d3 rel = {(double)vh.x - (double)p.vprev.x, (double)vh.y - (double)p.vprev.y,
                        (double)vh.z - (double)p.vprev.z};
        p.rate = p.rate + dot(p.d, rel);
        p.vprev = {vh.x, vh.y, vh.z};
    }
    const double tg = dot(nc, v0 - p.o) / dot(nc, p.d);
    p.cosTheta = cosT;
    p.r = p.r + tg;
    p.time = p.time + tg / kC;
    if (!a.geometricalOptics) {
        const float tF = (float)tg;
        const float fourPi = 4.0f * kPiF;
        if (p.depth == 0) {
            p.H = (1.0f / (fourPi * tF)) * p.H;
        } else {
            const float k0 = ((2.0f * kPiF) * p.freq) / 299792448.0f;
            const cf green = cf{1.0f / fmaxf(tF, 1e-3f), k0} / (fourPi * tF);
            p.H = p.H * green;
        }
        p.H = (float)(tg * tg / cosT) * p.H;
        if (is_zero(p.H)) return false;
    }
    p.n_in = n;
    p.H_in = p.H;
    p.d_in = p.d;
    p.o = p.o + tg * p.d;
    p.depth++;
    if (!a.geometricalOptics) {
        const uint64_t index = (uint64_t)(a.offset + p.local);
        const int dimA = 3 + 2 * (p.depth - 1);
        const float ua = halton_dim(index, dimA), ub = halton_dim(index, dimA + 1);
        const float s0 = a.primaryOrder == 0 ? ub : ua, s1 = a.primaryOrder == 0 ? ua : ub;
        p.d = uniform_sample_sphere((double)s0, (double)s1);
        const cf3 J = 2.0f * cross(cast3<float>(n), p.H_in);
        p.H = cross(cast3<float>(p.d), J);
        bool below;
        if (kRef) {
            p.c32.r = p.c32.r + t32;
            p.c32.n_in = n32;
            p.c32.d_in = p.c32.d;
            p.c32.o = p.c32.o + t32 * p.c32.d;
            p.c32.d = uniform_sample_sphere_f(s0, s1);
            below = dot(p.c32.d, n32) < 0.0f;
        } else {
            below = dot(p.d, n) < 0.0;
        }
        if (below) p.H = cf3{};
        return true;
    }
    const d3 wi = -p.d;
    p.d = (2.0 * dot(n, wi)) * n - wi;
    const f3 nF = cast3<float>(n);
    p.H = rscale(2.0f * nF, rdot(nF, p.H)) - p.H;
    if (kRef) {
        p.c32.r = p.c32.r + t32;
        p.c32.n_in = n32;
        p.c32.d_in = p.c32.d;
        p.c32.o = p.c32.o + t32 * p.c32.d;
        const f3 wi32 = -p.c32.d;
        p.c32.d = (2.0f * dot(n32, wi32)) * n32 - wi32;
    }
    return true;
}

// Termination test before the extension ray (path.h:400-404); kRef: the range cut-off of the
// FP32 chain.
template <bool kRef = false>
__device__ inline bool path_ends(const TraceArgs &a, const PathState &p) {
    if (kRef) {
        const float maxDist32 = (a.adcRate / a.freqSlope) * 299792448.0f;
        return p.depth >= a.maxDepth || p.c32.r >= maxDist32 || is_zero(p.H);
    }
    const double maxDist = ((double)a.adcRate / (double)a.freqSlope) * kC;
    return p.depth >= a.maxDepth || p.r >= maxDist || is_zero(p.H);
}

// Interruptible traversal: a lane advances its ray by up to `steps` node visits per call, so
// that a persistent kernel can hand a finished lane the next ray while the others go on.
constexpr int32_t kDone = -1;  // never a valid leaf reference (a leaf holds at most 8 triangles)

struct Trav {
    f3 o, d, invd;
    float tfar;      // any-hit: the ray's extent; closest: the best hit so far (+inf)
    int32_t node;    // next node reference, kDone when finished
    int32_t best_id, best_slot;
};

__device__ inline void trav_begin(Trav &t, f3 o, f3 d, float tfar, Stack &st) {
    t.o = o;
    t.d = d;
    t.invd = {safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)};
    t.tfar = tfar;
    t.node = 0;
    t.best_id = -1;
    t.best_slot = -1;
    st.sp = 0;
}

template <bool kAny, bool kCount>
__device__ inline void trav_run(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, Trav &t,
                                Stack &st, int steps, uint32_t &cn, uint32_t &ct) {
    const SlabRay r{t.o, t.invd, kEps};
    for (int s = 0; s < steps && t.node != kDone; ++s) {
        if (t.node >= 0) {
            if (kCount) ++cn;
            uint32_t m;
            if (kAny) {
                const Node4 q = load_node4f(nodes, t.node);
                float te[4];
                m = slab4(q, r, t.tfar, te);
                if (m) t.node = any_children(q, m, st);
            } else {
                const QuadHit qh = slab4q(nodes, t.node, r, t.tfar);
                m = qh.m;
                if (m) t.node = order_children(qh, st);
            }
            if (!m) t.node = st.sp ? st.pop() : kDone;
        } else {
            int first, count;
            leaf_decode(t.node, first, count);
            bool hit = false;
            for (int i = first; i < first + count; ++i) {
                if (kCount) ++ct;
                const float4 p = tris[3 * i + 0];
                const float4 e1 = tris[3 * i + 1];
                const float4 e2 = tris[3 * i + 2];
                const float th = tri_intersect({p.x, p.y, p.z}, {e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, t.o, t.d,
                                               kEps, kAny ? t.tfar : __builtin_inff());
                if (kAny) {
                    if (th < __builtin_inff()) { hit = true; break; }
                } else {
                    const int32_t id = __float_as_int(p.w);
                    if (th < __builtin_inff() && (t.best_id < 0 || th < t.tfar || (th == t.tfar && id < t.best_id))) {
                        t.tfar = th;
                        t.best_id = id;
                        t.best_slot = i;
                    }
                }
            }
            if (kAny && hit) {
                t.best_id = 1;
                t.node = kDone;
            } else {
                t.node = st.sp ? st.pop() : kDone;
            }
        }
    }
}

// One NEE event ready for the frame: taps A·1/(s-k) at sample bins center+k (or one tap at
// center when delta), and with motion and M > 1 chirps the same along the chirp axis
// (A·1/(sc-l)·1/(s-k) at chirp cc+l), the separable kernel of Frame::splat (radar.h:564-627).
struct Event {
    int center;
    float s;
    cf A;
    int delta;
    int cc;
    float sc;
    int deltaC;
};

// Integrator::splat (integrator.h:113-135) up to the taps: fractional bin (setTime,
// radar.h:232-241), fractional chirp from the radial velocity (setVelocity, radar.h:248-251),
// IF phase (measureRay, integrator.h:107-110) and the splat's own e^{jπs}·sin(πs)/π factor
// per spread axis (radar.h:570-583). With one chirp the chirp taps all land in chirp 0 and
// are folded into A.
__device__ inline Event make_event(const TraceArgs &a, double tn, cf v, float weight, double dv) {
    Event e;
    e.cc = 0;
    e.sc = 0.0f;
    e.deltaC = 1;
    double wc = 1.0, pc_s = 0.0, pc_c = 1.0;
    if (a.motion) {
        const double dp = ((2.0 * (double)a.startFreq) * dv) / kC;
        const double q = dp / a.chirpFreq;
        const double xc = (double)a.M * (q - floor(q));
        const int cc = (int)round(xc);
        const double sc = xc - (double)cc;
        e.cc = cc;
        if (fabs(sc) >= 1e-4) {
            double ss;
            sincos_d(kPiD * sc, &ss, &pc_c);
            pc_s = ss;
            wc = ss / kPiD;
            e.sc = (float)sc;
            e.deltaC = 0;
            if (a.M == 1) {
                double f = 0.0;
                for (int l = -16; l <= 16; ++l) f += 1.0 / (sc - (double)l);
                wc = wc * f;
                e.deltaC = 1;
            }
        }
    }
    const double dt = tn + (double)a.antennaDelay;
    double x = 0.0;
    if (a.freqSlope != 0.0f) {
        const double q = (dt * (double)a.freqSlope) / (double)a.adcRate;
        x = (double)a.N * (q - floor(q));
    }
    cf ph;
    if (a.phaseMode == 1) {
        // the reference's FP32 evaluation of 2π(f0 - Δt S/2)Δt
        const float dtf = (float)dt;
        const float arg = ((2.0f * kPiF) * (a.startFreq - dtf * a.freqSlope / 2.0f)) * dtf;
        ph = {cosf(arg), sinf(arg)};
    } else if (a.phaseMode == 2) {
        // double-float: turns = (f0 - Δt S/2) Δt in float pairs, the fraction of a turn, then
        // an FP32 sin/cos of the reduced angle
        const float th = (float)dt;
        const df t{th, (float)(dt - (double)th)};
        const df half = df_mul(t, df{a.freqSlope * 0.5f, 0.0f});
        const df turns = df_mul(df_add(df{a.startFreq, 0.0f}, df{-half.hi, -half.lo}), t);
        const float whole = floorf(turns.hi);
        const df fr = df_add(df{turns.hi - whole, 0.0f}, df{turns.lo, 0.0f});
        float f = fr.hi + fr.lo;
        f = f - floorf(f);
        float sf, cf_;
        sincosf((2.0f * kPiF) * f, &sf, &cf_);
        ph = {cf_, sf};
    } else {
        const double turns = ((double)a.startFreq - dt * (double)a.freqSlope / 2.0) * dt;
        double s, c;
        sincos_d((2.0 * kPiD) * (turns - floor(turns)), &s, &c);
        ph = {(float)c, (float)s};
    }
    cf value = weight * (v * ph);
    if (pc_s != 0.0 || wc != 1.0) value = (value * cf{(float)pc_c, (float)pc_s}) * (float)wc;
    const int center = (int)round(x);
    const double s = x - (double)center;
    e.center = center;
    if (fabs(s) < 1e-4) {
        e.delta = 1;
        e.s = 0.0f;
        e.A = value;
        return e;
    }
    double ss, cs;
    sincos_d(kPiD * s, &ss, &cs);
    value = value * cf{(float)cs, (float)ss};
    const float w0 = (float)(ss / kPiD);
    e.delta = 0;
    e.s = (float)s;
    e.A = value * w0;
    return e;
}

__device__ inline int wrap_bin(int v, int N) {
    if (N >= 33) {
        if (v < 0) v += N;
        else if (v >= N) v -= N;
        return v;
    }
    v %= N;
    return v < 0 ? v + N : v;
}

// The workgroup's LDS frame tile of one chirp holds channel-major rows of bins, real parts
// then imaginary parts: re[c*N + b] at tile[c*N + b], im at tile[N*C + c*N + b]. The 33 taps of
// one event then fall on consecutive doubles of one row, so the lanes of one LDS atomic touch
// distinct, consecutive banks.
__device__ inline int tile_re(int b, int c, int N) { return c * N + b; }

// Splat of the wave's events (one per lane at most, all of frame channel ch) into the LDS tile:
// each lane adds its own event's 33 taps, starting at tap (lane mod 33) so that lanes whose
// events share a centre bin (neighbouring paths often do) touch different bins at each step.
// The tap weight 1/(s-k) uses the hardware reciprocal (1 ulp).
__device__ inline void flush_events(bool has, const Event &e, int ch, double *tile, int N, int C, int lane) {
    double *re = tile + ch * N, *im = tile + N * C + ch * N;
    if (!has) return;
    if (e.delta) {
        const int b = wrap_bin(e.center, N);
        atomicAdd(&re[b], (double)e.A.re);
        atomicAdd(&im[b], (double)e.A.im);
        return;
    }
    int k = lane % 33 - 16;
    for (int it = 0; it < 33; ++it) {
        const int b = wrap_bin(e.center + k, N);
        const float coef = __builtin_amdgcn_rcpf(e.s - (float)k);
        atomicAdd(&re[b], (double)(e.A.re * coef));
        atomicAdd(&im[b], (double)(e.A.im * coef));
        k = k == 16 ? -16 : k + 1;
    }
}

// Adds the LDS tile into the workgroup's partial frame [N][C] (complex, frame order).
__device__ inline void tile_flush(const double *tile, double2 *dst, int N, int C, int tid, int threads) {
    const int NC = N * C;
    for (int j = tid; j < NC; j += threads) {
        const int t = tile_re(j / C, j % C, N);
        double2 v = dst[j];
        v.x += tile[t];
        v.y += tile[NC + t];
        dst[j] = v;
    }
}

// The guard-band tile of the wavefront gather (N >= 16): each channel row holds N + 32 bins,
// bin b in [-16, N + 16) at row offset 16 + b, real parts then imaginary parts (rows of
// channel c at c * (N + 32) and (N + 32) * C + c * (N + 32)). An event's 33 taps around a
// centre in [0, N) then never wrap inside the splat; tile_flush_g adds the 16 guard bins on
// each side into bins b mod N.
constexpr int kGuard = 16;
// 0: the compiler unrolls the 33 taps and keeps the lane's tap offsets and their float values in
// registers across events (fewer instructions per tap, more registers)
#ifndef HSR_TAP_UNROLL
#define HSR_TAP_UNROLL 0
#endif
__device__ inline int guard_row(int N) { return N + 2 * kGuard; }

// flush_events into the guard-band tile: the same taps and weights, with the wrap of each tap
// replaced by the guard bins, in a bank-aligned order. The 32 taps k in [-16, 15] of an event
// cover 32 consecutive bins, one of each residue mod 32; lane L takes them in the order whose
// step `it` lands on the bin of residue L + it (mod 32). Every LDS atomic of the wave then puts
// the same number of lanes on each bank (two per 4-byte bank word for a wave of 64 doubles,
// the least it can), wherever the lanes' events are centred; a rotation by lane alone aligns
// only lanes whose events share a centre bin. Tap 16 follows on its own.
__device__ inline void flush_events_g(bool has, const Event &e, int ch, double *tile, int N, int C, int lane) {
    if (!has) return;
    const int R = guard_row(N);
    int c0 = e.center;  // in [0, N]: round(N * frac) reaches N
    if (c0 >= N) c0 -= N;
    // the event's centre bin in the LDS tile (tap k at re[k], k in [-16, 16] within the row)
    lds_double *re = (lds_double *)tile + (ch * R + kGuard + c0);
    const int imOff = R * C;
    if (e.delta) {
        lds_add(re, (double)e.A.re);
        lds_add(re + imOff, (double)e.A.im);
        return;
    }
    // tap of step it: k = the 5-bit two's-complement value of (lane - c0 + it), so that
    // c0 + k = lane + it (mod 32) and k runs from its first value up to 15, then from -16
    const int x0 = lane - c0;
#if HSR_TAP_UNROLL > 0
#pragma unroll HSR_TAP_UNROLL
#endif
    for (int it = 0; it < 32; ++it) {
        const int k = __builtin_amdgcn_sbfe(x0 + it, 0, 5);
        const float coef = __builtin_amdgcn_rcpf(e.s - (float)k);
        lds_add(re + k, (double)(e.A.re * coef));
        lds_add(re + k + imOff, (double)(e.A.im * coef));
    }
    const float coef = __builtin_amdgcn_rcpf(e.s - 16.0f);
    lds_add(re + 16, (double)(e.A.re * coef));
    lds_add(re + 16 + imOff, (double)(e.A.im * coef));
}

// Adds the guard-band tile into the workgroup's partial frame [N][C] (complex, frame order):
// bin b gets its own entry, the right guard entry of bin b + N (b < 16) and the left guard
// entry of bin b - N (b >= N - 16).
__device__ inline void tile_flush_g(const double *tile, double2 *dst, int N, int C, int tid, int threads) {
    const int NC = N * C, R = guard_row(N), RC = R * C;
    for (int j = tid; j < NC; j += threads) {
        const int b = j / C, c = j % C;
        const int t = c * R + kGuard + b;
        double re = tile[t], im = tile[RC + t];
        if (b < kGuard) {
            re += tile[t + N];
            im += tile[RC + t + N];
        }
        if (b >= N - kGuard) {
            re += tile[t - N];
            im += tile[RC + t - N];
        }
        double2 v = dst[j];
        v.x += re;
        v.y += im;
        dst[j] = v;
    }
}

// Wave-cooperative splat of this lane's event (if any) into the frame cube in device memory
// [M][N][C] with FP64 atomics: up to 33 x 33 taps per event, spread over the wave's lanes
// (tap weights with the hardware reciprocal, 1 ulp).
__device__ inline void flush_events_cube(bool has, const Event &e, int ch, double *frame, int M, int N, int C,
                                         int lane) {
    uint64_t m = __ballot(has);
    while (m) {
        const int src = __builtin_ctzll(m);
        m &= m - 1;
        const int center = __builtin_amdgcn_readlane(e.center, src);
        const float s = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.s), src));
        const float Are = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.A.re), src));
        const float Aim = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.A.im), src));
        const int delta = __builtin_amdgcn_readlane(e.delta, src);
        const int cc = __builtin_amdgcn_readlane(e.cc, src);
        const float sc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.sc), src));
        const int deltaC = __builtin_amdgcn_readlane(e.deltaC, src);
        const int ns = delta ? 1 : 33, nc = deltaC ? 1 : 33;
        for (int t = lane; t < ns * nc; t += 64) {
            const int l = deltaC ? 0 : t / ns - 16;
            const int k = delta ? 0 : t % ns - 16;
            float coef = 1.0f;
            if (!deltaC) coef = __builtin_amdgcn_rcpf(sc - (float)l);
            if (!delta) coef = coef * __builtin_amdgcn_rcpf(s - (float)k);
            const int bc = wrap_bin(cc + l, M);
            const int bs = wrap_bin(center + k, N);
            double *dst = frame + 2 * (((int64_t)bc * N + bs) * C + ch);
            unsafeAtomicAdd(dst + 0, (double)(Are * coef));
            unsafeAtomicAdd(dst + 1, (double)(Aim * coef));
        }
    }
}

// A cube event of this lane (if any) of frame channel ch (wave-uniform) into that channel's
// event list (list ch - chBase) for the LDS tile splat, or, without a list or past its
// capacity, straight into the frame cube with device atomics. All lanes of the wave call it.
__device__ inline void cube_event(bool has, const Event &e, int ch, const TraceArgs &a, int lane) {
    bool direct = has;
    if (a.events) {
        const uint64_t m = __ballot(has);
        if (m) {
            const int leader = __builtin_ctzll(m);
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(a.evCount + (ch - a.chBase), (unsigned)__popcll(m));
            base = __shfl(base, leader, 64);
            const unsigned pos = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
            if (has && pos < a.evCap) {
                CubeEvent c;
                c.center = e.center;
                c.s = e.s;
                c.are = e.A.re;
                c.aim = e.A.im;
                c.cc = e.cc;
                c.sc = e.sc;
                c.flags = (e.delta ? 1 : 0) | (e.deltaC ? 2 : 0) | (ch << 2);
                c.pad = 0;
                a.events[(size_t)(ch - a.chBase) * a.evCap + pos] = c;
                direct = false;
            }
        }
    }
    flush_events_cube(direct, e, ch, a.frameAcc, a.M, a.N, a.C, lane);
}

// Wave-aggregated FP64 add into the training statistics: lanes with the same leaf are summed
// with a butterfly and one lane issues the global atomic.
__device__ inline void record_stats(bool pending, int leaf, double val, double *stat, int lane) {
    uint64_t m = __ballot(pending);
    while (m) {
        const int src = __builtin_ctzll(m);
        const int L = __builtin_amdgcn_readlane(leaf, src);
        const bool mine = pending && leaf == L;
        double v = mine ? val : 0.0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == src) atomicAdd(&stat[L], v);
        m &= ~__ballot(mine);
        pending = pending && !mine;
    }
}

}  // namespace hsr

// Binned-SAH BVH builder (host). A BVH2 is built first (binned SAH over all three axes, boxes
// widened conservatively), then collapsed into four-wide nodes: starting from a node's two
// children, the inner child with the largest surface area is replaced by its two children
// until four children are reached or only leaves remain. The kernels fetch a four-wide node
// as eight float4 (child boxes as six structure-of-arrays rows, then the child references),
// so one dependent fetch tests four boxes; triangle slots stay in leaf order.
#include "bvh.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>

namespace hsr {
namespace {

#ifndef HSR_SAH_BINS
#define HSR_SAH_BINS 64
#endif
#ifndef HSR_SAH_TRAV
#define HSR_SAH_TRAV 0.5f
#endif
constexpr int kBins = HSR_SAH_BINS;
constexpr float kSahTrav = HSR_SAH_TRAV;  // cost of a node test relative to a triangle test
#ifndef HSR_MAX_LEAF
#define HSR_MAX_LEAF 8
#endif
constexpr int kMaxLeaf = HSR_MAX_LEAF;
constexpr int kMedianDepth = 48;
constexpr int kMaxDepth = 80;

struct Box {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::numeric_limits<float>::infinity();
            hi[k] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const Box &b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f)) return 0.0f;
        return dx * dy + dy * dz + dz * dx;
    }
};

// Early split clipping (Ernst and Greiner 2007): a triangle whose box is large next to the
// mean triangle box is entered in the build as several references, each with the box of the
// triangle's part inside one half of its box (split at the middle of the longest axis,
// recursively), so that thin or tilted triangles do not inflate the boxes above them. A leaf
// holds references, and its slots the referenced triangles (whole, so every test is the
// exact triangle test; a triangle met through two leaves gives the same t and id).
#ifndef HSR_PRESPLIT
#define HSR_PRESPLIT 4.0
#endif
constexpr double kPresplit = HSR_PRESPLIT;  // split while box area > kPresplit x the mean (0: off)
constexpr int kPresplitMaxRefs = 8;         // references per triangle at most
constexpr double kPresplitBudget = 0.5;     // extra references at most, per triangle

// Box of the part of triangle (a, b, c) inside box [lo, hi] (Sutherland-Hodgman clipping in
// double, bounds rounded outwards); false when that part is empty.
bool clipped_box(const double a[3], const double b[3], const double c[3], const double lo[3], const double hi[3],
                 float olo[3], float ohi[3]) {
    double poly[12][3], tmp[12][3];
    int n = 3;
    for (int k = 0; k < 3; ++k) {
        poly[0][k] = a[k];
        poly[1][k] = b[k];
        poly[2][k] = c[k];
    }
    for (int axis = 0; axis < 3 && n > 0; ++axis) {
        for (int side = 0; side < 2 && n > 0; ++side) {
            const double plane = side ? hi[axis] : lo[axis];
            auto inside = [&](const double *p) { return side ? p[axis] <= plane : p[axis] >= plane; };
            int m = 0;
            for (int i = 0; i < n; ++i) {
                const double *P = poly[i], *Q = poly[(i + 1) % n];
                const bool pin = inside(P), qin = inside(Q);
                if (pin && m < 12) {
                    for (int k = 0; k < 3; ++k) tmp[m][k] = P[k];
                    ++m;
                }
                if (pin != qin && m < 12) {
                    const double t = (plane - P[axis]) / (Q[axis] - P[axis]);
                    for (int k = 0; k < 3; ++k) tmp[m][k] = P[k] + t * (Q[k] - P[k]);
                    tmp[m][axis] = plane;
                    ++m;
                }
            }
            n = m;
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) poly[i][k] = tmp[i][k];
        }
    }
    if (n == 0) return false;
    for (int k = 0; k < 3; ++k) {
        double l = poly[0][k], h = poly[0][k];
        for (int i = 1; i < n; ++i) {
            l = std::min(l, poly[i][k]);
            h = std::max(h, poly[i][k]);
        }
        // the clipped part lies within the triangle's own box: never beyond it
        l = std::max(l, std::min({a[k], b[k], c[k]}));
        h = std::min(h, std::max({a[k], b[k], c[k]}));
        olo[k] = std::nextafter((float)l, -std::numeric_limits<float>::infinity());
        ohi[k] = std::nextafter((float)h, std::numeric_limits<float>::infinity());
    }
    return true;
}

struct Builder {
    const float *V;
    const int32_t *I;
    std::vector<Box> tbox;
    std::vector<float> cen;          // 3 per triangle
    std::vector<int32_t> order;      // slot -> triangle id
    // output (inner nodes), guarded for parallel subtrees
    std::mutex mtx;
    std::vector<std::vector<float>> chunks;  // per-subtree node arrays, merged at the end
    std::atomic<int32_t> maxDepth{0};
    std::atomic<int32_t> maxLeaf{0};

    // Widened float box of a child (lo rounded down, hi rounded up).
    static void widen(const Box &b, float lo[3], float hi[3]) {
        for (int k = 0; k < 3; ++k) {
            const double l = b.lo[k], h = b.hi[k];
            const double m = (std::fabs(l) + std::fabs(h) + (h - l)) * (1.0 / 1048576.0) + 1e-6;
            lo[k] = std::nextafter((float)(l - m), -std::numeric_limits<float>::infinity());
            hi[k] = std::nextafter((float)(h + m), std::numeric_limits<float>::infinity());
        }
    }

    Box range_box(int32_t b, int32_t e, Box *cbox) const {
        Box bb, cb;
        bb.reset();
        cb.reset();
        for (int32_t i = b; i < e; ++i) {
            const int32_t t = order[i];
            bb.grow(tbox[t]);
            for (int k = 0; k < 3; ++k) {
                cb.lo[k] = std::min(cb.lo[k], cen[3 * t + k]);
                cb.hi[k] = std::max(cb.hi[k], cen[3 * t + k]);
            }
        }
        if (cbox) *cbox = cb;
        return bb;
    }

    // Returns the split position (b < mid < e) or -1 for a leaf.
    int32_t split(int32_t b, int32_t e, int depth) {
        const int32_t n = e - b;
        if (n <= 1) return -1;
        Box cb;
        const Box bb = range_box(b, e, &cb);
        if (depth >= kMedianDepth) {
            int axis = 0;
            for (int k = 1; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > cb.hi[axis] - cb.lo[axis]) axis = k;
            const int32_t mid = b + n / 2;
            std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e,
                             [&](int32_t x, int32_t y) { return cen[3 * x + axis] < cen[3 * y + axis]; });
            return n <= kMaxLeaf && depth >= kMaxDepth - 1 ? -1 : mid;
        }
        float best = std::numeric_limits<float>::infinity();
        int best_axis = -1, best_bin = -1;
        for (int k = 0; k < 3; ++k) {
            const float ext = cb.hi[k] - cb.lo[k];
            if (!(ext > 0.0f)) continue;
            Box bins[kBins];
            int32_t cnt[kBins] = {0};
            for (auto &x : bins) x.reset();
            const float scale = (float)kBins / ext;
            for (int32_t i = b; i < e; ++i) {
                const int32_t t = order[i];
                int bi = (int)((cen[3 * t + k] - cb.lo[k]) * scale);
                bi = std::min(std::max(bi, 0), kBins - 1);
                cnt[bi]++;
                bins[bi].grow(tbox[t]);
            }
            float rarea[kBins];
            int32_t rcnt[kBins];
            Box acc;
            acc.reset();
            int32_t c = 0;
            for (int i = kBins - 1; i > 0; --i) {
                acc.grow(bins[i]);
                c += cnt[i];
                rarea[i] = acc.area();
                rcnt[i] = c;
            }
            acc.reset();
            c = 0;
            for (int i = 0; i < kBins - 1; ++i) {
                acc.grow(bins[i]);
                c += cnt[i];
                if (c == 0 || rcnt[i + 1] == 0) continue;
                const float cost = acc.area() * (float)c + rarea[i + 1] * (float)rcnt[i + 1];
                if (cost < best) {
                    best = cost;
                    best_axis = k;
                    best_bin = i;
                }
            }
        }
        const float leaf_cost = bb.area() * (float)n;
        if (best_axis < 0) {
            if (n <= kMaxLeaf) return -1;
            return b + n / 2;  // coincident centroids: split by count
        }
        // SAH: a leaf when its triangle tests cost no more than a node test and the split
        if (n <= kMaxLeaf && leaf_cost <= best + kSahTrav * bb.area()) return -1;
        const int k = best_axis;
        const float scale = (float)kBins / (cb.hi[k] - cb.lo[k]);
        auto it = std::partition(order.begin() + b, order.begin() + e, [&](int32_t t) {
            int bi = (int)((cen[3 * t + k] - cb.lo[k]) * scale);
            bi = std::min(std::max(bi, 0), kBins - 1);
            return bi <= best_bin;
        });
        int32_t mid = (int32_t)(it - order.begin());
        if (mid == b || mid == e) mid = b + n / 2;
        return mid;
    }

    static int32_t leaf_ref(int32_t first, int32_t count) {
        return (int32_t)(0x80000000u | ((uint32_t)(count - 1) << 27) | (uint32_t)first);
    }

    // Build the subtree over [b, e) into `nodes` (local indices); returns its root ref.
    // Inner node refs are local to `nodes` with `base` added.
    int32_t build(int32_t b, int32_t e, int depth, std::vector<float> &nodes, int32_t base) {
        const int32_t mid = split(b, e, depth);
        if (mid < 0) {
            int32_t d = depth, cur = maxDepth.load();
            while (d > cur && !maxDepth.compare_exchange_weak(cur, d)) {}
            int32_t l = e - b, cl = maxLeaf.load();
            while (l > cl && !maxLeaf.compare_exchange_weak(cl, l)) {}
            return leaf_ref(b, e - b);
        }
        const int32_t me = (int32_t)(nodes.size() / 16);
        nodes.resize(nodes.size() + 16, 0.0f);
        const int32_t r0 = build(b, mid, depth + 1, nodes, base);
        const int32_t r1 = build(mid, e, depth + 1, nodes, base);
        fill(nodes, me, b, mid, e, r0, r1, base);
        return base + me;
    }

    void fill(std::vector<float> &nodes, int32_t me, int32_t b, int32_t mid, int32_t e, int32_t r0, int32_t r1,
              int32_t) {
        float lo0[3], hi0[3], lo1[3], hi1[3];
        widen(range_box(b, mid, nullptr), lo0, hi0);
        widen(range_box(mid, e, nullptr), lo1, hi1);
        float *n = &nodes[16 * (size_t)me];
        n[0] = lo0[0]; n[1] = hi0[0]; n[2] = lo0[1]; n[3] = hi0[1];
        n[4] = lo1[0]; n[5] = hi1[0]; n[6] = lo1[1]; n[7] = hi1[1];
        n[8] = lo0[2]; n[9] = hi0[2]; n[10] = lo1[2]; n[11] = hi1[2];
        std::memcpy(&n[12], &r0, 4);
        std::memcpy(&n[13], &r1, 4);
        n[14] = 0.0f;
        n[15] = 0.0f;
    }
};

// Spatial-split build (HSR_SBVH; Stich, Friedrich and Dietrich 2009): besides the binned object
// split, a node whose two object-split children overlap can be cut by a plane, with the
// references that straddle it clipped into both sides; the cheaper of the two by SAH wins.
#ifndef HSR_SBVH
#define HSR_SBVH 1
#endif
#ifndef HSR_SPATIAL_BINS
#define HSR_SPATIAL_BINS 32
#endif
constexpr int kSpatialBins = HSR_SPATIAL_BINS;
#ifndef HSR_SPATIAL_ALPHA
#define HSR_SPATIAL_ALPHA 1e-5f
#endif
#ifndef HSR_SPATIAL_BUDGET
#define HSR_SPATIAL_BUDGET 0.3
#endif
constexpr float kSpatialAlpha = HSR_SPATIAL_ALPHA;    // overlap / root area above which spatial splits are tried
constexpr double kSpatialBudget = HSR_SPATIAL_BUDGET;  // extra references at most, per triangle (a soft cap)
constexpr int kSpatialParMin = 1 << 15;  // references of a subtree split over two threads
constexpr int kSpatialParDepth = 5;      // at most 2^5 threads

struct SRef {
    int32_t tri;
    Box box;
};

struct SpatialBuilder {
    const float *V;
    const int32_t *I;
    std::vector<float> nodes;
    std::vector<int32_t> slotTri;
    int64_t budget = 0;
    float rootArea = 1.0f;
    int maxDepth = 0, maxLeaf = 0;

    void tri(int32_t t, double v[3][3]) const {
        for (int q = 0; q < 3; ++q)
            for (int k = 0; k < 3; ++k) v[q][k] = V[3 * (int64_t)I[3 * t + q] + k];
    }
    // the part of reference r inside [lo, hi] on axis k (false: empty)
    bool clip(const SRef &r, int k, float lo, float hi, Box &out) const {
        double v[3][3];
        tri(r.tri, v);
        double blo[3], bhi[3];
        for (int e = 0; e < 3; ++e) {
            blo[e] = r.box.lo[e];
            bhi[e] = r.box.hi[e];
        }
        blo[k] = std::max(blo[k], (double)lo);
        bhi[k] = std::min(bhi[k], (double)hi);
        if (blo[k] > bhi[k]) return false;
        if (!clipped_box(v[0], v[1], v[2], blo, bhi, out.lo, out.hi)) return false;
        for (int e = 0; e < 3; ++e) {  // never beyond the reference's own box
            out.lo[e] = std::max(out.lo[e], r.box.lo[e]);
            out.hi[e] = std::min(out.hi[e], r.box.hi[e]);
        }
        return true;
    }

    bool overflow = false;  // a leaf beyond the 16 references a leaf reference can hold

    int32_t leaf(const std::vector<SRef> &refs, int depth) {
        if (refs.size() > 16) overflow = true;
        const int32_t first = (int32_t)slotTri.size();
        for (const SRef &r : refs) slotTri.push_back(r.tri);
        maxDepth = std::max(maxDepth, depth);
        maxLeaf = std::max(maxLeaf, (int)refs.size());
        return Builder::leaf_ref(first, (int32_t)refs.size());
    }

    static Box bounds(const std::vector<SRef> &refs) {
        Box b;
        b.reset();
        for (const SRef &r : refs) b.grow(r.box);
        return b;
    }

    // Splits refs into l and r; returns false for a leaf.
    bool split(std::vector<SRef> &refs, int depth, std::vector<SRef> &L, std::vector<SRef> &R) {
        const int n = (int)refs.size();
        const Box bb = bounds(refs);
        Box cb;
        cb.reset();
        for (const SRef &r : refs)
            for (int k = 0; k < 3; ++k) {
                const float c = 0.5f * (r.box.lo[k] + r.box.hi[k]);
                cb.lo[k] = std::min(cb.lo[k], c);
                cb.hi[k] = std::max(cb.hi[k], c);
            }
        // object split
        float best = std::numeric_limits<float>::infinity();
        int bAxis = -1, bBin = -1;
        Box bl, br;
        for (int k = 0; k < 3; ++k) {
            const float ext = cb.hi[k] - cb.lo[k];
            if (!(ext > 0.0f)) continue;
            Box bins[kBins];
            int cnt[kBins] = {0};
            for (auto &x : bins) x.reset();
            const float scale = (float)kBins / ext;
            for (const SRef &r : refs) {
                int bi = (int)((0.5f * (r.box.lo[k] + r.box.hi[k]) - cb.lo[k]) * scale);
                bi = std::min(std::max(bi, 0), kBins - 1);
                cnt[bi]++;
                bins[bi].grow(r.box);
            }
            Box racc[kBins];
            int rc[kBins];
            Box acc;
            acc.reset();
            int c = 0;
            for (int i = kBins - 1; i > 0; --i) {
                acc.grow(bins[i]);
                c += cnt[i];
                racc[i] = acc;
                rc[i] = c;
            }
            acc.reset();
            c = 0;
            for (int i = 0; i < kBins - 1; ++i) {
                acc.grow(bins[i]);
                c += cnt[i];
                if (c == 0 || rc[i + 1] == 0) continue;
                const float cost = acc.area() * (float)c + racc[i + 1].area() * (float)rc[i + 1];
                if (cost < best) {
                    best = cost;
                    bAxis = k;
                    bBin = i;
                    bl = acc;
                    br = racc[i + 1];
                }
            }
        }
        // spatial split, where the object split's children overlap
        float sBest = std::numeric_limits<float>::infinity();
        int sAxis = -1;
        float sPlane = 0.0f;
        if (bAxis >= 0 && budget > 0 && depth < 32) {
            Box ov;
            for (int k = 0; k < 3; ++k) {
                ov.lo[k] = std::max(bl.lo[k], br.lo[k]);
                ov.hi[k] = std::min(bl.hi[k], br.hi[k]);
            }
            if (ov.area() > kSpatialAlpha * rootArea) {
                for (int k = 0; k < 3; ++k) {
                    const float ext = bb.hi[k] - bb.lo[k];
                    if (!(ext > 0.0f)) continue;
                    Box bins[kSpatialBins];
                    int enter[kSpatialBins] = {0}, leave[kSpatialBins] = {0};
                    for (auto &x : bins) x.reset();
                    const float w = ext / (float)kSpatialBins;
                    auto binOf = [&](float x) {
                        return std::min(std::max((int)((x - bb.lo[k]) / w), 0), kSpatialBins - 1);
                    };
                    for (const SRef &r : refs) {
                        const int b0 = binOf(r.box.lo[k]), b1 = binOf(r.box.hi[k]);
                        enter[b0]++;
                        leave[b1]++;
                        if (b0 == b1) {
                            bins[b0].grow(r.box);
                            continue;
                        }
                        for (int b = b0; b <= b1; ++b) {
                            Box part;
                            if (clip(r, k, bb.lo[k] + w * b, b == kSpatialBins - 1 ? bb.hi[k] : bb.lo[k] + w * (b + 1),
                                     part))
                                bins[b].grow(part);
                        }
                    }
                    Box racc[kSpatialBins];
                    int rc[kSpatialBins];
                    Box acc;
                    acc.reset();
                    int c = 0;
                    for (int i = kSpatialBins - 1; i > 0; --i) {
                        acc.grow(bins[i]);
                        c += leave[i];
                        racc[i] = acc;
                        rc[i] = c;
                    }
                    acc.reset();
                    c = 0;
                    for (int i = 0; i < kSpatialBins - 1; ++i) {
                        acc.grow(bins[i]);
                        c += enter[i];
                        if (c == 0 || rc[i + 1] == 0) continue;
                        const float cost = acc.area() * (float)c + racc[i + 1].area() * (float)rc[i + 1];
                        if (cost < sBest) {
                            sBest = cost;
                            sAxis = k;
                            sPlane = bb.lo[k] + w * (i + 1);
                        }
                    }
                }
            }
        }
        const float leafCost = bb.area() * (float)n;
        const float bestAny = std::min(best, sBest);
        if (n <= kMaxLeaf && (bAxis < 0 && sAxis < 0 || leafCost <= bestAny + kSahTrav * bb.area())) return false;
        L.clear();
        R.clear();
        if (sAxis >= 0 && sBest < best) {
            const int k = sAxis;
            for (const SRef &r : refs) {
                if (r.box.hi[k] <= sPlane) L.push_back(r);
                else if (r.box.lo[k] >= sPlane) R.push_back(r);
                else {
                    SRef a{r.tri, {}}, b{r.tri, {}};
                    const bool ha = clip(r, k, r.box.lo[k], sPlane, a.box);
                    const bool hb = clip(r, k, sPlane, r.box.hi[k], b.box);
                    if (ha) L.push_back(a);
                    if (hb) R.push_back(b);
                    if (ha && hb) --budget;
                    if (!ha && !hb) L.push_back(r);
                }
            }
            if (!L.empty() && !R.empty()) return true;
            L.clear();
            R.clear();
        }
        if (bAxis >= 0) {
            const int k = bAxis;
            const float scale = (float)kBins / (cb.hi[k] - cb.lo[k]);
            for (const SRef &r : refs) {
                int bi = (int)((0.5f * (r.box.lo[k] + r.box.hi[k]) - cb.lo[k]) * scale);
                bi = std::min(std::max(bi, 0), kBins - 1);
                (bi <= bBin ? L : R).push_back(r);
            }
            if (!L.empty() && !R.empty()) return true;
        }
        // coincident centroids: by count
        L.assign(refs.begin(), refs.begin() + n / 2);
        R.assign(refs.begin() + n / 2, refs.end());
        return true;
    }

    // Appends a subtree built by another builder (node indices and leaf slots local to it):
    // its inner references move by this builder's node count, its leaves by its slot count.
    int32_t append(SpatialBuilder &sub, int32_t root) {
        const int32_t nodeBase = (int32_t)(nodes.size() / 16), slotBase = (int32_t)slotTri.size();
        auto move = [&](int32_t ref) {
            if (ref >= 0) return ref + nodeBase;
            const int32_t first = ref & 0x07FFFFFF, count = ((ref >> 27) & 0xF) + 1;
            return Builder::leaf_ref(first + slotBase, count);
        };
        for (size_t q = 0; q < sub.nodes.size() / 16; ++q) {
            float *nd = &sub.nodes[16 * q];
            for (int c = 0; c < 2; ++c) {
                int32_t ref;
                std::memcpy(&ref, &nd[12 + c], 4);
                ref = move(ref);
                std::memcpy(&nd[12 + c], &ref, 4);
            }
        }
        nodes.insert(nodes.end(), sub.nodes.begin(), sub.nodes.end());
        slotTri.insert(slotTri.end(), sub.slotTri.begin(), sub.slotTri.end());
        maxDepth = std::max(maxDepth, sub.maxDepth);
        maxLeaf = std::max(maxLeaf, sub.maxLeaf);
        overflow = overflow || sub.overflow;
        return move(root);
    }

    // Large subtrees are built by their own builders, the left one on another thread. The
    // spatial-split budget is shared out by reference count, so the tree does not depend on
    // the threads' timing.
    int32_t build(std::vector<SRef> &refs, int depth) {
        std::vector<SRef> L, R;
        const bool inner = (int)refs.size() > 1 && depth < kMaxDepth - 1 && split(refs, depth, L, R);
        if (!inner || ((int)refs.size() <= 16 && depth >= kMaxDepth - 1)) return leaf(refs, depth);
        std::vector<SRef>().swap(refs);
        const int32_t me = (int32_t)(nodes.size() / 16);
        nodes.resize(nodes.size() + 16, 0.0f);
        const Box b0 = bounds(L), b1 = bounds(R);
        const int64_t bl = (int64_t)((double)budget * (double)L.size() / (double)(L.size() + R.size()));
        const int64_t br = budget - bl;
        int32_t r0, r1;
        if (L.size() + R.size() >= (size_t)kSpatialParMin && depth < kSpatialParDepth) {
            SpatialBuilder sl{V, I, {}, {}, bl, rootArea, 0, 0}, sr{V, I, {}, {}, br, rootArea, 0, 0};
            int32_t rl = 0;
            std::thread t([&] { rl = sl.build(L, depth + 1); });
            const int32_t rr = sr.build(R, depth + 1);
            t.join();
            r0 = append(sl, rl);
            r1 = append(sr, rr);
            budget = sl.budget + sr.budget;
        } else {
            budget = bl;
            r0 = build(L, depth + 1);
            const int64_t leftover = budget;
            budget = br + leftover;
            r1 = build(R, depth + 1);
        }
        float lo0[3], hi0[3], lo1[3], hi1[3];
        Builder::widen(b0, lo0, hi0);
        Builder::widen(b1, lo1, hi1);
        float *nd = &nodes[16 * (size_t)me];
        nd[0] = lo0[0]; nd[1] = hi0[0]; nd[2] = lo0[1]; nd[3] = hi0[1];
        nd[4] = lo1[0]; nd[5] = hi1[0]; nd[6] = lo1[1]; nd[7] = hi1[1];
        nd[8] = lo0[2]; nd[9] = hi0[2]; nd[10] = lo1[2]; nd[11] = hi1[2];
        std::memcpy(&nd[12], &r0, 4);
        std::memcpy(&nd[13], &r1, 4);
        nd[14] = 0.0f;
        nd[15] = 0.0f;
        return me;
    }
};

// BVH2 child c of node n: its widened box and reference
void child2(const std::vector<float> &n2, int32_t n, int c, float lo[3], float hi[3], int32_t &ref) {
    const float *p = &n2[16 * (size_t)n];
    const int o = c ? 4 : 0;
    lo[0] = p[o + 0];
    hi[0] = p[o + 1];
    lo[1] = p[o + 2];
    hi[1] = p[o + 3];
    lo[2] = p[8 + 2 * c];
    hi[2] = p[9 + 2 * c];
    std::memcpy(&ref, &p[12 + c], 4);
}

// Collapses the BVH2 subtree of node n2 into a four-wide node appended to n4 (depth-first);
// returns its index.
int32_t collapse_node(const std::vector<float> &n2, int32_t node, std::vector<float> &n4, int depth, int &maxd) {
    struct E {
        float lo[3], hi[3];
        int32_t ref;
        float area() const {
            const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
            return dx * dy + dy * dz + dz * dx;
        }
    };
    E e[4];
    int n = 2;
    child2(n2, node, 0, e[0].lo, e[0].hi, e[0].ref);
    child2(n2, node, 1, e[1].lo, e[1].hi, e[1].ref);
    while (n < 4) {
        int bi = -1;
        float ba = -1.0f;
        for (int i = 0; i < n; ++i)
            if (e[i].ref >= 0 && e[i].area() > ba) {
                ba = e[i].area();
                bi = i;
            }
        if (bi < 0) break;
        const int32_t r = e[bi].ref;
        child2(n2, r, 0, e[bi].lo, e[bi].hi, e[bi].ref);
        child2(n2, r, 1, e[n].lo, e[n].hi, e[n].ref);
        ++n;
    }
    maxd = std::max(maxd, depth);
    const int32_t me = (int32_t)(n4.size() / 32);
    n4.resize(n4.size() + 32, 0.0f);
    int32_t refs[4];
    for (int i = 0; i < 4; ++i) {
        if (i < n) refs[i] = e[i].ref >= 0 ? collapse_node(n2, e[i].ref, n4, depth + 1, maxd) : e[i].ref;
        else refs[i] = kNoChild;
    }
    float *q = &n4[32 * (size_t)me];
    for (int i = 0; i < 4; ++i) {
        const bool used = i < n;
        q[0 + i] = used ? e[i].lo[0] : 0.0f;
        q[4 + i] = used ? e[i].hi[0] : 0.0f;
        q[8 + i] = used ? e[i].lo[1] : 0.0f;
        q[12 + i] = used ? e[i].hi[1] : 0.0f;
        q[16 + i] = used ? e[i].lo[2] : 0.0f;
        q[20 + i] = used ? e[i].hi[2] : 0.0f;
        std::memcpy(&q[24 + i], &refs[i], 4);
    }
    return me;
}

// SAH-optimal collapse (HSR_COLLAPSE_DP; off: on the vehicle workload it makes 12 % fewer
// nodes and 1 % fewer node visits than the greedy collapse, but one level more, and the frame
// takes 1 % longer). Dynamic programming over the BVH2, as in the wide-BVH collapse of
// Ylitie, Karras and Laine 2017, here for four children). For a BVH2 subtree n and k in 1..4,
// D(n, k) is the least SAH cost of covering n with at most k four-wide children: one child
// (n itself as a four-wide node, or n's leaf) or its two subtrees covered by i + j <= k. A
// four-wide node costs its area times kCollapseNode and the cheapest cover of its subtree
// with up to four children; a leaf its area times its triangle count. Areas are those of the
// boxes the parent holds for its children.
#ifndef HSR_COLLAPSE_NODE
#define HSR_COLLAPSE_NODE 1.5f
#endif
constexpr float kCollapseNode = HSR_COLLAPSE_NODE;  // four-wide node visit / triangle test

struct Cover {
    float cost[5];     // D(n, k), k = 1..4 (index 0 unused)
    int8_t left[5];    // the left subtree's share of the cover with D(n, k), 0: n as one child
};

struct Collapser {
    const std::vector<float> &n2;
    std::vector<Cover> dp;      // per BVH2 node
    std::vector<float> area;    // per BVH2 node: area of its box
    std::vector<int8_t> split;  // per BVH2 node: the left share of its four-wide node's cover

    static float box_area(const float lo[3], const float hi[3]) {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return dx * dy + dy * dz + dz * dx;
    }
    static int32_t leaf_count(int32_t ref) { return ((ref >> 27) & 0xF) + 1; }

    // D(child, k) of child slot c of BVH2 node n
    float child_cost(int32_t n, int c, int k) const {
        float lo[3], hi[3];
        int32_t ref;
        child2(n2, n, c, lo, hi, ref);
        if (ref < 0) return box_area(lo, hi) * (float)leaf_count(ref);
        return dp[ref].cost[k];
    }

    void solve(int32_t n) {  // children are solved first (post-order)
        for (int c = 0; c < 2; ++c) {
            float lo[3], hi[3];
            int32_t ref;
            child2(n2, n, c, lo, hi, ref);
            if (ref >= 0) {
                area[ref] = box_area(lo, hi);
                solve(ref);
            }
        }
        Cover &d = dp[n];
        // n as one four-wide node: its two subtrees covered by i + j <= 4 children
        float best = std::numeric_limits<float>::infinity();
        int8_t bi = 1;
        for (int i = 1; i <= 3; ++i) {
            const float v = child_cost(n, 0, i) + child_cost(n, 1, 4 - i);
            if (v < best) {
                best = v;
                bi = (int8_t)i;
            }
        }
        split[n] = bi;
        d.cost[1] = area[n] * kCollapseNode + best;
        d.left[1] = 0;
        for (int k = 2; k <= 4; ++k) {
            d.cost[k] = d.cost[k - 1];
            d.left[k] = d.left[k - 1];
            for (int i = 1; i < k; ++i) {
                const float v = child_cost(n, 0, i) + child_cost(n, 1, k - i);
                if (v < d.cost[k]) {
                    d.cost[k] = v;
                    d.left[k] = (int8_t)i;
                }
            }
        }
    }

    // The four-wide children covering BVH2 child slot c of node n with at most k children.
    void gather(int32_t n, int c, int k, float (*lo)[3], float (*hi)[3], int32_t *refs, bool *wide, int &m) const {
        float blo[3], bhi[3];
        int32_t ref;
        child2(n2, n, c, blo, bhi, ref);
        const int share = ref >= 0 ? dp[ref].left[k] : 0;
        if (share == 0) {  // one child: a leaf or ref as a four-wide node
            std::memcpy(lo[m], blo, sizeof(blo));
            std::memcpy(hi[m], bhi, sizeof(bhi));
            refs[m] = ref;
            wide[m] = ref >= 0;
            ++m;
            return;
        }
        gather(ref, 0, share, lo, hi, refs, wide, m);
        gather(ref, 1, k - share, lo, hi, refs, wide, m);
    }

    int32_t emit(int32_t n, std::vector<float> &n4, int depth, int &maxd) const {
        float lo[4][3], hi[4][3];
        int32_t refs[4];
        bool wide[4];
        int m = 0;
        gather(n, 0, split[n], lo, hi, refs, wide, m);
        gather(n, 1, 4 - split[n], lo, hi, refs, wide, m);
        maxd = std::max(maxd, depth);
        const int32_t me = (int32_t)(n4.size() / 32);
        n4.resize(n4.size() + 32, 0.0f);
        int32_t out[4];
        for (int i = 0; i < 4; ++i) out[i] = i < m ? (wide[i] ? emit(refs[i], n4, depth + 1, maxd) : refs[i]) : kNoChild;
        float *q = &n4[32 * (size_t)me];
        for (int i = 0; i < 4; ++i) {
            const bool used = i < m;
            q[0 + i] = used ? lo[i][0] : 0.0f;
            q[4 + i] = used ? hi[i][0] : 0.0f;
            q[8 + i] = used ? lo[i][1] : 0.0f;
            q[12 + i] = used ? hi[i][1] : 0.0f;
            q[16 + i] = used ? lo[i][2] : 0.0f;
            q[20 + i] = used ? hi[i][2] : 0.0f;
            std::memcpy(&q[24 + i], &out[i], 4);
        }
        return me;
    }
};

}  // namespace

void collapse4(const std::vector<float> &n2, std::vector<float> &n4, int32_t &depth4) {
    n4.clear();
    n4.reserve(n2.size());
    int maxd = 0;
#ifndef HSR_COLLAPSE_DP
    collapse_node(n2, 0, n4, 0, maxd);
#else
    Collapser cl{n2, {}, {}, {}};
    const size_t count = n2.size() / 16;
    cl.dp.resize(count);
    cl.area.assign(count, 0.0f);
    cl.split.assign(count, 1);
    float lo0[3], hi0[3], lo1[3], hi1[3];
    int32_t r0, r1;
    child2(n2, 0, 0, lo0, hi0, r0);
    child2(n2, 0, 1, lo1, hi1, r1);
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo0[k], lo1[k]);
        hi[k] = std::max(hi0[k], hi1[k]);
    }
    cl.area[0] = Collapser::box_area(lo, hi);
    cl.solve(0);
    cl.emit(0, n4, 0, maxd);
#endif
    depth4 = maxd;
}

float dequant(float origin, uint32_t q, float scale) {
    volatile float prod = (float)q * scale;  // exact: q < 256 and scale a power of two
    return origin + prod;
}

bool quantize4(const std::vector<float> &n4, std::vector<float> &out, int32_t emptyRef) {
    const size_t count = n4.size() / 32;
    out.assign(16 * count, 0.0f);
    for (size_t nd = 0; nd < count; ++nd) {
        const float *q = &n4[32 * nd];
        int32_t refs[4];
        std::memcpy(refs, &q[24], 16);
        uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
        float origin[3], scale[3];
        for (int k = 0; k < 3; ++k) {
            // rows: lo.x 0, hi.x 4, lo.y 8, hi.y 12, lo.z 16, hi.z 20
            const float *lo = &q[8 * k], *hi = &q[8 * k + 4];
            float o = std::numeric_limits<float>::infinity(), top = -std::numeric_limits<float>::infinity();
            for (int i = 0; i < 4; ++i)
                if (refs[i] != kNoChild) {
                    o = std::min(o, lo[i]);
                    top = std::max(top, hi[i]);
                }
            if (!(o <= top)) o = top = 0.0f;  // no child
            origin[k] = o;
            // smallest power-of-two step with which 255 steps reach the top
            int e = -120;
            const double ext = (double)top - (double)o;
            if (ext > 0.0) e = std::max(-120, (int)std::ceil(std::log2(ext / 255.0)));
            for (;; ++e) {
                if (e > 120) return false;
                const float s = std::ldexp(1.0f, e);
                if (dequant(o, 255, s) >= top) break;
            }
            const float s = std::ldexp(1.0f, e);
            scale[k] = s;
            for (int i = 0; i < 4; ++i) {
                uint32_t a = 255, b = 0;  // an unused slot: the inverted box
                if (refs[i] != kNoChild) {
                    double fa = std::floor(((double)lo[i] - (double)o) / (double)s);
                    double fb = std::ceil(((double)hi[i] - (double)o) / (double)s);
                    a = (uint32_t)std::min(255.0, std::max(0.0, fa));
                    b = (uint32_t)std::min(255.0, std::max(0.0, fb));
                    while (a > 0 && dequant(o, a, s) > lo[i]) --a;
                    while (b < 255 && dequant(o, b, s) < hi[i]) ++b;
                    if (dequant(o, a, s) > lo[i] || dequant(o, b, s) < hi[i]) return false;
                }
                qlo[k] |= a << (8 * i);
                qhi[k] |= b << (8 * i);
            }
        }
        float *d = &out[16 * nd];
        d[0] = origin[0];
        d[1] = origin[1];
        d[2] = origin[2];
        d[3] = scale[0];
        const uint32_t w[6] = {qlo[0], qhi[0], qlo[1], qhi[1], qlo[2], qhi[2]};
        std::memcpy(&d[4], w, 24);
        d[10] = scale[1];
        d[11] = scale[2];
        for (int i = 0; i < 4; ++i)
            if (refs[i] == kNoChild) refs[i] = emptyRef;
        std::memcpy(&d[12], refs, 16);
    }
    return true;
}

bool build_bvh(const float *V, int64_t nv, const int32_t *I, int64_t nt, FlatBVH &out, int threads) {
    (void)threads;
    if (nt < 0 || nt >= (1 << 27)) return false;
    for (int64_t t = 0; t < 3 * nt; ++t)
        if (I[t] < 0 || I[t] >= nv) return false;
    Builder B;
    B.V = V;
    B.I = I;
    std::vector<Box> tb(nt);
    double meanArea = 0.0;
    for (int64_t t = 0; t < nt; ++t) {
        Box bx;
        bx.reset();
        for (int q = 0; q < 3; ++q) {
            const float *p = V + 3 * (int64_t)I[3 * t + q];
            for (int k = 0; k < 3; ++k) {
                bx.lo[k] = std::min(bx.lo[k], p[k]);
                bx.hi[k] = std::max(bx.hi[k], p[k]);
            }
        }
        tb[t] = bx;
        meanArea += bx.area();
    }
    meanArea = nt ? meanArea / (double)nt : 0.0;
    // references: (triangle, box); the triangle ids of the slots follow from them
    std::vector<int32_t> refTri;
    refTri.reserve((size_t)nt + (size_t)(kPresplitBudget * (double)nt));
    B.tbox.reserve(refTri.capacity());
    const int64_t budgetExtra = (int64_t)(kPresplitBudget * (double)nt);
    int64_t extra = 0;
    for (int64_t t = 0; t < nt; ++t) {
        const Box &bx = tb[t];
        if (kPresplit <= 0.0 || bx.area() <= kPresplit * meanArea || extra >= budgetExtra) {
            refTri.push_back((int32_t)t);
            B.tbox.push_back(bx);
            continue;
        }
        double v[3][3];
        for (int q = 0; q < 3; ++q)
            for (int k = 0; k < 3; ++k) v[q][k] = V[3 * (int64_t)I[3 * t + q] + k];
        // split the largest pending box at the middle of its longest axis until the boxes are
        // small enough or the triangle has kPresplitMaxRefs references
        std::vector<Box> parts{bx};
        bool grew = true;
        while (grew && (int)parts.size() < kPresplitMaxRefs && extra < budgetExtra) {
            grew = false;
            int bi = -1;
            float ba = (float)(kPresplit * meanArea);
            for (int i = 0; i < (int)parts.size(); ++i)
                if (parts[i].area() > ba) {
                    ba = parts[i].area();
                    bi = i;
                }
            if (bi < 0) break;
            const Box pb = parts[bi];
            int axis = 0;
            for (int k = 1; k < 3; ++k)
                if (pb.hi[k] - pb.lo[k] > pb.hi[axis] - pb.lo[axis]) axis = k;
            const double mid = 0.5 * ((double)pb.lo[axis] + (double)pb.hi[axis]);
            Box out[2];
            int got = 0;
            for (int side = 0; side < 2; ++side) {
                double lo[3], hi[3];
                for (int k = 0; k < 3; ++k) {
                    lo[k] = pb.lo[k];
                    hi[k] = pb.hi[k];
                }
                if (side) lo[axis] = mid;
                else hi[axis] = mid;
                Box cbx;
                if (clipped_box(v[0], v[1], v[2], lo, hi, cbx.lo, cbx.hi)) out[got++] = cbx;
            }
            if (got == 2) {
                parts[bi] = out[0];
                parts.push_back(out[1]);
                ++extra;
                grew = true;
            } else if (got == 1 && out[0].area() < pb.area()) {
                parts[bi] = out[0];
                grew = true;
            }
        }
        for (const Box &pb : parts) {
            refTri.push_back((int32_t)t);
            B.tbox.push_back(pb);
        }
    }
    const int64_t nr = (int64_t)refTri.size();
    if (nr >= (1 << 27)) return false;
    B.cen.resize(3 * nr);
    B.order.resize(nr);
    for (int64_t r = 0; r < nr; ++r) {
        for (int k = 0; k < 3; ++k) B.cen[3 * r + k] = 0.5f * (B.tbox[r].lo[k] + B.tbox[r].hi[k]);
        B.order[r] = (int32_t)r;
    }

    std::vector<float> nodes;
    nodes.reserve(16 * (size_t)std::max<int64_t>(2, nr));
    int32_t slotCount = (int32_t)nr;
    std::vector<int32_t> slotTri;  // spatial-split build: triangle of each slot
    bool spatial = false;
    if (HSR_SBVH && nr >= 2) {
        SpatialBuilder S;
        S.V = V;
        S.I = I;
        S.budget = (int64_t)(kSpatialBudget * (double)nt);
        std::vector<SRef> refs((size_t)nr);
        for (int64_t r = 0; r < nr; ++r) refs[r] = SRef{refTri[r], B.tbox[r]};
        S.rootArea = std::max(SpatialBuilder::bounds(refs).area(), 1e-30f);
        S.nodes.reserve(16 * (size_t)nr);
        const int32_t root = S.build(refs, 0);
        // The root must be an inner node at index 0, and no leaf may hold more than 16
        // references. A mesh of a few triangles whose split is not worth it by SAH comes back
        // as a single leaf, and a degenerate subtree can overflow a leaf: both are built by the
        // object builder below instead (its root over a single leaf has that leaf twice, and it
        // splits by the median below depth kMedianDepth).
        if (root == 0 && !S.overflow && S.slotTri.size() < (size_t)(1 << 27)) {
            nodes.swap(S.nodes);
            slotTri.swap(S.slotTri);
            slotCount = (int32_t)slotTri.size();
            B.maxDepth = S.maxDepth;
            B.maxLeaf = S.maxLeaf;
            spatial = true;
        }
    }
    if (!spatial && nr == 0) {
        // a root whose two children are a far point box over one degenerate triangle slot
        slotCount = 1;
        nodes.assign(16, 1e30f);
        const int32_t ref = Builder::leaf_ref(0, 1);
        std::memcpy(&nodes[12], &ref, 4);
        std::memcpy(&nodes[13], &ref, 4);
        nodes[14] = nodes[15] = 0.0f;
    } else if (!spatial) {
        // the root must be an inner node: its children are the two halves of the first split,
        // or the single leaf twice
        const int32_t mid = nr > 1 ? B.split(0, (int32_t)nr, 0) : -1;
        nodes.resize(16, 0.0f);
        if (mid < 0) {
            const int32_t ref = Builder::leaf_ref(0, (int32_t)nr);
            B.fill(nodes, 0, 0, (int32_t)nr, (int32_t)nr, ref, ref, 0);
            float *n = &nodes[0];
            for (int k = 0; k < 4; ++k) { n[4 + k] = n[k]; }
            n[10] = n[8];
            n[11] = n[9];
            B.maxLeaf = (int32_t)nr;
        } else {
            const int32_t r0 = B.build(0, mid, 1, nodes, 0);
            const int32_t r1 = B.build(mid, (int32_t)nr, 1, nodes, 0);
            B.fill(nodes, 0, 0, mid, (int32_t)nr, r0, r1, 0);
        }
    }
    collapse4(nodes, out.nodesF, out.depth4);
    // the sentinel leaf: one slot after the last, a degenerate triangle (its slot must fit the
    // 27 bits of a leaf's first slot)
    if (slotCount >= (1 << 27) - 1) return false;
    const int32_t sentinel = (int32_t)(0x80000000u | (uint32_t)slotCount);
    if (!quantize4(out.nodesF, out.nodes, sentinel)) return false;
    out.nodeCount = (int32_t)(out.nodes.size() / 16);
    out.slotCount = slotCount;
    out.maxDepth = B.maxDepth.load();
    out.maxLeaf = B.maxLeaf.load();
    out.tris.assign(12 * ((size_t)slotCount + 1), 0.0f);
    const int32_t noId = -1;
    std::memcpy(&out.tris[12 * (size_t)slotCount + 3], &noId, 4);
    for (int32_t s = 0; s < slotCount && nr > 0; ++s) {
        const int32_t t = slotTri.empty() ? refTri[B.order[s]] : slotTri[s];
        const float *a = V + 3 * (int64_t)I[3 * t + 0];
        const float *b = V + 3 * (int64_t)I[3 * t + 1];
        const float *c = V + 3 * (int64_t)I[3 * t + 2];
        float *r = &out.tris[12 * (size_t)s];
        r[0] = a[0]; r[1] = a[1]; r[2] = a[2];
        std::memcpy(&r[3], &t, 4);
        r[4] = b[0] - a[0]; r[5] = b[1] - a[1]; r[6] = b[2] - a[2]; r[7] = 0.0f;
        r[8] = c[0] - a[0]; r[9] = c[1] - a[1]; r[10] = c[2] - a[2]; r[11] = 0.0f;
    }
    // the closest-hit traversal keeps at most three pending children per level of the path
    return out.maxDepth < kMaxDepth && 3 * (out.depth4 + 1) + 1 <= kStackEntries;
}

}  // namespace hsr

// Primary visibility lists: the geometry shared by the device build (primary.hip), the depth-0
// closest-hit kernel (wavefront.hip) and the host check (tests/cpp/primary_lists.cpp).
//
// Every depth-0 ray of a batch starts at the transmitter's position o. Around o the directions
// are split into the six faces of a cube map: face f has major axis m = f / 2 (sign + for even
// f), face coordinates u = p_a / |p_m|, v = p_b / |p_m| with a = (m + 1) % 3, b = (m + 2) % 3.
// On each face a rectangle [u0, u0 + R / su) x [v0, v0 + R / sv) holding every triangle's
// footprint is cut into R x R cells. A cell lists the triangles whose footprint (the central
// projection of the triangle from o, clipped to the face's pyramid) overlaps it, each with a
// lower bound of its distance from o, in nondecreasing order of that bound. A ray then tests
// the triangles of its cell in list order and stops at the first whose bound exceeds its best
// hit so far: every triangle it skips lies farther than that hit, so the closest hit (smallest
// t, ties to the smaller id) is the one a full traversal finds, with the same triangle test.
//
// The footprints are widened so that a ray the float triangle test accepts is never outside
// its triangle's cells: each face's pyramid by a relative kFaceMargin, each footprint
// rectangle by kUvMargin in face coordinates (the float test accepts rays within a few ulps
// of the triangle's edges, and the ray's own u, v are rounded once). The distance bound is
// the exact point-triangle distance (double) shrunk by a relative 2^-10 and 1e-5 m, far more
// than the float triangle test's error in t. Triangles closer to o than kNear (which could
// surround it) are tested by every ray, before its cell.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define HSR_PV_HD __host__ __device__
#else
#define HSR_PV_HD
#endif

namespace hsr {
namespace pv {

constexpr int kFaces = 6;
constexpr double kFaceMargin = 1e-4;  // relative widening of each face's pyramid
constexpr double kUvMargin = 1e-4;    // widening of a footprint in face coordinates
constexpr double kNear = 1e-2;        // [m] triangles nearer than this are tested by every ray

struct V3 {
    double x, y, z;
};
HSR_PV_HD inline double at(const V3 &p, int i) { return i == 0 ? p.x : (i == 1 ? p.y : p.z); }
HSR_PV_HD inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
HSR_PV_HD inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
HSR_PV_HD inline V3 scale(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
HSR_PV_HD inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HSR_PV_HD inline double len(V3 a) { return sqrt(dot(a, a)); }

// Distance from the origin to the triangle (a, b, c): the closest point by Voronoi regions
// (Ericson, Real-Time Collision Detection, 5.1.5). A degenerate triangle (no finite result)
// falls back to the distance to its bounding box, also a lower bound.
HSR_PV_HD inline double point_triangle_distance(V3 a, V3 b, V3 c) {
    const V3 ab = sub(b, a), ac = sub(c, a);
    const V3 ap = scale(-1.0, a), bp = scale(-1.0, b), cp = scale(-1.0, c);
    double r;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    const double vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
    if (d1 <= 0.0 && d2 <= 0.0) {
        r = len(a);
    } else if (d3 >= 0.0 && d4 <= d3) {
        r = len(b);
    } else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        r = len(add(a, scale(d1 / (d1 - d3), ab)));
    } else if (d6 >= 0.0 && d5 <= d6) {
        r = len(c);
    } else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        r = len(add(a, scale(d2 / (d2 - d6), ac)));
    } else if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0) {
        r = len(add(b, scale((d4 - d3) / ((d4 - d3) + (d5 - d6)), sub(c, b))));
    } else {
        const double den = 1.0 / (va + vb + vc);
        r = len(add(a, add(scale(vb * den, ab), scale(vc * den, ac))));
    }
    if (!(r >= 0.0 && r < 1e300)) {
        double q = 0.0;
        for (int i = 0; i < 3; ++i) {
            const double lo = fmin(fmin(at(a, i), at(b, i)), at(c, i));
            const double hi = fmax(fmax(at(a, i), at(b, i)), at(c, i));
            const double g = lo > 0.0 ? lo : (hi < 0.0 ? -hi : 0.0);
            q += g * g;
        }
        r = sqrt(q);
    }
    return r;
}

// The list key of a triangle at distance r: a float no larger than r (1 - 2^-10) - 1e-5.
HSR_PV_HD inline float distance_key(double r) {
    double k = r * (1.0 - 1.0 / 1024.0) - 1e-5;
    if (!(k > 0.0)) return 0.0f;
    float f = (float)k;
    if ((double)f > k) f = nextafterf(f, 0.0f);
    return f;
}

// Axes of face f: major m, sign s (+1 or -1), face coordinates a, b.
HSR_PV_HD inline void face_axes(int f, int &m, double &s, int &ia, int &ib) {
    m = f >> 1;
    s = (f & 1) ? -1.0 : 1.0;
    ia = (m + 1) % 3;
    ib = (m + 2) % 3;
}

// Keeps the part of the convex polygon p[0..n) with w_a * p[ia] + w_m * p[im] <= 0; returns
// its vertex count (at most n + 1).
HSR_PV_HD inline int clip_half(const V3 *p, int n, V3 *out, int ia, double wa, int im, double wm) {
    int k = 0;
    for (int i = 0; i < n; ++i) {
        const V3 &P = p[i], &Q = p[(i + 1) % n];
        const double dp = wa * at(P, ia) + wm * at(P, im), dq = wa * at(Q, ia) + wm * at(Q, im);
        if (dp <= 0.0) out[k++] = P;
        if ((dp < 0.0 && dq > 0.0) || (dp > 0.0 && dq < 0.0)) {
            const double t = dp / (dp - dq);
            out[k++] = add(P, scale(t, sub(Q, P)));
        }
    }
    return k;
}

// Footprint of the triangle (a, b, c), given relative to o and at least kNear from it, on face
// f: the bounding box (u, v) of the triangle clipped to the face's widened pyramid. Returns
// false when the triangle does not reach the face.
HSR_PV_HD inline bool face_footprint(V3 a, V3 b, V3 c, int f, double box[4]) {
    int m, ia, ib;
    double s;
    face_axes(f, m, s, ia, ib);
    const double w = 1.0 + kFaceMargin;
    V3 p0[8], p1[8];
    p0[0] = a;
    p0[1] = b;
    p0[2] = c;
    int n = 3;
    // |p_a| <= w s p_m and |p_b| <= w s p_m (these imply s p_m >= 0)
    n = clip_half(p0, n, p1, ia, 1.0, m, -w * s);
    if (n == 0) return false;
    n = clip_half(p1, n, p0, ia, -1.0, m, -w * s);
    if (n == 0) return false;
    n = clip_half(p0, n, p1, ib, 1.0, m, -w * s);
    if (n == 0) return false;
    n = clip_half(p1, n, p0, ib, -1.0, m, -w * s);
    if (n == 0) return false;
    double u0 = 1e300, u1 = -1e300, v0 = 1e300, v1 = -1e300;
    for (int i = 0; i < n; ++i) {
        const double pm = s * at(p0[i], m);
        if (!(pm > 0.0)) continue;  // only a vertex at o itself, excluded by kNear
        const double u = at(p0[i], ia) / pm, v = at(p0[i], ib) / pm;
        u0 = fmin(u0, u);
        u1 = fmax(u1, u);
        v0 = fmin(v0, v);
        v1 = fmax(v1, v);
    }
    if (!(u0 <= u1 && v0 <= v1)) return false;
    box[0] = u0;
    box[1] = u1;
    box[2] = v0;
    box[3] = v1;
    return true;
}

// Float bounds of a footprint box for the face grid: lower ends rounded down, upper ends up.
HSR_PV_HD inline float round_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}
HSR_PV_HD inline float round_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}

// Face grid: cells (x, y) = floor((u - u0) su), floor((v - v0) sv) in [0, R); su = 0 marks a
// face without triangles.
struct FaceGrid {
    float u0, v0, su, sv;
};

// The grid of a face whose footprints span [umin, umax] x [vmin, vmax] (float bounds of the
// double boxes; empty when umin > umax): the span widened by twice kUvMargin, its origin
// rounded down and the cells per unit rounded down, so that R cells cover at least the span.
HSR_PV_HD inline FaceGrid face_grid(float umin, float umax, float vmin, float vmax, int R) {
    FaceGrid g{0.0f, 0.0f, 0.0f, 0.0f};
    if (!(umin <= umax && vmin <= vmax)) return g;
    const double lo_u = (double)umin - 2.0 * kUvMargin, lo_v = (double)vmin - 2.0 * kUvMargin;
    const double hi_u = (double)umax + 2.0 * kUvMargin, hi_v = (double)vmax + 2.0 * kUvMargin;
    float fu = (float)lo_u, fv = (float)lo_v;
    if ((double)fu > lo_u) fu = nextafterf(fu, -INFINITY);
    if ((double)fv > lo_v) fv = nextafterf(fv, -INFINITY);
    const double su = (double)R / (hi_u - (double)fu), sv = (double)R / (hi_v - (double)fv);
    float gu = (float)su, gv = (float)sv;
    if ((double)gu > su) gu = nextafterf(gu, 0.0f);
    if ((double)gv > sv) gv = nextafterf(gv, 0.0f);
    g.u0 = fu;
    g.v0 = fv;
    g.su = gu;
    g.sv = gv;
    return g;
}

// Cell rectangle [x0, x1] x [y0, y1] of a footprint box on a face grid (the box widened by
// kUvMargin, clamped to the grid); false for an empty grid.
HSR_PV_HD inline bool cell_rect(const double box[4], const FaceGrid &g, int R, int r[4]) {
    if (!(g.su > 0.0f)) return false;
    const double su = g.su, sv = g.sv, u0 = g.u0, v0 = g.v0;
    auto cl = [R](double x) {
        const double fx = floor(x);
        return fx < 0.0 ? 0 : (fx > (double)(R - 1) ? R - 1 : (int)fx);
    };
    r[0] = cl((box[0] - kUvMargin - u0) * su);
    r[1] = cl((box[1] + kUvMargin - u0) * su);
    r[2] = cl((box[2] - kUvMargin - v0) * sv);
    r[3] = cl((box[3] + kUvMargin - v0) * sv);
    return true;
}

// Face and face coordinates of a float ray direction, as the depth-0 kernel computes them.
HSR_PV_HD inline int ray_face(float dx, float dy, float dz, float &u, float &v) {
    const float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    float dm, da, db;
    int m;
    if (ax >= ay && ax >= az) {
        m = 0;
        dm = dx;
        da = dy;
        db = dz;
    } else if (ay >= az) {
        m = 1;
        dm = dy;
        da = dz;
        db = dx;
    } else {
        m = 2;
        dm = dz;
        da = dx;
        db = dy;
    }
    const float am = fabsf(dm);
    u = da / am;
    v = db / am;
    return 2 * m + (dm < 0.0f ? 1 : 0);
}

// Cell of a ray's face coordinates on its face grid, or -1 when it lies outside every footprint.
HSR_PV_HD inline int ray_cell(int f, float u, float v, const FaceGrid &g, int R) {
    if (!(g.su > 0.0f)) return -1;
    const float x = (u - g.u0) * g.su, y = (v - g.v0) * g.sv;
    if (!(x >= 0.0f && x < (float)R && y >= 0.0f && y < (float)R)) return -1;
    return (f * R + (int)x) * R + (int)y;
}

}  // namespace pv
}  // namespace hsr

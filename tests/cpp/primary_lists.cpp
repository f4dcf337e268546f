// Host check of the primary visibility lists (hussar_amd/csrc/primary.h): the lists built on
// the host with the same geometry functions as the device build (distance keys, face
// footprints, face grids, cell rectangles, lists in key order with ties in triangle order),
// then the closest hit of rays from the origin through the lists (near triangles first, then
// the ray's cell, stopping at the first key beyond the best hit) against brute force over
// every triangle slot with the same float triangle test and tie rule. Rays are aimed at
// triangle vertices (edge-heavy), at random points inside triangles and in uniform random
// directions. Prints "rays R hits H mismatches M entries E near N tested T" (T: mean
// triangles tested per ray).
// Usage: primary_lists mesh.bin ox oy oz R rays   (mesh.bin as for bvh_brute_force)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../hussar_amd/csrc/bvh.h"
#include "../../hussar_amd/csrc/primary.h"

namespace {
namespace pv = hsr::pv;

// Moeller-Trumbore as device_math.h's tri_intersect (tnear 5e-5 < t < tfar = inf)
float tri_test(const float *p, const float o[3], const float d[3]) {
    const float *v0 = p, *e1 = p + 4, *e2 = p + 8;
    const float pv_[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    const float det = (e1[0] * pv_[0] + e1[1] * pv_[1]) + e1[2] * pv_[2];
    if (det == 0.0f) return INFINITY;
    const float inv = 1.0f / det;
    const float tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    const float u = ((tv[0] * pv_[0] + tv[1] * pv_[1]) + tv[2] * pv_[2]) * inv;
    if (u < 0.0f || u > 1.0f) return INFINITY;
    const float qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    const float v = ((d[0] * qv[0] + d[1] * qv[1]) + d[2] * qv[2]) * inv;
    if (v < 0.0f || u + v > 1.0f) return INFINITY;
    const float t = ((e2[0] * qv[0] + e2[1] * qv[1]) + e2[2] * qv[2]) * inv;
    return (t > 5e-5f && t < INFINITY) ? t : INFINITY;
}

int id_of(const float *p) {
    int id;
    std::memcpy(&id, p + 3, 4);
    return id;
}

void test(const hsr::FlatBVH &B, int32_t s, const float o[3], const float d[3], float &best, int &bid, int &bslot) {
    const float *p = &B.tris[12 * (size_t)s];
    const float t = tri_test(p, o, d);
    const int id = id_of(p);
    if (t < INFINITY && (bid < 0 || t < best || (t == best && id < bid))) {
        best = t;
        bid = id;
        bslot = s;
    }
}

struct TriD {
    pv::V3 a, b, c;
};
TriD tri_d(const hsr::FlatBVH &B, int32_t s, const float o[3]) {
    const float *p = &B.tris[12 * (size_t)s];
    TriD t;
    t.a = {(double)p[0] - (double)o[0], (double)p[1] - (double)o[1], (double)p[2] - (double)o[2]};
    t.b = pv::add(t.a, pv::V3{(double)p[4], (double)p[5], (double)p[6]});
    t.c = pv::add(t.a, pv::V3{(double)p[8], (double)p[9], (double)p[10]});
    return t;
}
}  // namespace

int main(int argc, char **argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s mesh.bin ox oy oz R rays\n", argv[0]);
        return 2;
    }
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t nv = 0, nt = 0;
    if (std::fread(&nv, 8, 1, f) != 1 || std::fread(&nt, 8, 1, f) != 1) return 2;
    std::vector<float> V(3 * nv);
    std::vector<int32_t> I(3 * nt);
    if (nv && std::fread(V.data(), 4, V.size(), f) != V.size()) return 2;
    if (nt && std::fread(I.data(), 4, I.size(), f) != I.size()) return 2;
    std::fclose(f);
    const float o[3] = {(float)std::atof(argv[2]), (float)std::atof(argv[3]), (float)std::atof(argv[4])};
    const int R = std::atoi(argv[5]);
    const int nrays = std::atoi(argv[6]);
    hsr::FlatBVH B;
    if (!hsr::build_bvh(V.data(), nv, I.data(), nt, B)) {
        std::printf("build failed\n");
        return 1;
    }
    // one slot per triangle (its first in leaf order)
    std::vector<int32_t> first(nt, -1), triSlot;
    for (int32_t s = 0; s < B.slotCount; ++s) {
        const int id = id_of(&B.tris[12 * (size_t)s]);
        if (first[id] < 0) first[id] = s;
    }
    for (int32_t v : first)
        if (v >= 0) triSlot.push_back(v);
    const int64_t T = (int64_t)triSlot.size();
    // keys, near triangles and face boxes
    std::vector<uint32_t> keys(T);
    std::vector<char> isNear(T, 0);
    float box[6][4];
    for (auto &b : box) b[0] = b[2] = INFINITY, b[1] = b[3] = -INFINITY;
    for (int64_t t = 0; t < T; ++t) {
        const TriD q = tri_d(B, triSlot[t], o);
        const double r = pv::point_triangle_distance(q.a, q.b, q.c);
        const float k = pv::distance_key(r);
        std::memcpy(&keys[t], &k, 4);
        if (r < pv::kNear) {
            isNear[t] = 1;
            continue;
        }
        for (int fc = 0; fc < 6; ++fc) {
            double bx[4];
            if (!pv::face_footprint(q.a, q.b, q.c, fc, bx)) continue;
            box[fc][0] = std::min(box[fc][0], pv::round_down(bx[0]));
            box[fc][1] = std::max(box[fc][1], pv::round_up(bx[1]));
            box[fc][2] = std::min(box[fc][2], pv::round_down(bx[2]));
            box[fc][3] = std::max(box[fc][3], pv::round_up(bx[3]));
        }
    }
    pv::FaceGrid grid[6];
    for (int fc = 0; fc < 6; ++fc) grid[fc] = pv::face_grid(box[fc][0], box[fc][1], box[fc][2], box[fc][3], R);
    std::vector<int64_t> order(T);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return keys[x] < keys[y]; });
    std::vector<std::vector<std::pair<int32_t, uint32_t>>> cells(6 * (size_t)R * R);
    std::vector<std::pair<int32_t, uint32_t>> nearList;
    int64_t entries = 0;
    for (int64_t t : order) {
        if (isNear[t]) {
            nearList.push_back({triSlot[t], keys[t]});
            continue;
        }
        const TriD q = tri_d(B, triSlot[t], o);
        for (int fc = 0; fc < 6; ++fc) {
            double bx[4];
            int r[4];
            if (!pv::face_footprint(q.a, q.b, q.c, fc, bx) || !pv::cell_rect(bx, grid[fc], R, r)) continue;
            for (int x = r[0]; x <= r[1]; ++x)
                for (int y = r[2]; y <= r[3]; ++y) {
                    cells[((size_t)fc * R + x) * R + y].push_back({triSlot[t], keys[t]});
                    ++entries;
                }
        }
    }
    // rays
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    int64_t hits = 0, mism = 0, tested = 0;
    for (int i = 0; i < nrays; ++i) {
        double dir[3];
        const int kind = i % 3;
        if (kind < 2 && nt > 0) {
            const int64_t s = triSlot[(int64_t)(U(rng) * T) % T];
            const float *p = &B.tris[12 * (size_t)s];
            double w1, w2;
            if (kind == 0) {  // a vertex or a point on an edge
                const int e = (int)(U(rng) * 3);
                const double a = U(rng) < 0.5 ? 0.0 : U(rng);
                w1 = e == 0 ? a : (e == 1 ? 1.0 - a : 0.0);
                w2 = e == 0 ? 0.0 : (e == 1 ? a : a);
            } else {
                w1 = U(rng);
                w2 = U(rng);
                if (w1 + w2 > 1.0) w1 = 1.0 - w1, w2 = 1.0 - w2;
            }
            for (int k = 0; k < 3; ++k) dir[k] = (double)p[k] + w1 * p[4 + k] + w2 * p[8 + k] - (double)o[k];
        } else {
            const double z = 2.0 * U(rng) - 1.0, ph = 2.0 * M_PI * U(rng), rr = std::sqrt(std::max(0.0, 1.0 - z * z));
            dir[0] = rr * std::cos(ph);
            dir[1] = rr * std::sin(ph);
            dir[2] = z;
        }
        const double n = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        if (!(n > 0.0)) continue;
        const float d[3] = {(float)(dir[0] / n), (float)(dir[1] / n), (float)(dir[2] / n)};
        // brute force
        float tb = INFINITY;
        int bb = -1, sb = -1;
        for (int32_t s = 0; s < B.slotCount; ++s) test(B, s, o, d, tb, bb, sb);
        // lists, as k_wf_trav0
        float tl = INFINITY;
        int bl = -1, sl = -1;
        for (auto &e : nearList) {
            float k;
            std::memcpy(&k, &e.second, 4);
            if (k > tl) break;
            ++tested;
            test(B, e.first, o, d, tl, bl, sl);
        }
        float u, v;
        const int fc = pv::ray_face(d[0], d[1], d[2], u, v);
        const int c = pv::ray_cell(fc, u, v, grid[fc], R);
        if (c >= 0)
            for (auto &e : cells[c]) {
                float k;
                std::memcpy(&k, &e.second, 4);
                if (k > tl) break;
                ++tested;
                test(B, e.first, o, d, tl, bl, sl);
            }
        if (bb >= 0) ++hits;
        if (bb != bl || (bb >= 0 && tb != tl)) {
            if (mism < 10)
                std::fprintf(stderr, "ray %d kind %d: brute %d t %.9g, lists %d t %.9g\n", i, kind, bb, tb, bl, tl);
            ++mism;
        }
    }
    std::printf("rays %d hits %lld mismatches %lld entries %lld near %lld tested %.3f\n", nrays, (long long)hits,
                (long long)mism, (long long)entries, (long long)nearList.size(),
                nrays ? (double)tested / nrays : 0.0);
    return 0;
}

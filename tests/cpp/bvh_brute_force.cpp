// Host check of the product's BVH: closest hits through the quantised four-wide nodes (the
// box test of slab4q in trace_common.h, in float on the host, FMA as written) against brute
// force over every triangle slot, for rays aimed at triangle vertices (edge-heavy), from
// points on triangles in random directions and along the axes, and that every triangle sits
// in at least one slot. Prints
// "nodes N slots S depth4 D rays R hits H mismatches M missing K".
// Usage: bvh_brute_force mesh.bin rays   (mesh.bin: int64 nv, int64 nt, float v[3nv], int32 i[3nt])
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../hussar_amd/csrc/bvh.h"

namespace {
struct Ray {
    float o[3], d[3];
};

// Moeller-Trumbore as device_math.h's tri_intersect (tnear 5e-5 < t < tfar = inf)
float tri_test(const float *p, const Ray &r) {
    const float *v0 = p, *e1 = p + 4, *e2 = p + 8;
    const float pv[3] = {r.d[1] * e2[2] - r.d[2] * e2[1], r.d[2] * e2[0] - r.d[0] * e2[2],
                         r.d[0] * e2[1] - r.d[1] * e2[0]};
    const float det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
    if (det == 0.0f) return INFINITY;
    const float inv = 1.0f / det;
    const float tv[3] = {r.o[0] - v0[0], r.o[1] - v0[1], r.o[2] - v0[2]};
    const float u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
    if (u < 0.0f || u > 1.0f) return INFINITY;
    const float qv[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2], tv[0] * e1[1] - tv[1] * e1[0]};
    const float v = (r.d[0] * qv[0] + r.d[1] * qv[1] + r.d[2] * qv[2]) * inv;
    if (v < 0.0f || u + v > 1.0f) return INFINITY;
    const float t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
    return t > 5e-5f ? t : INFINITY;
}

bool better(float t, int id, float tb, int best) { return t < INFINITY && (best < 0 || t < tb || (t == tb && id < best)); }

long long sentinelVisits = 0;  // visits of the sentinel leaf of unused child slots

int closest(const hsr::FlatBVH &B, const Ray &r, float &tb) {
    float inv[3];
    for (int k = 0; k < 3; ++k)
        inv[k] = std::fabs(r.d[k]) > 7.888609052210118e-31f ? 1.0f / r.d[k] : std::copysign(1.2676506002282294e30f, r.d[k]);
    std::vector<int> st;
    int node = 0, best = -1;
    tb = INFINITY;
    while (true) {
        if (node >= 0) {
            const float *q = &B.nodes[16 * (size_t)node];
            const float scale[3] = {q[3], q[10], q[11]};
            float A[3], Bn[3], Bf[3];
            uint32_t nw[3], fw[3], w[6];
            std::memcpy(w, &q[4], 24);
            for (int k = 0; k < 3; ++k) {
                A[k] = scale[k] * inv[k];
                const float b = (q[k] - r.o[k]) * inv[k];
                const float err = std::fabs(b) * 9.5367431640625e-07f;
                Bn[k] = b - err;
                Bf[k] = b + err;
                const bool pos = inv[k] >= 0.0f;
                nw[k] = pos ? w[2 * k] : w[2 * k + 1];
                fw[k] = pos ? w[2 * k + 1] : w[2 * k];
            }
            float te[4];
            int ch[4];
            unsigned m = 0;
            for (int i = 0; i < 4; ++i) {
                std::memcpy(&ch[i], &q[12 + i], 4);
                float nn[3], ff[3];
                for (int k = 0; k < 3; ++k) {
                    nn[k] = std::fmaf((float)((nw[k] >> (8 * i)) & 0xFFu), A[k], Bn[k]);
                    ff[k] = std::fmaf((float)((fw[k] >> (8 * i)) & 0xFFu), A[k], Bf[k]);
                }
                const float t0 = std::fmax(std::fmax(nn[0], nn[1]), std::fmax(nn[2], 5e-5f));
                const float t1 = std::fmin(std::fmin(ff[0], ff[1]), ff[2]) * 1.0000004f;
                te[i] = t0;
                if (t0 <= t1 && t0 <= tb) m |= 1u << i;
            }
            if (m) {
                std::vector<std::pair<float, int>> hit;
                for (int i = 0; i < 4; ++i)
                    if (m >> i & 1) hit.push_back({te[i], ch[i]});
                std::sort(hit.begin(), hit.end(), [](auto a, auto b) { return a.first > b.first; });
                for (size_t i = 0; i + 1 < hit.size(); ++i) st.push_back(hit[i].second);
                node = hit.back().second;
                continue;
            }
        } else {
            const int first = node & 0x07FFFFFF, count = ((node >> 27) & 0xF) + 1;
            if (first == B.slotCount) ++sentinelVisits;
            for (int i = first; i < first + count; ++i) {
                const float *p = &B.tris[12 * (size_t)i];
                const float t = tri_test(p, r);
                int id;
                std::memcpy(&id, &p[3], 4);
                if (better(t, id, tb, best)) {
                    tb = t;
                    best = id;
                }
            }
        }
        if (st.empty()) return best;
        node = st.back();
        st.pop_back();
    }
}
}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    long long nv = 0, nt = 0;
    if (std::fread(&nv, 8, 1, f) != 1 || std::fread(&nt, 8, 1, f) != 1) return 2;
    std::vector<float> V(3 * nv);
    std::vector<int> I(3 * nt);
    if (std::fread(V.data(), 4, V.size(), f) != V.size() || std::fread(I.data(), 4, I.size(), f) != I.size()) return 2;
    std::fclose(f);
    const int NR = std::atoi(argv[2]);
    hsr::FlatBVH B;
    if (!hsr::build_bvh(V.data(), nv, I.data(), nt, B)) {
        std::printf("build failed\n");
        return 1;
    }
    std::mt19937 rng(3);
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    int bad = 0, hits = 0;
    for (int q = 0; q < NR && nt > 0; ++q) {
        Ray r{{U(rng) * 2.0f, 0.75f + U(rng), U(rng) * 2.0f}, {0, 0, 0}};
        const int kind = q % 3;
        const int t = (int)((U(rng) + 1.0f) * 0.5f * (float)(nt - 1));
        if (kind == 0) {  // aimed at a vertex
            const float *p = &V[3 * (size_t)I[3 * t + (q % 3)]];
            float n2 = 0.0f;
            for (int k = 0; k < 3; ++k) {
                r.d[k] = p[k] - r.o[k];
                n2 += r.d[k] * r.d[k];
            }
            for (int k = 0; k < 3; ++k) r.d[k] /= std::sqrt(n2);
        } else {  // from a point on a triangle: a random or an axis direction
            const float *p0 = &V[3 * (size_t)I[3 * t]], *p1 = &V[3 * (size_t)I[3 * t + 1]], *p2 = &V[3 * (size_t)I[3 * t + 2]];
            float a = (U(rng) + 1.0f) * 0.5f, b = (U(rng) + 1.0f) * 0.5f;
            if (a + b > 1.0f) {
                a = 1.0f - a;
                b = 1.0f - b;
            }
            for (int k = 0; k < 3; ++k) r.o[k] = p0[k] + a * (p1[k] - p0[k]) + b * (p2[k] - p0[k]);
            if (kind == 1) {
                float n2 = 0.0f;
                for (int k = 0; k < 3; ++k) {
                    r.d[k] = U(rng);
                    n2 += r.d[k] * r.d[k];
                }
                for (int k = 0; k < 3; ++k) r.d[k] /= std::sqrt(n2);
            } else {
                r.d[(q / 3) % 3] = (q / 9) % 2 ? 1.0f : -1.0f;
            }
        }
        float tb;
        const int id = closest(B, r, tb);
        float bt = INFINITY;
        int bid = -1;
        for (int s = 0; s < B.slotCount; ++s) {
            const float *p = &B.tris[12 * (size_t)s];
            const float tt = tri_test(p, r);
            int ii;
            std::memcpy(&ii, &p[3], 4);
            if (better(tt, ii, bt, bid)) {
                bt = tt;
                bid = ii;
            }
        }
        if (id != bid || (id >= 0 && tb != bt)) ++bad;
        if (id >= 0) ++hits;
    }
    std::vector<char> seen((size_t)nt, 0);
    for (int s = 0; s < B.slotCount && nt > 0; ++s) {
        int ii;
        std::memcpy(&ii, &B.tris[12 * (size_t)s + 3], 4);
        if (ii >= 0 && ii < nt) seen[(size_t)ii] = 1;
    }
    const long long missing = std::count(seen.begin(), seen.end(), (char)0);
    std::printf("nodes %d slots %d depth4 %d rays %d hits %d mismatches %d missing %lld sentinel %lld\n", B.nodeCount,
                B.slotCount, B.depth4, NR, hits, bad, missing, sentinelVisits);
    return 0;
}

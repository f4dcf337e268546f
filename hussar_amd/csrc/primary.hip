// Device build of the primary visibility lists (primary.h): the depth-0 closest hits of a batch
// all start at the transmitter's position, so instead of a BVH traversal per ray they scan the
// distance-ordered triangle list of their direction cell and stop at the first triangle that
// lies beyond their best hit (k_wf_trav0, wavefront.hip).
//
// Build, all on the device, once per transmitter position (about a dozen launches):
//   k_pv_prep    per triangle: its distance key (a lower bound of its distance from o) and its
//                footprint on each cube face; per workgroup the faces' footprint boxes and the
//                count of near triangles
//   sort         triangles by key (rocPRIM radix sort, stable: ties keep the triangle order)
//   k_pv_faces   the face grids from the boxes
//   k_pv_count   per chunk of consecutive sorted triangles: how many of them cover each cell
//                (LDS counters), one row of the chunk x cell count matrix
//   k_pv_scan    per cell: the chunks' offsets within the cell's list (exclusive scan over the
//                chunks) and the cell's total
//   k_pv_start   one workgroup: the cells' first entries (exclusive scan over the cells)
//   k_pv_fill    per chunk: its triangles written into the lists of the cells they cover in key
//                order, 16 at a time (one per wave), each placed after the earlier ones of its
//                group that cover the same cell
// Each cell's list is then in the sorted triangle order: chunks in order, and inside a chunk
// the triangles in order. With more entries than the list holds, the build marks the lists
// unusable (meta[1] = 1) and the depth-0 kernel traverses the BVH instead.
#include <algorithm>
#include <climits>

#include <rocprim/device/device_radix_sort.hpp>

#include "kernels.h"
#include "primary.h"

namespace hsr {

namespace {

constexpr int kPrepBlock = 256;
constexpr int kChunkBlock = 1024;  // threads of the count and fill workgroups: 16 waves
constexpr int kGroup = kChunkBlock / 64;

struct TriD {
    pv::V3 a, b, c;
};

// Triangle of slot s relative to o, in double: v0 - o, then the edges added exactly as the
// float triangle test's v0 + u e1 + v e2 describes it.
__device__ inline TriD load_tri(const float4 *__restrict__ tris, int32_t s, const float o[3]) {
    const float4 p = tris[3 * (int64_t)s + 0], e1 = tris[3 * (int64_t)s + 1], e2 = tris[3 * (int64_t)s + 2];
    TriD t;
    t.a = {(double)p.x - (double)o[0], (double)p.y - (double)o[1], (double)p.z - (double)o[2]};
    t.b = pv::add(t.a, pv::V3{(double)e1.x, (double)e1.y, (double)e1.z});
    t.c = pv::add(t.a, pv::V3{(double)e2.x, (double)e2.y, (double)e2.z});
    return t;
}

__device__ inline pv::FaceGrid grid_of(const float4 *faces, int f) {
    const float4 g = faces[f];
    return pv::FaceGrid{g.x, g.y, g.z, g.w};
}

// per triangle: key and slot; per workgroup: footprint boxes of the six faces (umin, -umax,
// vmin, -vmax, as minima) and the near count, in part[blockIdx.x * 25 ...]
__global__ __launch_bounds__(kPrepBlock) void k_pv_prep(const float4 *__restrict__ tris,
                                                        const int32_t *__restrict__ triSlot, int64_t T, float ox,
                                                        float oy, float oz, uint32_t *keys, int32_t *slots,
                                                        float *part) {
    __shared__ float box[pv::kFaces * 4][kPrepBlock / 64];
    __shared__ int nearCnt[kPrepBlock / 64];
    const float o[3] = {ox, oy, oz};
    const int64_t t = (int64_t)blockIdx.x * kPrepBlock + threadIdx.x;
    float mine[pv::kFaces * 4];
    for (int i = 0; i < pv::kFaces * 4; ++i) mine[i] = __builtin_inff();
    int near = 0;
    if (t < T) {
        const int32_t s = triSlot[t];
        const TriD q = load_tri(tris, s, o);
        const double r = pv::point_triangle_distance(q.a, q.b, q.c);
        keys[t] = __float_as_uint(pv::distance_key(r));
        slots[t] = s;
        if (r < pv::kNear) {
            near = 1;
        } else {
            for (int f = 0; f < pv::kFaces; ++f) {
                double bx[4];
                if (pv::face_footprint(q.a, q.b, q.c, f, bx)) {
                    mine[4 * f + 0] = pv::round_down(bx[0]);
                    mine[4 * f + 1] = -pv::round_up(bx[1]);
                    mine[4 * f + 2] = pv::round_down(bx[2]);
                    mine[4 * f + 3] = -pv::round_up(bx[3]);
                }
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = 0; i < pv::kFaces * 4; ++i) {
        float v = mine[i];
        for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
        if (lane == 0) box[i][wave] = v;
    }
    for (int off = 32; off >= 1; off >>= 1) near += __shfl_xor(near, off, 64);
    if (lane == 0) nearCnt[wave] = near;
    __syncthreads();
    if (threadIdx.x < pv::kFaces * 4) {
        float v = __builtin_inff();
        for (int w = 0; w < kPrepBlock / 64; ++w) v = fminf(v, box[threadIdx.x][w]);
        part[(int64_t)blockIdx.x * 25 + threadIdx.x] = v;
    } else if (threadIdx.x == pv::kFaces * 4) {
        int c = 0;
        for (int w = 0; w < kPrepBlock / 64; ++w) c += nearCnt[w];
        part[(int64_t)blockIdx.x * 25 + 24] = __int_as_float(c);
    }
}

// face grids from the workgroups' boxes; meta[0] = near triangles, meta[1] = 0 (lists valid)
__global__ __launch_bounds__(256) void k_pv_faces(const float *__restrict__ part, int blocks, int R, float4 *faces,
                                                  int32_t *meta) {
    __shared__ float red[25][256 / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = 0; i < 25; ++i) {
        float v = i < 24 ? __builtin_inff() : 0.0f;
        int c = 0;
        for (int b = threadIdx.x; b < blocks; b += 256) {
            const float x = part[(int64_t)b * 25 + i];
            if (i < 24) v = fminf(v, x);
            else c += __float_as_int(x);
        }
        for (int off = 32; off >= 1; off >>= 1) {
            v = fminf(v, __shfl_xor(v, off, 64));
            c += __shfl_xor(c, off, 64);
        }
        if (lane == 0) red[i][wave] = i < 24 ? v : __int_as_float(c);
    }
    __syncthreads();
    if (threadIdx.x < pv::kFaces) {
        const int f = threadIdx.x;
        float m[4];
        for (int i = 0; i < 4; ++i) {
            float v = __builtin_inff();
            for (int w = 0; w < 4; ++w) v = fminf(v, red[4 * f + i][w]);
            m[i] = v;
        }
        const pv::FaceGrid g = pv::face_grid(m[0], -m[1], m[2], -m[3], R);
        faces[f] = make_float4(g.u0, g.v0, g.su, g.sv);
    }
    if (threadIdx.x == 0) {
        int c = 0;
        for (int w = 0; w < 4; ++w) c += __float_as_int(red[24][w]);
        meta[0] = c;
        meta[1] = 0;
        meta[2] = 0;
    }
}

// sorted triangles [first, last) of chunk `chunk`: the near triangles lead the sorted order and
// are left out; the others are cut into gridDim.x chunks of equal size
__device__ inline void chunk_range(const int32_t *meta, int64_t T, int chunks, int chunk, int64_t &first,
                                   int64_t &last) {
    const int64_t near = meta[0];
    const int64_t rest = T - near;
    const int64_t K = (rest + chunks - 1) / chunks;
    first = near + (int64_t)chunk * K;
    last = first + K < T ? first + K : T;
    if (first > T) first = T;
}

// cell rectangles of a triangle on the six faces (r[f][0] > r[f][1] when absent)
__device__ inline void tri_rects(const TriD &q, const float4 *faces, int R, int r[pv::kFaces][4]) {
    for (int f = 0; f < pv::kFaces; ++f) {
        double bx[4];
        r[f][0] = 1;
        r[f][1] = 0;
        if (pv::face_footprint(q.a, q.b, q.c, f, bx)) {
            int rr[4];
            if (pv::cell_rect(bx, grid_of(faces, f), R, rr)) {
                r[f][0] = rr[0];
                r[f][1] = rr[1];
                r[f][2] = rr[2];
                r[f][3] = rr[3];
            }
        }
    }
}

__global__ __launch_bounds__(kChunkBlock) void k_pv_count(const float4 *__restrict__ tris,
                                                          const int32_t *__restrict__ slots, int64_t T, float ox,
                                                          float oy, float oz, const float4 *__restrict__ faces,
                                                          const int32_t *__restrict__ meta, int R,
                                                          int32_t *__restrict__ matrix) {
    extern __shared__ int32_t cnt[];
    const int cells = pv::kFaces * R * R;
    for (int c = threadIdx.x; c < cells; c += kChunkBlock) cnt[c] = 0;
    __syncthreads();
    int64_t first, last;
    chunk_range(meta, T, gridDim.x, blockIdx.x, first, last);
    const float o[3] = {ox, oy, oz};
    for (int64_t i = first + threadIdx.x; i < last; i += kChunkBlock) {
        const TriD q = load_tri(tris, slots[i], o);
        int r[pv::kFaces][4];
        tri_rects(q, faces, R, r);
        for (int f = 0; f < pv::kFaces; ++f)
            for (int x = r[f][0]; x <= r[f][1]; ++x)
                for (int y = r[f][2]; y <= r[f][3]; ++y) atomicAdd(&cnt[(f * R + x) * R + y], 1);
    }
    __syncthreads();
    int32_t *row = matrix + (int64_t)blockIdx.x * cells;
    for (int c = threadIdx.x; c < cells; c += kChunkBlock) row[c] = cnt[c];
}

// per cell: exclusive prefix of the chunks' counts (in place), the cell's total into start[c]
__global__ __launch_bounds__(256) void k_pv_scan(int32_t *__restrict__ matrix, int chunks, int cells,
                                                 int32_t *__restrict__ start) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= cells) return;
    int32_t s = 0;
    int k = 0;
    for (; k + 8 <= chunks; k += 8) {
        int32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = matrix[(int64_t)(k + j) * cells + c];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            matrix[(int64_t)(k + j) * cells + c] = s;
            s += v[j];
        }
    }
    for (; k < chunks; ++k) {
        const int32_t v = matrix[(int64_t)k * cells + c];
        matrix[(int64_t)k * cells + c] = s;
        s += v;
    }
    start[c] = s;
}

// exclusive scan of the cells' totals into their first entries, start[cells] = all entries;
// meta[2] = the entries, meta[1] = 1 when they exceed the capacity
__global__ __launch_bounds__(1024) void k_pv_start(int32_t *__restrict__ start, int cells, int64_t cap,
                                                   int32_t *meta) {
    __shared__ int64_t sums[1024];
    const int per = (cells + 1023) / 1024;
    const int c0 = threadIdx.x * per, c1 = min(cells, c0 + per);
    int64_t s = 0;
    for (int c = c0; c < c1; ++c) s += start[c];
    sums[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = threadIdx.x >= off ? sums[threadIdx.x - off] : 0;
        __syncthreads();
        sums[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t run = threadIdx.x > 0 ? sums[threadIdx.x - 1] : 0;
    const bool over = sums[1023] > cap;
    if (!over)
        for (int c = c0; c < c1; ++c) {
            const int32_t v = start[c];
            start[c] = (int32_t)run;
            run += v;
        }
    if (threadIdx.x == 1023) {
        if (!over) start[cells] = (int32_t)sums[1023];
        meta[1] = over ? 1 : 0;
        meta[2] = (int32_t)(sums[1023] < INT_MAX ? sums[1023] : INT_MAX);
    }
}

// the lists: chunk blockIdx.x's triangles in sorted order, kGroup at a time (one per wave); a
// triangle's entry in a cell goes after those of the earlier triangles of its group that
// cover the cell, and the group's last triangle in a cell advances the cell's position
__global__ __launch_bounds__(kChunkBlock) void k_pv_fill(const float4 *__restrict__ tris,
                                                         const int32_t *__restrict__ slots,
                                                         const uint32_t *__restrict__ keys, int64_t T, float ox,
                                                         float oy, float oz, const float4 *__restrict__ faces,
                                                         const int32_t *__restrict__ meta, int R,
                                                         const int32_t *__restrict__ matrix,
                                                         const int32_t *__restrict__ start, int64_t cap,
                                                         int2 *__restrict__ list) {
    extern __shared__ int32_t pos[];
    __shared__ int rects[kGroup][pv::kFaces][4];
    if (meta[1] != 0) return;  // over capacity: the lists are not used
    const int cells = pv::kFaces * R * R;
    const int32_t *row = matrix + (int64_t)blockIdx.x * cells;
    for (int c = threadIdx.x; c < cells; c += kChunkBlock) pos[c] = start[c] + row[c];
    int64_t first, last;
    chunk_range(meta, T, gridDim.x, blockIdx.x, first, last);
    const float o[3] = {ox, oy, oz};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t g0 = first; g0 < last; g0 += kGroup) {
        const int64_t i = g0 + wave;
        const bool valid = i < last;
        int32_t slot = 0;
        uint32_t key = 0;
        if (valid) {
            slot = slots[i];
            key = keys[i];
        }
        // lane f < 6 finds the triangle's rectangle on face f
        if (lane < pv::kFaces) {
            int r0 = 1, r1 = 0, r2 = 0, r3 = 0;
            if (valid) {
                const TriD q = load_tri(tris, slot, o);
                double bx[4];
                if (pv::face_footprint(q.a, q.b, q.c, lane, bx)) {
                    int rr[4];
                    if (pv::cell_rect(bx, grid_of(faces, lane), R, rr)) {
                        r0 = rr[0];
                        r1 = rr[1];
                        r2 = rr[2];
                        r3 = rr[3];
                    }
                }
            }
            rects[wave][lane][0] = r0;
            rects[wave][lane][1] = r1;
            rects[wave][lane][2] = r2;
            rects[wave][lane][3] = r3;
        }
        __syncthreads();  // rectangles of the group; the positions of the previous group
        for (int f = 0; f < pv::kFaces; ++f) {
            const int x0 = rects[wave][f][0], x1 = rects[wave][f][1], y0 = rects[wave][f][2], y1 = rects[wave][f][3];
            if (x0 > x1) continue;
            const int h = y1 - y0 + 1, n = (x1 - x0 + 1) * h;
            for (int j = lane; j < n; j += 64) {
                const int x = x0 + j / h, y = y0 + j % h;
                int before = 0;
                for (int w = 0; w < wave; ++w)
                    before += (x >= rects[w][f][0] && x <= rects[w][f][1] && y >= rects[w][f][2] &&
                               y <= rects[w][f][3])
                                  ? 1
                                  : 0;
                const int c = (f * R + x) * R + y;
                const int64_t at = (int64_t)pos[c] + before;
                if (at < cap) list[at] = make_int2(slot, (int32_t)key);
            }
        }
        __syncthreads();  // every position of this group read
        for (int f = 0; f < pv::kFaces; ++f) {
            const int x0 = rects[wave][f][0], x1 = rects[wave][f][1], y0 = rects[wave][f][2], y1 = rects[wave][f][3];
            if (x0 > x1) continue;
            const int h = y1 - y0 + 1, n = (x1 - x0 + 1) * h;
            for (int j = lane; j < n; j += 64) {
                const int x = x0 + j / h, y = y0 + j % h;
                int before = 0;
                bool lastOne = true;
                for (int w = 0; w < kGroup; ++w) {
                    const bool in =
                        x >= rects[w][f][0] && x <= rects[w][f][1] && y >= rects[w][f][2] && y <= rects[w][f][3];
                    if (w < wave) before += in ? 1 : 0;
                    if (w > wave && in) lastOne = false;
                }
                if (lastOne) pos[(f * R + x) * R + y] += before + 1;
            }
        }
        __syncthreads();  // positions advanced, rectangles free
    }
}

}  // namespace

size_t pv_sort_bytes(int64_t T) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                    (const int32_t *)nullptr, (int32_t *)nullptr, (size_t)T, 0, 31);
    return bytes;
}

int pv_prep_blocks(int64_t T) { return (int)std::max<int64_t>(1, (T + kPrepBlock - 1) / kPrepBlock); }

hipError_t launch_pv_build(const PvBuild &p, hipStream_t s) {
    if (p.T <= 0) return hipSuccess;
    const int cells = pv::kFaces * p.R * p.R;
    const size_t lds = (size_t)cells * sizeof(int32_t);
    hipError_t e = hipFuncSetAttribute((const void *)k_pv_count, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_pv_fill, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    const int blocks = pv_prep_blocks(p.T);
    hipLaunchKernelGGL(k_pv_prep, dim3(blocks), dim3(kPrepBlock), 0, s, p.tris, p.triSlot, p.T, p.o[0], p.o[1],
                       p.o[2], p.keys, p.slots, p.part);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = p.sortBytes;
    e = rocprim::radix_sort_pairs(p.sortTemp, bytes, p.keys, p.keys2, p.slots, p.slots2, (size_t)p.T, 0, 31, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pv_faces, dim3(1), dim3(256), 0, s, p.part, blocks, p.R, p.faces, p.meta);
    hipLaunchKernelGGL(k_pv_count, dim3(p.chunks), dim3(kChunkBlock), lds, s, p.tris, p.slots2, p.T, p.o[0], p.o[1],
                       p.o[2], p.faces, p.meta, p.R, p.matrix);
    hipLaunchKernelGGL(k_pv_scan, dim3((cells + 255) / 256), dim3(256), 0, s, p.matrix, p.chunks, cells, p.start);
    hipLaunchKernelGGL(k_pv_start, dim3(1), dim3(1024), 0, s, p.start, cells, p.cap, p.meta);
    hipLaunchKernelGGL(k_pv_fill, dim3(p.chunks), dim3(kChunkBlock), lds, s, p.tris, p.slots2, p.keys2, p.T, p.o[0],
                       p.o[1], p.o[2], p.faces, p.meta, p.R, p.matrix, p.start, p.cap, p.list);
    return hipGetLastError();
}

}  // namespace hsr

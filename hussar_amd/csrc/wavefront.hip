// Wavefront pipeline of the radar integrator (PathTracer::sample, path.h:236-451, split into
// stages over a batch of paths whose state lives in HBM):
//
//   k_wf_gen      Halton sample, guiding warp, Tx sample; the first extension queue.
//   k_wf_nee0     when onlyIndirect is off: one thread per (path, receiver), the direct
//                 Tx -> Rx NEE set-up of depth 0
//   per depth k:
//   k_wf_shadow   any-hit BVH traversal of the depth-k shadow rays, a packet of up to four
//                 receivers of one path per lane (persistent, lane refill)
//   k_wf_gather   per path: guiding weight of the visible NEE terms, splat of the visible
//                 events into the workgroup's LDS frame tile, guiding record of paths that end
//   k_wf_trav     closest-hit traversal of the extension rays (persistent, lane refill)
//   k_wf_shade    hit update and reflection, the next extension queue and gather list, the
//                 guiding record of paths that miss or graze, and the NEE set-up of depth
//                 k+1 towards every receiver with its shadow packets
//
// Queue sizes stay on the device, so a batch is enqueued without host synchronisation.
// Device-scope atomics on one counter are serialised (about 12 ns apiece on MI355X), so they
// are kept rare: producers stage their queue entries in LDS and append a workgroup's range of
// items with one atomic per queue, and the persistent traversals claim their queue in chunks
// of kClaim items per atomic. The traversal kernels hold only the ray and stack, which keeps
// their register use low and their occupancy high; a lane whose ray is done takes the next
// one from its wave's claim while the others continue.
#include <algorithm>
#include <utility>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "primary.h"
#include "trace_common.h"

namespace hsr {

// Path state rows. The vertex of a bounce (normal, incoming direction and field, cosθ) is
// produced and consumed inside k_wf_shade and never stored.
enum : int { D_OX = 0, D_OY, D_OZ, D_DX, D_DY, D_DZ, D_TIME, D_R, D_RATE, D_N };
enum : int { F_H = 0, F_GWR = 6, F_GWI, F_X0, F_X1, F_PDF, F_FREQ, F_VX, F_VY, F_VZ, F_N };
enum : int { I_FLAGS = 0, I_LOCAL, I_N };
enum : int { S_SETUP = 1, S_TEST = 2, S_GW = 4, S_SPLAT = 8, S_GW1 = 16 };

constexpr int kWfBlock = 256;
// Threads per workgroup of the cube splats: their LDS tile (up to half a CU's LDS) limits them to
// two workgroups per CU, so each holds 16 waves to keep the CU's SIMDs busy (they need few
// registers).
constexpr int kCubeBlock = 1024;
#ifndef HSR_TRAV_STEPS
#define HSR_TRAV_STEPS 256
#endif
#ifndef HSR_REFILL_MIN
#define HSR_REFILL_MIN 32
#endif
// minimum resident waves per SIMD asked of the register allocator for the shading stages
#ifndef HSR_HIT_WAVES
#define HSR_HIT_WAVES 1
#endif
#ifndef HSR_GATHER_WAVES
#define HSR_GATHER_WAVES 1
#endif
#ifndef HSR_NEE_WAVES
#define HSR_NEE_WAVES 4
#endif
// A wave refills its idle lanes only after kTravSteps node visits (or leaves) per lane and when
// at least kRefillMin lanes are idle: in practice once all of its rays are done. Rays claimed
// together are neighbours in the queue and stay coherent; refilling early mixes them with
// others and costs more than the idle lanes (closest-hit traversal 10.4 ms per frame at these
// settings, 11.8 ms refilling at 16 idle lanes every 16 visits).
constexpr int kTravSteps = HSR_TRAV_STEPS;
constexpr int kRefillMin = HSR_REFILL_MIN;
constexpr int kPacket = 4;       // shadow rays (receivers of one path) traced together per lane
constexpr int kMaxGroups = (kMaxChannels + kPacket - 1) / kPacket;  // shadow packets per path
#ifndef HSR_CLAIM
#define HSR_CLAIM 64
#endif
#ifndef HSR_STAGE_ITERS
#define HSR_STAGE_ITERS 4
#endif
constexpr unsigned kClaim = HSR_CLAIM;          // queue items a traversal wave claims per atomic
constexpr int kStageIters = HSR_STAGE_ITERS;    // workgroup rounds per producer range
constexpr int kRange = kStageIters * kWfBlock;  // items of one producer range

// Sum of a per-lane count over the wave, added to a device counter by one lane.
__device__ inline void wave_count(unsigned long long *counter, uint32_t v, int lane) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) atomicAdd(counter, (unsigned long long)v);
}

// ------------------------------------------------------------------------ queues
// Producer side: a workgroup stages the entries it appends to a queue over one range of items
// in an LDS list (one LDS atomic per wave and append), then reserves the list's place in the
// device queue with one atomic and copies it out. Entries keep the order of their range.
__device__ inline void stage_append(int32_t *list, unsigned *sn, bool pred, int32_t value, int lane) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const int leader = __builtin_ctzll(m);
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(sn, (unsigned)__popcll(m));
    base = __shfl(base, leader, 64);
    if (pred) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}
// After a barrier that ends the range's appends: thread `who` reserves the list's place
// (sb = position, sc = count) and empties the list's count for the next range.
__device__ inline void stage_reserve(unsigned *sn, unsigned *sb, unsigned *sc, unsigned *counter, int who) {
    if ((int)threadIdx.x == who) {
        const unsigned c = *sn;
        *sc = c;
        *sb = c ? atomicAdd(counter, c) : 0u;
        *sn = 0;
    }
}
// After the barrier that follows stage_reserve: the workgroup copies the list out.
__device__ inline void stage_copy(const int32_t *list, const unsigned *sb, const unsigned *sc, int32_t *queue) {
    const unsigned c = *sc, b = *sb;
    for (unsigned j = threadIdx.x; j < c; j += kWfBlock) queue[b + j] = list[j];
}

// Consumer side: a queue of n items is cut into kQueueSegments contiguous segments, each with
// its own head counter. A persistent traversal wave claims positions [next, end) of the
// segment of its workgroup's XCD (workgroups are dispatched to the XCDs in turn), kClaim at a
// time with one atomic on that segment's head, and hands them to its idle lanes in order;
// when the segment is used up it takes from the following segments. The XCDs then work on
// distinct parts of the coherence-ordered queue, and the heads' atomics spread over
// kQueueSegments lines.
struct Claim {
    unsigned next, end;
    int seg, tried;
    bool dry;  // every segment is used up: nothing is left to claim
};

__device__ inline Claim claim_begin() {
    return Claim{0u, 0u, (int)(blockIdx.x % kQueueSegments), 0, false};
}

__device__ inline unsigned seg_lo(unsigned n, int s) {
    return (unsigned)(((uint64_t)n * (uint64_t)s) / kQueueSegments);
}

// Positions for the lanes of `idle` (wave-uniform) from the wave's claim, claiming further
// chunks as needed; returns this lane's position, or ~0u. Every lane of the wave calls it.
__device__ inline unsigned claim_items(WfCounter *heads, unsigned n, uint64_t idle, int lane, Claim &c) {
    unsigned my = ~0u;
    while (idle) {
        while (c.next >= c.end && !c.dry) {
            const unsigned lo = seg_lo(n, c.seg), hi = seg_lo(n, c.seg + 1);
            // in the segments of other XCDs, a plain load first: a used-up segment then costs
            // no atomic (the waves that find nothing left would otherwise queue up on the
            // heads at the end of every launch)
            unsigned base = hi - lo;
            if (lane == 0 &&
                (c.tried == 0 ||
                 __hip_atomic_load(&heads[c.seg].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < hi - lo))
                base = atomicAdd(&heads[c.seg].v, kClaim);
            base = __shfl(base, 0, 64);
            if (base < hi - lo) {
                c.next = lo + base;
                c.end = min(lo + base + kClaim, hi);
            } else if (++c.tried == kQueueSegments) {
                c.dry = true;
            } else {
                c.seg = (c.seg + 1) % kQueueSegments;
            }
        }
        if (c.next >= c.end) break;
        const unsigned take = min(c.end - c.next, (unsigned)__popcll(idle));
        const unsigned rank = (unsigned)__popcll(idle & ((1ull << lane) - 1ull));
        const bool mine = ((idle >> lane) & 1ull) && rank < take;
        if (mine) my = c.next + rank;
        idle &= ~__ballot(mine);
        c.next += take;
    }
    return my;
}

// A new path's state. What every path has at depth 0 is not stored: the origin (the Tx),
// time, length, guiding weight and motion terms are zero or the scene's, and the depth of a
// path is that of the queue it is in (the stages take it as their depth argument); the
// loaders below supply them at depth 0.
__device__ inline void store_state(const WfArgs &w, int64_t i, const PathState &p) {
    double *d = w.sd;
    float *f = w.sf;
    const int64_t P = w.P;
    d[D_DX * P + i] = p.d.x; d[D_DY * P + i] = p.d.y; d[D_DZ * P + i] = p.d.z;
    f[(F_H + 0) * P + i] = p.H.x.re; f[(F_H + 1) * P + i] = p.H.x.im;
    f[(F_H + 2) * P + i] = p.H.y.re; f[(F_H + 3) * P + i] = p.H.y.im;
    f[(F_H + 4) * P + i] = p.H.z.re; f[(F_H + 5) * P + i] = p.H.z.im;
    f[F_X0 * P + i] = p.x0;
    f[F_X1 * P + i] = p.x1;
    f[F_PDF * P + i] = p.pdf;
    f[F_FREQ * P + i] = p.freq;
    w.si[I_LOCAL * P + i] = (int32_t)(p.local - w.base);
}

// motion: the path-length rate and the current vertex's velocity (0 and the platform's at
// depth 0)
__device__ inline void load_motion(const TraceArgs &a, const WfArgs &w, int64_t i, PathState &p, int depth) {
    const int64_t P = w.P;
    if (depth == 0) {
        p.rate = 0.0;
        p.vprev = {a.ego[0], a.ego[1], a.ego[2]};
        return;
    }
    p.rate = w.sd[D_RATE * P + i];
    p.vprev = {w.sf[F_VX * P + i], w.sf[F_VY * P + i], w.sf[F_VZ * P + i]};
}
__device__ inline void store_motion(const WfArgs &w, int64_t i, const PathState &p) {
    const int64_t P = w.P;
    w.sd[D_RATE * P + i] = p.rate;
    w.sf[F_VX * P + i] = p.vprev.x;
    w.sf[F_VY * P + i] = p.vprev.y;
    w.sf[F_VZ * P + i] = p.vprev.z;
}

// The fields a stage needs, loaded separately to keep register use and traffic down.
__device__ inline void load_chain(const TraceArgs &a, const WfArgs &w, int64_t i, PathState &p, int depth) {
    const double *d = w.sd;
    const int64_t P = w.P;
    if (depth == 0) {
        p.o = posd(a.tx);
        p.time = 0.0;
        p.r = 0.0;
    } else {
        p.o = {d[D_OX * P + i], d[D_OY * P + i], d[D_OZ * P + i]};
        p.time = d[D_TIME * P + i];
        p.r = d[D_R * P + i];
    }
    p.d = {d[D_DX * P + i], d[D_DY * P + i], d[D_DZ * P + i]};
    p.depth = depth;
    p.local = w.base + w.si[I_LOCAL * P + i];
}
__device__ inline void load_field(const WfArgs &w, int64_t i, PathState &p) {
    const float *f = w.sf;
    const int64_t P = w.P;
    p.H = {{f[(F_H + 0) * P + i], f[(F_H + 1) * P + i]},
           {f[(F_H + 2) * P + i], f[(F_H + 3) * P + i]},
           {f[(F_H + 4) * P + i], f[(F_H + 5) * P + i]}};
}
// gwZero: the guiding weight is still zero (never stored: depth 0 before any NEE)
__device__ inline void load_record(const WfArgs &w, int64_t i, PathState &p, bool gwZero) {
    const float *f = w.sf;
    const int64_t P = w.P;
    p.gw = gwZero ? cf{0.0f, 0.0f} : cf{f[F_GWR * P + i], f[F_GWI * P + i]};
    p.x0 = f[F_X0 * P + i];
    p.x1 = f[F_X1 * P + i];
    p.pdf = f[F_PDF * P + i];
}
// after a hit: the chain, the field and the depth for the next bounce
__device__ inline void store_hit(const WfArgs &w, int64_t i, const PathState &p) {
    double *d = w.sd;
    float *f = w.sf;
    const int64_t P = w.P;
    d[D_OX * P + i] = p.o.x; d[D_OY * P + i] = p.o.y; d[D_OZ * P + i] = p.o.z;
    d[D_DX * P + i] = p.d.x; d[D_DY * P + i] = p.d.y; d[D_DZ * P + i] = p.d.z;
    d[D_TIME * P + i] = p.time;
    d[D_R * P + i] = p.r;
    f[(F_H + 0) * P + i] = p.H.x.re; f[(F_H + 1) * P + i] = p.H.x.im;
    f[(F_H + 2) * P + i] = p.H.y.re; f[(F_H + 3) * P + i] = p.H.y.im;
    f[(F_H + 4) * P + i] = p.H.z.re; f[(F_H + 5) * P + i] = p.H.z.im;
}

// Guiding record of a finished path (path.h:448-450): its training statistic and leaf at the
// path's own slot (statV / statL; statV is cleared when a training batch starts). After the
// batch, k_stat_reduce sums them per leaf in a fixed order, so the statistics do not depend on
// the order in which the paths of a batch finish.
__device__ inline void finish_record(const TraceArgs &a, const WfArgs &w, bool finishing, int64_t i,
                                     const PathState &p) {
    if (!(a.doGuiding && a.guideOn && !a.isFinal)) return;  // uniform
    if (finishing && p.pdf > 0.0f) {
        const float density = cabsf_(p.gw) * p.pdf;
        const float weight = 1.0f / p.pdf;
        const double dd = (double)density;
        w.statV[i] = (a.secondMoment ? dd * dd : dd) * (double)weight;
        w.statL[i] = tree_leaf(a.gtChild, p.x0, p.x1);
    }
}

// A live path's vertex rounded to float (the origin of its shadow rays and of its extension
// ray), the extension direction of a path that goes on, and the flag of a path that ends
// after its NEE. kRef: the vertex and direction of the FP32 chain, its length in ray_o.w.
template <bool kRef = false>
__device__ inline void write_extend(const WfArgs &w, bool alive, bool ends, int64_t i, const PathState &p) {
    if (alive) {
        const f3 o = kRef ? p.c32.o : cast3<float>(p.o);
        w.ray_o[i] = make_float4(o.x, o.y, o.z, kRef ? p.c32.r : 0.0f);
        w.si[I_FLAGS * w.P + i] = ends ? 1 : 0;
    }
    if (alive && !ends) {
        const f3 d = kRef ? p.c32.d : cast3<float>(p.d);
        w.ray_d[i] = make_float4(d.x, d.y, d.z, 0.0f);
    }
}

// ------------------------------------------------------------------------ coherence order
// The batch's paths are laid out in slot order by the cell of their primary sample (before
// the guiding warp, which maps sample-space neighbourhoods to direction-space neighbourhoods
// at the scale of the guiding tree): neighbouring lanes then trace nearly parallel rays from
// the same antenna, and the queues of later bounces, appended in range order, keep much of
// that coherence. No sort is needed: the primary sample is (RI_5(h), RI_3(h)) (or swapped)
// for Halton index h, and the first 3 base-5 digits of RI_5(h) are the last 3 base-5 digits
// of h reversed (likewise 5 base-3 digits), so the cell of [0,1)^2 cut into 125 x 243 is a
// function of h mod 125 and h mod 243, i.e. of the residue of h mod kCells = 30375 (Chinese
// remainder theorem). The cells are visited in Morton order of (u, v), each with the residue
// of its indices (kCells residues, made once per backend by coherence_cells); cell c holds
// the indices h = B0 + t_c + j * kCells of the batch [B0, B0 + n), t_c = (r_c - B0) mod
// kCells, and its slots follow those of the cells before it. The order is deterministic.
// With n = q kCells + rem, cell c holds q + (t_c < rem) paths.
__device__ inline int32_t cell_first(int32_t r, int32_t b0mod) {
    const int32_t t = r - b0mod;
    return t < 0 ? t + kCells : t;
}

// The batch's set-up in one launch: workgroup 0 writes cellStart[c] = the first slot of cell c
// (exclusive prefix of the cells' path counts) and cellStart[kCells] = n, the counts first to
// LDS with coalesced loads of the residues, then each thread summing and writing a contiguous
// run of cells; the other workgroups clear the batch's queue counters (nCounts depths), the
// first nPartials doubles of the partial frames and the first nStat training records.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_wf_setup(WfArgs w, int nCounts, double *partials,
                                                           int64_t nPartials, int64_t nStat) {
    __shared__ uint16_t cnt[kCells];  // n / kCells + 1 paths per cell at most: 553 for a batch of 2^24
    __shared__ int32_t part[kScanThreads];
    if (blockIdx.x > 0) {
        const int64_t t = (int64_t)(blockIdx.x - 1) * kScanThreads + threadIdx.x;
        const int64_t T = (int64_t)(gridDim.x - 1) * kScanThreads;
        if (t < nCounts) {
            WfCounts &c = w.cnt[t];
            c.n_ext.v = 0;
            c.n_shd.v = 0;
            c.n_gat.v = 0;
            for (int k = 0; k < kQueueSegments; ++k) {
                c.head_ext[k].v = 0;
                c.head_shd[k].v = 0;
            }
        }
        double2 *p2 = (double2 *)partials;  // nPartials is even (whole complex cells)
        for (int64_t i = t; i < nPartials / 2; i += T) p2[i] = make_double2(0.0, 0.0);
        for (int64_t i = t; i < nStat; i += T) w.statV[i] = 0.0;
        return;
    }
    constexpr int per = (kCells + kScanThreads - 1) / kScanThreads;
    const int32_t q = w.n / kCells, rem = w.n % kCells;
    for (int c = threadIdx.x; c < kCells; c += kScanThreads)
        cnt[c] = (uint16_t)(q + (cell_first(w.cells[c], w.b0mod) < rem ? 1 : 0));
    __syncthreads();
    const int c0 = threadIdx.x * per, c1 = min(kCells, c0 + per);
    int32_t sum = 0;
    for (int c = c0; c < c1; ++c) sum += cnt[c];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int32_t run = part[threadIdx.x] - sum;
    for (int c = c0; c < c1; ++c) {
        w.cellStart[c] = run;
        run += cnt[c];
    }
    if (threadIdx.x == kScanThreads - 1) w.cellStart[kCells] = run;
}

// ------------------------------------------------------------------------ generation
// Workgroup w generates the paths of cells [w * K, (w + 1) * K) of the coherence order into
// their slots (cellStart), in ranges of kRange paths whose depth-0 gather and extension entries it
// stages. A sampling tree of up to kGuideLds nodes is copied into LDS first: the guiding warp
// descends it with one dependent load per level.
constexpr int kGuideLds = 2048;
// Order of one cell's paths. The cell's Halton indices are h = H0 + j kCells (j < cnt), and the
// next base-5 digit of RI_5(h) (floor(h / 125) mod 5 = (A + 3 j) mod 5, kCells = 125 * 243) and
// the next base-3 digit of RI_3(h) (floor(h / 243) mod 3 = (B + 2 j) mod 3) cut the cell into
// 5 x 3 sub-cells that depend on j mod 15 alone. Slot p of the cell takes the j of the p-th
// path when the sub-cells are taken in a serpentine order (rows of the base-3 digit, the base-5
// digit alternately up and down) and each sub-cell's paths in index order, so that a wave's
// rays come from one or two neighbouring sub-cells.
#ifndef HSR_SUBCELLS
#define HSR_SUBCELLS 1
#endif
__device__ inline int32_t subcell_j(int32_t p, int32_t cnt, int64_t H0) {
#if HSR_SUBCELLS
    const int A = (int)((H0 / 125) % 5), Bv = (int)((H0 / 243) % 3);
    int32_t base = 0;
    for (int s = 0; s < 15; ++s) {
        const int dv = s / 5, du = (dv & 1) ? 4 - s % 5 : s % 5;
        // j mod 5 with (A + 3 j) mod 5 = du (3^-1 = 2 mod 5), j mod 3 with (B + 2 j) mod 3 = dv
        // (2^-1 = 2 mod 3), combined into j mod 15 (6 = 1 mod 5 and 0 mod 3, 10 = 0 mod 5 and 1 mod 3)
        const int r5 = (2 * (du - A + 5)) % 5, r3 = (2 * (dv - Bv + 3)) % 3;
        const int r = (6 * r5 + 10 * r3) % 15;
        const int32_t n = r < cnt ? (cnt - r + 14) / 15 : 0;
        if (p < base + n) return r + 15 * (p - base);
        base += n;
    }
#endif
    (void)cnt;
    (void)H0;
    return p;
}

constexpr int kGenCellsMax = kRange;  // cells per workgroup (K)
template <bool kRef>
__global__ __launch_bounds__(kWfBlock) void k_wf_gen(TraceArgs a, WfArgs w, int K) {
    __shared__ int32_t l_gat[kRange], l_ext[kRange];
    __shared__ unsigned sn[2], sb[2], sc[2];
    __shared__ int32_t g_child[kGuideLds];
    __shared__ float g_p0[kGuideLds];
    __shared__ int32_t cstart[kGenCellsMax + 1];  // first slots of this workgroup's cells
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < 2) sn[threadIdx.x] = 0;
    const bool guideLds = a.doGuiding && a.guideOn && a.gsSize <= kGuideLds;
    if (guideLds) {
        for (int j = threadIdx.x; j < a.gsSize; j += kWfBlock) {
            g_child[j] = a.gsChild[j];
            g_p0[j] = a.gsP0[j];
        }
    }
    const int32_t *gsChild = guideLds ? g_child : a.gsChild;
    const float *gsP0 = guideLds ? g_p0 : a.gsP0;
    const int64_t B0 = a.offset + w.base;
    const int c0 = blockIdx.x * K, nc = min(K, kCells - c0);
    const int32_t s0 = w.cellStart[c0];
    for (int c = threadIdx.x; c <= nc; c += kWfBlock) cstart[c] = w.cellStart[c0 + c] - s0;
    __syncthreads();
    const int32_t total = cstart[nc];
    WfCounts *cnt = &w.cnt[0];
    for (int32_t r0 = 0; r0 < total; r0 += kRange) {
        for (int r = 0; r < kStageIters; ++r) {
            const int32_t item = r0 + r * kWfBlock + threadIdx.x;
            const bool valid = item < total;
            const int64_t i = s0 + item;  // slot
            PathState p;
            if (valid) {
                // the cell of this item: the last c with cstart[c] <= item
                int lo = 0, hi = nc - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (cstart[mid] <= item) lo = mid;
                    else hi = mid - 1;
                }
                const int64_t t = cell_first(w.cells[c0 + lo], w.b0mod);
                const int32_t j = subcell_j(item - cstart[lo], cstart[lo + 1] - cstart[lo], B0 + t);
                const int64_t m = t + (int64_t)j * kCells;
                p.local = w.base + m;
                generate<kRef>(a, (uint64_t)(B0 + m), p, gsChild, gsP0);
            }
            const bool nee = valid && !a.onlyIndirect;
            stage_append(l_gat, &sn[0], nee, (int32_t)i, lane);
            const bool ends = valid && path_ends<kRef>(a, p);
            stage_append(l_ext, &sn[1], valid && !ends, (int32_t)i, lane);
            write_extend<kRef>(w, valid, ends, i, p);
            if (valid) store_state(w, i, p);
            finish_record(a, w, valid && ends && !nee, i, p);
        }
        __syncthreads();
        stage_reserve(&sn[0], &sb[0], &sc[0], &cnt->n_gat.v, 0);
        stage_reserve(&sn[1], &sb[1], &sc[1], &cnt->n_ext.v, 64);
        __syncthreads();
        stage_copy(l_gat, &sb[0], &sc[0], w.q_gat);
        stage_copy(l_ext, &sb[1], &sc[1], w.q_ext0);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------ traversal
// Closest-hit traversal of the extension queue. Persistent: each wave keeps its lanes on
// rays, refilling idle lanes from its claim on the queue.
#ifndef HSR_TRAV_OCC
#define HSR_TRAV_OCC 8
#endif
// resident waves per SIMD the register allocator must allow (workgroups of 4 waves per CU)
template <bool kCount>
__global__ __launch_bounds__(kWfBlock, HSR_TRAV_OCC) void k_wf_trav(TraceArgs a, WfArgs w, int depth) {
    __shared__ int32_t lds_stack[kStackShort * kWfBlock];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    int32_t spill[kStackSpill];
    Stack st;
    st.spill = spill;
    st.set_column(lds_stack + tid);
    st.sp = 0;
    WfCounts *cnt = &w.cnt[depth];
    const unsigned int n = cnt->n_ext.v;
    WfCounter *heads = cnt->head_ext;
    const int32_t *queue = (depth & 1) ? w.q_ext1 : w.q_ext0;
    uint32_t cn = 0, ct = 0, cr = 0;
    Trav t;
    int32_t item = -1;
    bool has = false;
    Claim cl = claim_begin();
    while (true) {
        const uint64_t idle = __ballot(!has);
        if (!cl.dry && __popcll(idle) >= kRefillMin) {
            const unsigned int my = claim_items(heads, n, idle, lane, cl);
            if (my != ~0u) {
                item = queue[my];
                const float4 o = w.tray_o[item];
                const float4 d = w.tray_d[item];
                trav_begin(t, {o.x, o.y, o.z}, {d.x, d.y, d.z}, __builtin_inff(), st);
                has = true;
                if (kCount) ++cr;
            }
        }
        if (!__any(has)) {
            if (cl.dry) break;
            continue;
        }
        if (has) {
            trav_run<false, kCount>(a.nodes, a.tris, t, st, kTravSteps, cn, ct);
            if (t.node == kDone) {
                w.hit[item] = make_int2(t.best_slot, __float_as_int(t.tfar));
                has = false;
            }
        }
    }
    if (kCount) {
        wave_count(&a.counters[0], cn, lane);
        wave_count(&a.counters[1], ct, lane);
        wave_count(&a.counters[2], cr, lane);
    }
}

// Closest hits of the depth-0 extension rays through the traced transmitter's primary
// visibility lists (primary.h, built by primary.hip): the triangles near the transmitter, then
// the list of the ray's direction cell, each in key order until a key exceeds the best hit so
// far. The triangle test and the tie rule are those of the BVH traversal, so the hits are the
// same. Lists over their capacity (pvMeta[1] != 0) fall back to the BVH traversal.
__device__ inline void pv_test(const float4 *__restrict__ tris, int32_t s, f3 o, f3 d, float &best, int32_t &bid,
                               int32_t &bslot) {
    const float4 p = tris[3 * (int64_t)s + 0];
    const float4 e1 = tris[3 * (int64_t)s + 1];
    const float4 e2 = tris[3 * (int64_t)s + 2];
    const float t = tri_intersect({p.x, p.y, p.z}, {e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, o, d, kEps,
                                  __builtin_inff());
    const int32_t id = __float_as_int(p.w);
    if (t < __builtin_inff() && (bid < 0 || t < best || (t == best && id < bid))) {
        best = t;
        bid = id;
        bslot = s;
    }
}

template <bool kCount>
__global__ __launch_bounds__(kWfBlock) void k_wf_trav0(TraceArgs a, WfArgs w, int depth) {
    __shared__ int32_t lds_stack[kStackShort * kWfBlock];
    const int lane = threadIdx.x & 63;
    const unsigned n = w.cnt[depth].n_ext.v;
    const int32_t *queue = (depth & 1) ? w.q_ext1 : w.q_ext0;
    const bool lists = a.pvMeta[1] == 0;
    const int near = a.pvMeta[0];
    int32_t spill[kStackSpill];
    Stack st;
    st.spill = spill;
    st.set_column(lds_stack + threadIdx.x);
    st.sp = 0;
    uint32_t cn = 0, ct = 0, cr = 0;
    for (unsigned i = blockIdx.x * kWfBlock + threadIdx.x; i < n; i += gridDim.x * kWfBlock) {
        const int32_t item = queue[i];
        const float4 o4 = w.tray_o[item];
        const float4 d4 = w.tray_d[item];
        const f3 o{o4.x, o4.y, o4.z}, d{d4.x, d4.y, d4.z};
        float best = __builtin_inff();
        int32_t bid = -1, bslot = -1;
        if (kCount) ++cr;
        if (lists) {
            for (int j = 0; j < near; ++j) {
                if (__uint_as_float(a.pvKeys[j]) > best) break;
                if (kCount) ++ct;
                pv_test(a.tris, a.pvSlots[j], o, d, best, bid, bslot);
            }
            float u, v;
            const int f = pv::ray_face(d.x, d.y, d.z, u, v);
            const float4 g = a.pvFaces[f];
            const int c = pv::ray_cell(f, u, v, pv::FaceGrid{g.x, g.y, g.z, g.w}, a.pvR);
            if (c >= 0) {
                const int32_t j1 = a.pvStart[c + 1];
                for (int32_t j = a.pvStart[c]; j < j1; ++j) {
                    const int2 e = a.pvList[j];
                    if (__int_as_float(e.y) > best) break;
                    if (kCount) ++ct;
                    pv_test(a.tris, e.x, o, d, best, bid, bslot);
                }
            }
        } else {
            bid = closest_hit<kCount>(a.nodes, a.tris, o, d, st, best, bslot, cn, ct);
        }
        w.hit[item] = make_int2(bslot, __float_as_int(best));
    }
    if (kCount) {
        wave_count(&a.counters[0], cn, lane);
        wave_count(&a.counters[1], ct, lane);
        wave_count(&a.counters[2], cr, lane);
    }
}

// Any-hit traversal of up to kPacket shadow rays that share their origin (one path's vertex,
// several receivers): a node is entered when any still-unoccluded ray hits one of its child
// boxes and every leaf triangle is tested against each such ray, so each ray's outcome is its
// own exact any-hit result while the node fetches are shared.
struct Packet {
    f3 o;
    f3 d[kPacket];
    float tfar[kPacket];
    f3 ivlo, ivhi;   // per axis the least and greatest reciprocal direction of the rays
    float tmax;      // the greatest extent
    uint32_t alive;  // rays not yet occluded
    int32_t node;
    bool same;       // all rays share their direction signs, given by sgn (bit k: axis k >= 0)
    uint32_t sgn;
};


__device__ inline bool slab_diff(float dlx, float dhx, float dly, float dhy, float dlz, float dhz, f3 invd,
                                 float tfar, float &tentry) {
    const float ax = dlx * invd.x, bx = dhx * invd.x;
    const float ay = dly * invd.y, by = dhy * invd.y;
    const float az = dlz * invd.z, bz = dhz * invd.z;
    const float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), kEps));
    float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    t1 = t1 * 1.0000004f;
    tentry = t0;
    return t0 <= t1 && t0 <= tfar;
}

template <bool kCount>
__device__ inline void packet_run(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, Packet &pk,
                                  Stack &st, int steps, uint32_t &cn, uint32_t &ct) {
    for (int s = 0; s < steps && pk.node != kDone; ++s) {
        if (pk.node >= 0) {
            if (kCount) ++cn;
            const Node4 q = load_node4f(nodes, pk.node);
            // the children's faces relative to the packet's shared origin, then per ray a slab
            // test with its own direction: a child is visited when any live ray hits it
            float dl[4][6];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float lx = i == 0 ? q.lx.x : (i == 1 ? q.lx.y : (i == 2 ? q.lx.z : q.lx.w));
                const float hx = i == 0 ? q.hx.x : (i == 1 ? q.hx.y : (i == 2 ? q.hx.z : q.hx.w));
                const float ly = i == 0 ? q.ly.x : (i == 1 ? q.ly.y : (i == 2 ? q.ly.z : q.ly.w));
                const float hy = i == 0 ? q.hy.x : (i == 1 ? q.hy.y : (i == 2 ? q.hy.z : q.hy.w));
                const float lz = i == 0 ? q.lz.x : (i == 1 ? q.lz.y : (i == 2 ? q.lz.z : q.lz.w));
                const float hz = i == 0 ? q.hz.x : (i == 1 ? q.hz.y : (i == 2 ? q.hz.z : q.hz.w));
                dl[i][0] = lx - pk.o.x;
                dl[i][1] = hx - pk.o.x;
                dl[i][2] = ly - pk.o.y;
                dl[i][3] = hy - pk.o.y;
                dl[i][4] = lz - pk.o.z;
                dl[i][5] = hz - pk.o.z;
            }
            uint32_t m = 0;
            if (pk.same) {
                // one interval test for the whole packet: the direction signs pick each axis'
                // near and far face, and a face's distance over the packet's reciprocal
                // directions [ivlo, ivhi] lies between its two products. The entry is bounded
                // below and the exit above by those of every ray (rounding is monotonic), so
                // a child any ray enters within its extent passes; the leaves test each ray.
                const bool sx = pk.sgn & 1u, sy = pk.sgn & 2u, sz = pk.sgn & 4u;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float nx = sx ? dl[i][0] : dl[i][1], fx = sx ? dl[i][1] : dl[i][0];
                    const float ny = sy ? dl[i][2] : dl[i][3], fy = sy ? dl[i][3] : dl[i][2];
                    const float nz = sz ? dl[i][4] : dl[i][5], fz = sz ? dl[i][5] : dl[i][4];
                    const float tnx = fminf(nx * pk.ivlo.x, nx * pk.ivhi.x);
                    const float tny = fminf(ny * pk.ivlo.y, ny * pk.ivhi.y);
                    const float tnz = fminf(nz * pk.ivlo.z, nz * pk.ivhi.z);
                    const float tfx = fmaxf(fx * pk.ivlo.x, fx * pk.ivhi.x);
                    const float tfy = fmaxf(fy * pk.ivlo.y, fy * pk.ivhi.y);
                    const float tfz = fmaxf(fz * pk.ivlo.z, fz * pk.ivhi.z);
                    const float t0 = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, kEps));
                    const float t1 = fminf(fminf(tfx, tfy), tfz) * 1.0000004f;
                    if (t0 <= t1 && t0 <= pk.tmax) m |= 1u << i;
                }
            } else {
#pragma unroll
                for (int r = 0; r < kPacket; ++r) {
                    if (pk.alive & (1u << r)) {
                        const f3 iv = {safe_rcp(pk.d[r].x), safe_rcp(pk.d[r].y), safe_rcp(pk.d[r].z)};
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            float e;
                            if (slab_diff(dl[i][0], dl[i][1], dl[i][2], dl[i][3], dl[i][4], dl[i][5], iv, pk.tfar[r],
                                          e))
                                m |= 1u << i;
                        }
                    }
                }
            }
            if (q.ch.x == kNoChild) m &= ~1u;
            if (q.ch.y == kNoChild) m &= ~2u;
            if (q.ch.z == kNoChild) m &= ~4u;
            if (q.ch.w == kNoChild) m &= ~8u;
            if (m) pk.node = any_children(q, m, st);
            else pk.node = st.sp ? st.pop() : kDone;
        } else {
            int first, count;
            leaf_decode(pk.node, first, count);
            for (int i = first; i < first + count; ++i) {
                if (kCount) ++ct;
                const float4 p = tris[3 * i + 0];
                const float4 e1 = tris[3 * i + 1];
                const float4 e2 = tris[3 * i + 2];
#pragma unroll
                for (int r = 0; r < kPacket; ++r) {
                    if ((pk.alive & (1u << r)) &&
                        tri_intersect({p.x, p.y, p.z}, {e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, pk.o, pk.d[r], kEps,
                                      pk.tfar[r]) < __builtin_inff())
                        pk.alive &= ~(1u << r);
                }
            }
            pk.node = (pk.alive && st.sp) ? st.pop() : kDone;
        }
    }
}

// A shadow packet from shadow queue entry e (path * 4 + group): the path's vertex as the shared
// origin, the group's receivers' directions and extents from their slots (slot0 = the first),
// the range of their reciprocal directions and whether they share their direction signs.
__device__ inline void packet_begin(const TraceArgs &a, const WfArgs &w, int32_t e, Packet &pk, int64_t &slot0,
                                    int &rays) {
    const int nrx = min(a.rxCount, a.C);
    const int64_t i = e >> 2;
    const int c0 = (e & 3) * kPacket;
    rays = min(kPacket, nrx - c0);
    slot0 = i * a.C + c0;
    pk.alive = 0;
    uint32_t sg[kPacket];
    const float4 o = w.ray_o[i];
    pk.o = {o.x, o.y, o.z};
    const float inf = __builtin_inff();
    pk.ivlo = {inf, inf, inf};
    pk.ivhi = {-inf, -inf, -inf};
    pk.tmax = 0.0f;
#pragma unroll
    for (int r = 0; r < kPacket; ++r) {
        f3 iv = {0.0f, 0.0f, 1.0f};
        if (r < rays) {
            const float4 d = w.s_d[slot0 + r];
            pk.d[r] = {d.x, d.y, d.z};
            iv = {safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)};
            pk.tfar[r] = d.w;
            pk.alive |= 1u << r;
            pk.ivlo = {fminf(pk.ivlo.x, iv.x), fminf(pk.ivlo.y, iv.y), fminf(pk.ivlo.z, iv.z)};
            pk.ivhi = {fmaxf(pk.ivhi.x, iv.x), fmaxf(pk.ivhi.y, iv.y), fmaxf(pk.ivhi.z, iv.z)};
            pk.tmax = fmaxf(pk.tmax, d.w);
        } else {
            pk.d[r] = {0.0f, 0.0f, 1.0f};
            pk.tfar[r] = 0.0f;
        }
        sg[r] = (iv.x >= 0.0f ? 1u : 0u) | (iv.y >= 0.0f ? 2u : 0u) | (iv.z >= 0.0f ? 4u : 0u);
    }
    pk.sgn = sg[0];
    pk.same = true;
#pragma unroll
    for (int r = 1; r < kPacket; ++r)
        if (r < rays && sg[r] != sg[0]) pk.same = false;
    pk.node = 0;
}

template <bool kCount>
__global__ __launch_bounds__(kWfBlock) void k_wf_shadow(TraceArgs a, WfArgs w, int depth) {
    __shared__ int32_t lds_stack[kStackShort * kWfBlock];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    int32_t spill[kStackSpill];
    Stack st;
    st.spill = spill;
    st.set_column(lds_stack + tid);
    st.sp = 0;
    WfCounts *cnt = &w.cnt[depth];
    const unsigned int n = cnt->n_shd.v;
    WfCounter *heads = cnt->head_shd;
    uint32_t cn = 0, ct = 0, cr = 0;
    Packet pk;
    int64_t slot0 = 0;
    int rays = 0;
    bool has = false;
    Claim cl = claim_begin();
    while (true) {
        const uint64_t idle = __ballot(!has);
        if (!cl.dry && __popcll(idle) >= kRefillMin) {
            const unsigned int my = claim_items(heads, n, idle, lane, cl);
            if (my != ~0u) {
                packet_begin(a, w, w.q_shd[my], pk, slot0, rays);
                st.sp = 0;
                has = true;
                if (kCount) cr += rays;
            }
        }
        if (!__any(has)) {
            if (cl.dry) break;
            continue;
        }
        if (has) {
            packet_run<kCount>(a.nodesF, a.tris, pk, st, kTravSteps, cn, ct);
            if (pk.node == kDone) {
                for (int r = 0; r < rays; ++r) w.s_vis[slot0 + r] = (pk.alive & (1u << r)) ? 2 : 1;
                has = false;
            }
        }
    }
    if (kCount) {
        wave_count(&a.counters[5], cn, lane);
        wave_count(&a.counters[6], ct, lane);
        wave_count(&a.counters[3], cr, lane);
    }
}

// ------------------------------------------------------------------------ hit
// One shadow slot of the NEE set-up ns in slot j; returns whether it needs a visibility query.
// The ray's origin is the path's vertex in ray_o (write_extend); s_g1 is written only when
// set (S_GW1) and s_vis by the shadow traversal of the slots it tests (S_TEST).
__device__ inline bool write_slot(const TraceArgs &a, const WfArgs &w, int64_t j, const NeeSetup &ns) {
    if (a.motion) w.s_dv[j] = ns.dv;
    w.s_d[j] = make_float4(ns.d.x, ns.d.y, ns.d.z, ns.tfar);
    w.s_tn[j] = ns.tn;
    w.s_v[j] = make_float2(ns.v.re, ns.v.im);
    const bool g1 = ns.gw1.re != 0.0f || ns.gw1.im != 0.0f;
    if (g1) w.s_g1[j] = make_float2(ns.gw1.re, ns.gw1.im);
    const bool test = !a.poDiffraction;
    const int st = S_SETUP | (test ? S_TEST : 0) | (ns.gwAdd ? S_GW : 0) | (ns.splat ? S_SPLAT : 0) | (g1 ? S_GW1 : 0);
    w.s_st[j] = (int8_t)st;
    return test;
}

// ------------------------------------------------------------------------ shade
// Per path of the depth-k extension queue: the hit update and reflection, the next extension
// queue and the gather list of depth k+1, the guiding record of paths that miss or graze, and
// the NEE set-up of the new vertex towards every receiver (its shadow slots and one shadow
// packet per four receivers). The vertex (normal, incoming direction and field, cosθ) stays
// in registers from the hit update to the NEE set-up.
template <bool kRef>
__global__ __launch_bounds__(kWfBlock, HSR_HIT_WAVES) void k_wf_shade(TraceArgs a, WfArgs w, int depth) {
    __shared__ int32_t l_gat[kRange], l_ext[kRange], l_shd[kRange * kMaxGroups];
    __shared__ unsigned sn[3], sb[3], sc[3];
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < 3) sn[threadIdx.x] = 0;
    __syncthreads();
    const unsigned int n = w.cnt[depth].n_ext.v;
    const int32_t *q = (depth & 1) ? w.q_ext1 : w.q_ext0;
    int32_t *qnext = (depth & 1) ? w.q_ext0 : w.q_ext1;
    WfCounts *next = &w.cnt[depth + 1];
    const int nrx = min(a.rxCount, a.C);
    const int groups = (nrx + kPacket - 1) / kPacket;
    const bool recording = a.doGuiding && a.guideOn && !a.isFinal;  // finish_record's condition
    for (int64_t r0 = (int64_t)blockIdx.x * kRange; r0 < n; r0 += (int64_t)gridDim.x * kRange) {
        for (int r = 0; r < kStageIters; ++r) {
            const int64_t k = r0 + r * kWfBlock + threadIdx.x;
            const bool valid = k < n;
            const int64_t i = valid ? q[k] : 0;
            PathState p;
            bool alive = false;
            if (valid) {
                // a missed path needs none of the chain and field (most rays of a batch's
                // first bounce miss): only its depth and index, and its record when training
                const int2 h = w.hit[i];
                p.depth = depth;
                p.local = w.base + w.si[I_LOCAL * w.P + i];
                if (a.recHits && p.local < a.recPaths)
                    a.recHits[p.local * a.maxDepth + p.depth] = h.x >= 0 ? __float_as_int(a.tris[3 * h.x].w) : -1;
                if (h.x >= 0) {
                    load_chain(a, w, i, p, depth);
                    load_field(w, i, p);
                    if (a.motion) load_motion(a, w, i, p, depth);
                    if (kRef) {  // the FP32 chain is the ray that was traced
                        const float4 o = w.tray_o[i], d = w.tray_d[i];
                        p.c32.o = {o.x, o.y, o.z};
                        p.c32.r = o.w;
                        p.c32.d = {d.x, d.y, d.z};
                    }
                    // without geometrical optics the hit update weighs the field by the surface
                    // Green's function, whose wave number needs the frequency
                    if (!a.geometricalOptics) p.freq = w.sf[F_FREQ * w.P + i];
                    alive = apply_hit<kRef>(a, p, h.x, __int_as_float(h.y));
                }
                if (!alive && recording) load_record(w, i, p, depth == 0 && a.onlyIndirect);
            }
            stage_append(l_gat, &sn[0], alive, (int32_t)i, lane);
            const bool ends = alive && path_ends<kRef>(a, p);
            stage_append(l_ext, &sn[1], alive && !ends, (int32_t)i, lane);
            write_extend<kRef>(w, alive, ends, i, p);
            if (alive) store_hit(w, i, p);
            if (alive && a.motion) store_motion(w, i, p);
            finish_record(a, w, valid && !alive, i, p);
            // NEE set-up at the new vertex (path.h:290-397 of the next bounce)
            if (alive) p.freq = w.sf[F_FREQ * w.P + i];
            for (int g = 0; g < groups; ++g) {
                bool lead = false;
                if (alive) {
                    const int c1 = min(nrx, (g + 1) * kPacket);
                    for (int c = g * kPacket; c < c1; ++c)
                        lead = write_slot(a, w, i * a.C + c, nee_setup<kRef>(a, p, c));
                }
                stage_append(l_shd, &sn[2], lead, (int32_t)(i * 4 + g), lane);
            }
        }
        __syncthreads();
        stage_reserve(&sn[0], &sb[0], &sc[0], &next->n_gat.v, 0);
        stage_reserve(&sn[1], &sb[1], &sc[1], &next->n_ext.v, 64);
        stage_reserve(&sn[2], &sb[2], &sc[2], &next->n_shd.v, 128);
        __syncthreads();
        stage_copy(l_gat, &sb[0], &sc[0], w.q_gat);
        stage_copy(l_ext, &sb[1], &sc[1], qnext);
        stage_copy(l_shd, &sb[2], &sc[2], w.q_shd);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------ NEE set-up, depth 0
// The direct paths Tx -> Rx (onlyIndirect off): items j < n_gat * nrx of the depth-0 gather
// list, path q_gat[j / nrx], receiver j % nrx.
template <bool kRef>
__global__ __launch_bounds__(kWfBlock, HSR_NEE_WAVES) void k_wf_nee0(TraceArgs a, WfArgs w) {
    __shared__ int32_t l_shd[kRange];
    __shared__ unsigned sn, sb, sc;
    const int lane = threadIdx.x & 63;
    if (threadIdx.x == 0) sn = 0;
    __syncthreads();
    const int nrx = min(a.rxCount, a.C);
    WfCounts *cnt = &w.cnt[0];
    const int64_t n = (int64_t)cnt->n_gat.v * nrx;
    for (int64_t r0 = (int64_t)blockIdx.x * kRange; r0 < n; r0 += (int64_t)gridDim.x * kRange) {
        for (int r = 0; r < kStageIters; ++r) {
            const int64_t k = r0 + r * kWfBlock + threadIdx.x;
            const bool valid = k < n;
            bool test = false;
            int64_t j = 0;
            if (valid) {
                const int64_t i = w.q_gat[k / nrx];
                const int ch = (int)(k % nrx);
                PathState p;
                load_chain(a, w, i, p, 0);
                p.freq = w.sf[F_FREQ * w.P + i];
                if (a.motion) load_motion(a, w, i, p, 0);
                if (kRef) p.c32.o = posf(a.tx);
                j = i * a.C + ch;
                test = write_slot(a, w, j, nee_setup<kRef>(a, p, ch));
            }
            // one shadow packet per four receivers of a path: entry = path * 4 + group
            const bool lead = test && (k % nrx) % kPacket == 0;
            stage_append(l_shd, &sn, lead, (int32_t)((j / a.C) * 4 + (int32_t)((k % nrx) / kPacket)), lane);
        }
        __syncthreads();
        stage_reserve(&sn, &sb, &sc, &cnt->n_shd.v, 0);
        __syncthreads();
        stage_copy(l_shd, &sb, &sc, w.q_shd);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------ wave event lists
// A wave's events waiting for the LDS tile splat. An LDS atomic instruction costs the same for
// one active lane as for 64, so the fused tail, whose waves have few lanes with an event at a
// time, lists its events per wave and splats them 64 at a time, one per lane (flush_events_g
// takes each lane's own channel row). The gather stage splats at once: nearly all of its lanes
// have an event per receiver, and a list there measured slower (more instructions per event).
constexpr int kWaveEvents = 128;  // entries per wave: 64 pending at most, plus one append of 64
// One listed event in 16 bytes (one LDS store and one load): the centre bin (bits 0-19), the
// frame channel (bits 20-27) and the delta flag (bit 28), the fractional offset, the amplitude.
struct alignas(16) WaveEvent {
    int32_t key;
    float s, are, aim;
};
constexpr size_t kWaveEventBytes = sizeof(WaveEvent) * kWaveEvents * (kWfBlock / 64);  // per workgroup

__device__ inline void splat_listed(const WaveEvent &t, double *tile, int N, int C, bool guard, int lane) {
    const Event e{t.key & 0xFFFFF, t.s, {t.are, t.aim}, (t.key >> 28) & 1};
    const int ch = (t.key >> 20) & 0xFF;
    if (guard)
        flush_events_g(true, e, ch, tile, N, C, lane);
    else
        flush_events(true, e, ch, tile, N, C, lane);
}

// Appends the lane's event (if has) of frame channel ch to the wave's list and splats 64 of
// them once 64 are pending; pending is wave-uniform. All lanes of the wave call it.
__device__ inline void list_event(bool has, const Event &ev, int ch, WaveEvent *list, int &pending, double *tile,
                                  int N, int C, bool guard, int lane) {
    const uint64_t m = __ballot(has);
    if (m == 0) return;
    if (has)
        list[pending + __popcll(m & ((1ull << lane) - 1ull))] =
            WaveEvent{ev.center | (ch << 20) | (ev.delta ? 1 << 28 : 0), ev.s, ev.A.re, ev.A.im};
    pending += __popcll(m);
    if (pending >= 64) {
        pending -= 64;
        splat_listed(list[pending + lane], tile, N, C, guard, lane);
    }
}
// The wave's remaining events (fewer than 64).
__device__ inline void list_flush(const WaveEvent *list, int pending, double *tile, int N, int C, bool guard,
                                  int lane) {
    if (lane < pending) splat_listed(list[lane], tile, N, C, guard, lane);
}

// ------------------------------------------------------------------------ shadow + splat
// The shadow stage of a batch that takes no guiding records (its gather would only splat and
// record the NEE outcomes), one-chirp frames in an LDS tile: as each packet's traversal ends,
// its receivers' visible contributions become events (the gather's make_event of the same
// slots) listed per wave and splatted 64 at a time into the workgroup's frame tile, which is
// added into its partial frame at the end, and the NEE records are written. Every NEE slot of
// such a batch is tested in exactly one packet (without PO diffraction), so no gather stage
// follows, and the splat work overlaps the traversal instead of waiting for its end.
// LDS: the traversal stack, then (dynamic) the frame tile and the waves' event lists.
size_t shadow_splat_lds_bytes(int N, int C) {
    return gather_lds_bytes(N, C) + kWaveEventBytes;
}
// the one-chirp instance's waves per SIMD: 3 when three workgroups fit a CU's 160 KiB of LDS
static int shadow_splat_waves(int N, int C) {
    return 3 * (shadow_splat_lds_bytes(N, C) + kStackShort * kWfBlock * sizeof(int32_t)) <= 163840 ? 3 : 2;
}

// kWaves: resident waves per SIMD the instance is held to. One-chirp frames: 3 (168
// registers, a few spills) when the frame tile lets 3 workgroups share a CU's LDS (256 bins x
// 4 channels: 14.5 ms per vehicle frame against 14.8 at the natural 213 registers and 2
// waves), else 2, which larger tiles cannot exceed anyway (shadow_splat_waves).
// kCube: the instance for cube frames (their events go to the event lists): no frame tile and
// no wave lists, held to 128 registers for 4 resident waves per SIMD (a few spill), which
// measured 555 ms per urban frame against 578 at its natural 143 registers (3 waves)
template <bool kCount, bool kCube, int kWaves>
__global__ __launch_bounds__(kWfBlock, kWaves) void k_wf_shadow_splat(TraceArgs a, WfArgs w, int depth, double2 *partials,
                                                             bool count) {
    __shared__ int32_t lds_stack[kStackShort * kWfBlock];
    extern __shared__ double tile[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    constexpr bool cube = kCube;  // events go to the cube's event lists, no tile
    const bool guard = a.N >= 2 * kGuard;
    const int tileDoubles = cube ? 0 : 2 * (guard ? guard_row(a.N) : a.N) * a.C;
    for (int j = tid; j < tileDoubles; j += kWfBlock) tile[j] = 0.0;
    WaveEvent *list = (WaveEvent *)(tile + tileDoubles) + (tid >> 6) * kWaveEvents;
    int pending = 0;
    __syncthreads();
    const int nrx = min(a.rxCount, a.C);
    int32_t spill[kStackSpill];
    Stack st;
    st.spill = spill;
    st.set_column(lds_stack + tid);
    st.sp = 0;
    WfCounts *cnt = &w.cnt[depth];
    const unsigned int n = cnt->n_shd.v;
    WfCounter *heads = cnt->head_shd;
    const int64_t P = w.P;
    uint32_t cn = 0, ct = 0, cr = 0, cev = 0;
    Packet pk;
    int64_t slot0 = 0;
    int rays = 0, c0 = 0;
    int64_t path = 0;
    bool has = false;
    Claim cl = claim_begin();
    while (true) {
        const uint64_t idle = __ballot(!has);
        if (!cl.dry && __popcll(idle) >= kRefillMin) {
            const unsigned int my = claim_items(heads, n, idle, lane, cl);
            if (my != ~0u) {
                const int32_t e = w.q_shd[my];
                packet_begin(a, w, e, pk, slot0, rays);
                path = e >> 2;
                c0 = (e & 3) * kPacket;
                st.sp = 0;
                has = true;
                if (kCount) cr += rays;
            }
        }
        if (!__any(has)) {
            if (cl.dry) break;
            continue;
        }
        bool done = false;
        if (has) {
            packet_run<kCount>(a.nodesF, a.tris, pk, st, kTravSteps, cn, ct);
            if (pk.node == kDone) {
                done = true;
                has = false;
            }
        }
        // the finished packets' receivers, one receiver per step (wave-uniform calls)
        if (__any(done)) {
            int64_t local = 0;
            bool rec = false;
            if (done && a.recNee && depth >= 1) {
                local = w.base + w.si[I_LOCAL * P + path];
                rec = local < a.recPaths;
            }
#pragma unroll
            for (int r = 0; r < kPacket; ++r) {
                bool ev_has = false;
                Event ev{0, 0.0f, {0.0f, 0.0f}, 0};
                if (done && r < rays) {
                    const int64_t j = slot0 + r;
                    const int stt = (int)w.s_st[j];
                    const bool vis = (pk.alive >> r) & 1u;
                    int8_t flag = vis ? 2 : 1;
                    if (vis && (stt & S_SPLAT)) {
                        flag = 3;
                        const float2 v2 = w.s_v[j];
                        ev = make_event(a, w.s_tn[j], cf{v2.x, v2.y}, a.sampleWeight, a.motion ? w.s_dv[j] : 0.0);
                        ev_has = true;
                        if (count) ++cev;
                    }
                    if (rec) a.recNee[(local * a.maxDepth + (depth - 1)) * a.C + a.chBase + c0 + r] = flag;
                }
                if (cube) {
                    // cube_event appends a wave's events of one channel at a time
                    for (int p0 = 0; p0 < nrx; p0 += kPacket)
                        cube_event(ev_has && c0 == p0, ev, a.chBase + p0 + r, a, lane);
                } else {
                    list_event(ev_has, ev, a.chBase + c0 + r, list, pending, tile, a.N, a.C, guard, lane);
                }
            }
        }
    }
    if (cube) {
        if (kCount) {
            wave_count(&a.counters[5], cn, lane);
            wave_count(&a.counters[6], ct, lane);
            wave_count(&a.counters[3], cr, lane);
        }
        if (count) wave_count(&a.counters[4], cev, lane);
        return;
    }
    list_flush(list, pending, tile, a.N, a.C, guard, lane);
    if (kCount) {
        wave_count(&a.counters[5], cn, lane);
        wave_count(&a.counters[6], ct, lane);
        wave_count(&a.counters[3], cr, lane);
    }
    if (count) wave_count(&a.counters[4], cev, lane);
    __syncthreads();
    if (guard)
        tile_flush_g(tile, partials + (size_t)blockIdx.x * a.N * a.C, a.N, a.C, tid, kWfBlock);
    else
        tile_flush(tile, partials + (size_t)blockIdx.x * a.N * a.C, a.N, a.C, tid, kWfBlock);
}

// ------------------------------------------------------------------------ gather + splat
__global__ __launch_bounds__(kWfBlock, HSR_GATHER_WAVES) void k_wf_gather(TraceArgs a, WfArgs w, int depth,
                                                                          double2 *partials, bool count) {
    // one chirp's frame tile in LDS: the guard-band layout (flush_events_g) for N >= 16, else
    // N*C complex doubles (tile_re); unused with M > 1
    extern __shared__ double tile[];
    const int NC = a.N * a.C;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool cube = a.cube != 0;  // uniform: splat into the frame cube in device memory
    const bool guard = a.N >= 2 * kGuard;
    const bool splat = !a.noSplat;  // uniform
    if (!cube && splat) {
        const int tileDoubles = 2 * (guard ? guard_row(a.N) : a.N) * a.C;
        for (int j = tid; j < tileDoubles; j += kWfBlock) tile[j] = 0.0;
        __syncthreads();
    }
    const int64_t wave = ((int64_t)blockIdx.x * kWfBlock + tid) >> 6;
    const int64_t waves = ((int64_t)gridDim.x * kWfBlock) >> 6;
    const unsigned int n = w.cnt[depth].n_gat.v;
    const int nrx = min(a.rxCount, a.C);
    const int64_t P = w.P;
    uint32_t cev = 0;
    for (int64_t b = wave * 64; b < n; b += waves * 64) {
        const int64_t k = b + lane;
        const bool valid = k < n;
        const int64_t i = valid ? w.q_gat[k] : 0;
        PathState p;
        if (valid) {
            // the guiding weight is in memory once a gather has stored it: from depth 1 on, or
            // depth 2 when the direct paths are not traced (no gather at depth 0)
            const bool gwZero = depth == 0 || (depth == 1 && a.onlyIndirect);
            p.gw = gwZero ? cf{0.0f, 0.0f} : cf{w.sf[F_GWR * P + i], w.sf[F_GWI * P + i]};
            p.x0 = w.sf[F_X0 * P + i];
            p.x1 = w.sf[F_X1 * P + i];
            p.pdf = w.sf[F_PDF * P + i];
            p.depth = depth;
            p.local = w.base + w.si[I_LOCAL * P + i];
        }
        const bool rec = valid && a.recNee && p.local < a.recPaths && p.depth >= 1;
        for (int ch = 0; ch < nrx; ++ch) {
            bool has = false;
            Event ev{0, 0.0f, {0.0f, 0.0f}, 0};
            const int64_t j = i * a.C + ch;
            const int stt = valid ? (int)w.s_st[j] : 0;
            if (stt & S_SETUP) {
                const bool vis = (stt & S_TEST) ? w.s_vis[j] == 2 : true;
                int8_t flag = 1;
                if (vis) {
                    flag = 2;
                    const float2 v2 = w.s_v[j];
                    const cf v = {v2.x, v2.y};
                    if (stt & S_GW1) {
                        const float2 g = w.s_g1[j];
                        p.gw = p.gw + cf{g.x, g.y};
                    }
                    if (stt & S_GW) p.gw = p.gw + v;
                    if (stt & S_SPLAT) {
                        flag = 3;
                        if (splat) {
                            ev = make_event(a, w.s_tn[j], v, a.sampleWeight, a.motion ? w.s_dv[j] : 0.0);
                            has = true;
                            if (count) ++cev;
                        }
                    }
                }
                if (rec) a.recNee[(p.local * a.maxDepth + (p.depth - 1)) * a.C + a.chBase + ch] = flag;
            }
            if (!splat)
                continue;
            if (cube)
                cube_event(has, ev, a.chBase + ch, a, lane);
            else if (guard)
                flush_events_g(has, ev, a.chBase + ch, tile, a.N, a.C, lane);
            else
                flush_events(has, ev, a.chBase + ch, tile, a.N, a.C, lane);
        }
        if (valid) {
            w.sf[F_GWR * P + i] = p.gw.re;
            w.sf[F_GWI * P + i] = p.gw.im;
        }
        const bool finishing = valid && (w.si[I_FLAGS * P + i] & 1);
        finish_record(a, w, finishing, i, p);
    }
    if (count) wave_count(&a.counters[4], cev, lane);
    if (cube || !splat) return;
    __syncthreads();
    if (guard)
        tile_flush_g(tile, partials + (size_t)blockIdx.x * NC, a.N, a.C, tid, kWfBlock);
    else
        tile_flush(tile, partials + (size_t)blockIdx.x * NC, a.N, a.C, tid, kWfBlock);
}

// ------------------------------------------------------------------------ fused tail
// The last bounces of a batch in one launch. Each lane takes a path of the depth-K extension
// queue (its ray traced, its hit in w.hit) and carries it to its end: the hit update, the NEE at
// the new vertex towards every receiver (set-up, any-hit shadow rays, guiding weight, splat,
// records), the next extension ray (closest hit), bounce after bounce, and the guiding record
// of the finished path. This replaces the stages from the depth-K shading on, few paths each
// and every one waiting on the longest ray of the one before, by one launch whose length is
// that of its longest path. Every value comes from the device functions the stages use (hit
// update, termination, NEE set-up, closest-hit traversal, the shadow stage's packets of four
// receivers, event), in the stages' order, so hits, visibility, records, guiding weights and
// events equal the stages'. The loop is wave-uniform per bounce; idle lanes take further paths
// from the queue through the consumer heads of depth K + 1, whose stages do not run. Events go
// to the wave's event list (list_event); the workgroup's frame tile is then added into its
// partial frame (blockIdx.x < the gather grid) like a gather stage's.

// Any-hit visibility of receivers [c0, c0 + rays) of the lane's vertex p: one packet of the
// shadow stage (the same rays, the same traversal), bit r of the result = receiver c0 + r is
// visible.
template <bool kRef, bool kCount>
__device__ inline uint32_t tail_visibility(const TraceArgs &a, const PathState &p, int c0, int rays, Stack &st,
                                           uint32_t &cn, uint32_t &ct) {
    Packet pk;
    pk.alive = 0;
    uint32_t sg[kPacket];
    const float inf = __builtin_inff();
    pk.ivlo = {inf, inf, inf};
    pk.ivhi = {-inf, -inf, -inf};
    pk.tmax = 0.0f;
#pragma unroll
    for (int r = 0; r < kPacket; ++r) {
        f3 iv = {0.0f, 0.0f, 1.0f};
        if (r < rays) {
            const NeeSetup ns = nee_setup<kRef>(a, p, c0 + r);
            pk.o = ns.o;  // the same vertex for every receiver
            pk.d[r] = ns.d;
            iv = {safe_rcp(ns.d.x), safe_rcp(ns.d.y), safe_rcp(ns.d.z)};
            pk.tfar[r] = ns.tfar;
            pk.alive |= 1u << r;
            pk.ivlo = {fminf(pk.ivlo.x, iv.x), fminf(pk.ivlo.y, iv.y), fminf(pk.ivlo.z, iv.z)};
            pk.ivhi = {fmaxf(pk.ivhi.x, iv.x), fmaxf(pk.ivhi.y, iv.y), fmaxf(pk.ivhi.z, iv.z)};
            pk.tmax = fmaxf(pk.tmax, ns.tfar);
        } else {
            pk.d[r] = {0.0f, 0.0f, 1.0f};
            pk.tfar[r] = 0.0f;
        }
        sg[r] = (iv.x >= 0.0f ? 1u : 0u) | (iv.y >= 0.0f ? 2u : 0u) | (iv.z >= 0.0f ? 4u : 0u);
    }
    pk.sgn = sg[0];
    pk.same = true;
#pragma unroll
    for (int r = 1; r < kPacket; ++r)
        if (r < rays && sg[r] != sg[0]) pk.same = false;
    pk.node = 0;
    st.sp = 0;
    while (pk.node != kDone) packet_run<kCount>(a.nodesF, a.tris, pk, st, kTravSteps, cn, ct);
    return pk.alive;
}

template <bool kRef, bool kCount>
__global__ __launch_bounds__(kWfBlock, 2) void k_wf_tail(TraceArgs a, WfArgs w, int K, double2 *partials) {
    // [frame tile (splat, no cube)] [kStackShort][kWfBlock] stack [4][kTailEvents] event lists
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const bool cube = a.cube != 0, splat = !a.noSplat;  // uniform
    const bool guard = a.N >= 2 * kGuard;
    const bool tiled = splat && !cube;
    const int tileDoubles = tiled ? 2 * (guard ? guard_row(a.N) : a.N) * a.C : 0;
    double *tile = smem;
    int32_t *lds_stack = (int32_t *)(smem + tileDoubles);
    WaveEvent *list = (WaveEvent *)(lds_stack + kStackShort * kWfBlock) + (tid >> 6) * kWaveEvents;
    for (int j = tid; j < tileDoubles; j += kWfBlock) tile[j] = 0.0;
    __syncthreads();
    int32_t spill[kStackSpill];
    Stack st;
    st.spill = spill;
    st.set_column(lds_stack + tid);
    st.sp = 0;
    uint32_t cn = 0, ct = 0, cev = 0;
    int pending = 0;  // events in the wave's list (wave-uniform)
    const unsigned int n = w.cnt[K].n_ext.v;
    const int32_t *queue = (K & 1) ? w.q_ext1 : w.q_ext0;
    WfCounter *heads = w.cnt[K + 1].head_ext;  // depth K + 1 has no stages of its own
    const int nrx = min(a.rxCount, a.C);
    const int64_t P = w.P;
    PathState p;
    int64_t i = 0;      // the lane's path slot
    int32_t hx = -1;    // slot of the triangle its current ray hit, -1 none
    float ht = 0.0f;    // the hit's distance (the traversal's)
    bool active = false;
    Claim cl = claim_begin();
    while (true) {
        const uint64_t idle = __ballot(!active);
        if (!cl.dry && __popcll(idle) >= kRefillMin) {
            const unsigned my = claim_items(heads, n, idle, lane, cl);
            if (my != ~0u) {
                i = queue[my];
                load_chain(a, w, i, p, K);
                load_field(w, i, p);
                if (a.motion) load_motion(a, w, i, p, K);
                load_record(w, i, p, false);  // K >= 1: the depth-K gather stored the guiding weight
                p.freq = w.sf[F_FREQ * P + i];
                if (kRef) {  // the FP32 chain is the ray that was traced
                    const float4 o = w.tray_o[i], d = w.tray_d[i];
                    p.c32.o = {o.x, o.y, o.z};
                    p.c32.r = o.w;
                    p.c32.d = {d.x, d.y, d.z};
                }
                const int2 h = w.hit[i];
                hx = h.x;
                ht = __int_as_float(h.y);
                active = true;
            }
        }
        if (!__any(active)) {
            if (cl.dry) break;
            continue;
        }
        // the hit of this bounce and the new vertex (the shading stage)
        bool vertex = false, ends = false;
        const bool rec = active && p.local < a.recPaths;
        if (active) {
            if (a.recHits && rec)
                a.recHits[p.local * a.maxDepth + p.depth] = hx >= 0 ? __float_as_int(a.tris[3 * hx].w) : -1;
            vertex = hx >= 0 && apply_hit<kRef>(a, p, hx, ht);
            ends = vertex && path_ends<kRef>(a, p);
        }
        // NEE at the new vertex towards every receiver (the shading stage's set-up, the shadow
        // stage's packets of four receivers, the gather stage)
        for (int c0 = 0; c0 < nrx; c0 += kPacket) {
            const int rays = min(kPacket, nrx - c0);
            uint32_t vis = (1u << rays) - 1u;
            if (vertex && !a.poDiffraction) vis = tail_visibility<kRef, kCount>(a, p, c0, rays, st, cn, ct);
            for (int r = 0; r < rays; ++r) {
                const int ch = c0 + r;
                bool has = false;
                Event ev{0, 0.0f, {0.0f, 0.0f}, 0};
                if (vertex) {
                    const NeeSetup ns = nee_setup<kRef>(a, p, ch);
                    int8_t flag = 1;
                    if (vis & (1u << r)) {
                        flag = 2;
                        if (ns.gw1.re != 0.0f || ns.gw1.im != 0.0f) p.gw = p.gw + ns.gw1;
                        if (ns.gwAdd) p.gw = p.gw + ns.v;
                        if (ns.splat) {
                            flag = 3;
                            if (splat) {
                                ev = make_event(a, ns.tn, ns.v, a.sampleWeight, ns.dv);
                                has = true;
                                if (kCount) ++cev;
                            }
                        }
                    }
                    if (a.recNee && rec)
                        a.recNee[(p.local * a.maxDepth + (p.depth - 1)) * a.C + a.chBase + ch] = flag;
                }
                if (!splat) continue;
                if (cube) {
                    cube_event(has, ev, a.chBase + ch, a, lane);
                    continue;
                }
                list_event(has, ev, a.chBase + ch, list, pending, tile, a.N, a.C, guard, lane);
            }
        }
        // the path's end (its guiding record) or its next ray
        if (active) {
            if (!vertex || ends) {
                finish_record(a, w, true, i, p);
                active = false;
            } else {
                const f3 ro = kRef ? p.c32.o : cast3<float>(p.o);
                const f3 rd = kRef ? p.c32.d : cast3<float>(p.d);
                float t;
                int32_t slot;
                const int32_t id = closest_hit<kCount>(a.nodes, a.tris, ro, rd, st, t, slot, cn, ct);
                hx = id >= 0 ? slot : -1;
                ht = t;
            }
        }
    }
    if (kCount) {
        wave_count(&a.counters[7], cn, lane);
        wave_count(&a.counters[8], ct, lane);
        wave_count(&a.counters[4], cev, lane);
    }
    if (!tiled) return;
    list_flush(list, pending, tile, a.N, a.C, guard, lane);
    __syncthreads();
    const int NC = a.N * a.C;
    if (guard)
        tile_flush_g(tile, partials + (size_t)blockIdx.x * NC, a.N, a.C, tid, kWfBlock);
    else
        tile_flush(tile, partials + (size_t)blockIdx.x * NC, a.N, a.C, tid, kWfBlock);
}

// ------------------------------------------------------------------------ cube splat
// The listed events of a cube frame into the frame through LDS: workgroup blockIdx / G owns
// the tile of list r (frame channel c = chBase + r) and chirps [m0, m0 + Mt) (all samples),
// blockIdx % G = g its share of the list. A wave loads 64 events and adds each one's taps inside the tile with
// LDS atomics, spread over its lanes (the taps and their weights are those of
// flush_events_cube). An event spread over both axes has 33 x 33 taps: each lane holds one
// sample tap k = (lane mod 32) - 16 and its weight 1/(s - k), and the wave covers two chirp
// rows per instruction (lanes 0-31 row l, lanes 32-63 row l + 1), then the sample tap 16 of all
// 33 rows; the weight of a tap is 1/(sc - l) times 1/(s - k), as in flush_events_cube. The tile
// is then stored into partial cube g, of which every cell is written by exactly one workgroup,
// and k_cube_reduce adds the G partial cubes into the frame in a fixed order: no device atomics
// at all.
__global__ __launch_bounds__(kCubeBlock) void k_cube_splat(const CubeEvent *__restrict__ events,
                                                         const unsigned *counts, unsigned cap, int chBase, int M,
                                                         int N, int C, int Mt, int G, double2 *__restrict__ parts) {
    extern __shared__ double t[];  // re [Mt * N], then im [Mt * N]
    const int mTiles = (M + Mt - 1) / Mt;
    const int tile = blockIdx.x / G, part = blockIdx.x % G;
    const int r = tile / mTiles, c = chBase + r, m0 = (tile % mTiles) * Mt;
    const int mt = min(Mt, M - m0);
    const int TN = Mt * N;
    for (int j = threadIdx.x; j < 2 * TN; j += kCubeBlock) t[j] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const CubeEvent *ev = events + (size_t)r * cap;  // channel c's list
    const unsigned n = min(counts[r], cap);
    const unsigned e0 = (unsigned)(((uint64_t)n * part) / G), e1 = (unsigned)(((uint64_t)n * (part + 1)) / G);
    const int half = lane >> 5, kl = (lane & 31) - 16;  // two-row layout: row of the pair, sample tap
    for (unsigned b = e0 + wave * 64; b < e1; b += kCubeBlock) {
        const unsigned i = b + lane;
        CubeEvent e{};
        const bool mine = i < e1;
        if (mine) e = ev[i];
        uint64_t m = __ballot(mine);
        while (m) {
            const int src = __builtin_ctzll(m);
            m &= m - 1;
            const int center = __builtin_amdgcn_readlane(e.center, src);
            const float s = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.s), src));
            const float Are = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.are), src));
            const float Aim = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.aim), src));
            const int cc = __builtin_amdgcn_readlane(e.cc, src);
            const float sc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.sc), src));
            const int flags = __builtin_amdgcn_readlane(e.flags, src);
            const bool delta = flags & 1, deltaC = flags & 2;
            if (!delta && !deltaC) {
                const float vk = __builtin_amdgcn_rcpf(s - (float)kl);
                const int bk = wrap_bin(center + kl, N);
                for (int it = 0; it < 17; ++it) {
                    // a pair of rows with neither in this tile is skipped by the whole wave
                    const int r0 = wrap_bin(cc + 2 * it - 16, M) - m0, r1 = wrap_bin(cc + 2 * it - 15, M) - m0;
                    if ((r0 < 0 || r0 >= mt) && (r1 < 0 || r1 >= mt || it == 16)) continue;
                    const int l = 2 * it + half - 16;  // rows -16 .. 17; row 17 is none
                    const int rel = half ? r1 : r0;
                    if (l > 16 || rel < 0 || rel >= mt) continue;
                    const float coef = __builtin_amdgcn_rcpf(sc - (float)l) * vk;
                    const int j = rel * N + bk;
                    atomicAdd(&t[j], (double)(Are * coef));
                    atomicAdd(&t[TN + j], (double)(Aim * coef));
                }
                if (lane < 33) {  // sample tap 16 of row lane - 16
                    const int l = lane - 16;
                    const int rel = wrap_bin(cc + l, M) - m0;
                    if (rel >= 0 && rel < mt) {
                        const float coef = __builtin_amdgcn_rcpf(sc - (float)l) * __builtin_amdgcn_rcpf(s - 16.0f);
                        const int j = rel * N + wrap_bin(center + 16, N);
                        atomicAdd(&t[j], (double)(Are * coef));
                        atomicAdd(&t[TN + j], (double)(Aim * coef));
                    }
                }
                continue;
            }
            // one axis or none spread (a fractional offset within 1e-4 of a bin): at most 33 taps
            const int ns = delta ? 1 : 33, nc = deltaC ? 1 : 33;
            for (int tt = lane; tt < ns * nc; tt += 64) {
                const int l = deltaC ? 0 : tt / ns - 16;
                const int k = delta ? 0 : tt % ns - 16;
                const int rel = wrap_bin(cc + l, M) - m0;
                if (rel < 0 || rel >= mt) continue;
                float coef = 1.0f;
                if (!deltaC) coef = __builtin_amdgcn_rcpf(sc - (float)l);
                if (!delta) coef = coef * __builtin_amdgcn_rcpf(s - (float)k);
                const int j = rel * N + wrap_bin(center + k, N);
                atomicAdd(&t[j], (double)(Are * coef));
                atomicAdd(&t[TN + j], (double)(Aim * coef));
            }
        }
    }
    __syncthreads();
    double2 *dst = parts + (int64_t)part * M * N * C;
    for (int j = threadIdx.x; j < mt * N; j += kCubeBlock)
        dst[((int64_t)(m0 + j / N) * N + j % N) * C + c] = make_double2(t[j], t[TN + j]);
}

// frame += the G partial cubes (in order) on the cells of channels [chBase, chBase + lists) of
// the cube [M][N][C] (cells: M * N * lists of them), and their event lists emptied
__global__ void k_cube_reduce(double2 *__restrict__ frame, const double2 *__restrict__ parts, int G, int64_t cells,
                              unsigned *counts, int C, int chBase, int lists) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < lists) counts[q] = 0u;
    if (q >= cells) return;
    const int64_t i = (q / lists) * C + chBase + q % lists;  // the cell in the cube
    const int64_t stride = cells / lists * C;                // cells of one partial cube
    double2 s = make_double2(0.0, 0.0);
    for (int g = 0; g < G; ++g) {
        const double2 v = parts[(int64_t)g * stride + i];
        s.x += v.x;
        s.y += v.y;
    }
    double2 f = frame[i];
    f.x += s.x;
    f.y += s.y;
    frame[i] = f;
}

// The sample-tiled splat. Workgroup blockIdx / G owns the tile of list r (frame channel
// c = chBase + r) and sample slab j: every chirp of the samples [s0, s1) plus kGuard guard bins
// on each side; blockIdx % G = g its share of the list. It takes the events centred in its
// slab, so each event is splatted by one tile, all of its 33 x 33 taps inside it (the chirp
// axis wraps within the tile, the sample axis runs into the guard bins), in the two-row layout
// of k_cube_splat. The slab's cells go to partial cube g (each cell written by one
// workgroup), the guard bins to their own partial rows, which k_cube_reduce_s adds to the
// neighbouring slabs' cells (circularly) in a fixed order.
__device__ inline int slab_lo(int j, int N, int T) { return (int)(((int64_t)j * N) / T); }

__global__ __launch_bounds__(kCubeBlock) void k_cube_splat_s(const CubeEvent *__restrict__ events,
                                                           const unsigned *counts, unsigned cap, int chBase, int lists,
                                                           int M, int N, int C, int T, int G,
                                                           double2 *__restrict__ parts, double2 *__restrict__ guards) {
    extern __shared__ double t[];  // re [M][W], then im [M][W]
    const int tile = blockIdx.x / G, part = blockIdx.x % G;
    const int r = tile / T, j = tile % T, c = chBase + r;
    const int s0 = slab_lo(j, N, T), s1 = slab_lo(j + 1, N, T), ns = s1 - s0, W = ns + 2 * kGuard;
    const int TW = M * W;
    for (int q = threadIdx.x; q < 2 * TW; q += kCubeBlock) t[q] = 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const CubeEvent *ev = events + (size_t)r * cap;
    const unsigned n = min(counts[r], cap);
    const unsigned e0 = (unsigned)(((uint64_t)n * part) / G), e1 = (unsigned)(((uint64_t)n * (part + 1)) / G);
    const int half = lane >> 5, kl = (lane & 31) - 16;
    for (unsigned b = e0 + wave * 64; b < e1; b += kCubeBlock) {
        const unsigned i = b + lane;
        CubeEvent e{};
        bool mine = false;
        if (i < e1) {
            e = ev[i];
            const int c0 = e.center >= N ? e.center - N : e.center;  // round(N * frac) reaches N
            mine = c0 >= s0 && c0 < s1;
        }
        uint64_t m = __ballot(mine);
        while (m) {
            const int src = __builtin_ctzll(m);
            m &= m - 1;
            const int center = __builtin_amdgcn_readlane(e.center, src);
            const float s = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.s), src));
            const float Are = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.are), src));
            const float Aim = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.aim), src));
            const int cc = __builtin_amdgcn_readlane(e.cc, src);
            const float sc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.sc), src));
            const int flags = __builtin_amdgcn_readlane(e.flags, src);
            const bool delta = flags & 1, deltaC = flags & 2;
            const int gc = (center >= N ? center - N : center) - s0 + kGuard;  // the centre's column
            if (!delta && !deltaC) {
                const float vk = __builtin_amdgcn_rcpf(s - (float)kl);
                const int g = gc + kl;
                for (int it = 0; it < 17; ++it) {
                    const int l = 2 * it + half - 16;  // rows -16 .. 17; row 17 is none
                    if (l > 16) continue;
                    const float coef = __builtin_amdgcn_rcpf(sc - (float)l) * vk;
                    const int q = wrap_bin(cc + l, M) * W + g;
                    atomicAdd(&t[q], (double)(Are * coef));
                    atomicAdd(&t[TW + q], (double)(Aim * coef));
                }
                if (lane < 33) {  // sample tap 16 of row lane - 16
                    const int l = lane - 16;
                    const float coef = __builtin_amdgcn_rcpf(sc - (float)l) * __builtin_amdgcn_rcpf(s - 16.0f);
                    const int q = wrap_bin(cc + l, M) * W + gc + 16;
                    atomicAdd(&t[q], (double)(Are * coef));
                    atomicAdd(&t[TW + q], (double)(Aim * coef));
                }
                continue;
            }
            const int nsamp = delta ? 1 : 33, nc = deltaC ? 1 : 33;
            for (int tt = lane; tt < nsamp * nc; tt += 64) {
                const int l = deltaC ? 0 : tt / nsamp - 16;
                const int k = delta ? 0 : tt % nsamp - 16;
                float coef = 1.0f;
                if (!deltaC) coef = __builtin_amdgcn_rcpf(sc - (float)l);
                if (!delta) coef = coef * __builtin_amdgcn_rcpf(s - (float)k);
                const int q = wrap_bin(cc + l, M) * W + gc + k;
                atomicAdd(&t[q], (double)(Are * coef));
                atomicAdd(&t[TW + q], (double)(Aim * coef));
            }
        }
    }
    __syncthreads();
    const int64_t cells = (int64_t)M * N * C;
    double2 *dst = parts + (int64_t)part * cells;
    double2 *gl = guards + ((((int64_t)part * lists + r) * T + j) * 2) * M * kGuard;  // [2][M][kGuard]
    for (int q = threadIdx.x; q < TW; q += kCubeBlock) {
        const int mm = q / W, g = q % W;
        const double2 v = make_double2(t[q], t[TW + q]);
        if (g < kGuard)
            gl[mm * kGuard + g] = v;  // left guard: samples s0 - 16 .. s0 - 1
        else if (g >= kGuard + ns)
            gl[(M + mm) * kGuard + (g - kGuard - ns)] = v;  // right guard: samples s1 .. s1 + 15
        else
            dst[((int64_t)mm * N + s0 + g - kGuard) * C + c] = v;
    }
}

// frame += the G partial cubes of the sample-tiled splat, each with the guard bins its
// neighbouring slabs hold for its cells (in a fixed order), on the cells of channels
// [chBase, chBase + lists) (cells: M * N * lists of them); their event lists emptied.
__global__ void k_cube_reduce_s(double2 *__restrict__ frame, const double2 *__restrict__ parts,
                                const double2 *__restrict__ guards, int G, int M, int N, int C, int T, int chBase,
                                int lists, unsigned *counts) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < lists) counts[q] = 0u;
    const int64_t cells = (int64_t)M * N * lists;
    if (q >= cells) return;
    const int r = (int)(q % lists);
    const int64_t mb = q / lists;
    const int mm = (int)(mb / N), b = (int)(mb % N);
    const int64_t i = mb * C + chBase + r;  // the cell in the cube
    int j = (int)(((int64_t)b * T) / N);
    while (j + 1 < T && slab_lo(j + 1, N, T) <= b) ++j;
    while (j > 0 && slab_lo(j, N, T) > b) --j;
    const int s0 = slab_lo(j, N, T), s1 = slab_lo(j + 1, N, T);
    const int jl = (j + T - 1) % T, jr = (j + 1) % T;
    const bool fromLeft = b - s0 < kGuard, fromRight = s1 - b <= kGuard;
    const int64_t stride = (int64_t)M * N * C;
    double2 s = make_double2(0.0, 0.0);
    for (int g = 0; g < G; ++g) {
        double2 v = parts[(int64_t)g * stride + i];
        const int64_t base = ((int64_t)g * lists + r) * T;
        if (fromLeft) {  // the right guard of the slab before
            const double2 u = guards[((base + jl) * 2 + 1) * M * kGuard + (int64_t)mm * kGuard + (b - s0)];
            v.x += u.x;
            v.y += u.y;
        }
        if (fromRight) {  // the left guard of the slab after
            const double2 u = guards[((base + jr) * 2) * M * kGuard + (int64_t)mm * kGuard + (b - (s1 - kGuard))];
            v.x += u.x;
            v.y += u.y;
        }
        s.x += v.x;
        s.y += v.y;
    }
    double2 f = frame[i];
    f.x += s.x;
    f.y += s.y;
    frame[i] = f;
}

// ------------------------------------------------------------------------ training statistics
// The per-slot records of a batch (statV, statL over slots [0, n)) summed per leaf into
// rows[blockIdx.x][nodes]: workgroup g takes the slots [g·n/G, (g+1)·n/G), each of its waves a
// contiguous quarter, 64 slots at a time in slot order; the records of one leaf among the 64
// are summed by a butterfly (zeros for the other lanes) and added by lane 0 to the wave's LDS
// accumulator. The four accumulators are then added in wave order into the row. Every sum has
// a fixed order, so the result does not depend on scheduling.
__global__ __launch_bounds__(kWfBlock) void k_stat_reduce(const double *__restrict__ statV,
                                                          const int32_t *__restrict__ statL, int64_t n, int nodes,
                                                          double *__restrict__ rows) {
    extern __shared__ double acc[];  // [4][nodes]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = threadIdx.x; j < 4 * nodes; j += kWfBlock) acc[j] = 0.0;
    __syncthreads();
    const int64_t lo = n * blockIdx.x / gridDim.x, hi = n * (blockIdx.x + 1) / gridDim.x;
    const int64_t wlo = lo + (hi - lo) * wave / 4, whi = lo + (hi - lo) * (wave + 1) / 4;
    double *my = acc + (size_t)wave * nodes;
    for (int64_t b = wlo; b < whi; b += 64) {
        const int64_t i = b + lane;
        double val = 0.0;
        int32_t leaf = 0;
        if (i < whi) {
            val = statV[i];
            leaf = statL[i];
        }
        bool pend = val != 0.0;
        uint64_t m = __ballot(pend);
        while (m) {
            const int src = __builtin_ctzll(m);
            const int lw = __builtin_amdgcn_readlane(leaf, src);
            const bool mine = pend && leaf == lw;
            double sum = mine ? val : 0.0;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
            if (lane == 0) my[lw] += sum;
            m &= ~__ballot(mine);
            pend = pend && !mine;
        }
    }
    __syncthreads();
    double *row = rows + (int64_t)blockIdx.x * nodes;
    for (int j = threadIdx.x; j < nodes; j += kWfBlock)
        row[j] = ((acc[j] + acc[nodes + j]) + acc[2 * nodes + j]) + acc[3 * nodes + j];
}

// The same for trees too large for LDS accumulators: one wave per workgroup, accumulating into
// its own row in device memory (lane 0 is the only writer, in slot order).
__global__ __launch_bounds__(64) void k_stat_reduce_mem(const double *__restrict__ statV,
                                                        const int32_t *__restrict__ statL, int64_t n, int nodes,
                                                        double *__restrict__ rows) {
    const int lane = threadIdx.x;
    double *row = rows + (int64_t)blockIdx.x * nodes;
    for (int j = lane; j < nodes; j += 64) row[j] = 0.0;
    __syncthreads();
    const int64_t lo = n * blockIdx.x / gridDim.x, hi = n * (blockIdx.x + 1) / gridDim.x;
    for (int64_t b = lo; b < hi; b += 64) {
        const int64_t i = b + lane;
        double val = 0.0;
        int32_t leaf = 0;
        if (i < hi) {
            val = statV[i];
            leaf = statL[i];
        }
        bool pend = val != 0.0;
        uint64_t m = __ballot(pend);
        while (m) {
            const int src = __builtin_ctzll(m);
            const int lw = __builtin_amdgcn_readlane(leaf, src);
            const bool mine = pend && leaf == lw;
            double sum = mine ? val : 0.0;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
            if (lane == 0) row[lw] += sum;
            m &= ~__ballot(mine);
            pend = pend && !mine;
        }
    }
}

// stat[j] += the sum of rows[g][j] over g < G: workgroups of 32 columns, each column's rows in
// eight ranges summed in row order by eight threads, then the eight sums in range order (a
// fixed order; the loads of 32 columns are coalesced).
__global__ __launch_bounds__(256) void k_stat_rows_fold(const double *__restrict__ rows, int G, int nodes,
                                                        double *__restrict__ stat) {
    __shared__ double part[8][32];
    const int col = threadIdx.x & 31, seg = threadIdx.x >> 5;
    const int j = blockIdx.x * 32 + col;
    const int g0 = G * seg / 8, g1 = G * (seg + 1) / 8;
    double sum = 0.0;
    if (j < nodes)
        for (int g = g0; g < g1; ++g) sum += rows[(int64_t)g * nodes + j];
    part[seg][col] = sum;
    __syncthreads();
    if (seg == 0 && j < nodes) {
        double t = part[0][col];
        for (int k = 1; k < 8; ++k) t += part[k][col];
        stat[j] += t;
    }
}

// ------------------------------------------------------------------------ host side
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// one vertex set: rays, NEE slots, shadow and gather queues
static size_t wf_set_bytes(int64_t P, int64_t S) {
    return 2 * align256(sizeof(float4) * P) + align256(sizeof(float4) * S) + 2 * align256(sizeof(double) * S) +
           2 * align256(sizeof(float2) * S) + 2 * align256(sizeof(int8_t) * S) + align256(sizeof(int32_t) * S) +
           align256(sizeof(int32_t) * P);
}

size_t wavefront_bytes(int64_t P, int C, int maxDepth) {
    size_t b = 0;
    b += align256(sizeof(double) * D_N * P);
    b += align256(sizeof(float) * F_N * P);
    b += align256(sizeof(int32_t) * I_N * P);
    b += align256(sizeof(int2) * P);
    b += 2 * wf_set_bytes(P, P * C);
    b += 2 * align256(sizeof(int32_t) * P);
    b += align256(sizeof(WfCounts) * (maxDepth + 2));
    b += align256(sizeof(int32_t) * (kCells + 1));
    b += align256(sizeof(double) * P) + align256(sizeof(int32_t) * P);
    return b;
}

void wavefront_layout(WfArgs &w, WfSet set[2], void *base, int64_t P, int C, int maxDepth) {
    char *p = (char *)base;
    auto take = [&](size_t bytes) {
        char *r = p;
        p += align256(bytes);
        return (void *)r;
    };
    const int64_t S = P * C;
    w.P = P;
    w.sd = (double *)take(sizeof(double) * D_N * P);
    w.sf = (float *)take(sizeof(float) * F_N * P);
    w.si = (int32_t *)take(sizeof(int32_t) * I_N * P);
    w.hit = (int2 *)take(sizeof(int2) * P);
    for (int v = 0; v < 2; ++v) {
        WfSet &t = set[v];
        t.ray_o = (float4 *)take(sizeof(float4) * P);
        t.ray_d = (float4 *)take(sizeof(float4) * P);
        t.s_d = (float4 *)take(sizeof(float4) * S);
        t.s_tn = (double *)take(sizeof(double) * S);
        t.s_dv = (double *)take(sizeof(double) * S);
        t.s_v = (float2 *)take(sizeof(float2) * S);
        t.s_g1 = (float2 *)take(sizeof(float2) * S);
        t.s_st = (int8_t *)take(sizeof(int8_t) * S);
        t.s_vis = (int8_t *)take(sizeof(int8_t) * S);
        t.q_shd = (int32_t *)take(sizeof(int32_t) * S);
        t.q_gat = (int32_t *)take(sizeof(int32_t) * P);
    }
    w.q_ext0 = (int32_t *)take(sizeof(int32_t) * P);
    w.q_ext1 = (int32_t *)take(sizeof(int32_t) * P);
    w.cnt = (WfCounts *)take(sizeof(WfCounts) * (maxDepth + 2));
    w.cellStart = (int32_t *)take(sizeof(int32_t) * (kCells + 1));
    w.statV = (double *)take(sizeof(double) * P);
    w.statL = (int32_t *)take(sizeof(int32_t) * P);
    w = wf_with_sets(w, set[0], set[0]);
}

WfArgs wf_with_sets(const WfArgs &w, const WfSet &v, const WfSet &t) {
    WfArgs r = w;
    r.ray_o = v.ray_o;
    r.ray_d = v.ray_d;
    r.s_d = v.s_d;
    r.s_tn = v.s_tn;
    r.s_dv = v.s_dv;
    r.s_v = v.s_v;
    r.s_g1 = v.s_g1;
    r.s_st = v.s_st;
    r.s_vis = v.s_vis;
    r.q_shd = v.q_shd;
    r.q_gat = v.q_gat;
    r.tray_o = t.ray_o;
    r.tray_d = t.ray_d;
    return r;
}

size_t tile_lds_bytes(int N, int C) {
    return (size_t)(N >= 2 * kGuard ? N + 2 * kGuard : N) * C * 2 * sizeof(double);
}

size_t gather_lds_bytes(int N, int C) { return tile_lds_bytes(N, C); }

std::vector<int32_t> coherence_cells() {
    // cell (u, v): the first 3 base-5 digits and the first 5 base-3 digits of the primary
    // sample; its indices h satisfy h mod 125 = the 3 digits of u reversed and h mod 243 =
    // the 5 digits of v reversed, i.e. one residue mod 30375
    constexpr int U = 125, V = 243;
    static_assert(U * V == kCells, "cells");
    auto rev = [](int x, int base, int digits) {
        int r = 0;
        for (int k = 0; k < digits; ++k) {
            r = r * base + x % base;
            x /= base;
        }
        return r;
    };
    auto spread = [](uint32_t x) {  // bits of x to the even bit positions
        uint32_t r = 0;
        for (int b = 0; b < 16; ++b) r |= ((x >> b) & 1u) << (2 * b);
        return r;
    };
    std::vector<std::pair<uint32_t, int32_t>> cells;
    cells.reserve(kCells);
    for (int u = 0; u < U; ++u)
        for (int v = 0; v < V; ++v) {
            const int a5 = rev(u, 5, 3), a3 = rev(v, 3, 5);
            int r = a5;  // r = a5 (mod 125), r = a3 (mod 243)
            while (r % V != a3) r += U;
            cells.push_back({spread((uint32_t)u) | (spread((uint32_t)v) << 1), (int32_t)r});
        }
    std::sort(cells.begin(), cells.end());
    std::vector<int32_t> out(kCells);
    for (int c = 0; c < kCells; ++c) out[c] = cells[c].second;
    return out;
}

CubeSplat cube_splat_plan(int M, int N, size_t ldsPerBlock, size_t ldsPerCU) {
    CubeSplat p{};
    const int64_t col = (int64_t)(2 * sizeof(double)) * M;  // LDS of one sample of every chirp
    const int64_t cols = (int64_t)(ldsPerCU / 2) / col;      // two tiles per CU
    if (N >= 2 * kGuard && cols >= 3 * kGuard) {
        const int64_t T = (N + (cols - 2 * kGuard) - 1) / (cols - 2 * kGuard);
        if (N / T >= kGuard) {  // every slab as wide as a guard
            p.samples = true;
            p.T = (int)T;
            p.tilesPerList = (int)T;
            p.lds = (size_t)(col * ((N + T - 1) / T + 2 * kGuard));
            return p;
        }
    }
    const int64_t row = (int64_t)(2 * sizeof(double)) * N;  // LDS of one chirp of every sample
    int64_t rows = (int64_t)(ldsPerCU / 2) / row;
    if (rows < 1) rows = (int64_t)ldsPerBlock / row;
    if (rows < 1) return p;  // a chirp does not fit
    const int64_t tiles = (M + rows - 1) / rows;
    p.Mt = (int)((M + tiles - 1) / tiles);  // tiles of even height
    p.tilesPerList = (M + p.Mt - 1) / p.Mt;
    p.lds = (size_t)(row * p.Mt);
    return p;
}

int64_t cube_guard_doubles(const CubeSplat &p, int M, int lists, int G) {
    return p.samples ? (int64_t)G * lists * p.T * 2 * M * kGuard * 2 : 0;
}

hipError_t launch_cube_splat(const CubeSplat &p, const CubeEvent *events, unsigned *counts, unsigned cap, int chBase,
                             int lists, int M, int N, int C, int G, double2 *parts, double *frame, hipStream_t s) {
    (void)hipGetLastError();
    const void *f = p.samples ? (const void *)k_cube_splat_s : (const void *)k_cube_splat;
    if (p.lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds);
        if (e != hipSuccess) return e;
    }
    const int64_t cells = (int64_t)M * N * lists;
    const dim3 grid((unsigned)(lists * p.tilesPerList * G)), rgrid((unsigned)((cells + kWfBlock - 1) / kWfBlock));
    if (p.samples) {
        double2 *guards = parts + (int64_t)G * M * N * C;  // after the partial cubes
        hipLaunchKernelGGL(k_cube_splat_s, grid, dim3(kCubeBlock), p.lds, s, events, counts, cap, chBase, lists, M, N,
                           C, p.T, G, parts, guards);
        hipLaunchKernelGGL(k_cube_reduce_s, rgrid, dim3(kWfBlock), 0, s, (double2 *)frame, (const double2 *)parts,
                           (const double2 *)guards, G, M, N, C, p.T, chBase, lists, counts);
    } else {
        hipLaunchKernelGGL(k_cube_splat, grid, dim3(kCubeBlock), p.lds, s, events, counts, cap, chBase, M, N, C, p.Mt,
                           G, parts);
        hipLaunchKernelGGL(k_cube_reduce, rgrid, dim3(kWfBlock), 0, s, (double2 *)frame, (const double2 *)parts, G,
                           cells, counts, C, chBase, lists);
    }
    return hipGetLastError();
}

hipError_t launch_stat_reduce(const double *statV, const int32_t *statL, int64_t n, int nodes, double *rows,
                              int64_t rowsCap, double *stat, hipStream_t s) {
    (void)hipGetLastError();
    if (n <= 0 || nodes <= 0) return hipSuccess;
    const size_t lds = (size_t)4 * nodes * sizeof(double);
    const bool inLds = nodes <= kStatLdsNodes;
    const int64_t perBlock = inLds ? 4096 : 1024;
    const int G = (int)std::max<int64_t>(1, std::min<int64_t>({(n + perBlock - 1) / perBlock, rowsCap / nodes, 1024}));
    if (inLds) {
        if (lds > 65536) {
            const hipError_t e = hipFuncSetAttribute((const void *)k_stat_reduce,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_stat_reduce, dim3(G), dim3(kWfBlock), lds, s, statV, statL, n, nodes, rows);
    } else {
        hipLaunchKernelGGL(k_stat_reduce_mem, dim3(G), dim3(64), 0, s, statV, statL, n, nodes, rows);
    }
    hipLaunchKernelGGL(k_stat_rows_fold, dim3((nodes + 31) / 32), dim3(256), 0, s, rows, G, nodes, stat);
    return hipGetLastError();
}

int wavefront_trav_blocks_per_cu(bool count) {
    int n = 0;
    hipError_t e = count ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_trav<true>, kWfBlock, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_wf_trav<false>, kWfBlock, 0);
    return e == hipSuccess && n > 0 ? n : 1;
}

size_t tail_lds_bytes(int N, int C, bool noTile) {
    return (noTile ? 0 : tile_lds_bytes(N, C)) + sizeof(int32_t) * kStackShort * kWfBlock + kWaveEventBytes;
}

hipError_t prepare_wavefront(int N, int C, bool cube, bool noTile) {
    const size_t glds = cube ? 0 : gather_lds_bytes(N, C);
    if (glds > 65536) {
        const hipError_t e =
            hipFuncSetAttribute((const void *)k_wf_gather, hipFuncAttributeMaxDynamicSharedMemorySize, (int)glds);
        if (e != hipSuccess) return e;
    }
    const size_t slds = cube ? 0 : shadow_splat_lds_bytes(N, C);
    if (slds > 65536 && slds + kStackShort * kWfBlock * sizeof(int32_t) <= 163840) {
        const void *fs[4] = {(const void *)k_wf_shadow_splat<false, false, 3>, (const void *)k_wf_shadow_splat<true, false, 3>,
                             (const void *)k_wf_shadow_splat<false, false, 2>, (const void *)k_wf_shadow_splat<true, false, 2>};
        for (const void *f : fs) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)slds);
            if (e != hipSuccess) return e;
        }
    }
    const size_t tlds = tail_lds_bytes(N, C, noTile);
    if (tlds > 65536) {
        const void *fs[4] = {(const void *)k_wf_tail<false, false>, (const void *)k_wf_tail<false, true>,
                             (const void *)k_wf_tail<true, false>, (const void *)k_wf_tail<true, true>};
        for (const void *f : fs) {
            const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)tlds);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// ------------------------------------------------------------------------ batch plans
// The by-value parameters of one kernel step, for the batch arguments h.
struct StepParams {
    int i0 = 0;
    int64_t l1 = 0, l2 = 0;
    bool b = false;
    double *dp = nullptr;
    double2 *d2p = nullptr;
    WfArgs w;  // the batch's buffers with the step's vertex and traced-ray sets
    void *p[5] = {};
};
static void step_params(const BatchStep &st, const BatchArgs &h, StepParams &sp) {
    sp.w = wf_with_sets(h.w, h.set[st.vset], h.set[st.tset]);
    void *a = (void *)&h.a, *w = (void *)&sp.w;
    switch (st.sig) {
        case BatchStep::kSetup:  // (WfArgs, int nCounts, double *partials, int64 nPartials, int64 nStat)
            sp.i0 = (int)st.x[0];
            sp.dp = (double *)h.partials;
            sp.l1 = st.x[1];
            sp.l2 = st.x[2];
            sp.p[0] = w;
            sp.p[1] = &sp.i0;
            sp.p[2] = &sp.dp;
            sp.p[3] = &sp.l1;
            sp.p[4] = &sp.l2;
            break;
        case BatchStep::kAW:  // (TraceArgs, WfArgs)
            sp.p[0] = a;
            sp.p[1] = w;
            break;
        case BatchStep::kAWi:  // (TraceArgs, WfArgs, int)
            sp.i0 = (int)st.x[0];
            sp.p[0] = a;
            sp.p[1] = w;
            sp.p[2] = &sp.i0;
            break;
        case BatchStep::kTail:  // (TraceArgs, WfArgs, int K, double2 *partials): its rows from x[1]
            sp.i0 = (int)st.x[0];
            sp.d2p = h.partials + st.x[1];
            sp.p[0] = a;
            sp.p[1] = w;
            sp.p[2] = &sp.i0;
            sp.p[3] = &sp.d2p;
            break;
        default:  // kGather: (TraceArgs, WfArgs, int depth, double2 *partials, bool count)
            sp.i0 = (int)st.x[0];
            sp.d2p = h.partials;
            sp.b = st.x[1] != 0;
            sp.p[0] = a;
            sp.p[1] = w;
            sp.p[2] = &sp.i0;
            sp.p[3] = &sp.d2p;
            sp.p[4] = &sp.b;
            break;
    }
}

BatchPlan plan_wavefront(const TraceArgs &a, const WfArgs &w, int gridTrav, int gridWork, int gridGather,
                         int64_t partialDoubles, bool zeroStats, bool count, hipEvent_t *evClosest,
                         hipEvent_t *evShadow, int tailDepth, bool decouple, bool fuseSplat) {
    BatchPlan plan;
    auto &S = plan.steps;
    const int maxDepth = a.maxDepth;
    // grids no larger than the batch's items can use (queue sizes are only known on the
    // device; the batch's path count bounds them), so that small batches launch small grids
    const int64_t n = w.n;
    const int nrx = min(a.rxCount, a.C);
    auto fit = [](int grid, int64_t items, int64_t per) {
        return (int)std::max<int64_t>(1, std::min<int64_t>(grid, (items + per - 1) / per));
    };
    const int gClosest = fit(gridTrav, n, kWfBlock);
    const int gShadow = fit(gridTrav, n * ((nrx + kPacket - 1) / kPacket), kWfBlock);
    const int gShade = fit(gridWork, n, kRange);
    const int gNee = fit(gridWork, n * nrx, kRange);
    // a kernel step using the buffers of vertex set vset and the traced rays of set tset
    auto kernel = [&](int stream, const void *f, int grid, unsigned block, size_t lds, int8_t sig, int vset,
                      int tset, int64_t x0 = 0, int64_t x1 = 0, int64_t x2 = 0) {
        BatchStep st;
        st.kind = BatchStep::kKernel;
        st.stream = (int8_t)stream;
        st.func = f;
        st.grid = dim3(grid);
        st.block = dim3(block);
        st.lds = (unsigned)lds;
        st.sig = sig;
        st.vset = (int8_t)vset;
        st.tset = (int8_t)tset;
        st.x[0] = x0;
        st.x[1] = x1;
        st.x[2] = x2;
        S.push_back(st);
    };
    auto marker = [&](int8_t kind, int stream, hipEvent_t e) {
        BatchStep st;
        st.kind = kind;
        st.stream = (int8_t)stream;
        st.event = e;
        S.push_back(st);
    };
    auto named = [&](const char *name) { S.back().name = name; };
    auto signal = [&](int stream) {
        BatchStep st;
        st.kind = BatchStep::kSignal;
        st.stream = (int8_t)stream;
        st.x[0] = plan.nSync[stream]++;
        S.push_back(st);
        return (int)st.x[0];
    };
    auto wait = [&](int stream, int from, int idx) {
        BatchStep st;
        st.kind = BatchStep::kWait;
        st.stream = (int8_t)stream;
        st.x[0] = idx;
        st.x[1] = from;
        S.push_back(st);
    };
    // set-up: one workgroup for the scan, the clears spread over up to 256 more
    {
        const int64_t nStat = zeroStats ? n : 0;
        const int64_t clears = std::max<int64_t>(maxDepth + 2, partialDoubles / 2 + nStat);
        const int gSetup = 1 + (int)std::min<int64_t>(256, (clears + 4 * kScanThreads - 1) / (4 * kScanThreads));
        kernel(0, (const void *)k_wf_setup, gSetup, kScanThreads, 0, BatchStep::kSetup, 0, 0, maxDepth + 2,
               partialDoubles, nStat);
        named("hsr setup");
    }
    // generation (vertex 0's rays and gather list): K cells per workgroup, about kRange paths each
    const int64_t q = n / kCells;
    const int K = (int)std::max<int64_t>(1, std::min<int64_t>(kGenCellsMax, kRange / (q + 1)));
    kernel(0, a.refChain ? (const void *)k_wf_gen<true> : (const void *)k_wf_gen<false>, (kCells + K - 1) / K,
           kWfBlock, 0, BatchStep::kAWi, 0, 0, K);
    named("hsr generation");
    const size_t glds = a.cube ? 0 : gather_lds_bytes(a.N, a.C);  // prepare_wavefront allowed it
    if (!a.onlyIndirect) {
        kernel(0, a.refChain ? (const void *)k_wf_nee0<true> : (const void *)k_wf_nee0<false>, gNee, kWfBlock, 0,
               BatchStep::kAW, 0, 0);
        named("hsr nee0");
    }
    const void *fTrav = count ? (const void *)k_wf_trav<true> : (const void *)k_wf_trav<false>;
    const void *fShadow = count ? (const void *)k_wf_shadow<true> : (const void *)k_wf_shadow<false>;
    const void *fShade = a.refChain ? (const void *)k_wf_shade<true> : (const void *)k_wf_shade<false>;
    const void *fTail = a.refChain ? (count ? (const void *)k_wf_tail<true, true> : (const void *)k_wf_tail<true, false>)
                                   : (count ? (const void *)k_wf_tail<false, true> : (const void *)k_wf_tail<false, false>);
    // vertex k: its closest-hit traversal reads the rays of set k & 1, its shadow and gather
    // stages that set's slots and queues, and its shading writes vertex k + 1's set
    // depth 0 through the transmitter's primary visibility lists when the batch has them
    const void *fTrav0 = count ? (const void *)k_wf_trav0<true> : (const void *)k_wf_trav0<false>;
    auto trav = [&](int stream, int k) {
        if (evClosest) marker(BatchStep::kRecord, stream, evClosest[2 * plan.nClosest]);
        kernel(stream, k == 0 && a.pvList ? fTrav0 : fTrav, gClosest, kWfBlock, 0, BatchStep::kAWi, k & 1, k & 1, k);
        named("hsr closest-hit traversal");
        if (evClosest) marker(BatchStep::kRecord, stream, evClosest[2 * plan.nClosest + 1]);
        ++plan.nClosest;
    };
    // decoupled batches whose shadow stage splats (shadow_splat_eligible): one fused launch per
    // vertex, no gather stage
    const bool fused = decouple && fuseSplat && !a.noSplat && !a.poDiffraction;
    const bool w3 = shadow_splat_waves(a.N, a.C) == 3;
    const void *fShadowSplat =
        a.cube ? (count ? (const void *)k_wf_shadow_splat<true, true, 4> : (const void *)k_wf_shadow_splat<false, true, 4>)
        : w3   ? (count ? (const void *)k_wf_shadow_splat<true, false, 3> : (const void *)k_wf_shadow_splat<false, false, 3>)
               : (count ? (const void *)k_wf_shadow_splat<true, false, 2> : (const void *)k_wf_shadow_splat<false, false, 2>);
    auto shadow_gather = [&](int k) {
        if (fused) {
            if (evShadow) marker(BatchStep::kRecord, 0, evShadow[2 * plan.nShadow]);
            // cube frames: no frame tile, the shadow stage's persistent grid
            kernel(0, fShadowSplat, a.cube ? gShadow : gridGather, kWfBlock,
                   a.cube ? 0 : shadow_splat_lds_bytes(a.N, a.C), BatchStep::kGather, k & 1, k & 1, k, count ? 1 : 0);
            named("hsr shadow + splat");
            if (evShadow) marker(BatchStep::kRecord, 0, evShadow[2 * plan.nShadow + 1]);
            ++plan.nShadow;
            return;
        }
        if (!a.poDiffraction) {
            if (evShadow) marker(BatchStep::kRecord, 0, evShadow[2 * plan.nShadow]);
            kernel(0, fShadow, gShadow, kWfBlock, 0, BatchStep::kAWi, k & 1, k & 1, k);
            named("hsr shadow");
            if (evShadow) marker(BatchStep::kRecord, 0, evShadow[2 * plan.nShadow + 1]);
            ++plan.nShadow;
        }
        kernel(0, (const void *)k_wf_gather, gridGather, kWfBlock, glds, BatchStep::kGather, k & 1, k & 1, k,
               count ? 1 : 0);
        named("hsr gather");
    };
    // the fused tail from the depth-k hits on, in place of the stages from this depth's shading
    // on; its frame tiles go to the partials from row `row`
    auto tail = [&](int stream, int k, int64_t row) {
        kernel(stream, fTail, gridGather, kWfBlock, tail_lds_bytes(a.N, a.C, a.cube || a.noSplat), BatchStep::kTail,
               k & 1, k & 1, k, row * a.N * a.C);
        named("hsr tail");
    };
    auto shade = [&](int stream, int k) {
        kernel(stream, fShade, gShade, kWfBlock, 0, BatchStep::kAWi, (k + 1) & 1, k & 1, k);
        named("hsr shade");
    };
    if (!decouple) {
        for (int k = 0; k <= maxDepth; ++k) {
            // the depth's closest-hit traversal on stream 1, beside its shadow and gather stages:
            // it reads the extension queue and rays of depth k and writes only the hits, which the
            // shading stage reads after the join. A depth without those stages (depth 0 when the
            // direct paths are not traced) keeps it on stream 0: no fork and join to wait for.
            const bool nee = k >= 1 || !a.onlyIndirect;
            if (k < maxDepth) {
                if (nee) marker(BatchStep::kFork, 0, nullptr);
                trav(nee ? 1 : 0, k);
            }
            if (nee) shadow_gather(k);
            if (k < maxDepth) {
                if (nee) marker(BatchStep::kJoin, 0, nullptr);
                if (tailDepth >= 1 && k == tailDepth) {
                    tail(0, k, 0);
                    break;
                }
                shade(0, k);
            }
        }
        return plan;
    }
    // decoupled: the path chain on stream 1, the NEE of each vertex on stream 0
    wait(1, 0, signal(0));  // generation (and the depth-0 NEE set-up) done
    std::vector<int> gathered(maxDepth + 1, -1);  // stream-0 sync after vertex k's gather
    int shaded = -1;                              // stream-1 sync after the last shading
    for (int k = 0; k <= maxDepth; ++k) {
        const bool nee = k >= 1 || !a.onlyIndirect;
        if (k < maxDepth) trav(1, k);
        if (nee) {
            if (shaded >= 0) wait(0, 1, shaded);  // vertex k's slots, origins and queues
            shadow_gather(k);
            gathered[k] = signal(0);
        }
        if (k < maxDepth) {
            if (tailDepth >= 1 && k == tailDepth) {
                tail(1, k, gridGather);
                break;
            }
            if (k >= 1 && gathered[k - 1] >= 0) wait(1, 0, gathered[k - 1]);  // set (k + 1) & 1 free
            shade(1, k);
            shaded = signal(1);
        }
    }
    wait(0, 1, signal(1));
    return plan;
}

hipError_t launch_plan(const BatchPlan &plan, const BatchArgs &h, hipStream_t s, hipStream_t s2, hipEvent_t *evFork,
                       hipEvent_t *evJoin, bool markers) {
    (void)hipGetLastError();
    const hipStream_t st[2] = {s, s2};
    hipEvent_t *const sync[2] = {evFork, evJoin};
    int forks = 0, joins = 0;
    hipError_t r = hipSuccess;
    for (const BatchStep &x : plan.steps) {
        switch (x.kind) {
            case BatchStep::kKernel: {
                StepParams sp;
                step_params(x, h, sp);
                if (markers && x.name) roctxRangePushA(x.name);
                r = hipLaunchKernel(x.func, x.grid, x.block, sp.p, x.lds, st[x.stream]);
                if (markers && x.name) roctxRangePop();
                break;
            }
            case BatchStep::kFork:
                r = hipEventRecord(evFork[forks], s);
                if (r == hipSuccess) r = hipStreamWaitEvent(s2, evFork[forks], 0);
                ++forks;
                break;
            case BatchStep::kJoin:
                r = hipEventRecord(evJoin[joins], s2);
                if (r == hipSuccess) r = hipStreamWaitEvent(s, evJoin[joins], 0);
                ++joins;
                break;
            case BatchStep::kSignal: r = hipEventRecord(sync[x.stream][x.x[0]], st[x.stream]); break;
            case BatchStep::kWait: r = hipStreamWaitEvent(st[x.stream], sync[x.x[1]][x.x[0]], 0); break;
            default: r = hipEventRecord(x.event, st[x.stream]); break;
        }
        if (r != hipSuccess) return r;
    }
    return hipGetLastError();
}

hipError_t graph_plan(BatchGraph *out, const BatchPlan &plan, const BatchArgs &h) {
    *out = BatchGraph{};
    hipGraph_t g = nullptr;
    hipError_t r = hipGraphCreate(&g, 0);
    if (r != hipSuccess) return r;
    // each node depends on the last node of its stream; a fork hands stream 0's to stream 1,
    // a join adds stream 1's to stream 0's, a wait the nodes its sync event followed
    std::vector<hipGraphNode_t> deps[2];
    std::vector<hipGraphNode_t> nodes;
    std::vector<std::vector<hipGraphNode_t>> sync[2];  // the stream's last nodes at each sync event
    for (int i = 0; i < 2; ++i) sync[i].resize(plan.nSync[i]);
    for (const BatchStep &x : plan.steps) {
        hipGraphNode_t nd = nullptr;
        if (x.kind == BatchStep::kKernel) {
            StepParams sp;
            step_params(x, h, sp);
            hipKernelNodeParams kp{};
            kp.blockDim = x.block;
            kp.func = const_cast<void *>(x.func);
            kp.gridDim = x.grid;
            kp.kernelParams = sp.p;
            kp.sharedMemBytes = x.lds;
            r = hipGraphAddKernelNode(&nd, g, deps[x.stream].data(), deps[x.stream].size(), &kp);
            nodes.push_back(nd);
        } else if (x.kind == BatchStep::kFork) {
            deps[1] = deps[0];
        } else if (x.kind == BatchStep::kJoin) {
            deps[0].insert(deps[0].end(), deps[1].begin(), deps[1].end());
        } else if (x.kind == BatchStep::kSignal) {
            sync[x.stream][x.x[0]] = deps[x.stream];
        } else if (x.kind == BatchStep::kWait) {
            for (hipGraphNode_t d : sync[x.x[1]][x.x[0]])
                if (std::find(deps[x.stream].begin(), deps[x.stream].end(), d) == deps[x.stream].end())
                    deps[x.stream].push_back(d);
        } else {
            r = hipGraphAddEventRecordNode(&nd, g, deps[x.stream].data(), deps[x.stream].size(), x.event);
        }
        if (r != hipSuccess) break;
        if (nd) deps[x.stream].assign(1, nd);
    }
    for (int i = 0; i < 2 && r == hipSuccess; ++i) r = hipGraphInstantiate(&out->exec[i], g, nullptr, nullptr, 0);
    for (int i = 0; i < 2 && r == hipSuccess; ++i) r = hipEventCreateWithFlags(&out->done[i], hipEventDisableTiming);
    out->graph = g;
    out->nodes = std::move(nodes);
    if (r != hipSuccess) destroy_graph(*out);
    return r;
}

hipError_t replay_plan(BatchGraph &g, const BatchPlan &plan, const BatchArgs &h, hipStream_t s) {
    // the instance that ran two replays ago: its parameters change only once that run is over
    const int i = g.next;
    g.next ^= 1;
    hipError_t r;
    if (g.used[i]) {
        while ((r = hipEventQuery(g.done[i])) == hipErrorNotReady) {
        }
        if (r != hipSuccess) return r;
    }
    size_t j = 0;
    for (const BatchStep &x : plan.steps) {
        if (x.kind != BatchStep::kKernel) continue;
        StepParams sp;
        step_params(x, h, sp);
        hipKernelNodeParams kp{};
        kp.blockDim = x.block;
        kp.func = const_cast<void *>(x.func);
        kp.gridDim = x.grid;
        kp.kernelParams = sp.p;
        kp.sharedMemBytes = x.lds;
        if ((r = hipGraphExecKernelNodeSetParams(g.exec[i], g.nodes[j++], &kp)) != hipSuccess) return r;
    }
    if ((r = hipGraphLaunch(g.exec[i], s)) != hipSuccess) return r;
    g.used[i] = true;
    return hipEventRecord(g.done[i], s);
}

void destroy_graph(BatchGraph &g) {
    for (int i = 0; i < 2; ++i) {
        if (g.exec[i]) (void)hipGraphExecDestroy(g.exec[i]);
        if (g.done[i]) (void)hipEventDestroy(g.done[i]);
    }
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = BatchGraph{};
}

}  // namespace hsr

// gfx950 kernels of the Monte Carlo FMCW radar integrator.
//
// k_trace runs PathTracer::sample (libhussar/include/hussar/integrators/path.h:236-451) for
// a range of Halton indices. Each wave64 owns a contiguous index chunk and keeps its 64
// lanes busy by regenerating a path in any lane whose path ended (the loop is wave-uniform,
// one bounce per iteration). Rays are traced through a host-built BVH2 with an LDS-staged
// short stack. NEE events are splatted into the workgroup's frame tile in LDS: the wave
// walks its events in lane order and, for each, 33 lanes add the 33 taps of
// radar::Frame::splat (libradar/include/radar/radar.h:564-627) into distinct bins with
// ds_add_f64 — no global atomics. At the end every workgroup stores its tile; k_reduce_*
// sums the tiles in a fixed order with Kahan-compensated FP64 adds into the frame.
//
// Precision model: FP32 traversal, triangle tests, fields and weights; FP64 geometry chain
// (directions, hit points, segment lengths re-evaluated against the hit triangle's plane),
// FP64 path time, IF phase and frame sums. Compiled without FP contraction.
#include "trace_common.h"

namespace hsr {

// ------------------------------------------------------------------------ the path kernel
template <bool kCount>
__global__ __launch_bounds__(kTraceBlock) void k_trace(TraceArgs a) {
    extern __shared__ double smem[];
    const int NC = a.N * a.C;
    double *tile = smem;                                   // [N*C] complex double
    int32_t *lds_stack = (int32_t *)(smem + 2 * (size_t)NC);  // [kStackShort][kTraceBlock]
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    for (int i = tid; i < 2 * NC; i += kTraceBlock) tile[i] = 0.0;
    __syncthreads();

    int32_t spill[kStackSpill];
    Stack st;
    st.spill = spill;
    st.set_column(lds_stack + tid);
    st.sp = 0;
    uint32_t cn = 0, ct = 0, crc = 0, crs = 0, cev = 0;

    const int64_t wave = (int64_t)blockIdx.x * (kTraceBlock / 64) + (tid >> 6);
    const int64_t begin = wave * a.chunk;
    const int64_t end = min(a.count, begin + a.chunk);
    int64_t next = begin;  // wave-uniform
    const int nrx = min(a.rxCount, a.C);

    PathState p;
    bool active = false;

    while (true) {
        // ---- regenerate: idle lanes take the next indices of the chunk, in lane order
        const uint64_t idle = __ballot(!active);
        if (next < end && __popcll(idle) >= 16) {
            const uint64_t below = idle & ((1ull << lane) - 1ull);
            const int64_t mine = next + (int64_t)__popcll(below);
            if (!active && mine < end) {
                p.local = mine;
                generate(a, (uint64_t)(a.offset + mine), p, a.gsChild, a.gsP0);
                active = true;
            }
            next = min(end, next + (int64_t)__popcll(idle));
        }
        if (!__any(active)) {
            if (next >= end) break;
            continue;
        }
        const bool rec = a.recHits && active && p.local < a.recPaths;

        // ---- NEE (path.h:290-397)
        const bool doNee = active && !(p.depth == 0 && a.onlyIndirect);
        for (int ch = 0; ch < nrx; ++ch) {
            bool has = false;
            Event ev{0, 0.0f, {0.0f, 0.0f}, 0};
            if (doNee) {
                const NeeSetup ns = nee_setup(a, p, ch);
                bool vis = true;
                if (!a.poDiffraction) {
                    if (kCount) ++crs;
                    vis = !occluded<kCount>(a.nodesF, a.tris, ns.o, ns.d, ns.tfar, st, cn, ct);
                }
                int8_t flag = 1;
                if (vis) {
                    flag = 2;
                    if (ns.gw1.re != 0.0f || ns.gw1.im != 0.0f) p.gw = p.gw + ns.gw1;
                    if (ns.gwAdd) p.gw = p.gw + ns.v;
                    if (ns.splat) {
                        ev = make_event(a, ns.tn, ns.v, a.sampleWeight, ns.dv);
                        has = true;
                        flag = 3;
                        if (kCount) ++cev;
                    }
                }
                if (rec && p.depth >= 1) a.recNee[(p.local * a.maxDepth + (p.depth - 1)) * a.C + a.chBase + ch] = flag;
            }
            flush_events(has, ev, a.chBase + ch, tile, a.N, a.C, lane);
        }

        // ---- extend (path.h:399-442)
        bool finished = false;
        if (active) {
            if (path_ends(a, p)) {
                finished = true;
            } else {
                if (kCount) ++crc;
                float t;
                int32_t slot;
                const int32_t id = closest_hit<kCount>(a.nodes, a.tris, cast3<float>(p.o), cast3<float>(p.d), st, t,
                                                       slot, cn, ct);
                if (rec) a.recHits[p.local * a.maxDepth + p.depth] = id;
                finished = id < 0 || !apply_hit(a, p, slot);
            }
        }
        // ---- finished paths: guiding record (path.h:448-450)
        bool pend = false;
        int32_t leaf = 0;
        double val = 0.0;
        if (finished) {
            active = false;
            if (a.doGuiding && a.guideOn && !a.isFinal && p.pdf > 0.0f) {
                const float density = cabsf_(p.gw) * p.pdf;
                const float weight = 1.0f / p.pdf;
                leaf = tree_leaf(a.gtChild, p.x0, p.x1);
                const double dd = (double)density;
                val = (a.secondMoment ? dd * dd : dd) * (double)weight;
                pend = true;
            }
        }
        if (a.doGuiding && a.guideOn && !a.isFinal) record_stats(pend, leaf, val, a.gtStat + (blockIdx.x % kStatCopies) * a.gtStride, lane);
    }

    // ---- store the workgroup's frame tile
    __syncthreads();
    double2 *dst = a.partials + (size_t)blockIdx.x * NC;
    for (int i = tid; i < NC; i += kTraceBlock) {
        const int t = tile_re(i / a.C, i % a.C, a.N);
        dst[i] = make_double2(tile[t], tile[NC + t]);
    }
    if (kCount) {
        atomicAdd(&a.counters[0], (unsigned long long)cn);
        atomicAdd(&a.counters[1], (unsigned long long)ct);
        atomicAdd(&a.counters[2], (unsigned long long)crc);
        atomicAdd(&a.counters[3], (unsigned long long)crs);
        atomicAdd(&a.counters[4], (unsigned long long)cev);
    }
}

// ------------------------------------------------------------------------ frame reduction
// Stage 1: slice s of the workgroup tiles is summed per frame double with Kahan compensation.
__global__ void k_reduce_stage(const double *__restrict__ partials, int W, int M, int S, double *stage_sum,
                               double *stage_comp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (j >= M) return;
    const int w0 = (int)((int64_t)W * s / S), w1 = (int)((int64_t)W * (s + 1) / S);
    double sum = 0.0, c = 0.0;
    for (int w = w0; w < w1; ++w) {
        const double y = partials[(size_t)w * M + j] - c;
        const double t = sum + y;
        c = (t - sum) - y;
        sum = t;
    }
    stage_sum[(size_t)s * M + j] = sum;
    stage_comp[(size_t)s * M + j] = c;
}

// Stage 2: the slices in order, then a compensated add into the persistent frame.
__global__ void k_reduce_final(const double *__restrict__ stage_sum, const double *__restrict__ stage_comp, int S,
                               int M, double *frame_sum, double *frame_comp) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    double sum = 0.0, c = 0.0;
    for (int s = 0; s < S; ++s) {
        const double y = (stage_sum[(size_t)s * M + j] - stage_comp[(size_t)s * M + j]) - c;
        const double t = sum + y;
        c = (t - sum) - y;
        sum = t;
    }
    const double fy = (sum - c) - frame_comp[j];
    const double ft = frame_sum[j] + fy;
    frame_comp[j] = (ft - frame_sum[j]) - fy;
    frame_sum[j] = ft;
}

__global__ void k_zero(double *p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_zero2(double *p, double *q, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = 0.0;
        q[i] = 0.0;
    }
}

// ------------------------------------------------------------------------ launchers
size_t trace_lds_bytes(int N, int C) {
    return (size_t)N * C * 2 * sizeof(double) + (size_t)kStackShort * kTraceBlock * sizeof(int32_t);
}

int trace_blocks_per_cu(int N, int C, bool count) {
    int n = 0;
    const size_t lds = trace_lds_bytes(N, C);
    hipError_t e = count ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace<true>, kTraceBlock, lds)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace<false>, kTraceBlock, lds);
    return e == hipSuccess ? n : 1;
}

// Each launcher first clears a stale error of an earlier, already reported runtime call, so
// that hipGetLastError() after the launch reports this launch only.
hipError_t launch_trace(const TraceArgs &args, int grid, bool count, hipStream_t stream) {
    (void)hipGetLastError();
    const size_t lds = trace_lds_bytes(args.N, args.C);
    if (count)
        hipLaunchKernelGGL(k_trace<true>, dim3(grid), dim3(kTraceBlock), lds, stream, args);
    else
        hipLaunchKernelGGL(k_trace<false>, dim3(grid), dim3(kTraceBlock), lds, stream, args);
    return hipGetLastError();
}

hipError_t launch_reduce(const double *partials, int W, int M, double *stage_sum, double *stage_comp, int S,
                         double *frame_sum, double *frame_comp, hipStream_t stream) {
    (void)hipGetLastError();
    const int bs = 256;
    hipLaunchKernelGGL(k_reduce_stage, dim3((M + bs - 1) / bs, S), dim3(bs), 0, stream, partials, W, M, S,
                       stage_sum, stage_comp);
    hipLaunchKernelGGL(k_reduce_final, dim3((M + bs - 1) / bs), dim3(bs), 0, stream, stage_sum, stage_comp, S, M,
                       frame_sum, frame_comp);
    return hipGetLastError();
}

// stat[i] += stat[k * stride + i] for the copies k = 1 .. kStatCopies-1, in order; the
// copies are cleared
__global__ void k_stat_fold(double *stat, int64_t stride, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = stat[i];
    for (int k = 1; k < kStatCopies; ++k) {
        s += stat[k * stride + i];
        stat[k * stride + i] = 0.0;
    }
    stat[i] = s;
}

hipError_t launch_stat_fold(double *stat, int64_t stride, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_stat_fold, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, stat, stride, n);
    return hipGetLastError();
}

hipError_t launch_zero2(double *p, double *q, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    const int bs = 256;
    hipLaunchKernelGGL(k_zero2, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, stream, p, q, n);
    return hipGetLastError();
}

hipError_t launch_zero(double *p, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    const int bs = 256;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace hsr

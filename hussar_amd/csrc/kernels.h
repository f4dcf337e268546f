// Launch interface of the gfx950 integrator kernels (host side of kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsr {

constexpr int kMaxChannels = 16;
constexpr int kTraceBlock = 256;      // threads per workgroup of the path kernel (4 waves)
constexpr int kStackShort = 16;       // LDS-staged traversal stack entries per thread
constexpr int kStackSpill = 64;       // further entries per thread in private memory (BVH depth <= 80)

struct DevAntenna {
    float pos[3];
    float rot[9];
};

// Everything one launch of the path kernel needs, passed by value.
struct TraceArgs {
    // scene (radar::RFConfig + antennas)
    float startFreq, freqSlope, adcRate, rampTime, antennaDelay;
    int32_t rxCount;
    DevAntenna tx;
    DevAntenna rx[kMaxChannels];
    // PathTracer knobs
    int32_t onlyIndirect, maxDepth, poDiffraction, doGuiding, doFiltering, filteringSphere, filterGuiding,
        primaryOrder, produceDebugImage;
    float filteringMin, filteringMax, filteringRadius;
    int32_t phaseMode;  // 0 FP64 chain (default), 1 FP32 as the reference
    // frame
    int32_t N, C;
    // work: Halton indices offset + [0, count); wave w takes [w*chunk, (w+1)*chunk)
    int64_t offset, count, chunk;
    float sampleWeight;
    int32_t isFinal;
    // mesh: BVH2 nodes (4 float4 each) and triangles (3 float4 each)
    const float4 *nodes;
    const float4 *tris;
    // per-workgroup partial frames [gridDim.x][N*C] (complex double)
    double2 *partials;
    // guiding: sampling tree (child, p0) and training tree (child, stat)
    const int32_t *gsChild;
    const float *gsP0;
    const int32_t *gtChild;
    double *gtStat;
    float uniformProb;
    int32_t secondMoment;
    int32_t guideOn;
    // optional per-path record (first recPaths paths of the launch range)
    int32_t *recHits;
    int8_t *recNee;
    int64_t recPaths;
    // counters (kCount instance): nodes, tris, closest rays, shadow rays, events
    unsigned long long *counters;
};

hipError_t launch_trace(const TraceArgs &args, int grid, bool count, hipStream_t stream);
// partials [W][M] doubles -> stage [S][M] (sum, comp) -> frame (sum, comp) += Kahan
hipError_t launch_reduce(const double *partials, int W, int M, double *stage_sum, double *stage_comp, int S,
                         double *frame_sum, double *frame_comp, hipStream_t stream);
hipError_t launch_zero(double *p, int64_t n, hipStream_t stream);
size_t trace_lds_bytes(int N, int C);
// resident workgroups of the path kernel per CU for this frame size
int trace_blocks_per_cu(int N, int C, bool count);

}  // namespace hsr

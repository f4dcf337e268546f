// Launch interface of the gfx950 integrator kernels (host side of kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace hsr {

constexpr int kMaxChannels = 16;
constexpr int kTraceBlock = 256;      // threads per workgroup of the path kernel (4 waves)
constexpr int kStackShort = 16;       // LDS-staged traversal stack entries per thread
constexpr int kStackSpill = 128;      // further entries per thread in private memory (16 + 128 =
                                      // kStackEntries of bvh.h: three per level of the 4-wide tree)

// One splat event of a cube frame, listed by the gather kernel for the LDS tile splat
// (k_cube_splat): the fields of Event plus the frame channel.
struct CubeEvent {
    int32_t center;  // sample bin of the centre tap
    float s;         // fractional sample offset (0 with delta)
    float are, aim;  // amplitude
    int32_t cc;      // chirp of the centre tap
    float sc;        // fractional chirp offset (0 with deltaC)
    int32_t flags;   // bit 0 delta, bit 1 deltaC, bits 2.. frame channel
    int32_t pad;
};
constexpr unsigned kCubeEvents = 1u << 24;  // capacity of a backend's event lists (512 MiB)

// Copies of the guiding tree's training statistics: records of one leaf from every workgroup
// would otherwise queue on one address (device atomics on one address are serialised); the
// copies are summed when the statistics are read back.
constexpr int kStatCopies = 8;

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
        primaryOrder, produceDebugImage, correctPhase, geometricalOptics;
    float filteringMin, filteringMax, filteringRadius;
    int32_t phaseMode;  // IF phase: 0 FP64 (default), 1 FP32 as the reference, 2 double-float
    int32_t refChain;   // 1: trace with the reference's FP32 geometry chain (wavefront only)
    // frame
    int32_t N, C;
    // work: Halton indices offset + [0, count); wave w takes [w*chunk, (w+1)*chunk)
    int64_t offset, count, chunk;
    float sampleWeight;
    int32_t isFinal;
    // 1: nothing reads this run's frame before it is cleared (a training iteration of the guiding
    // schedule): the wavefront gather keeps the guiding weights and records and splats nothing
    int32_t noSplat;
    // mesh: four-wide BVH nodes, quantised (4 float4 each) and with float boxes (8 float4
    // each, same node indices), and triangles (3 float4 each)
    const float4 *nodes;
    const float4 *nodesF;
    const float4 *tris;
    // per-workgroup partial frames [gridDim.x][N*C] (complex double)
    double2 *partials;
    // guiding: sampling tree (child, p0) and training tree (child, stat)
    const int32_t *gsChild;
    const float *gsP0;
    int32_t gsSize;     // nodes of the sampling tree
    const int32_t *gtChild;
    double *gtStat;     // kStatCopies copies, gtStride apart: workgroup g adds into copy g % kStatCopies
    int64_t gtStride;
    float uniformProb;
    int32_t secondMoment;
    int32_t guideOn;
    // optional per-path record (first recPaths paths of the launch range)
    int32_t *recHits;
    int8_t *recNee;
    int64_t recPaths;
    // counters (kCount instance): nodes, tris, closest rays, shadow rays, events
    unsigned long long *counters;
    // MIMO and motion
    int32_t chBase;     // frame channel of receiver 0 (transmitter index * rxCount)
    int32_t M;          // chirps per frame
    int32_t motion;     // 1: every contribution carries its radial velocity (chirp index)
    float ego[3];       // platform velocity of all antennas [m/s]
    double chirpFreq;   // RFConfig::chirpFrequency, 1 / (idleTime + rampTime) [Hz]
    const float4 *vel;  // per triangle slot velocity (xyz) or null
    double *frameAcc;   // cube: the frame sum [M][N][C] complex, accumulated with atomics
    int32_t cube;       // 1: splat into frameAcc (M > 1, or one chirp too large for an LDS tile)
    // cube: the event lists for k_cube_splat, one per receiver r of the launch (frame channel
    // chBase + r) at events + r * evCap with its count at evCount[r] (null: every event splats
    // with device atomics); events past a list's evCap are splatted with device atomics at once
    CubeEvent *events;
    unsigned *evCount;
    unsigned evCap;
    // primary visibility lists of the traced transmitter (primary.h), read by the depth-0
    // closest-hit kernel; null: depth 0 traverses the BVH
    const int32_t *pvMeta;   // [0] triangles near the transmitter (they lead pvKeys / pvSlots),
                             // [1] 1 = the lists exceeded their capacity (the BVH is used)
    const float4 *pvFaces;   // [6] face grids (u0, v0, su, sv)
    const uint32_t *pvKeys;  // triangles in key order: keys (float bits) and slots
    const int32_t *pvSlots;
    const int32_t *pvStart;  // [6 R R + 1] first entry of each cell
    const int2 *pvList;      // (slot, key bits) per cell in key order
    int32_t pvR;
};

hipError_t launch_trace(const TraceArgs &args, int grid, bool count, hipStream_t stream);

// ---------------------------------------------------------------- wavefront pipeline
// Per-depth queue counters (device memory, zeroed per batch).
// Queue sizes and consumer heads of one depth, each counter on a 256-byte line of its own:
// device-scope atomics on one address are serialised (about 12 ns apiece on MI355X), while
// atomics on distinct lines proceed in parallel.
struct alignas(256) WfCounter {
    unsigned int v;
    unsigned int pad[63];
};
constexpr int kQueueSegments = 8;  // consumer heads per queue: one per XCD of MI355X
struct WfCounts {
    WfCounter n_ext, n_shd, n_gat;
    WfCounter head_ext[kQueueSegments], head_shd[kQueueSegments];
};

// Structure-of-arrays path state and queues for a batch of up to P paths.
//  sd: 9 double fields x P   (o.xyz, d.xyz, time, r, path-length rate)
//  sf: 15 float fields x P   (H 6, guidingWeight 2, x0, x1, pdf, freq, vertex velocity 3)
//  si: 2 int fields x P      (flags: bit 0 = the path ends after its pending NEE, the path's
//                            index within the batch); a path's depth is its queue's
//  (the origin, time, length, guiding weight and motion rows are not written at depth 0)
//  ray_o / ray_d [P]: the path's current vertex rounded to float (origin of its shadow and
//  extension rays) and its extension direction
//  shadow slots [P * C]: ray direction and extent (d.xyz, tfar), NEE payload (tn, v,
//  plane-filter guiding term), status bits (1 set-up, 2 needs a visibility query, 4 add v to
//  the guiding weight, 8 splat, 16 add the plane-filter term) and the visibility (1 occluded,
//  2 visible).
struct WfArgs {
    int64_t P;
    int64_t base;   // index (relative to the run's first path) of slot 0
    int32_t n;      // paths in this batch
    double *sd;
    float *sf;
    int32_t *si;
    float4 *ray_o, *ray_d;  // the current vertex's rays: written by generation and shading, read by shadow
    const float4 *tray_o, *tray_d;  // the rays being traced (closest-hit traversal, shading, tail)
    int2 *hit;
    float4 *s_d;
    double *s_tn;
    double *s_dv;  // radial velocity of the contribution (motion only)
    float2 *s_v, *s_g1;
    int8_t *s_st, *s_vis;
    int32_t *q_ext0, *q_ext1, *q_shd, *q_gat;
    WfCounts *cnt;  // [maxDepth + 2]
    // coherence order: the residue mod kCells of the Halton indices of each cell of the
    // primary sample square, cells in Morton order (coherence_cells)
    const int32_t *cells;
    int32_t *cellStart;  // [kCells + 1] first slot of each cell in this batch
    int32_t b0mod;       // Halton index of the batch's first path (offset + base) mod kCells
    // guiding record of a training batch: each path's statistic and leaf at its slot
    double *statV;
    int32_t *statL;
};
constexpr int kCells = 125 * 243;  // cells of the coherence order: 5^3 x 3^5
// The kCells residues of the coherence order, cells in Morton order (host).
std::vector<int32_t> coherence_cells();

// The buffers of one vertex parity: a vertex's rays (its origin, the extension ray leaving it),
// NEE slots and gather and shadow queues. Vertex k uses set k & 1, so that the shading of
// vertex k (which writes vertex k + 1's set) may run beside the shadow and gather stages of
// vertex k.
struct WfSet {
    float4 *ray_o, *ray_d, *s_d;
    double *s_tn, *s_dv;
    float2 *s_v, *s_g1;
    int8_t *s_st, *s_vis;
    int32_t *q_shd, *q_gat;
};
// Bytes of device memory the wavefront buffers need for P paths and C channels.
size_t wavefront_bytes(int64_t P, int C, int maxDepth);
// Carves the buffers out of one allocation: w with the buffers of set 0 in place, both sets in
// set[2].
void wavefront_layout(WfArgs &w, WfSet set[2], void *base, int64_t P, int C, int maxDepth);
// w with the vertex buffers of set v and the traced rays of set t.
WfArgs wf_with_sets(const WfArgs &w, const WfSet &v, const WfSet &t);
// One wavefront batch's arguments: the stages receive TraceArgs and WfArgs by value (their
// pointers then stay in the global address space), the gather and set-up stages the partials.
struct BatchArgs {
    TraceArgs a;
    WfArgs w;
    WfSet set[2];
    double2 *partials;  // [gatherGrid][N*C] partial frames, accumulated by the gather stages
};
// Allows the gather stage and the fused tail the LDS of an N x C tile (beyond 64 KiB); before
// the first batch of a frame shape. noTile: the batch splats into no LDS tile (cube frames, or
// nothing splatted).
hipError_t prepare_wavefront(int N, int C, bool cube, bool noTile);
// LDS of the fused tail: the frame tile (unless noTile) and the traversal stack.
size_t tail_lds_bytes(int N, int C, bool noTile);
// One step of a batch's launch sequence: a kernel on stream 0 or 1 (its by-value parameters
// built from the batch's arguments and the step's own scalars x, after the layout sig), a
// fork (stream 1 waits for stream 0's work so far), a join (the reverse) or a timing event.
struct BatchStep {
    // kSignal: record sync event x[0] of this stream; kWait: this stream waits for sync event
    // x[0] of stream x[1] (recorded earlier in the plan)
    enum : int8_t { kKernel, kFork, kJoin, kRecord, kSignal, kWait };
    enum : int8_t { kSetup, kAW, kAWi, kGather, kTail };
    int8_t kind = kKernel, stream = 0, sig = kAW;
    int8_t vset = 0, tset = 0;  // kernels: the vertex set and the traced-ray set (WfSet) they use
    hipEvent_t event = nullptr;
    const char *name = nullptr;  // kernels: the stage's name (ROCTx ranges)
    const void *func = nullptr;
    dim3 grid, block;
    unsigned lds = 0;
    int64_t x[3] = {};
};
// A batch's launch sequence: set-up, generation, then per depth shadow / gather / extend /
// shade, with all queue sizes on the device (no host synchronisation). It depends on the
// launch shape only (paths, grids, depth, receivers, frame tile, options, timing events), not
// on the argument values. The closest-hit traversal of each depth runs on stream 1 beside the
// depth's shadow and gather stages on stream 0; the batch ends on stream 0. evClosest /
// evShadow (optional): room for 2 * (maxDepth + 1) events each, a pair recorded around every
// closest-hit / shadow traversal launch. The set-up computes the coherence order's slot
// ranges and clears the queue counters, the first partialDoubles doubles of the partials and
// (zeroStats) the training records. tailDepth K in [1, maxDepth): the fused tail (k_wf_tail)
// takes the paths from the depth-K hits on, in place of the stages from the depth-K shading on
// (its node visits and triangle tests are counted in counters[7] and [8]); otherwise (0) every
// depth has its stages.
// decouple (batches that take no guiding records): the chain of each path (closest-hit
// traversal, shading, tail) runs on stream 1 and the NEE of each vertex (shadow, gather) on
// stream 0, ordered by sync events only where they share data: a vertex's shadow stage after
// the shading that set up its slots, a shading after the gather of two vertices back (whose
// buffer set it reuses). A depth's gather then overlaps the next depth's shading and
// traversal and the tail; the tail adds its frame tiles into the partials after the gather
// stages' (from row gridGather on), and the batch ends on stream 0 once stream 1 is done.
// Without records the shading stage needs nothing the gather computes (the guiding weight
// only feeds the records).
struct BatchPlan {
    std::vector<BatchStep> steps;
    int nClosest = 0, nShadow = 0;  // traversal launches (timed pairs)
    int nSync[2] = {0, 0};          // sync events recorded on stream 0 / 1
};
BatchPlan plan_wavefront(const TraceArgs &a, const WfArgs &w, int gridTrav, int gridWork, int gridGather,
                         int64_t partialDoubles, bool zeroStats, bool count, hipEvent_t *evClosest,
                         hipEvent_t *evShadow, int tailDepth, bool decouple, bool fuseSplat = false);
// LDS (dynamic) of the fused shadow + splat stage of decoupled batches (frame tile and per-wave
// event lists; the traversal stack is static); the stage is used when it and the stack fit a
// workgroup's LDS.
size_t shadow_splat_lds_bytes(int N, int C);
// The plan launched with the arguments h on stream (0) and stream2 (1), ordered by the events
// evFork / evJoin (maxDepth + 2 each; in a decoupled plan the sync events of stream 0 and 1);
// markers: a ROCTx range around each kernel launch, named after its stage.
hipError_t launch_plan(const BatchPlan &plan, const BatchArgs &h, hipStream_t stream, hipStream_t stream2,
                       hipEvent_t *evFork, hipEvent_t *evJoin, bool markers = false);
// The plan as a graph: its kernels as kernel nodes, its timing events as event-record nodes,
// each depending on the last node of its stream (built node by node: no stream is captured, so
// other threads' HIP calls are unaffected), instantiated twice. A replay sets the kernel
// nodes' parameters of one instance to the batch's arguments and launches it; the two
// instances alternate, each updated only once its previous run has ended (done[i]).
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    int next = 0;
    std::vector<hipGraphNode_t> nodes;  // the kernel nodes, in plan order
};
hipError_t graph_plan(BatchGraph *out, const BatchPlan &plan, const BatchArgs &h);
hipError_t replay_plan(BatchGraph &g, const BatchPlan &plan, const BatchArgs &h, hipStream_t stream);
void destroy_graph(BatchGraph &g);
int wavefront_trav_blocks_per_cu(bool count);
// Training statistics of a batch: stat[leaf] += the records (statV, statL) of slots [0, n) of
// each leaf, summed in a fixed order through up to rowsCap / nodes partial rows (rows: rowsCap
// doubles) of a training tree of `nodes` nodes.
constexpr int kStatLdsNodes = 5120;  // trees up to this size are summed in LDS accumulators
constexpr int64_t kStatRows = int64_t(1) << 21;  // doubles of partial rows a backend holds
hipError_t launch_stat_reduce(const double *statV, const int32_t *statL, int64_t n, int nodes, double *rows,
                              int64_t rowsCap, double *stat, hipStream_t stream);
// LDS of one chirp's frame tile of N samples x C channels, and of a gather stage's workgroup
// (the tile).
size_t tile_lds_bytes(int N, int C);
size_t gather_lds_bytes(int N, int C);
// partials [W][M] doubles -> stage [S][M] (sum, comp) -> frame (sum, comp) += Kahan
hipError_t launch_reduce(const double *partials, int W, int M, double *stage_sum, double *stage_comp, int S,
                         double *frame_sum, double *frame_comp, hipStream_t stream);
hipError_t launch_zero(double *p, int64_t n, hipStream_t stream);
hipError_t launch_zero2(double *p, double *q, int64_t n, hipStream_t stream);  // both arrays
// Folds the kStatCopies copies of n training statistics (stride apart) into the first.
hipError_t launch_stat_fold(double *stat, int64_t stride, int64_t n, hipStream_t stream);
// The LDS tiles of a cube frame's splat, two per CU where they fit. Sample-tiled (samples): a
// tile holds every chirp of a slab of about N / T samples and 16 guard bins on each side, and
// splats the events centred in its slab, each event in one tile; used when slabs at least 16
// samples wide fit (M small enough). Chirp-tiled otherwise: a tile holds Mt chirps of every
// sample, and each event is splatted by every tile its chirp taps reach. lds = 0: no tile fits.
struct CubeSplat {
    bool samples = false;
    int T = 0;             // sample slabs per channel (sample-tiled)
    int Mt = 0;            // chirps per tile (chirp-tiled)
    int tilesPerList = 0;  // tiles per channel
    size_t lds = 0;        // LDS bytes of a tile
};
CubeSplat cube_splat_plan(int M, int N, size_t ldsPerBlock, size_t ldsPerCU);
// Doubles of the guard rows a sample-tiled splat of `lists` channels and G partial cubes holds
// after its G partial cubes (0 when chirp-tiled).
int64_t cube_guard_doubles(const CubeSplat &plan, int M, int lists, int G);
// Adds the listed events of frame channels [chBase, chBase + lists) (list r at events + r * cap,
// its count on the device at counts[r]) into the frame cube [M][N][C] through the LDS tiles of
// plan, G workgroups per tile and G partial cubes in parts (G * M * N * C complex, the listed
// channels' cells written, then the guard rows), then empties the lists.
hipError_t launch_cube_splat(const CubeSplat &plan, const CubeEvent *events, unsigned *counts, unsigned cap,
                             int chBase, int lists, int M, int N, int C, int G, double2 *parts, double *frame,
                             hipStream_t stream);
size_t trace_lds_bytes(int N, int C);

// ---------------------------------------------------------------- primary visibility lists
// Build of the lists of one transmitter position o (primary.hip); every buffer is the caller's:
// keys, keys2, slots, slots2 [T]; part [pv_prep_blocks(T) * 25]; faces [6]; meta [4]; matrix
// [chunks * 6 R R]; start [6 R R + 1]; list [cap]; sortTemp [sortBytes >= pv_sort_bytes(T)].
// R <= kPvMaxRes (the count and fill workgroups hold 6 R R counters in LDS).
constexpr int kPvMaxRes = 80;
struct PvBuild {
    const float4 *tris;
    const int32_t *triSlot;  // [T] one slot of each triangle
    int64_t T;
    float o[3];
    int R, chunks;
    int64_t cap;
    uint32_t *keys, *keys2;
    int32_t *slots, *slots2;
    float *part;
    float4 *faces;
    int32_t *meta;
    int32_t *matrix, *start;
    int2 *list;
    void *sortTemp;
    size_t sortBytes;
};
size_t pv_sort_bytes(int64_t T);
int pv_prep_blocks(int64_t T);
hipError_t launch_pv_build(const PvBuild &p, hipStream_t stream);

// ---------------------------------------------------------------- frame post-processing
// post = (sum - comp) * inv, n doubles
hipError_t launch_post_normalize(const double *sum, const double *comp, double inv, int64_t n, double *post,
                                 hipStream_t stream);
// in-place forward 3-D DFT of post [M][N][C] (complex double); tmp: a second buffer of that size
hipError_t launch_post_fft3(double2 *post, double2 *tmp, int M, int N, int C, hipStream_t stream);
// index of the largest |v| (first on ties) into *out; scratch: post_scratch_bytes()
hipError_t launch_post_argmax(const double2 *v, int64_t n, void *scratch, int64_t *out, hipStream_t stream);
size_t post_scratch_bytes();
constexpr int kPostMaxAxis = 4096;  // longest DFT axis (line + twiddles in LDS)
// Allocation-free peak search: a static device array per device holding two frame buffers of
// kNoAllocCells complex cells and the argmax scratch; noalloc_scratch returns this device's
// copy (the device must be current). Frame buffer b starts at cell b * kNoAllocCells, the
// argmax scratch at cell 2 * kNoAllocCells.
constexpr int64_t kNoAllocCells = 1 << 19;  // 8 MiB per buffer
constexpr int kNoAllocMaxDevices = 64;
hipError_t noalloc_scratch(int device, double2 **out);
// resident workgroups of the path kernel per CU for this frame size
int trace_blocks_per_cu(int N, int C, bool count);

}  // namespace hsr

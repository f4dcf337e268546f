/*
 * hussar_hip.h — C ABI of the MI355X (gfx950) backend for libhussar's Monte Carlo
 * FMCW radar integrator.
 *
 * The reference selects a backend at compile time through a two-method "Backend
 * concept": a constructor `Backend(const TriangleMesh &, Integrator &)` and
 * `void run(const Scene &, long budget, bool *interruptFlag)`
 * (libhussar/include/hussar/arch/cpu.h:17-46, libhussar/include/hussar/arch/gpu.h:17-32).
 * `PathTracer::run` (libhussar/include/hussar/integrators/path.h:171-224) drives it.
 * The entry points below are what a binding for that concept needs once the per-path
 * work (`PathTracer::sample`, path.h:235-451) runs as HIP kernels instead of as a C++
 * template executed by the backend: every input crosses as plain-old-data, every
 * buffer as a pointer plus a count, and every failure as a status code with a message
 * retrievable through hsr_last_error() (exceptions cannot cross a C ABI).
 *
 * Threading: a backend is not re-entrant (as in the reference, gpu.cpp:35-37); calls on
 * one backend must be serialised by the caller. Different backends are independent.
 */
#ifndef HUSSAR_HIP_H
#define HUSSAR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSR_ABI_VERSION 5
#define HSR_MAX_CHANNELS 16
#define HSR_MAX_TX 4

typedef enum hsr_status {
    HSR_OK = 0,
    HSR_ERR_INVALID = 1,     /* bad argument (null pointer, size out of range, ...) */
    HSR_ERR_HIP = 2,         /* a HIP runtime call failed */
    HSR_ERR_OOM = 3,         /* device memory exhausted / admission refused */
    HSR_ERR_INTERRUPTED = 4, /* the caller's interrupt flag was raised */
    HSR_ERR_UNSUPPORTED = 5, /* knob combination the backend does not implement */
    HSR_ERR_BUSY = 6         /* serving queue full */
} hsr_status;

/* radar::RFConfig (libradar/include/radar/radar.h:113-136). SI units, float like the reference. */
typedef struct hsr_rf_config {
    float startFreq;    /* [Hz] */
    float freqSlope;    /* [Hz/s] */
    float adcRate;      /* [Hz] */
    float idleTime;     /* [s] */
    float rampTime;     /* [s] */
    float antennaDelay; /* [s] */
} hsr_rf_config;

/* radar::FrameConfig (radar.h:141-164). The frame is laid out [chirp][sample][channel]
 * row-major (Frame::makeIndex, radar.h:487-494), complex interleaved (re, im). Like the
 * reference's, it is the Fourier-domain cube: a contribution lands at the fractional sample
 * index of its delay (PIndex::setTime, radar.h:232-241) and the fractional chirp index of its
 * radial velocity (PIndex::setVelocity, radar.h:248-251), spread by the separable truncated
 * Dirichlet kernel (Frame::splat, radar.h:564-627). The reference integrator splats at
 * velocity 0 (integrator.h:130); here a contribution's velocity comes from the motion of the
 * scene (hsr_scene.velocity, hsr_hip_set_velocities). chirpCount > 1 needs the wavefront
 * pipeline (HSR_OPT_PIPELINE = 1). */
typedef struct hsr_frame_config {
    int32_t chirpCount;
    int32_t samplesPerChirp;
    int32_t channelCount;
} hsr_frame_config;

/* hussar::NFAntenna with its AWRAngularDistribution (libhussar/include/hussar/core/emitter.h:159-214).
 * rotation is the 3x3 matrix, row-major: rotation[3*row + col]. Its columns are the
 * antenna's local axes in world space; local +z is the boresight. */
typedef struct hsr_antenna {
    float position[3];
    float rotation[9];
} hsr_antenna;

/* hussar::Scene (libhussar/include/hussar/core/scene.h:10-16): one RF config, one Tx and
 * rxCount receivers. The reference has exactly one transmitter and one receiver and fills
 * channel 0 only (path.h:291-292). Extensions:
 *  - rx[1..rxCount-1] feed frame channels 1..rxCount-1;
 *  - MIMO: txCount transmitters (0 or 1 = `tx` only; up to HSR_MAX_TX), tx 0 = `tx`, tx t =
 *    txMore[t-1]. Every transmitter emits the same Halton paths; the pair (tx t, rx r) feeds
 *    virtual channel t*rxCount + r, so txCount*rxCount must not exceed channelCount. The
 *    total weight counts each path index once;
 *  - velocity: the platform velocity [m/s] shared by all antennas (ego motion). */
typedef struct hsr_scene {
    hsr_rf_config rfConfig;
    hsr_antenna tx;
    hsr_antenna rx[HSR_MAX_CHANNELS];
    int32_t rxCount;
    int32_t txCount;
    hsr_antenna txMore[HSR_MAX_TX - 1];
    float velocity[3];
    int32_t reserved0;
} hsr_scene;

/* PathTracer public knobs (path.h:153-169), same names and defaults (see
 * hsr_knobs_default). useGeometricalOptics=0 (path.h:277-279, 420-427: a uniform sphere sample
 * per bounce) needs filteringSphere=0 and correctPhase=0 and at most maxDepth=3, as the
 * reference asserts (path.h:317, 344; its Halton sampler's dimensions 0-9, halton.h:58-67),
 * else HSR_ERR_INVALID; correctPhase=1 (path.h:316-329, 360-368) needs
 * filteringSphere=1, as the reference asserts (path.h:318), else HSR_ERR_INVALID; its
 * corrected values are real for real antenna patterns, so the splat skip rule
 * (integrator.h:121-124) drops them unless produceDebugImage is set, as in the reference.
 * primaryOrder fixes the unspecified evaluation order of
 * `Vector2f(get1D(), get1D())` (libhussar/include/hussar/samplers/halton.h:24-27):
 * 0 = right-to-left as g++ evaluates it (primary = (RI_5, RI_3)), 1 = left-to-right as
 * clang does (primary = (RI_3, RI_5)). */
typedef struct hsr_knobs {
    int32_t onlyIndirect;
    int32_t maxDepth;
    int32_t useGeometricalOptics;
    int32_t poDiffraction;
    int32_t doGuiding;
    int32_t clearBeforeIteration;
    int32_t correctPhase;
    int32_t filterGuiding;
    int32_t doFiltering;
    int32_t filteringSphere;
    float filteringMin;
    float filteringMax;
    float filteringRadius;
    int32_t primaryOrder;
    /* Integrator::produceDebugImage (integrator.h:58). The debug image itself is not
     * produced; the flag keeps its effect on the frame: when it is 0, contributions with a
     * zero real or imaginary part (or zero weight) are not splatted (integrator.h:121-124). */
    int32_t produceDebugImage;
} hsr_knobs;

/* Fills the reference defaults of path.h:153-169 (maxDepth 10, guiding on, ...). */
void hsr_knobs_default(hsr_knobs *knobs);

typedef struct hsr_backend hsr_backend;
typedef struct hsr_guide hsr_guide;

/* Replaces the backend constructor Backend(const TriangleMesh &, Integrator &)
 * (cpu.h:18-20 -> cpu.cpp:32-47 Embree scene build; gpu.h:18-20 -> gpu.cpp:482-493 OptiX
 * GAS build). vertices: float[3*vertexCount]; indices: int32[3*triangleCount] (the
 * TriangleMesh vertexBuffer / indexBuffer, mesh.h:16-28). The mesh is copied, a BVH is
 * built on the host and uploaded; the caller may free its buffers afterwards.
 * device: HIP device ordinal. */
hsr_status hsr_hip_create(const float *vertices, int64_t vertexCount,
                          const int32_t *indices, int64_t triangleCount,
                          int32_t device, hsr_backend **out);
void hsr_hip_destroy(hsr_backend *backend);

/* Scene motion (no reference counterpart: the reference splats at velocity 0,
 * integrator.h:130). velocities: float[3*triangleCount], the velocity [m/s] of every mesh
 * triangle in the order given to hsr_hip_create; NULL (or triangleCount 0) makes the mesh
 * static again. A contribution's radial velocity is -1/2 of the rate of change of its path
 * length, with every path vertex moving with its triangle and the antennas with the scene's
 * platform velocity. */
hsr_status hsr_hip_set_velocities(hsr_backend *backend, const float *velocities, int64_t triangleCount);

/* Integrator::configureFrame (libhussar/include/hussar/core/integrator.h:60-62). Allocates
 * the device-resident accumulation frame and clears it. */
hsr_status hsr_hip_configure_frame(hsr_backend *backend, const hsr_frame_config *config);

/* PathTracer::clearFrame (path.h:147-150): zero the frame and totalWeight. */
hsr_status hsr_hip_clear_frame(hsr_backend *backend);

/* Backend::run (cpu.h:44-46, gpu.h:26-32) with the PathTracer state the reference
 * backends read from the integrator object (path.h:138-141,458-464): the per-path
 * work of PathTracer::sample (path.h:236-451) for Halton indices
 * sampleIndexOffset .. sampleIndexOffset+budget-1, accumulated into the frame.
 * Blocking, like the reference. guide may be NULL when knobs->doGuiding is 0.
 * interruptFlag (may be NULL) is the reference's `bool *` (one byte), polled between
 * device launches; when it reads non-zero
 * the call returns HSR_ERR_INTERRUPTED, with a partial frame (cpu.h:25 semantics). */
hsr_status hsr_hip_run(hsr_backend *backend, const hsr_scene *scene, const hsr_knobs *knobs,
                       int64_t sampleIndexOffset, int64_t budget, float sampleWeight,
                       int32_t isFinalIteration, hsr_guide *guide,
                       const volatile uint8_t *interruptFlag);

/* PathTracer::fetchFrame (path.h:453-455): frame / totalWeight.
 * out_c64: float[2*sampleCount] (complex64, the reference's output type) or NULL.
 * out_c128: double[2*sampleCount], same values before rounding to float, or NULL.
 * totalWeight: receives the sum of sample weights, or NULL. */
hsr_status hsr_hip_fetch_frame(hsr_backend *backend, float *out_c64, double *out_c128,
                               double *totalWeight);

/* PathTracer::run (path.h:171-224) including the doubling guiding schedule
 * (milestones 16384 * 2^k, clearBeforeIteration, final-iteration rule). The frame must
 * have been configured. guide may be NULL when knobs->doGuiding is 0. */
hsr_status hsr_path_tracer_run(hsr_backend *backend, const hsr_scene *scene,
                               const hsr_knobs *knobs, int64_t samples, hsr_guide *guide,
                               const volatile uint8_t *interruptFlag);

/* An explicit training schedule (an extension): trainingPasses passes of trainingPaths
 * paths each (each pass restarts at Halton index 0, records into the guide and is
 * followed by a guide step), then one final pass of finalPaths paths into a cleared
 * frame, which is what hsr_hip_fetch_frame returns afterwards. */
hsr_status hsr_path_tracer_run_schedule(hsr_backend *backend, const hsr_scene *scene,
                                        const hsr_knobs *knobs, int32_t trainingPasses,
                                        int64_t trainingPaths, int64_t finalPaths,
                                        hsr_guide *guide,
                                        const volatile uint8_t *interruptFlag);

/* GuidingWrapper<BTree<2,...>> (path.h:17-131) with the settings PathTracer::setup
 * assigns (path.h:229-232: uniformProb 0.1, splitThreshold 0.005, secondMoment true).
 * The guide lives on the backend's device. */
hsr_status hsr_guide_create(hsr_backend *backend, float uniformProb, float splitThreshold,
                            int32_t secondMoment, hsr_guide **out);
/* A guide is used only while its backend lives; destroying it after the backend is allowed. */
void hsr_guide_destroy(hsr_guide *guide);
/* GuidingWrapper::reset (path.h:48-51): both trees back to a single uniform leaf. */
hsr_status hsr_guide_reset(hsr_guide *guide);
/* GuidingWrapper::step (path.h:119-126): build the training tree from its statistics,
 * copy it to the sampling tree, refine the training tree and clear its statistics. Returns
 * once the statistics are read; the new trees reach the device asynchronously, on the
 * backend's stream ahead of every later run (as do reset and set_sampling). */
hsr_status hsr_guide_step(hsr_guide *guide);

/* Inspection of a guide (no reference counterpart; for tests and persistence).
 * Sampling tree: nodeCount nodes; child[i] is 0 for a leaf or the index of the first of
 * two consecutive children; p0[i] is the probability of child 0 (inner nodes).
 * Training tree: child[] plus one double statistic per node (leaves carry data). */
hsr_status hsr_guide_sampling_size(hsr_guide *guide, int32_t *nodeCount);
hsr_status hsr_guide_get_sampling(hsr_guide *guide, int32_t *child, float *p0, int32_t capacity);
hsr_status hsr_guide_set_sampling(hsr_guide *guide, const int32_t *child, const float *p0,
                                  int32_t nodeCount);
hsr_status hsr_guide_training_size(hsr_guide *guide, int32_t *nodeCount);
hsr_status hsr_guide_get_training(hsr_guide *guide, int32_t *child, double *stat,
                                  int32_t capacity);

/* Frame::fft (libradar/include/radar/radar.h:420-435: FFTW's forward 3-D transform over
 * [chirp][sample][channel], unnormalised) of the frame (frame / totalWeight), computed on the
 * device in FP64 by direct DFTs along each axis (any axis length up to 4096). The
 * accumulation frame is unchanged. out_c64 / out_c128 as in hsr_hip_fetch_frame. */
hsr_status hsr_hip_fetch_frame_fft(hsr_backend *backend, float *out_c64, double *out_c128);

/* Frame::argmax (radar.h:534-546) and Frame::frequencyEstimation (radar.h:517-531,631-643)
 * of the frame (fourier = 0) or of its 3-D DFT (fourier = 1), the search on the device:
 * index = (chirp, sample, channel) of the largest |value| (the first one on ties), refined =
 * the interpolated position per axis (b - l/(l+m) when the left neighbour is larger, else
 * b + r/(r+m), neighbours wrapping), magnitude = |value| at index (may be NULL). */
hsr_status hsr_hip_frame_peak(hsr_backend *backend, int32_t fourier, int32_t index[3], double refined[3],
                              double *magnitude);

/* hsr_hip_frame_peak without any device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments and results (bit for bit)
 * as hsr_hip_frame_peak, but the normalised frame, its DFT's second buffer and the argmax
 * candidates live in a static device array of the library (16 MiB + 16 KiB, one copy per
 * device, held until the process ends) instead of buffers the backend allocates. Frames of
 * more than 2^19 complex cells (chirps * samples * channels) return HSR_ERR_INVALID. Calls
 * on one device share that array: calls that overlap on one device (different threads or
 * backends) race, and the caller must serialise them. */
hsr_status hsr_hip_frame_peak_noalloc(hsr_backend *backend, int32_t fourier, int32_t index[3],
                                      double refined[3], double *magnitude);

/* Statistics of the last top-level call (hsr_hip_run, or all the runs of one
 * hsr_path_tracer_run / hsr_path_tracer_run_schedule), measured with HIP events on the
 * stream the kernels run on. A guided training run returns once its guide statistics have
 * arrived, before its frame reduction ends; hsr_hip_last_stats (or the next run) waits for
 * that reduction before it adds its reduce_ms and total_ms. */
typedef struct hsr_run_stats {
    double trace_ms;        /* device time of the path-kernel launches */
    double reduce_ms;       /* device time of the frame reduction */
    double total_ms;        /* device time of the whole last run (first to last launch) */
    int64_t paths;          /* paths traced by the last run */
    int64_t launches;       /* path-kernel launches (each followed by two reduction kernels) */
    int64_t nodes_visited;  /* BVH nodes fetched (only when counters are enabled) */
    int64_t tris_tested;    /* triangles tested (only when counters are enabled) */
    int64_t rays_closest;   /* closest-hit rays traced (counters enabled) */
    int64_t rays_shadow;    /* shadow rays traced (counters enabled) */
    int64_t events;         /* frame splat events (counters enabled) */
    double bvh_build_ms;    /* host BVH build time at create */
    /* wavefront pipeline: HIP-event time and launches of the two traversal kernels, and the
     * share of nodes / triangles of the shadow (any-hit) traversal (counters enabled) */
    double closest_ms;
    double shadow_ms;
    int64_t closest_launches;
    int64_t shadow_launches;
    int64_t nodes_shadow;
    int64_t tris_shadow;
    /* wavefront batches launched by replaying a captured graph (HSR_OPT_GRAPH), and graphs
     * captured, since the statistics were last reset */
    int64_t graph_launches;
    int64_t graph_captures;
    /* the fused tail's share of nodes_visited / tris_tested (HSR_OPT_TAIL, counters enabled) */
    int64_t nodes_tail;
    int64_t tris_tail;
    /* wavefront batches by the depth their fused tail started at (0: none; 7: 7 or deeper) */
    int64_t tail_batches[8];
    /* primary visibility lists (HSR_OPT_PRIMARY) built, and their device time */
    int64_t primary_builds;
    double primary_ms;
} hsr_run_stats;

hsr_status hsr_hip_last_stats(hsr_backend *backend, hsr_run_stats *out);

/* Options (no reference counterpart):
 *  HSR_OPT_COUNTERS      1 = count nodes / triangles / rays / events in the next runs
 *                        (a separate kernel instance; deterministic counts).
 *  HSR_OPT_RECORD_HITS   >0 = record, for the first N paths of each run, the triangle
 *                        id hit at every depth and the shadow-ray outcome of every NEE;
 *                        read back with hsr_hip_get_hit_record.
 *  HSR_OPT_PATHS_PER_LAUNCH  maximum paths per kernel launch (interrupt granularity).
 *  HSR_OPT_PIPELINE      0 = fused path kernel, 1 = wavefront stages (default: path
 *                        state in device memory, persistent traversal kernels). Both
 *                        give the same hits, NEE outcomes and frame up to summation order.
 *  HSR_OPT_PHASE_MODE    IF phase of every contribution (measureRay, integrator.h:107-110):
 *                        0 = FP64 (default), 1 = FP32 as in the reference, 2 = double-float
 *                        (pairs of floats, the arithmetic of GPUs without fast FP64). */
#define HSR_OPT_COUNTERS 1
#define HSR_OPT_RECORD_HITS 2
#define HSR_OPT_PATHS_PER_LAUNCH 3
#define HSR_OPT_PIPELINE 4
#define HSR_OPT_PHASE_MODE 5
/*  HSR_OPT_TRAV_WAVES    wavefront pipeline: resident waves per CU given to each traversal
 *                        kernel (0 = as many as fit, the default); fewer waves trace more
 *                        rays each, which shortens the tail of a launch. */
#define HSR_OPT_TRAV_WAVES 6
/*  HSR_OPT_BATCH_PATHS   wavefront pipeline: paths per batch (default and maximum 2^24, at
 *                        least 1024); larger batches order their rays more coherently but hold
 *                        more path state in device memory (about 0.5 KB per path with 4
 *                        receivers). */
#define HSR_OPT_BATCH_PATHS 7
/*  HSR_OPT_CUBE_EVENTS   wavefront pipeline, frames with several chirps (or one chirp above
 *                        the LDS frame tile): capacity of the event lists whose events are
 *                        accumulated through LDS tiles, split evenly into one list per receiver
 *                        of the transmitter being traced (default 2^24 = 512 MiB in all; 0 =
 *                        none). Events beyond their receiver's list are added to the frame
 *                        directly with device atomics. */
#define HSR_OPT_CUBE_EVENTS 8
/*  HSR_OPT_CHAIN         geometry chain of the rays (wavefront pipeline): 0 = FP64 (default:
 *                        sampled and reflected directions and hit points in double, each
 *                        segment re-evaluated against the hit triangle's plane, the traversal
 *                        and shadow rays rounded from it); 1 = the reference's FP32 chain
 *                        (path.h:236-442 and cpu.cpp:9-18,73-86 in float: direction, hit point
 *                        o + t·d, face-forwarded normal, reflection, receiver legs) gives the
 *                        traversal and shadow rays and every discrete decision (hit ids,
 *                        visibility, grazing and range cut-offs, back-face cull, filter
 *                        thresholds), so hit ids follow the reference's arithmetic, while the
 *                        FP64 chain refined on the same triangles still gives the path lengths,
 *                        time, IF phase and GO weights. */
#define HSR_OPT_CHAIN 9
/*  HSR_OPT_GRAPH         wavefront pipeline: 1 = each batch's launch sequence (set-up,
 *                        generation, the per-depth traversal, shading and gather stages, with
 *                        their stream fork/join and timing events) is built into a HIP graph
 *                        (node by node, no stream capture) at the second batch of its launch
 *                        shape (paths, depth, receivers, frame tile, options) and replayed from
 *                        then on with each batch's arguments set into its kernel nodes (up to 8
 *                        shapes per backend); 0 = launched stage by stage (default). A graph that
 *                        fails to build turns the option off for the backend; hsr_run_stats
 *                        counts the replays. */
#define HSR_OPT_GRAPH 10
/*  HSR_OPT_TAIL          wavefront pipeline: K >= 1 = the paths still going after the depth-K
 *                        extension rays are finished by one fused kernel (hit update, NEE with
 *                        the shadow stage's receiver packets, splat and further bounces per
 *                        lane) instead of a stage launch per depth, when maxDepth > K; -1 =
 *                        adaptive (default): K is the first depth >= 2 whose extension rays in
 *                        the backend's last batch, scaled to the batch's paths, were at most
 *                        2^17 (2 at the first batch); 0 = every depth has its stage launches.
 *                        Results are those of the stages; hsr_run_stats counts the batches by
 *                        their tail depth. */
#define HSR_OPT_TAIL 11
/*  HSR_OPT_MARKERS       1 = the host marks its work with ROCTx ranges (rocprofv3
 *                        --marker-trace shows them): one per PathTracer::run iteration
 *                        ("hsr pass"), per wavefront batch ("hsr batch"), around each stage
 *                        launch of a batch (the kernel's name) and around the frame reduction
 *                        ("hsr reduce"); 0 = none (default). */
#define HSR_OPT_MARKERS 12
/*  HSR_OPT_PRIMARY       wavefront pipeline: 1 = the depth-0 closest hits of a run of
 *                        at least HSR_OPT_PRIMARY_MIN paths come from primary visibility lists
 *                        of the transmitter's position instead of a BVH traversal: the
 *                        directions around the transmitter cut into the cells of a cube map
 *                        (HSR_OPT_PRIMARY_RES^2 per face over the part each face sees of the
 *                        mesh), each cell listing the triangles it sees in order of a lower
 *                        bound of their distance; a ray tests its cell's triangles in that order
 *                        until the bound passes its best hit. The same triangle test and tie
 *                        rule give the same hits as the traversal. The lists are built on the
 *                        device once per transmitter position and PathTracer::run
 *                        (hsr_path_tracer_run, hsr_path_tracer_run_schedule); plain runs
 *                        (hsr_hip_run) reuse the lists of the same position since the last such
 *                        run. 0 = the BVH traversal at every depth (the default). Setting 1 is refused with
 *                        HSR_ERR_UNSUPPORTED in this build: the lists' first device run faulted
 *                        and the cause has not been found. */
#define HSR_OPT_PRIMARY 13
/*  HSR_OPT_PRIMARY_MIN   fewest paths of a run that builds or uses the lists (default 2^20). */
#define HSR_OPT_PRIMARY_MIN 14
/*  HSR_OPT_PRIMARY_RES   cells per side of each cube face of the lists, 4 to 80 (default 64). */
#define HSR_OPT_PRIMARY_RES 15
/*  HSR_OPT_PRIMARY_CAP   list entries the lists may hold (0 = the default: at least 2^23 and
 *                        64 per triangle). A build that needs more leaves the BVH traversal to
 *                        that run's depth-0 rays, and the capacity grows for the next build. */
#define HSR_OPT_PRIMARY_CAP 16
/*  HSR_OPT_SHADOW_SPLAT  wavefront pipeline: 1 (default) = in a batch that takes no guiding
 *                        records (a final pass, an unguided run) with a one-chirp frame tile, each
 *                        vertex's shadow stage also splats its visible contributions (and writes
 *                        the NEE records) as its packets finish, and no gather stage runs;
 *                        0 = a gather stage after each shadow stage, as in training passes. The
 *                        results are the same up to the order of the frame sums. */
#define HSR_OPT_SHADOW_SPLAT 17
hsr_status hsr_hip_set_option(hsr_backend *backend, int32_t option, int64_t value);

/* Hit record of the last run (see HSR_OPT_RECORD_HITS): hit_ids int32[paths*maxDepth]
 * (-1 = no hit), nee int8[paths*maxDepth*rxCount] (0 = not attempted, 1 = occluded,
 * 2 = visible, 3 = visible and splatted). */
hsr_status hsr_hip_get_hit_record(hsr_backend *backend, int32_t *hit_ids, int8_t *nee,
                                  int64_t paths, int32_t maxDepth, int32_t rxCount);

/* Device information for roofline reporting. */
typedef struct hsr_device_info {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t xcds;
    int64_t hbm_bytes;
    int32_t lds_bytes_per_cu;
    int32_t clock_mhz;
    int32_t lds_bytes_per_block;
    int32_t reserved0;
} hsr_device_info;
hsr_status hsr_hip_device_info(int32_t device, hsr_device_info *out);

/* ---------------------------------------------------------------------------------------
 * Serving runtime (no reference counterpart; the reference's only analogue is the
 * visualizer's interruptible re-render loop, visualizer/src/Simulator.h:128-183).
 *
 * A server keeps scenes resident on one device (BVH built once per scene, and a guiding tree
 * once trained) and runs frame requests on `slots` workers, each with its own HIP stream,
 * wavefront buffers, frame and guide. Requests wait in one FIFO queue; a free slot takes the
 * next one, so requests of different scenes and frame shapes run concurrently and a short
 * request does not wait for a long one to finish.
 *
 * Admission by footprint: a request needs the device bytes of its own wavefront state (its
 * batch of min(paths, pathsPerBatch) paths with its channels and depth), frame, tiles and
 * event list. It starts when a slot is free and those bytes fit in hbmBudget beside the
 * resident scenes and the other slots' buffers (a slot's buffers grow to the largest request
 * it ran; idle slots give theirs back when the head of the queue needs the room); otherwise it
 * waits, and the requests behind it wait too (FIFO), until memory frees up. A request that
 * would not fit beside the resident scenes even alone is refused at submit with HSR_ERR_OOM
 * (counted as rejected), as is one whose frame exceeds frameReserveBytes when that is set;
 * hsr_server_open_scene returns HSR_ERR_OOM for a scene that does not fit; a full queue
 * returns HSR_ERR_BUSY. All functions are thread-safe.
 *
 * Guiding (knobs.doGuiding = 1), by hsr_request.guideMode:
 *  HSR_GUIDE_PER_REQUEST  the reference's PathTracer::run (path.h:171-224) on a cleared frame
 *                         with the slot's guide reset: the doubling schedule's training
 *                         iterations, then the final iteration;
 *  HSR_GUIDE_TRAIN        the same, and the trained sampling tree becomes the scene's resident
 *                         tree (replacing any);
 *  HSR_GUIDE_SCENE        with a resident tree: the whole budget is one final iteration
 *                         (sampleWeight 1, no training) with that tree, so the result depends
 *                         only on the request and the tree, not on the mix it is served in;
 *                         while the scene has none: as HSR_GUIDE_TRAIN.
 * Training statistics are summed in a fixed order (wavefront pipeline), so the same training
 * request gives the same tree wherever it runs. */
#define HSR_GUIDE_PER_REQUEST 0
#define HSR_GUIDE_SCENE 1
#define HSR_GUIDE_TRAIN 2
typedef struct hsr_server hsr_server;

typedef struct hsr_server_config {
    int32_t device;
    int32_t slots;             /* concurrent requests; 0 = 4 */
    int32_t queueCapacity;     /* queued requests before submit returns HSR_ERR_BUSY; 0 = 4096 */
    int32_t maxDepth;          /* largest knobs.maxDepth a request may use; 0 = 16 */
    int64_t hbmBudget;         /* device bytes for slots and scenes; 0 = 90% of the free memory */
    int64_t pathsPerBatch;     /* wavefront batch per slot; 0 = 2^22 */
    int64_t frameReserveBytes; /* largest frame a request may have (sum + compensation bytes); 0 = no
                                  limit beyond the HBM budget */
    int32_t graphs;            /* 1 = the slots replay their batches' graphs (HSR_OPT_GRAPH 1);
                                  0 = they launch them stage by stage (default) */
    int32_t reserved0;
} hsr_server_config;

typedef struct hsr_request {
    int32_t sceneId;
    int32_t guideMode;         /* HSR_GUIDE_PER_REQUEST (0), HSR_GUIDE_SCENE, HSR_GUIDE_TRAIN */
    int64_t paths;             /* the samples argument of PathTracer::run */
    hsr_scene scene;
    hsr_knobs knobs;
    hsr_frame_config frame;
    int32_t reserved1;
} hsr_request;

typedef struct hsr_request_info {
    int32_t status;            /* hsr_status of the request's run */
    int32_t slot;
    double queue_ms;           /* submit -> start */
    double run_ms;             /* start -> frame copied out */
    double latency_ms;         /* submit -> done */
    double totalWeight;
} hsr_request_info;

typedef struct hsr_server_stats {
    int64_t submitted, completed, failed, queued, running;
    double p50_ms, p99_ms, mean_ms, max_ms; /* latency (submit -> done) of completed requests */
    double frames_per_s;                    /* completed / (last completion - first submission) */
    int64_t hbm_used, hbm_budget;           /* scenes + slots' buffers, and the budget */
    int32_t scenes, slots;
    int64_t rejected;                       /* submits refused by admission (HSR_ERR_OOM) */
    int64_t memory_waits;                   /* requests that waited at the head of the queue for memory */
} hsr_server_stats;

hsr_status hsr_server_create(const hsr_server_config *config, hsr_server **out);
/* Interrupts running requests, drops queued ones and frees everything. */
void hsr_server_destroy(hsr_server *server);
/* Builds the scene's BVH and makes it resident (the mesh is copied). */
hsr_status hsr_server_open_scene(hsr_server *server, const float *vertices, int64_t vertexCount,
                                 const int32_t *indices, int64_t triangleCount, int32_t *sceneId);
/* Per-triangle velocities of a resident scene (see hsr_hip_set_velocities); only while no
 * request of the scene is queued or running. */
hsr_status hsr_server_set_velocities(hsr_server *server, int32_t sceneId, const float *velocities,
                                     int64_t triangleCount);
/* Releases a scene once its queued and running requests are done; later submits fail. */
hsr_status hsr_server_close_scene(hsr_server *server, int32_t sceneId);
hsr_status hsr_server_submit(hsr_server *server, const hsr_request *request, int64_t *ticket);
/* The device bytes a request needs on a slot that holds no buffers yet: its admission
 * footprint (wavefront state of its batch, frame, tiles, event list, training rows). */
hsr_status hsr_server_request_bytes(hsr_server *server, const hsr_request *request, int64_t *bytes);
/* done = 1 once the request has finished (successfully or not). */
hsr_status hsr_server_poll(hsr_server *server, int64_t ticket, int32_t *done);
/* Blocks until the request is done, copies its frame (complex64, frame.sampleCount()
 * values; frame_c64 may be NULL) and releases the ticket. Returns the request's status. */
hsr_status hsr_server_wait(hsr_server *server, int64_t ticket, float *frame_c64, int64_t capacity,
                           hsr_request_info *info);
hsr_status hsr_server_get_stats(hsr_server *server, hsr_server_stats *out);
hsr_status hsr_server_reset_stats(hsr_server *server);

const char *hsr_last_error(void);
int32_t hsr_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* HUSSAR_HIP_H */

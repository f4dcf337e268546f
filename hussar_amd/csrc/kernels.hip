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
#include "kernels.h"
#include "device_math.h"

namespace hsr {

static constexpr double kC = 299792448.0;  // radar.h:107, 299792458f after float rounding
static constexpr float kPiF = 3.14159265358979323846f;
static constexpr float kInv4PiF = 0.07957747154594766788f;
static constexpr double kPiD = 3.14159265358979323846;
static constexpr float kEps = 5e-5f;  // hussar.h:52
static constexpr float kOneMinusEps = 0x1.fffffep-1f;

__device__ inline mat3<float> rotf(const DevAntenna &a) {
    mat3<float> m;
#pragma unroll
    for (int i = 0; i < 9; ++i) m.m[i] = a.rot[i];
    return m;
}
__device__ inline mat3<double> rotd(const DevAntenna &a) {
    mat3<double> m;
#pragma unroll
    for (int i = 0; i < 9; ++i) m.m[i] = (double)a.rot[i];
    return m;
}
__device__ inline d3 posd(const DevAntenna &a) { return {(double)a.pos[0], (double)a.pos[1], (double)a.pos[2]}; }

// ------------------------------------------------------------------------ traversal stack
struct Stack {
    int32_t *lds;  // column of this thread: lds[i * kTraceBlock]
    int sp;
    int32_t spill[kStackSpill];
    __device__ inline void push(int32_t v) {
        if (sp < kStackShort) lds[sp * kTraceBlock] = v;
        else spill[sp - kStackShort] = v;
        ++sp;
    }
    __device__ inline int32_t pop() {
        --sp;
        return sp < kStackShort ? lds[sp * kTraceBlock] : spill[sp - kStackShort];
    }
};

// BVH2 node (64 B): c0.x lo/hi, c0.y lo/hi | c1.x lo/hi, c1.y lo/hi | c0.z lo/hi, c1.z lo/hi |
// child0, child1. A child >= 0 is an inner node; < 0 encodes a leaf: bit 31 set, bits 27-30
// count-1, bits 0-26 first triangle slot.
__device__ inline void leaf_decode(int32_t c, int &first, int &count) {
    first = c & 0x07FFFFFF;
    count = ((c >> 27) & 0xF) + 1;
}

struct SlabRay {
    f3 o, invd;
    float tnear;
};

__device__ inline bool slab(float lx, float hx, float ly, float hy, float lz, float hz, const SlabRay &r,
                            float tbest, float &tentry) {
    const float ax = (lx - r.o.x) * r.invd.x, bx = (hx - r.o.x) * r.invd.x;
    const float ay = (ly - r.o.y) * r.invd.y, by = (hy - r.o.y) * r.invd.y;
    const float az = (lz - r.o.z) * r.invd.z, bz = (hz - r.o.z) * r.invd.z;
    const float t0 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), r.tnear));
    float t1 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    t1 = t1 * 1.0000004f;  // 1 + 2*gamma(3): robust traversal (Ize 2013)
    tentry = t0;
    return t0 <= t1 && t0 <= tbest;
}

template <bool kCount>
__device__ inline int32_t closest_hit(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, f3 o,
                                      f3 d, Stack &st, float &tout, int32_t &slot, uint32_t &cn, uint32_t &ct) {
    SlabRay r{o, {1.0f / d.x, 1.0f / d.y, 1.0f / d.z}, kEps};
    float best = __builtin_inff();
    int32_t best_id = -1, best_slot = -1;
    st.sp = 0;
    int32_t node = 0;
    while (true) {
        while (node >= 0) {
            if (kCount) ++cn;
            const float4 a = nodes[4 * node + 0];
            const float4 b = nodes[4 * node + 1];
            const float4 c = nodes[4 * node + 2];
            const int4 ch = __builtin_bit_cast(int4, nodes[4 * node + 3]);
            float te0, te1;
            const bool h0 = slab(a.x, a.y, a.z, a.w, c.x, c.y, r, best, te0);
            const bool h1 = slab(b.x, b.y, b.z, b.w, c.z, c.w, r, best, te1);
            if (h0 && h1) {
                const bool first0 = te0 <= te1;
                st.push(first0 ? ch.y : ch.x);
                node = first0 ? ch.x : ch.y;
            } else if (h0) {
                node = ch.x;
            } else if (h1) {
                node = ch.y;
            } else {
                if (st.sp == 0) goto done;
                node = st.pop();
            }
        }
        {
            int first, count;
            leaf_decode(node, first, count);
            for (int i = first; i < first + count; ++i) {
                if (kCount) ++ct;
                const float4 p = tris[3 * i + 0];
                const float4 e1 = tris[3 * i + 1];
                const float4 e2 = tris[3 * i + 2];
                const float t = tri_intersect({p.x, p.y, p.z}, {e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, o, d, kEps,
                                              __builtin_inff());
                const int32_t id = __float_as_int(p.w);
                if (t < __builtin_inff() && (best_id < 0 || t < best || (t == best && id < best_id))) {
                    best = t;
                    best_id = id;
                    best_slot = i;
                }
            }
        }
        if (st.sp == 0) break;
        node = st.pop();
    }
done:
    tout = best;
    slot = best_slot;
    return best_id;
}

template <bool kCount>
__device__ inline bool occluded(const float4 *__restrict__ nodes, const float4 *__restrict__ tris, f3 o, f3 d,
                                float tfar, Stack &st, uint32_t &cn, uint32_t &ct) {
    SlabRay r{o, {1.0f / d.x, 1.0f / d.y, 1.0f / d.z}, kEps};
    st.sp = 0;
    int32_t node = 0;
    while (true) {
        while (node >= 0) {
            if (kCount) ++cn;
            const float4 a = nodes[4 * node + 0];
            const float4 b = nodes[4 * node + 1];
            const float4 c = nodes[4 * node + 2];
            const int4 ch = __builtin_bit_cast(int4, nodes[4 * node + 3]);
            float te0, te1;
            const bool h0 = slab(a.x, a.y, a.z, a.w, c.x, c.y, r, tfar, te0);
            const bool h1 = slab(b.x, b.y, b.z, b.w, c.z, c.w, r, tfar, te1);
            if (h0 && h1) {
                const bool first0 = te0 <= te1;
                st.push(first0 ? ch.y : ch.x);
                node = first0 ? ch.x : ch.y;
            } else if (h0) {
                node = ch.x;
            } else if (h1) {
                node = ch.y;
            } else {
                if (st.sp == 0) return false;
                node = st.pop();
            }
        }
        int first, count;
        leaf_decode(node, first, count);
        for (int i = first; i < first + count; ++i) {
            if (kCount) ++ct;
            const float4 p = tris[3 * i + 0];
            const float4 e1 = tris[3 * i + 1];
            const float4 e2 = tris[3 * i + 2];
            if (tri_intersect({p.x, p.y, p.z}, {e1.x, e1.y, e1.z}, {e2.x, e2.y, e2.z}, o, d, kEps, tfar) <
                __builtin_inff())
                return true;
        }
        if (st.sp == 0) return false;
        node = st.pop();
    }
}

// ------------------------------------------------------------------------ guiding tree
// Node 0 covers [0,1)^2; an inner node halves its cell along axis depth%2; children are
// stored consecutively (child, child+1), child 0 the lower half.
__device__ inline float tree_pdf(const int32_t *__restrict__ child, const float *__restrict__ p0, float x0,
                                 float x1) {
    float pdf = 1.0f, lo0 = 0.0f, lo1 = 0.0f, sz0 = 1.0f, sz1 = 1.0f;
    int32_t node = 0;
    int dep = 0;
    int32_t c;
    while ((c = child[node]) != 0) {
        const float p = p0[node];
        if ((dep & 1) == 0) {
            const float mid = lo0 + sz0 * 0.5f;
            if (x0 < mid) { pdf = pdf * (2.0f * p); node = c; }
            else { pdf = pdf * (2.0f * (1.0f - p)); lo0 = mid; node = c + 1; }
            sz0 = sz0 * 0.5f;
        } else {
            const float mid = lo1 + sz1 * 0.5f;
            if (x1 < mid) { pdf = pdf * (2.0f * p); node = c; }
            else { pdf = pdf * (2.0f * (1.0f - p)); lo1 = mid; node = c + 1; }
            sz1 = sz1 * 0.5f;
        }
        ++dep;
    }
    return pdf;
}

__device__ inline float tree_sample(const int32_t *__restrict__ child, const float *__restrict__ p0, float &x0,
                                    float &x1) {
    float pdf = 1.0f, lo0 = 0.0f, lo1 = 0.0f, sz0 = 1.0f, sz1 = 1.0f;
    int32_t node = 0;
    int dep = 0;
    int32_t c;
    while ((c = child[node]) != 0) {
        const float p = p0[node];
        if ((dep & 1) == 0) {
            if (x0 < p) { x0 = fminf(x0 / p, kOneMinusEps); pdf = pdf * (2.0f * p); node = c; }
            else {
                x0 = fminf((x0 - p) / (1.0f - p), kOneMinusEps);
                pdf = pdf * (2.0f * (1.0f - p));
                lo0 = lo0 + sz0 * 0.5f;
                node = c + 1;
            }
            sz0 = sz0 * 0.5f;
        } else {
            if (x1 < p) { x1 = fminf(x1 / p, kOneMinusEps); pdf = pdf * (2.0f * p); node = c; }
            else {
                x1 = fminf((x1 - p) / (1.0f - p), kOneMinusEps);
                pdf = pdf * (2.0f * (1.0f - p));
                lo1 = lo1 + sz1 * 0.5f;
                node = c + 1;
            }
            sz1 = sz1 * 0.5f;
        }
        ++dep;
    }
    x0 = lo0 + x0 * sz0;
    x1 = lo1 + x1 * sz1;
    return pdf;
}

__device__ inline int32_t tree_leaf(const int32_t *__restrict__ child, float x0, float x1) {
    float lo0 = 0.0f, lo1 = 0.0f, sz0 = 1.0f, sz1 = 1.0f;
    int32_t node = 0;
    int dep = 0;
    int32_t c;
    while ((c = child[node]) != 0) {
        if ((dep & 1) == 0) {
            const float mid = lo0 + sz0 * 0.5f;
            if (x0 < mid) node = c;
            else { lo0 = mid; node = c + 1; }
            sz0 = sz0 * 0.5f;
        } else {
            const float mid = lo1 + sz1 * 0.5f;
            if (x1 < mid) node = c;
            else { lo1 = mid; node = c + 1; }
            sz1 = sz1 * 0.5f;
        }
        ++dep;
    }
    return node;
}

// GuidingWrapper::sample (path.h:53-82)
__device__ inline float guide_sample(const TraceArgs &a, float &x0, float &x1) {
    const float u = a.uniformProb;
    if (u == 1.0f) return 1.0f;
    float pdf = 1.0f - u;
    if (x0 < u) {
        x0 /= u;
        pdf *= tree_pdf(a.gsChild, a.gsP0, x0, x1);
    } else {
        x0 -= u;
        x0 /= 1.0f - u;
        const float gpdf = tree_sample(a.gsChild, a.gsP0, x0, x1);
        pdf *= gpdf;
    }
    pdf += u;
    return pdf;
}

// ------------------------------------------------------------------------ path state
struct PathState {
    d3 o, d;        // FP64 chain ray
    d3 n_in, d_in;  // normal and direction at the last hit (surface.incoming)
    cf3 H, H_in;    // field of the ray and at the last hit
    double time;    // path time [s]
    double r;       // path length [m]
    double cosTheta;
    cf gw;          // guidingWeight
    float x0, x1;   // primary sample after the guiding warp
    float pdf;      // primaryPdf
    float freq;
    int depth;
    int64_t local;  // index - offset
};

__device__ inline void generate(const TraceArgs &a, uint64_t index, PathState &p) {
    const float u_f = radical_inverse_2(index);
    const float u_a = radical_inverse_b<3>(index);
    const float u_b = radical_inverse_b<5>(index);
    float x0 = a.primaryOrder == 0 ? u_b : u_a;
    float x1 = a.primaryOrder == 0 ? u_a : u_b;
    float pdf = 1.0f;
    if (a.doGuiding && a.guideOn) pdf = guide_sample(a, x0, x1);
    const d3 dl = uniform_sample_sphere((double)x0, (double)x1);
    cf3 Hl = awr_evaluate(cast3<float>(dl));
    Hl = Hl / kInv4PiF;
    p.o = posd(a.tx);
    p.d = mulv(rotd(a.tx), dl);
    p.H = (1.0f / pdf) * mulv(rotf(a.tx), Hl);
    p.freq = a.startFreq + u_f * (a.freqSlope * a.rampTime);
    p.time = 0.0;
    p.r = 0.0;
    p.cosTheta = 1.0;
    p.gw = {0.0f, 0.0f};
    p.x0 = x0;
    p.x1 = x1;
    p.pdf = pdf;
    p.depth = 0;
}

// One NEE event ready for the frame: taps value·coef at bins center+k, coef = 1/(s-k), or a
// single tap at center (delta).
struct Event {
    int center;
    float s;
    cf A;
    int delta;
};

// Integrator::splat (integrator.h:113-135) up to the taps: fractional bin (setTime,
// radar.h:232-241), IF phase (measureRay, integrator.h:107-110) and the splat's own
// e^{jπs}·sin(πs)/π factor (radar.h:570-583).
__device__ inline Event make_event(const TraceArgs &a, double tn, cf v, float weight) {
    Event e;
    const double dt = tn + (double)a.antennaDelay;
    double x = 0.0;
    if (a.freqSlope != 0.0f) {
        const double q = (dt * (double)a.freqSlope) / (double)a.adcRate;
        x = (double)a.N * (q - floor(q));
    }
    cf ph;
    if (a.phaseMode == 1) {
        // the reference's FP32 evaluation of 2π(f0 - Δt S/2)Δt
        const float dtf = (float)dt;
        const float arg = ((2.0f * kPiF) * (a.startFreq - dtf * a.freqSlope / 2.0f)) * dtf;
        ph = {cosf(arg), sinf(arg)};
    } else {
        const double turns = ((double)a.startFreq - dt * (double)a.freqSlope / 2.0) * dt;
        double s, c;
        sincos_d((2.0 * kPiD) * (turns - floor(turns)), &s, &c);
        ph = {(float)c, (float)s};
    }
    cf value = weight * (v * ph);
    const int center = (int)round(x);
    const double s = x - (double)center;
    e.center = center;
    if (fabs(s) < 1e-4) {
        e.delta = 1;
        e.s = 0.0f;
        e.A = value;
        return e;
    }
    double ss, cs;
    sincos_d(kPiD * s, &ss, &cs);
    value = value * cf{(float)cs, (float)ss};
    const float w0 = (float)(ss / kPiD);
    e.delta = 0;
    e.s = (float)s;
    e.A = value * w0;
    return e;
}

__device__ inline int wrap_bin(int v, int N) {
    if (N >= 33) {
        if (v < 0) v += N;
        else if (v >= N) v -= N;
        return v;
    }
    v %= N;
    return v < 0 ? v + N : v;
}

// Wave-cooperative splat of this lane's event (if any) into the workgroup's LDS tile.
__device__ inline void flush_events(bool has, const Event &e, int ch, double *tile, int N, int C, int lane) {
    uint64_t m = __ballot(has);
    while (m) {
        const int src = __builtin_ctzll(m);
        m &= m - 1;
        const int center = __builtin_amdgcn_readlane(e.center, src);
        const float s = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.s), src));
        const float Are = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.A.re), src));
        const float Aim = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(e.A.im), src));
        const int delta = __builtin_amdgcn_readlane(e.delta, src);
        if (delta) {
            if (lane == 0) {
                const int b = wrap_bin(center, N);
                atomicAdd(&tile[2 * (b * C + ch) + 0], (double)Are);
                atomicAdd(&tile[2 * (b * C + ch) + 1], (double)Aim);
            }
        } else if (lane < 33) {
            const int k = lane - 16;
            const int b = wrap_bin(center + k, N);
            const float coef = 1.0f / (s - (float)k);
            atomicAdd(&tile[2 * (b * C + ch) + 0], (double)(Are * coef));
            atomicAdd(&tile[2 * (b * C + ch) + 1], (double)(Aim * coef));
        }
    }
}

// Wave-aggregated FP64 add into the training statistics: lanes with the same leaf are summed
// with a butterfly and one lane issues the global atomic.
__device__ inline void record_stats(bool pending, int leaf, double val, double *stat, int lane) {
    uint64_t m = __ballot(pending);
    while (m) {
        const int src = __builtin_ctzll(m);
        const int L = __builtin_amdgcn_readlane(leaf, src);
        const bool mine = pending && leaf == L;
        double v = mine ? val : 0.0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == src) atomicAdd(&stat[L], v);
        m &= ~__ballot(mine);
        pending = pending && !mine;
    }
}

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

    Stack st;
    st.lds = lds_stack + tid;
    st.sp = 0;
    uint32_t cn = 0, ct = 0, crc = 0, crs = 0, cev = 0;

    const int64_t wave = (int64_t)blockIdx.x * (kTraceBlock / 64) + (tid >> 6);
    const int64_t begin = wave * a.chunk;
    const int64_t end = min(a.count, begin + a.chunk);
    int64_t next = begin;  // wave-uniform
    const int nrx = min(a.rxCount, a.C);
    const double maxDist = ((double)a.adcRate / (double)a.freqSlope) * kC;
    const float fourPi = 4.0f * kPiF;

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
                generate(a, (uint64_t)(a.offset + mine), p);
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
                const DevAntenna &rxa = a.rx[ch];
                d3 dn = posd(rxa) - p.o;
                const double rr = norm(dn);
                dn = dn / rr;
                bool vis = true;
                if (!a.poDiffraction) {
                    if (kCount) ++crs;
                    vis = !occluded<kCount>(a.nodes, a.tris, cast3<float>(p.o), cast3<float>(dn), (float)rr, st, cn,
                                            ct);
                }
                int8_t flag = 1;
                if (vis) {
                    flag = 2;
                    const cf3 Hrx = mulv(rotf(rxa), awr_evaluate(cast3<float>(mulTv(rotd(rxa), -dn))));
                    const float rrF = (float)rr;
                    cf3 Hn;
                    if (p.depth == 0) {
                        Hn = mulv(rotf(a.tx), awr_evaluate(cast3<float>(mulTv(rotd(a.tx), dn))));
                        Hn = (1.0f / (fourPi * rrF)) * Hn;
                    } else {
                        const cf3 J = 2.0f * cross(cast3<float>(p.n_in), p.H_in);
                        Hn = cross(cast3<float>(dn), J);
                        if (dot(dn, p.n_in) < 0.0) Hn = cf3{};
                        const float k0 = ((2.0f * kPiF) * p.freq) / 299792448.0f;
                        const cf green = cf{1.0f / fmaxf(rrF, 1e-3f), k0} / (fourPi * rrF);
                        Hn = Hn * green;
                    }
                    const double tn = p.time + rr / kC;
                    cf v = cdot(Hn, Hrx);
                    bool keep = true;
                    if (p.depth > 0) {
                        v = v * (float)(p.r / p.cosTheta);
                        v = v / fourPi;
                        if (p.cosTheta < 1e-3) keep = false;
                        double dphase;
                        if (a.filteringSphere) {
                            const d3 rxPos = p.o + rr * dn;
                            const double lam = fmax(0.0, dot(p.d, rxPos - p.o));
                            dphase = norm((p.o + lam * p.d) - rxPos);
                        } else {
                            const double c = dot(normalized(dn - p.d_in), p.n_in);
                            dphase = p.r * sqrt(1.0 - c * c) / c;
                        }
                        dphase = dphase / (kC / (double)p.freq);
                        if (a.doFiltering) {
                            if (a.filteringSphere) {
                                const double R = (double)a.filteringRadius;
                                if (dphase > R) {
                                    const double f = 1.0 - 0.20 * (dphase / R - 1.0);
                                    v = v * (float)fmax(f, 0.0);
                                }
                            } else {
                                const double fmin_ = a.filteringMin, fmax_ = a.filteringMax;
                                if (dphase > fmax_) {
                                    if (a.filterGuiding) {
                                        const double q = dphase / fmax_;
                                        p.gw = p.gw + v / (float)(q * q + 1.0);
                                    }
                                    v = cf{0.0f, 0.0f};
                                }
                                if (dphase > fmin_) v = v * (float)((fmax_ - dphase) / (fmax_ - fmin_));
                            }
                        }
                        if (keep) p.gw = p.gw + v;
                    }
                    // splat skip rule (integrator.h:121-124); with the debug image on every
                    // contribution is splatted (an exactly zero one adds nothing and is skipped)
                    const bool skip = a.produceDebugImage
                                          ? (a.sampleWeight == 0.0f || (v.re == 0.0f && v.im == 0.0f))
                                          : (a.sampleWeight == 0.0f || v.re == 0.0f || v.im == 0.0f);
                    if (keep && !skip) {
                        ev = make_event(a, tn, v, a.sampleWeight);
                        has = true;
                        flag = 3;
                        if (kCount) ++cev;
                    }
                }
                if (rec && p.depth >= 1) a.recNee[(p.local * a.maxDepth + (p.depth - 1)) * a.C + ch] = flag;
            }
            flush_events(has, ev, ch, tile, a.N, a.C, lane);
        }

        // ---- extend (path.h:399-442)
        bool finished = false;
        if (active) {
            if (p.depth >= a.maxDepth || p.r >= maxDist || is_zero(p.H)) {
                finished = true;
            } else {
                if (kCount) ++crc;
                float t;
                int32_t slot;
                const int32_t id = closest_hit<kCount>(a.nodes, a.tris, cast3<float>(p.o), cast3<float>(p.d), st, t,
                                                       slot, cn, ct);
                if (rec) a.recHits[p.local * a.maxDepth + p.depth] = id;
                if (id < 0) {
                    finished = true;
                } else {
                    // geometric normal (kernel.cu:86) and, for the FP64 chain, the distance to
                    // the hit triangle's plane, both from the float edges promoted to double
                    const float4 P = a.tris[3 * slot + 0];
                    const float4 E1 = a.tris[3 * slot + 1];
                    const float4 E2 = a.tris[3 * slot + 2];
                    const d3 e1 = {(double)E1.x, (double)E1.y, (double)E1.z};
                    const d3 e2 = {(double)E2.x, (double)E2.y, (double)E2.z};
                    const d3 v0 = {(double)P.x, (double)P.y, (double)P.z};
                    const d3 nc = cross(e1, e2);
                    d3 n = normalized(nc);
                    if (dot(n, p.d) > 0.0) n = -n;  // face-forward (cpu.cpp:104-107)
                    const double cosT = fabs(dot(n, -p.d));
                    if (cosT < 1e-3) {
                        finished = true;
                    } else {
                        const double tg = dot(nc, v0 - p.o) / dot(nc, p.d);
                        p.cosTheta = cosT;
                        p.r = p.r + tg;
                        p.time = p.time + tg / kC;
                        p.n_in = n;
                        p.H_in = p.H;
                        p.d_in = p.d;
                        p.o = p.o + tg * p.d;
                        p.depth++;
                        // reflectRay (integrator.h:164-168), Intersection::R (intersection.h:38)
                        const d3 wi = -p.d;
                        p.d = (2.0 * dot(n, wi)) * n - wi;
                        const f3 nF = cast3<float>(n);
                        p.H = rscale(2.0f * nF, rdot(nF, p.H)) - p.H;
                    }
                }
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
        if (a.doGuiding && a.guideOn && !a.isFinal) record_stats(pend, leaf, val, a.gtStat, lane);
    }

    // ---- store the workgroup's frame tile
    __syncthreads();
    double2 *dst = a.partials + (size_t)blockIdx.x * NC;
    for (int i = tid; i < NC; i += kTraceBlock) dst[i] = make_double2(tile[2 * i], tile[2 * i + 1]);
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

hipError_t launch_zero(double *p, int64_t n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    (void)hipGetLastError();
    const int bs = 256;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace hsr

// Frame post-processing on the device: radar::Frame::fft (libradar/include/radar/radar.h:420-435,
// FFTW forward 3-D transform over [chirp][sample][channel], unnormalised) and the peak search
// of Frame::argmax (radar.h:534-546).
//
// The transform is done in FP64 as one direct DFT per axis (any axis length, not only powers
// of two: the frame sizes of the serving mix include prime bin counts). A workgroup owns one
// line at a time: the line and the axis' twiddle table e^{-2πi j/L} sit in LDS, and output k
// sums x[n]·T[(k·n) mod L] with the table index advanced by k modulo L (no multiplications or
// divisions in the loop, no accumulated angle error). The frame is small (at most 2^28
// doubles by configuration, typically a few MB), so each axis pass is bound by the L-fold
// re-reads of the line in LDS, not by HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <utility>

#include "device_math.h"
#include "kernels.h"

namespace hsr {

constexpr int kPostBlock = 256;
constexpr double kTwoPi = 6.28318530717958647692;

// post[i] = (sum[i] - comp[i]) * inv: the Kahan-compensated frame divided by the total weight.
__global__ void k_post_normalize(const double *__restrict__ sum, const double *__restrict__ comp, double inv,
                                 int64_t n, double *__restrict__ post) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) post[i] = (sum[i] - comp[i]) * inv;
}

// DFT along one axis of the [A][L][B] view of the cube (L the axis length, A the product of
// the outer axes, B of the inner ones): line (a, b) holds elements (a*L + n)*B + b.
__global__ __launch_bounds__(kPostBlock) void k_post_dft(const double2 *__restrict__ in, double2 *__restrict__ out,
                                                         int64_t A, int L, int64_t B) {
    extern __shared__ double2 sm[];
    double2 *tw = sm;       // [L] e^{-2πi j/L}
    double2 *line = sm + L;  // [L]
    for (int j = threadIdx.x; j < L; j += kPostBlock) {
        double s, c;
        sincos_d(kTwoPi * (double)j / (double)L, &s, &c);
        tw[j] = make_double2(c, -s);
    }
    const int64_t lines = A * B;
    for (int64_t ln = blockIdx.x; ln < lines; ln += gridDim.x) {
        const int64_t a = ln / B, b = ln % B;
        const int64_t base = a * (int64_t)L * B + b;
        __syncthreads();
        for (int n = threadIdx.x; n < L; n += kPostBlock) line[n] = in[base + (int64_t)n * B];
        __syncthreads();
        for (int k = threadIdx.x; k < L; k += kPostBlock) {
            double re = 0.0, im = 0.0;
            int j = 0;
            for (int n = 0; n < L; ++n) {
                const double2 x = line[n], t = tw[j];
                re += x.x * t.x - x.y * t.y;
                im += x.x * t.y + x.y * t.x;
                j += k;
                if (j >= L) j -= L;
            }
            out[base + (int64_t)k * B] = make_double2(re, im);
        }
    }
}

// Argmax of |value| with the first index winning ties: per-workgroup candidates, then one
// workgroup over the candidates. |v|^2 orders like |v|.
struct Cand {
    double m;
    int64_t i;
};

__device__ inline Cand better(Cand x, Cand y) { return (y.m > x.m || (y.m == x.m && y.i < x.i)) ? y : x; }

__device__ inline Cand block_best(Cand c) {
    __shared__ double sm[kPostBlock];
    __shared__ int64_t si[kPostBlock];
    sm[threadIdx.x] = c.m;
    si[threadIdx.x] = c.i;
    __syncthreads();
    for (int off = kPostBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const Cand o = better(Cand{sm[threadIdx.x], si[threadIdx.x]},
                                  Cand{sm[threadIdx.x + off], si[threadIdx.x + off]});
            sm[threadIdx.x] = o.m;
            si[threadIdx.x] = o.i;
        }
        __syncthreads();
    }
    return Cand{sm[0], si[0]};
}

__global__ __launch_bounds__(kPostBlock) void k_post_argmax(const double2 *__restrict__ v, int64_t n,
                                                            double *__restrict__ cm, int64_t *__restrict__ ci) {
    Cand best{-1.0, 0x7fffffffffffffffLL};
    for (int64_t i = (int64_t)blockIdx.x * kPostBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kPostBlock) {
        const double2 x = v[i];
        best = better(best, Cand{x.x * x.x + x.y * x.y, i});
    }
    const Cand b = block_best(best);
    if (threadIdx.x == 0) {
        cm[blockIdx.x] = b.m;
        ci[blockIdx.x] = b.i;
    }
}

__global__ __launch_bounds__(kPostBlock) void k_post_argmax_final(const double *__restrict__ cm,
                                                                  const int64_t *__restrict__ ci, int n,
                                                                  int64_t *__restrict__ out) {
    Cand best{-1.0, 0x7fffffffffffffffLL};
    for (int i = threadIdx.x; i < n; i += kPostBlock) best = better(best, Cand{cm[i], ci[i]});
    const Cand b = block_best(best);
    if (threadIdx.x == 0) out[0] = b.i;
}

size_t post_scratch_bytes() { return 1024 * (sizeof(double) + sizeof(int64_t)) + sizeof(int64_t); }

// Scratch of the allocation-free peak search (hsr_hip_frame_peak_noalloc): two frame buffers
// of kNoAllocCells complex cells each (the normalised frame and the DFT's second buffer), then
// the argmax candidates and result (post_scratch_bytes()). Every device the library runs on
// holds its own copy (16 MiB + 16 KiB) until the process ends. Nothing in it needs to start at
// zero: each call writes every element it reads. Calls on one device share it, so calls that
// overlap on one device (from different threads or backends) race.
constexpr int64_t kNoAllocArgmaxCells = 1025;  // >= post_scratch_bytes() / sizeof(double2)
static __device__ __align__(256) double2 scratch[2 * kNoAllocCells + kNoAllocArgmaxCells];

// The device address of this device's copy of `scratch`, looked up once per device with that
// device current (the array's name is a host-side stand-in, never a device address).
hipError_t noalloc_scratch(int device, double2 **out) {
    static std::mutex mtx;
    static double2 *ptrs[kNoAllocMaxDevices] = {};
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    static_assert(kNoAllocArgmaxCells * sizeof(double2) >= 1024 * (sizeof(double) + sizeof(int64_t)) + sizeof(int64_t),
                  "argmax scratch too small");
    std::lock_guard<std::mutex> lock(mtx);
    if (!ptrs[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        ptrs[device] = (double2 *)p;
    }
    *out = ptrs[device];
    return hipSuccess;
}

hipError_t launch_post_normalize(const double *sum, const double *comp, double inv, int64_t n, double *post,
                                 hipStream_t s) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_post_normalize, dim3((unsigned)((n + kPostBlock - 1) / kPostBlock)), dim3(kPostBlock), 0, s,
                       sum, comp, inv, n, post);
    return hipGetLastError();
}

// 3-D forward DFT of post [M][N][C] (complex double) in place, using tmp as the second buffer.
hipError_t launch_post_fft3(double2 *post, double2 *tmp, int M, int N, int C, hipStream_t s) {
    (void)hipGetLastError();
    const int dims[3] = {M, N, C};
    double2 *src = post, *dst = tmp;
    for (int ax = 0; ax < 3; ++ax) {
        const int L = dims[ax];
        if (L == 1) continue;  // a one-point DFT is the identity
        int64_t A = 1, B = 1;
        for (int k = 0; k < ax; ++k) A *= dims[k];
        for (int k = ax + 1; k < 3; ++k) B *= dims[k];
        const int64_t lines = A * B;
        const int grid = (int)std::min<int64_t>(lines, 4096);
        const size_t lds = 2 * (size_t)L * sizeof(double2);
        if (lds > 65536) {
            const hipError_t e =
                hipFuncSetAttribute((const void *)k_post_dft, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_post_dft, dim3(grid), dim3(kPostBlock), lds, s, src, dst, A, L, B);
        std::swap(src, dst);
    }
    if (src != post)
        return hipMemcpyAsync(post, src, (size_t)M * N * C * sizeof(double2), hipMemcpyDeviceToDevice, s);
    return hipGetLastError();
}

hipError_t launch_post_argmax(const double2 *v, int64_t n, void *scratch, int64_t *out, hipStream_t s) {
    (void)hipGetLastError();
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + kPostBlock - 1) / kPostBlock));
    double *cm = (double *)scratch;
    int64_t *ci = (int64_t *)(cm + 1024);
    hipLaunchKernelGGL(k_post_argmax, dim3(grid), dim3(kPostBlock), 0, s, v, n, cm, ci);
    hipLaunchKernelGGL(k_post_argmax_final, dim3(1), dim3(kPostBlock), 0, s, cm, ci, grid, out);
    return hipGetLastError();
}

}  // namespace hsr

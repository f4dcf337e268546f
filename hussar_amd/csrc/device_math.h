// Device-side math for the gfx950 radar integrator: small vector / complex types with a
// fixed evaluation order, the Halton radical inverse, the reproducible sin/cos, the AWR
// antenna pattern and the Möller–Trumbore triangle test.
//
// Everything that feeds BVH traversal (sampled and reflected directions, hit points, NEE
// legs, the triangle test) is evaluated in a fixed order with no fused multiply-add
// (this translation unit is compiled with -ffp-contract=off), so that the host restatement
// in the same order finds the same triangles bit for bit. Sums of three terms are
// (a0 + a1) + a2; a matrix-vector product is three such row sums.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsr {

// ----------------------------------------------------------------- vectors (float, double)
template <class T> struct vec3 {
    T x, y, z;
};
using f3 = vec3<float>;
using d3 = vec3<double>;

template <class T> __host__ __device__ inline vec3<T> mk3(T x, T y, T z) { return vec3<T>{x, y, z}; }
template <class A, class B> __host__ __device__ inline vec3<A> cast3(vec3<B> v) {
    return vec3<A>{(A)v.x, (A)v.y, (A)v.z};
}
template <class T> __host__ __device__ inline vec3<T> operator+(vec3<T> a, vec3<T> b) {
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T> __host__ __device__ inline vec3<T> operator-(vec3<T> a, vec3<T> b) {
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T> __host__ __device__ inline vec3<T> operator-(vec3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> __host__ __device__ inline vec3<T> operator*(T s, vec3<T> a) {
    return {s * a.x, s * a.y, s * a.z};
}
template <class T> __host__ __device__ inline vec3<T> operator/(vec3<T> a, T s) {
    return {a.x / s, a.y / s, a.z / s};
}
template <class T> __host__ __device__ inline T dot(vec3<T> a, vec3<T> b) {
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
template <class T> __host__ __device__ inline vec3<T> cross(vec3<T> a, vec3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> __host__ __device__ inline T norm(vec3<T> a) { return sqrt(dot(a, a)); }
template <class T> __host__ __device__ inline vec3<T> normalized(vec3<T> a) {
    const T n2 = dot(a, a);
    return n2 > T(0) ? a / (T)sqrt(n2) : a;
}

// 3x3 row-major matrix products: row sums ((m0 x + m1 y) + m2 z)
template <class T> struct mat3 {
    T m[9];
};
template <class T, class V> __host__ __device__ inline V mul(const mat3<T> &M, V v);
template <class T> __host__ __device__ inline vec3<T> mulv(const mat3<T> &M, vec3<T> v) {
    return {(M.m[0] * v.x + M.m[1] * v.y) + M.m[2] * v.z, (M.m[3] * v.x + M.m[4] * v.y) + M.m[5] * v.z,
            (M.m[6] * v.x + M.m[7] * v.y) + M.m[8] * v.z};
}
template <class T> __host__ __device__ inline vec3<T> mulTv(const mat3<T> &M, vec3<T> v) {
    return {(M.m[0] * v.x + M.m[3] * v.y) + M.m[6] * v.z, (M.m[1] * v.x + M.m[4] * v.y) + M.m[7] * v.z,
            (M.m[2] * v.x + M.m[5] * v.y) + M.m[8] * v.z};
}

// ----------------------------------------------------------------- complex float
// radar::complex<float> (libradar/include/radar/complex.h:31-265)
struct cf {
    float re, im;
};
__host__ __device__ inline cf operator+(cf a, cf b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cf operator-(cf a, cf b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cf operator*(cf a, cf b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cf operator*(float s, cf a) { return {s * a.re, s * a.im}; }
__host__ __device__ inline cf operator*(cf a, float s) { return {a.re * s, a.im * s}; }
__host__ __device__ inline cf operator/(cf a, float s) { return {a.re / s, a.im / s}; }
__host__ __device__ inline cf conjf_(cf a) { return {a.re, -a.im}; }
// |z| as the scaled hypot of complex.h:225-237
__host__ __device__ inline float cabsf_(cf z) {
    float x = fabsf(z.re), y = fabsf(z.im);
    const float s = x < y ? y : x;
    if (s == 0.0f) return s;
    x /= s;
    y /= s;
    return s * sqrtf(x * x + y * y);
}

struct cf3 {
    cf x, y, z;
};
__host__ __device__ inline cf3 operator-(cf3 a, cf3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline cf3 operator*(float s, cf3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline cf3 operator*(cf3 a, cf s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline cf3 operator/(cf3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// real × complex cross product (Vector3f.cross(Vector3c))
__host__ __device__ inline cf3 cross(f3 a, cf3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// real · complex (conjugating the real operand is a no-op)
__host__ __device__ inline cf rdot(f3 a, cf3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// Eigen's complex dot: Σ conj(a_i) b_i (geometry.h:138 measureH)
__host__ __device__ inline cf cdot(cf3 a, cf3 b) {
    return (conjf_(a.x) * b.x + conjf_(a.y) * b.y) + conjf_(a.z) * b.z;
}
__host__ __device__ inline cf3 rscale(f3 v, cf c) { return {v.x * c, v.y * c, v.z * c}; }
__host__ __device__ inline cf3 mulv(const mat3<float> &M, cf3 v) {
    return {(M.m[0] * v.x + M.m[1] * v.y) + M.m[2] * v.z, (M.m[3] * v.x + M.m[4] * v.y) + M.m[5] * v.z,
            (M.m[6] * v.x + M.m[7] * v.y) + M.m[8] * v.z};
}
// Eigen isZero(1e-20) on complex entries: |z|^2 <= (1e-20)^2
__host__ __device__ inline bool is_zero(cf3 h) {
    const float p = 1e-20f;
    const float thr = p * p;
    return (h.x.re * h.x.re + h.x.im * h.x.im <= thr) && (h.y.re * h.y.re + h.y.im * h.y.im <= thr) &&
           (h.z.re * h.z.re + h.z.im * h.z.im <= thr);
}

// ----------------------------------------------------------------- Halton (halton.h:38-54)
// base 2: the reversed digits are the bit-reversed index; invBaseN = 2^-digits exactly.
__device__ inline float radical_inverse_2(uint64_t v) {
    if (v == 0) return 0.0f;
    const int n = 64 - __clzll((long long)v);
    const uint64_t rev = __brevll(v) >> (64 - n);
    const float inv = __int_as_float((127 - n) << 23);  // 2^-n, n <= 64
    return fminf((float)rev * inv, 0x1.fffffep-1f);
}
// other bases: the digit loop with 32-bit division by a constant while the index fits
template <uint32_t B> __device__ inline float radical_inverse_b(uint64_t v) {
    const float invBase = 1.0f / float(B);
    uint64_t rev = 0;
    float invBaseN = 1.0f;
    while (v >> 32) {
        const uint64_t next = v / B;
        rev = rev * B + (v - next * B);
        invBaseN *= invBase;
        v = next;
    }
    uint32_t w = (uint32_t)v;
    while (w) {
        const uint32_t next = w / B;
        rev = rev * B + (w - next * B);
        invBaseN *= invBase;
        w = next;
    }
    return fminf((float)rev * invBaseN, 0x1.fffffep-1f);
}

// HaltonSampler::get1D of dimension dim >= 3 (halton.h:20-22, 56-67: bases 7 to 29 for
// dimensions 3 to 9, the reference's last). Dimension 10 and up is the reference's Assert
// (halton.h:1082-1084); the host refuses the knobs that would reach it, and 0 is returned.
__device__ inline float halton_dim(uint64_t v, int dim) {
    switch (dim) {
        case 3: return radical_inverse_b<7>(v);
        case 4: return radical_inverse_b<11>(v);
        case 5: return radical_inverse_b<13>(v);
        case 6: return radical_inverse_b<17>(v);
        case 7: return radical_inverse_b<19>(v);
        case 8: return radical_inverse_b<23>(v);
        case 9: return radical_inverse_b<29>(v);
        default: return 0.0f;
    }
}

// ----------------------------------------------------------------- sin/cos
// Cody–Waite reduction by pi/2 (fdlibm's two-part pio2_1 / pio2_1t, valid for |x| < 2^20)
// and Taylor polynomials of degree 17/18 on |r| <= pi/4, in double without fused
// multiply-add: identical bits on host and device.
__host__ __device__ inline void sincos_d(double x, double *s_out, double *c_out) {
    const double k = rint(x * 0.63661977236758134308);
    const double r = (x - k * 1.57079632673412561417e+00) - k * 6.07710050650619224932e-11;
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;
    ps = ps * r2 - 1.0 / 1307674368000.0;
    ps = ps * r2 + 1.0 / 6227020800.0;
    ps = ps * r2 - 1.0 / 39916800.0;
    ps = ps * r2 + 1.0 / 362880.0;
    ps = ps * r2 - 1.0 / 5040.0;
    ps = ps * r2 + 1.0 / 120.0;
    ps = ps * r2 - 1.0 / 6.0;
    const double sr = r + (r * r2) * ps;
    double pc = -1.0 / 6402373705728000.0;
    pc = pc * r2 + 1.0 / 20922789888000.0;
    pc = pc * r2 - 1.0 / 87178291200.0;
    pc = pc * r2 + 1.0 / 479001600.0;
    pc = pc * r2 - 1.0 / 3628800.0;
    pc = pc * r2 + 1.0 / 40320.0;
    pc = pc * r2 - 1.0 / 720.0;
    pc = pc * r2 + 1.0 / 24.0;
    pc = pc * r2 - 0.5;
    const double cr = 1.0 + r2 * pc;
    const int q = (int)((long long)k & 3);
    double s = sr, c = cr;
    if (q == 1) { s = cr; c = -sr; }
    else if (q == 2) { s = -sr; c = -cr; }
    else if (q == 3) { s = -cr; c = sr; }
    *s_out = s;
    *c_out = c;
}

// ----------------------------------------------------------------- double-float
// An unevaluated pair hi + lo of floats carrying about 48 significant bits: Knuth's two-sum,
// the fused-multiply-add two-product (Ogita, Rump and Oishi 2005) and the usual
// renormalisation. Used by the double-float variant of the IF phase (the arithmetic of GPUs
// without fast FP64).
struct df {
    float hi, lo;
};
__device__ inline df df_two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ inline df df_fast_two_sum(float a, float b) {
    const float s = a + b;
    return {s, b - (s - a)};
}
__device__ inline df df_add(df a, df b) {
    df s = df_two_sum(a.hi, b.hi);
    const df t = df_two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = df_fast_two_sum(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return df_fast_two_sum(s.hi, s.lo);
}
__device__ inline df df_mul(df a, df b) {
    const float p = a.hi * b.hi;
    const float e = __fmaf_rn(a.hi, b.hi, -p);
    return df_fast_two_sum(p, e + (a.hi * b.lo + a.lo * b.hi));
}

// UniformSampleSphere (libhussar/include/hussar/core/sampling.h:38-44) in double
__host__ __device__ inline d3 uniform_sample_sphere(double u0, double u1) {
    const double phi = (2.0 * 3.14159265358979323846) * u0;
    const double y = 2.0 * u1 - 1.0;
    const double r = sqrt(fmax(0.0, 1.0 - y * y));
    double s, c;
    sincos_d(phi, &s, &c);
    return {-r * s, y, -r * c};
}

// AWRAngularDistribution::evaluate (emitter.h:126-136): H = e_y × d scaled by the two lobe
// polynomials. The reference's double literals promote the lobes to double; here they are
// evaluated in FP32 like the rest of the field arithmetic (relative difference ~1e-7).
__device__ inline cf3 awr_evaluate(f3 d) {
    const float cos_h = sqrtf(1.0f - d.x * d.x);
    const float cos_e = sqrtf(1.0f - d.y * d.y);
    const float a = cos_h - 1.8f;
    const float b = cos_e - 1.5f;
    const float a2 = a * a, b2 = b * b;
    const float f1 = 2.622f / (a2 * a2 * a2);
    const float f2 = 0.625f / (b2 * b2);
    const float hx = (d.z * f1) * f2;
    const float hz = (-d.x * f1) * f2;
    return cf3{{hx, 0.0f}, {0.0f, 0.0f}, {hz, 0.0f}};
}

// ----------------------------------------------------------------- ray / triangle
// Möller–Trumbore; returns t in (tnear, tfar) or +inf.
__device__ inline float tri_intersect(f3 v0, f3 e1, f3 e2, f3 o, f3 d, float tnear, float tfar) {
    const f3 pv = cross(d, e2);
    const float det = dot(e1, pv);
    if (det == 0.0f) return __builtin_inff();
    const float inv = 1.0f / det;
    const f3 tv = o - v0;
    const float u = dot(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return __builtin_inff();
    const f3 qv = cross(tv, e1);
    const float v = dot(d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return __builtin_inff();
    const float t = dot(e2, qv) * inv;
    if (!(t > tnear && t < tfar)) return __builtin_inff();
    return t;
}

}  // namespace hsr

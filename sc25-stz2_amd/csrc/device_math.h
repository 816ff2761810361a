// device_math.h -- the STZ per-point arithmetic on gfx950.
//
// Bit-exact with the reference's CPU code: the library is compiled with
// -ffp-contract=off (no FMA contraction, F5 is built without -march so x86
// has none) and -fgpu-flush-denormals-to-zero (F5:3046-3047 runs with
// FTZ/DAZ on); every expression keeps the reference's operand order.
// F5 = src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace stz {

// Histogram counts: an LDS bin (workgroup scope) and a global bin (device
// scope).  The scopes differ on purpose: two identical atomics on either
// side of a branch get merged into one through a selected generic pointer,
// a flat atomic that waits on both the vector-memory and the LDS counters.
__device__ __forceinline__ void lds_add(uint32_t* p, uint32_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void glb_add(uint32_t* p, uint32_t v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The digits (a, b, c) of work items t = t0, t0 + step, t0 + 2 step, ... in
// the mixed radix t = (c nb + b) na + a, kept up to date by additions: a
// grid-stride item loop then divides once, not once per item (a 64-bit
// division by a run-time value is a long scalar routine).  Every quantity is
// wave-uniform.
struct Walk3 {
    uint32_t a, b, c;     // digits of the current item
    uint32_t sa, sb, sc;  // digits of the step
    uint32_t na, nb;
    __device__ __forceinline__ void init(uint64_t t0, uint64_t step, uint32_t na_, uint32_t nb_)
    {
        na = na_;
        nb = nb_;
        const uint64_t nab = (uint64_t)na * nb;
        a = (uint32_t)(t0 % na);
        b = (uint32_t)((t0 / na) % nb);
        c = (uint32_t)(t0 / nab);
        sa = (uint32_t)(step % na);
        sb = (uint32_t)((step / na) % nb);
        sc = (uint32_t)(step / nab);
    }
    __device__ __forceinline__ void next()
    {
        a += sa;
        const uint32_t ca = a >= na ? 1u : 0u;
        a -= ca ? na : 0u;
        b += sb + ca;
        const uint32_t cb = b >= nb ? 1u : 0u;
        b -= cb ? nb : 0u;
        c += sc + cb;
    }
};

// F5:141-146 cubic_line_pred
template <typename T>
__device__ __forceinline__ T cub(T a, T b, T c, T d)
{
    return T(-0.0625f) * a + T(0.5625f) * b + T(0.5625f) * c - T(0.0625f) * d;
}

// F5:163-175 cubic_line_pv_cse (default build)
template <typename T>
__device__ __forceinline__ void pv(T a, T b, T c, T d, T& p, T& v)
{
    const T eps = T(1e-12f);
    const T c0 = a - T(2.0f) * b + c;
    const T c1 = b - T(2.0f) * c + d;
    v = T(0.5f) * (c0 * c0 + c1 * c1) + eps;
    p = T(0.5f) * (b + c) - T(0.0625f) * (c0 + c1);
}


// F5:177-186
template <typename T>
__device__ __forceinline__ T blend2(T p0, T v0, T p1, T v1, int opt = 0)
{
    if (opt & kPredWpow2) {
        v0 *= v0;
        v1 *= v1;
    }
    return (p0 * v1 + p1 * v0) / (v0 + v1);
}

// F5:188-203
template <typename T>
__device__ __forceinline__ T blend4(T p0, T v0, T p1, T v1, T p2, T v2, T p3, T v3, int opt = 0)
{
    if (opt & kPredWpow2) {
        v0 *= v0;
        v1 *= v1;
        v2 *= v2;
        v3 *= v3;
    }
    const T w0 = v1 * v2 * v3;
    const T w1 = v0 * v2 * v3;
    const T w2 = v0 * v1 * v3;
    const T w3 = v0 * v1 * v2;
    return (w0 * p0 + w1 * p1 + w2 * p2 + w3 * p3) / (w0 + w1 + w2 + w3);
}

// F5:148-155 cubic_line_var with F5:141-146 cubic_line_pred: the plain
// (non-CSE) form of a line's prediction and variance, which the plane-axis
// and cube-axis candidates use (F5:566-573)
template <typename T>
__device__ __forceinline__ void pv_plain(T a, T b, T c, T d, T& p, T& v)
{
    const T eps = T(1e-12f);
    p = cub(a, b, c, d);
    const T c0 = a - T(2.0f) * b + c;
    const T c1 = b - T(2.0f) * c + d;
    v = T(0.5f) * (c0 * c0 + c1 * c1) + eps;
}

// F5:579-607 predict_plane_dyn under PRED_PLANE_AXIS; f(o1, o2) is the
// value at offset o1 along the plane's first axis and o2 along its second
template <typename T, typename F2>
__device__ __forceinline__ T plane_dyn(const F2& f)
{
    T p0, v0, p1, v1, pa0, va0, pa1, va1, pb0, vb0, pb1, vb1;
    pv_plain(f(-1, -1), f(0, 0), f(1, 1), f(2, 2), p0, v0);
    pv_plain(f(2, -1), f(1, 0), f(0, 1), f(-1, 2), p1, v1);
    pv_plain(f(-1, 0), f(0, 0), f(1, 0), f(2, 0), pa0, va0);
    pv_plain(f(-1, 1), f(0, 1), f(1, 1), f(2, 1), pa1, va1);
    const T p2 = T(0.5f) * (pa0 + pa1);
    const T v2 = T(0.5f) * (va0 + va1);
    pv_plain(f(0, -1), f(0, 0), f(0, 1), f(0, 2), pb0, vb0);
    pv_plain(f(1, -1), f(1, 0), f(1, 1), f(1, 2), pb1, vb1);
    const T p3 = T(0.5f) * (pb0 + pb1);
    const T v3 = T(0.5f) * (vb0 + vb1);
    const T w0 = T(1) / v0, w1 = T(1) / v1, w2 = T(1) / v2, w3 = T(1) / v3;
    return (w0 * p0 + w1 * p1 + w2 * p2 + w3 * p3) / (w0 + w1 + w2 + w3);
}

// F5:609-622 predict_cube_axis_pair: the cubic along one axis averaged over
// the lines at offsets {0, t1, t2, t1+t2} across the other two; g(s, u, w)
// is the value at offset s along the line and (u, w) across it; the sums
// start from zero as in F5
template <typename T, typename G3>
__device__ __forceinline__ T cube_axis(const G3& g, T& v)
{
    T psum = T(0), vsum = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int u = k & 1, w = k >> 1;  // offsets 0, t1, t2, t1 + t2
        T p, vv;
        pv_plain(g(-1, u, w), g(0, u, w), g(1, u, w), g(2, u, w), p, vv);
        psum += p;
        vsum += vv;
    }
    v = T(0.25f) * vsum;
    return T(0.25f) * psum;
}

// F5:675-686 block 6 under PRED_CUBE_AXIS: the four diagonal lines (their
// predictions and variances given) and three axis-pair candidates, all
// weighted by inverse variance; A(ox, oy, oz) is the value at that offset
template <typename T, typename A3>
__device__ __forceinline__ T cube7(const A3& A, T p0, T v0, T p1, T v1, T p2, T v2, T p3, T v3)
{
    T v4, v5, v6;
    const T p4 = cube_axis<T>([&](int s, int u, int w) { return A(s, u, w); }, v4);
    const T p5 = cube_axis<T>([&](int s, int u, int w) { return A(u, s, w); }, v5);
    const T p6 = cube_axis<T>([&](int s, int u, int w) { return A(u, w, s); }, v6);
    const T w0 = T(1) / v0, w1 = T(1) / v1, w2 = T(1) / v2, w3 = T(1) / v3;
    const T w4 = T(1) / v4, w5 = T(1) / v5, w6 = T(1) / v6;
    return (w0 * p0 + w1 * p1 + w2 * p2 + w3 * p3 + w4 * p4 + w5 * p5 + w6 * p6) /
           (w0 + w1 + w2 + w3 + w4 + w5 + w6);
}

// stz2.cpp:122-155 (block 2; 186-199 block 4, 222-235 block 5): the
// 12-point truncated bicubic of a plane, sums in the order written there;
// f(o1, o2) as for plane_dyn
template <typename T, typename F2>
__device__ __forceinline__ T bicubic12(const F2& f)
{
    const T w_core = T(0.31640625f);
    const T core = f(0, 0) + f(1, 0) + f(0, 1) + f(1, 1);
    const T w_edge = T(-0.033203125f);
    const T edge = f(-1, 0) + f(2, 0) + f(-1, 1) + f(2, 1) + f(0, -1) + f(1, -1) + f(0, 2) + f(1, 2);
    return w_core * core + w_edge * edge;
}

// stz2.cpp:271-293: block 6 from the 8 cell corners and the 8 points one
// step beyond them along the body diagonals, in the order written there;
// A(ox, oy, oz) the value at that offset
template <typename T, typename A3>
__device__ __forceinline__ T diag16(const A3& A)
{
    const T a = T(0.140625f), b = T(0.015625f);
    return a * A(0, 0, 0) + a * A(1, 1, 1) + a * A(0, 1, 1) + a * A(1, 0, 1) + a * A(1, 1, 0) + a * A(1, 0, 0) +
           a * A(0, 1, 0) + a * A(0, 0, 1) - b * A(-1, -1, -1) - b * A(2, 2, 2) - b * A(2, -1, 2) -
           b * A(-1, 2, 2) - b * A(2, 2, -1) - b * A(-1, -1, 2) - b * A(-1, 2, -1) - b * A(2, -1, -1);
}

// F5:690-748 predict_block_const + F5:624-688 dynamic lines; interior points
// evaluate exactly what predict_interior_const (F5:828-840) evaluates.
template <int B, typename T>
__device__ __forceinline__ T predict(const T* __restrict__ r, int i, int x, int y, int z, int dx,
                                     int dy, int dz, int opt = 0)
{
    const int sx = dx, sxy = dx * dy;
    auto A = [&](int ox, int oy, int oz) -> T { return r[i + ox + oy * sx + oz * sxy]; };
    if constexpr (B == 0) {
        if (x == dx - 1) return r[i];
        if (x == dx - 2 || x == 0) return T(0.5f) * (r[i] + r[i + 1]);
        return cub(r[i - 1], r[i], r[i + 1], r[i + 2]);
    } else if constexpr (B == 1) {
        if (y == dy - 1) return r[i];
        if (y == dy - 2 || y == 0) return T(0.5f) * (r[i] + r[i + sx]);
        return cub(r[i - sx], r[i], r[i + sx], r[i + 2 * sx]);
    } else if constexpr (B == 3) {
        if (z == dz - 1) return r[i];
        if (z == dz - 2 || z == 0) return T(0.5f) * (r[i] + r[i + sxy]);
        return cub(r[i - sxy], r[i], r[i + sxy], r[i + 2 * sxy]);
    } else if constexpr (B == 2) {
        if (x == dx - 1 || y == dy - 1) return r[i];
        if (y == dy - 2 || y == 0 || x == dx - 2 || x == 0)
            return T(0.25f) * (r[i] + r[i + sx + 1] + r[i + sx] + r[i + 1]);
        if (opt & kPredBicubic12) return bicubic12<T>([&](int o1, int o2) { return A(o1, o2, 0); });
        if (opt & kPredPlaneAxis) return plane_dyn<T>([&](int o1, int o2) { return A(o1, o2, 0); });
        T p0, v0, p1, v1;
        pv(r[i - sx - 1], r[i], r[i + sx + 1], r[i + 2 * sx + 2], p0, v0);
        pv(r[i - sx + 2], r[i + 1], r[i + sx], r[i + 2 * sx - 1], p1, v1);
        return blend2(p0, v0, p1, v1, opt);
    } else if constexpr (B == 4) {
        if (x == dx - 1 || z == dz - 1) return r[i];
        if (z == dz - 2 || z == 0 || x == dx - 2 || x == 0)
            return T(0.25f) * (r[i] + r[i + sxy + 1] + r[i + 1] + r[i + sxy]);
        if (opt & kPredBicubic12) return bicubic12<T>([&](int o1, int o2) { return A(o1, 0, o2); });
        if (opt & kPredPlaneAxis) return plane_dyn<T>([&](int o1, int o2) { return A(o1, 0, o2); });
        T p0, v0, p1, v1;
        pv(r[i - sxy - 1], r[i], r[i + sxy + 1], r[i + 2 * sxy + 2], p0, v0);
        pv(r[i - sxy + 2], r[i + 1], r[i + sxy], r[i + 2 * sxy - 1], p1, v1);
        return blend2(p0, v0, p1, v1, opt);
    } else if constexpr (B == 5) {
        if (z == dz - 1 || y == dy - 1) return r[i];
        if (y == dy - 2 || y == 0 || z == dz - 2 || z == 0)
            return T(0.25f) * (r[i] + r[i + sxy + sx] + r[i + sxy] + r[i + sx]);
        if (opt & kPredBicubic12) return bicubic12<T>([&](int o1, int o2) { return A(0, o1, o2); });
        if (opt & kPredPlaneAxis) return plane_dyn<T>([&](int o1, int o2) { return A(0, o1, o2); });
        T p0, v0, p1, v1;
        pv(r[i - sxy - sx], r[i], r[i + sxy + sx], r[i + 2 * sxy + 2 * sx], p0, v0);
        pv(r[i - sxy + 2 * sx], r[i + sx], r[i + sxy], r[i + 2 * sxy - sx], p1, v1);
        return blend2(p0, v0, p1, v1, opt);
    } else {
        if (x == dx - 1 || y == dy - 1 || z == dz - 1) return r[i];
        if (z == dz - 2 || z == 0 || y == dy - 2 || y == 0 || x == dx - 2 || x == 0)
            return T(0.125f) * (r[i] + r[i + sxy + sx + 1] + r[i + sxy + sx] + r[i + sxy + 1] +
                                r[i + sx + 1] + r[i + 1] + r[i + sx] + r[i + sxy]);
        if (opt & kPredDiag16) return diag16<T>(A);
        T p0, v0, p1, v1, p2, v2, p3, v3;
        pv(r[i - sxy - sx - 1], r[i], r[i + sxy + sx + 1], r[i + 2 * sxy + 2 * sx + 2], p0, v0);
        pv(r[i - sxy - sx + 2], r[i + 1], r[i + sxy + sx], r[i + 2 * sxy + 2 * sx - 1], p1, v1);
        pv(r[i - sxy + 2 * sx - 1], r[i + sx], r[i + sxy + 1], r[i + 2 * sxy - sx + 2], p2, v2);
        pv(r[i + 2 * sxy - sx - 1], r[i + sxy], r[i + sx + 1], r[i - sxy + 2 * sx + 2], p3, v3);
        if (opt & kPredCubeAxis) return cube7<T>(A, p0, v0, p1, v1, p2, v2, p3, v3);
        return blend4(p0, v0, p1, v1, p2, v2, p3, v3, opt);
    }
}

// runtime block index -> compile-time predictor
template <typename T>
__device__ __forceinline__ T predict_rt(int b, const T* __restrict__ r, int i, int x, int y, int z,
                                        int dx, int dy, int dz, int opt = 0)
{
    switch (b) {
    case 0: return predict<0>(r, i, x, y, z, dx, dy, dz, opt);
    case 1: return predict<1>(r, i, x, y, z, dx, dy, dz, opt);
    case 2: return predict<2>(r, i, x, y, z, dx, dy, dz, opt);
    case 3: return predict<3>(r, i, x, y, z, dx, dy, dz, opt);
    case 4: return predict<4>(r, i, x, y, z, dx, dy, dz, opt);
    case 5: return predict<5>(r, i, x, y, z, dx, dy, dz, opt);
    default: return predict<6>(r, i, x, y, z, dx, dy, dz, opt);
    }
}

// Linear quantizer, bit-exact with F5:1345-1397 / F5:1204-1253 (itself
// bit-exact with SZ3 LinearQuantizer::quantize_and_overwrite(d, 0)).
// Returns the code (0 = unpredictable) and the decompressed value (d itself
// when unpredictable).
// Branch-free: every step is evaluated and the two rejections (out of range,
// NaN/Inf included; error above eb after rounding) select at the end.
template <typename T>
__device__ __forceinline__ uint32_t quantize(T d, double eb, double recip, T& dec_out)
{
    const double ad = fabs((double)d) * recip;
    const bool in_range = ad < (double)(2 * kRadius - 1);
    const int half = ((int)(in_range ? ad : 0.0) + 1) >> 1;
    T dec = (T)((double)(2 * half) * eb);  // F5:1353-1356 dec_lut[h]
    if (d < T(0)) dec = -dec;
    const T derr = dec - d;
    const bool ok = in_range && fabs((double)derr) <= eb;
    dec_out = ok ? dec : d;
    return ok ? (uint32_t)((d < T(0)) ? kRadius - half : kRadius + half) : 0u;
}

// F5:1531-1535 rec_lut[q] = (T)(2*(q-r)*eb)
template <typename T>
__device__ __forceinline__ T dequantize(uint32_t q, double eb)
{
    return (T)((double)(2 * ((int)q - kRadius)) * eb);
}

}  // namespace stz

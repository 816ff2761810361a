// k_staged.hip -- the residual levels under dependency.cpp's staged
// inverse-variance predictors (option bit kPredStaged).
//
// dependency.cpp:594-636 (low level) and 681-704 (high level) predict the
// blocks in three stages: blocks 0, 1, 3 (one offset axis) from the coarser
// grid by F5's single-axis rules, then blocks 2, 4, 5 from the reconstructed
// blocks 0, 1, 3, each interpolated along its missing axis and blended by
// inverse variance, then block 6 from the reconstructed 2, 4, 5
// (dependency.cpp:162-292).  A block's neighbours in the earlier stages are
// reconstructions of other points, so a stage starts once the one before it
// is complete: one launch per stage, one thread per block point running the
// stage's blocks.  The level's reconstruction is its 2x grid `out` (the mid
// grid for the low level, the padded full grid for the high one): the
// coarser grid at parity 0, block b at the parity of sub-block b + 1.
// Encode quantizes as it predicts (F5's quantizer at the level's bound) and
// writes codes, unpredictable residuals and histograms like the other level
// kernels; decode reads them back.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_math.h"
#include "kernels.h"

namespace stz {

namespace {

constexpr int kSHW = 1024;  // LDS histogram window (the codes around the centre)
constexpr uint32_t kSLo = (uint32_t)kRadius - kSHW / 2;

template <typename T>
struct StagedK {
    const T* full;  // encode: the padded field (targets)
    int px, py;
    int tstep;      // target stride in the field: 2 (high level) or 4 (low level)
    const T* ref;   // the coarser grid, dims g.dx, g.dy, g.dz (stage 1)
    Geom g;         // block grid and stream chunking
    LevelArgs a;
    uint16_t* codes;  // encode: written; decode: read
    T* side;
    uint32_t* hist;   // encode: 7 x 65536 bins
    T* out;           // the level's 2x grid
};

// dependency.cpp:162-239 interp_axis_with_variance: the reconstructed block
// of parity (qx, qy, qz) along `axis` of the block grid at point (x, y, z);
// |b - a| squared is (b - a) squared bit for bit
template <typename T>
__device__ __forceinline__ void st_axis(const T* __restrict__ o, size_t ox, size_t oy, int qx, int qy, int qz, int x,
                                        int y, int z, int axis, int dx, int dy, int dz, T& p, T& v)
{
    const T eps = T(1e-12f);
    const int sx = axis == 0, sy = axis == 1, sz = axis == 2;
    auto sv = [&](int j) -> T {
        return o[((size_t)(2 * (z + sz * j) + qz) * oy + (size_t)(2 * (y + sy * j) + qy)) * ox +
                 (size_t)(2 * (x + sx * j) + qx)];
    };
    const int c = axis == 0 ? x : axis == 1 ? y : z;
    const int n = axis == 0 ? dx : axis == 1 ? dy : dz;
    if (c == n - 1) {
        p = sv(0);
        v = eps;
        return;
    }
    if (c == n - 2 || c == 0) {
        const T a = sv(0), b = sv(1);
        const T e = b - a;
        p = T(0.5f) * (a + b);
        v = e * e + eps;
        return;
    }
    const T a = sv(-1), b = sv(0), cc = sv(1), d = sv(2);
    const T c0 = a - T(2.0f) * b + cc;
    const T c1 = b - T(2.0f) * cc + d;
    p = T(-0.0625f) * a + T(0.5625f) * b + T(0.5625f) * cc + T(-0.0625f) * d;
    v = T(0.5f) * (c0 * c0 + c1 * c1) + eps;
}

// dependency.cpp:258-292 staged_prediction of block 2, 4, 5 or 6; the blocks
// x, y, z, xy, xz, yz (0, 1, 3, 2, 4, 5) sit at parities (1,0,0), (0,1,0),
// (0,0,1), (1,1,0), (1,0,1), (0,1,1) of the 2x grid
template <typename T>
__device__ __forceinline__ T st_predict(int block, const T* __restrict__ o, size_t ox, size_t oy, int x, int y, int z,
                                        int dx, int dy, int dz)
{
    T p0, v0, p1, v1, p2, v2;
    if (block == 2) {  // xy: block x along y, block y along x
        st_axis<T>(o, ox, oy, 1, 0, 0, x, y, z, 1, dx, dy, dz, p0, v0);
        st_axis<T>(o, ox, oy, 0, 1, 0, x, y, z, 0, dx, dy, dz, p1, v1);
        return (p0 * v1 + p1 * v0) / (v0 + v1);
    }
    if (block == 4) {  // xz: block x along z, block z along x
        st_axis<T>(o, ox, oy, 1, 0, 0, x, y, z, 2, dx, dy, dz, p0, v0);
        st_axis<T>(o, ox, oy, 0, 0, 1, x, y, z, 0, dx, dy, dz, p1, v1);
        return (p0 * v1 + p1 * v0) / (v0 + v1);
    }
    if (block == 5) {  // yz: block y along z, block z along y
        st_axis<T>(o, ox, oy, 0, 1, 0, x, y, z, 2, dx, dy, dz, p0, v0);
        st_axis<T>(o, ox, oy, 0, 0, 1, x, y, z, 1, dx, dy, dz, p1, v1);
        return (p0 * v1 + p1 * v0) / (v0 + v1);
    }
    // xyz: xy along z, xz along y, yz along x (dependency.cpp:250-256 blend3)
    st_axis<T>(o, ox, oy, 1, 1, 0, x, y, z, 2, dx, dy, dz, p0, v0);
    st_axis<T>(o, ox, oy, 1, 0, 1, x, y, z, 1, dx, dy, dz, p1, v1);
    st_axis<T>(o, ox, oy, 0, 1, 1, x, y, z, 0, dx, dy, dz, p2, v2);
    const T w0 = v1 * v2, w1 = v0 * v2, w2 = v0 * v1;
    return (w0 * p0 + w1 * p1 + w2 * p2) / (w0 + w1 + w2);
}

// the blocks of stage s: {0, 1, 3}, {2, 4, 5}, {6}
__device__ __forceinline__ int stage_block(int s, int k)
{
    return s == 0 ? (k == 2 ? 3 : k) : s == 1 ? (k == 0 ? 2 : k == 1 ? 4 : 5) : 6;
}

template <typename T, bool kEnc, int kStage>
__global__ __launch_bounds__(256) void k_staged(const StagedK<T> K)
{
    constexpr int kNB = kStage == 2 ? 1 : 3;
    __shared__ uint32_t sh[kEnc ? kNB : 1][kEnc ? kSHW : 1];
    if constexpr (kEnc) {
        for (int k = threadIdx.x; k < kNB * kSHW; k += blockDim.x) sh[k / kSHW][k % kSHW] = 0u;
        __syncthreads();
    }
    const Geom& g = K.g;
    const int dx = g.dx, dy = g.dy, dz = g.dz;
    const size_t ox = 2 * (size_t)dx, oy = 2 * (size_t)dy;
    const uint64_t n = (uint64_t)dx * dy * dz;
    const uint32_t hs = (uint32_t)(K.tstep / 2);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % (uint64_t)dx), y = (int)((i / (uint64_t)dx) % (uint64_t)dy);
        const int z = (int)(i / ((uint64_t)dx * dy));
        const uint32_t pos = stream_pos(g, x, y, z);
        if constexpr (kStage == 0) K.out[((size_t)(2 * z) * oy + (size_t)(2 * y)) * ox + (size_t)(2 * x)] = K.ref[i];
#pragma unroll
        for (int k = 0; k < kNB; ++k) {
            const int b = stage_block(kStage, k);
            const int sb = b + 1, qx = sb & 1, qy = (sb >> 1) & 1, qz = sb >> 2;
            // F5's single-axis rules for blocks 0, 1, 3 (dependency.cpp's
            // interp_axis, dependency.cpp:90-118, evaluates the same)
            const T pr = kStage == 0 ? predict_rt<T>(b, K.ref, (int)i, x, y, z, dx, dy, dz, 0)
                                     : st_predict<T>(b, K.out, ox, oy, x, y, z, dx, dy, dz);
            const uint64_t e = K.a.code_off[b] + pos;
            T r;
            if constexpr (kEnc) {
                const size_t t = ((size_t)(K.tstep * z + (int)hs * qz) * K.py + (size_t)(K.tstep * y + (int)hs * qy)) *
                                     K.px + (size_t)(K.tstep * x + (int)hs * qx);
                const T d = K.full[t] - pr;
                T dec;
                const uint32_t q = quantize(d, K.a.eb, K.a.recip, dec);
                K.codes[e] = (uint16_t)q;
                if (!q) K.side[e] = d;
                if (q - kSLo < (uint32_t)kSHW) atomicAdd(&sh[k][q - kSLo], 1u);
                else atomicAdd(&K.hist[(size_t)b * 65536 + q], 1u);
                r = dec + pr;  // dependency.cpp:600-614: the decoded residual + the prediction
            } else {
                const uint32_t q = K.codes[e];
                r = (q ? dequantize<T>(q, K.a.eb) : K.side[e]) + pr;
            }
            K.out[((size_t)(2 * z + qz) * oy + (size_t)(2 * y + qy)) * ox + (size_t)(2 * x + qx)] = r;
        }
    }
    if constexpr (kEnc) {
        __syncthreads();
        for (int k = threadIdx.x; k < kNB * kSHW; k += blockDim.x) {
            const uint32_t v = sh[k / kSHW][k % kSHW];
            if (v) atomicAdd(&K.hist[(size_t)stage_block(kStage, k / kSHW) * 65536 + kSLo + (k % kSHW)], v);
        }
    }
}

template <typename T, bool kEnc>
void launch_stages(hipStream_t s, const StagedK<T>& K)
{
    const uint64_t n = (uint64_t)K.g.dx * K.g.dy * K.g.dz;
    if (!n) return;
    // enough workgroups to fill the device a few times over; the loop
    // strides the rest (fewer workgroups also flush fewer histograms)
    const unsigned G = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    k_staged<T, kEnc, 0><<<dim3(G), dim3(256), 0, s>>>(K);
    k_staged<T, kEnc, 1><<<dim3(G), dim3(256), 0, s>>>(K);
    k_staged<T, kEnc, 2><<<dim3(G), dim3(256), 0, s>>>(K);
}

}  // namespace

template <typename T>
void launch_staged_encode(hipStream_t s, const T* full, int px, int py, int tstep, const T* ref, Geom g, LevelArgs a,
                          uint16_t* codes, T* side, uint32_t* hist, T* out)
{
    StagedK<T> K{};
    K.full = full;
    K.px = px;
    K.py = py;
    K.tstep = tstep;
    K.ref = ref;
    K.g = g;
    K.a = a;
    K.codes = codes;
    K.side = side;
    K.hist = hist;
    K.out = out;
    launch_stages<T, true>(s, K);
}

template <typename T>
void launch_staged_decode(hipStream_t s, const T* ref, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                          T* out)
{
    StagedK<T> K{};
    K.ref = ref;
    K.g = g;
    K.a = a;
    K.codes = const_cast<uint16_t*>(codes);
    K.side = const_cast<T*>(side);
    K.out = out;
    launch_stages<T, false>(s, K);
}

#define STZ_INST_STAGED(T)                                                                                     \
    template void launch_staged_encode<T>(hipStream_t, const T*, int, int, int, const T*, Geom, LevelArgs,      \
                                          uint16_t*, T*, uint32_t*, T*);                                        \
    template void launch_staged_decode<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, T*);
STZ_INST_STAGED(float)
STZ_INST_STAGED(double)

}  // namespace stz

// k_residual.hip -- the low and high STZ levels on gfx950.
//
// Encode: one lane per point of the block grid handles all 7 parity blocks
// of its 2x2x2 cell.  It reads the cell's full-resolution targets as
// coalesced 2-wide vectors straight from the field (the reference first
// slices into 8 sub-block copies, F5:2102-2157), predicts each block from the
// reconstructed coarser grid, quantizes, writes the codes at their stream
// (chunk-major) positions and counts them into LDS histograms.  The low level
// also writes its reconstruction (closed loop, F5:2207-2212) straight into the
// mid grid, fusing merge_sub_blocks_to_full (F5:2230).
// Decode: the mirror image -- dequantize + predict + parity scatter in one
// pass (F5:1096-1159 fused depreprocess+merge, F5:2424-2451).
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "kernels.h"

namespace stz {

namespace {

constexpr int kHW = 1024;                 // LDS histogram window
constexpr uint32_t kHLo = kRadius - kHW / 2;

template <typename T> struct V2;
template <> struct V2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct V2<double> { typedef double type __attribute__((ext_vector_type(2))); };

// blockIdx -> logical block so that each XCD (round-robin dispatch over 8)
// walks a contiguous range of tiles and keeps its halo rows in its own L2
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t G = gridDim.x;
    if (G % 8u) return blockIdx.x;
    return (blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u;
}

__device__ __forceinline__ void hist_put(uint32_t* sh, uint32_t* gh, uint32_t q, uint32_t n)
{
    if (q - kHLo < (uint32_t)kHW) atomicAdd(&sh[q - kHLo], n);
    else atomicAdd(&gh[q], n);
}

// wave-aggregated histogram update: the two most frequent codes of the wave
// go in with one atomic each, the rest per lane
__device__ __forceinline__ void hist_add(uint32_t* sh, uint32_t* gh, uint32_t q, bool valid)
{
    bool pending = valid;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const unsigned long long act = __ballot(pending);
        if (act == 0ull) break;
        const int leader = __ffsll((long long)act) - 1;
        const uint32_t lq = (uint32_t)__shfl((int)q, leader, 64);
        const bool same = pending && q == lq;
        const unsigned long long m = __ballot(same);
        if (__lane_id() == leader) hist_put(sh, gh, lq, (uint32_t)__popcll(m));
        if (same) pending = false;
    }
    if (pending) hist_put(sh, gh, q, 1u);
}

__device__ __forceinline__ void hist_flush(uint32_t (*sh)[kHW], int nstreams, uint32_t* gh)
{
    __syncthreads();
    for (int k = threadIdx.x; k < nstreams * kHW; k += blockDim.x) {
        const uint32_t v = sh[k / kHW][k % kHW];
        if (v) atomicAdd(&gh[(size_t)(k / kHW) * 65536 + kHLo + (k % kHW)], v);
    }
}

// ---------------------------------------------------------------- predictors

// Interior predictions of all 7 blocks at x on the row `row` (the block grid
// row of the point, so the point is row[x]); requires 1 <= x <= dx-3 and the
// same on y and z.  The 34 neighbours sit on 12 rows whose base pointers are
// wave-uniform (scalar), so every load is scalar base + lane offset + an
// immediate.  Same expressions, in the same order, as predict<B>'s interior.
template <typename T>
__device__ __forceinline__ void predict7_interior(const T* __restrict__ row, int x, int sx, int sxy, T p[7])
{
    const T* R00 = row;
    const T* Rm0 = row - sx;             // (dy, dz) = (-1, 0)
    const T* R10 = row + sx;
    const T* R20 = row + 2 * sx;
    const T* R0m = row - sxy;            // (0, -1)
    const T* R01 = row + sxy;
    const T* R02 = row + 2 * sxy;
    const T* Rmm = row - sxy - sx;       // (-1, -1)
    const T* R11 = row + sxy + sx;
    const T* R22 = row + 2 * sxy + 2 * sx;
    const T* R2m = row - sxy + 2 * sx;   // (2, -1)
    const T* Rm2 = row + 2 * sxy - sx;   // (-1, 2)
    const T c = R00[x];
    p[0] = cub(R00[x - 1], c, R00[x + 1], R00[x + 2]);
    p[1] = cub(Rm0[x], c, R10[x], R20[x]);
    p[3] = cub(R0m[x], c, R01[x], R02[x]);
    T p0, v0, p1, v1, p2, v2, p3, v3;
    pv(Rm0[x - 1], c, R10[x + 1], R20[x + 2], p0, v0);
    pv(Rm0[x + 2], R00[x + 1], R10[x], R20[x - 1], p1, v1);
    p[2] = blend2(p0, v0, p1, v1);
    pv(R0m[x - 1], c, R01[x + 1], R02[x + 2], p0, v0);
    pv(R0m[x + 2], R00[x + 1], R01[x], R02[x - 1], p1, v1);
    p[4] = blend2(p0, v0, p1, v1);
    pv(Rmm[x], c, R11[x], R22[x], p0, v0);
    pv(R2m[x], R10[x], R01[x], Rm2[x], p1, v1);
    p[5] = blend2(p0, v0, p1, v1);
    pv(Rmm[x - 1], c, R11[x + 1], R22[x + 2], p0, v0);
    pv(Rmm[x + 2], R00[x + 1], R11[x], R22[x - 1], p1, v1);
    pv(R2m[x - 1], R10[x], R01[x + 1], Rm2[x + 2], p2, v2);
    pv(Rm2[x - 1], R01[x], R10[x + 1], R2m[x + 2], p3, v3);
    p[6] = blend4(p0, v0, p1, v1, p2, v2, p3, v3);
}

// all 7 predictions at (x,y,z): the interior form when `inner` (wave-uniform), else checked
template <typename T>
__device__ __forceinline__ void predict7(bool inner, const T* __restrict__ r, int x, int y, int z, int dx, int dy,
                                         int dz, T p[7])
{
    if (inner) {
        predict7_interior(r + ((size_t)z * dy + y) * dx, x, dx, dx * dy, p);
    } else {
        const int i = (z * dy + y) * dx + x;
        p[0] = predict<0>(r, i, x, y, z, dx, dy, dz);
        p[1] = predict<1>(r, i, x, y, z, dx, dy, dz);
        p[2] = predict<2>(r, i, x, y, z, dx, dy, dz);
        p[3] = predict<3>(r, i, x, y, z, dx, dy, dz);
        p[4] = predict<4>(r, i, x, y, z, dx, dy, dz);
        p[5] = predict<5>(r, i, x, y, z, dx, dy, dz);
        p[6] = predict<6>(r, i, x, y, z, dx, dy, dz);
    }
}

// Work split of a block grid: the main kernels cover x in [1, dx-3] in
// 64 x 4 tiles (a wave is one row, so y and z are wave-uniform and rows away
// from the y/z faces take the interior form); the edge kernels cover the
// columns x in {0, dx-2, dx-1} (every x when dx < 4) with checked predictors.
struct EdgeMap {
    int ncol;       // columns handled by the edge kernel
    int cols[4];
};

__host__ __device__ inline EdgeMap edge_map(int dx)
{
    EdgeMap e{};
    if (dx < 4) {
        e.ncol = dx;
        for (int k = 0; k < dx; ++k) e.cols[k] = k;
    } else {
        e.ncol = 3;
        e.cols[0] = 0;
        e.cols[1] = dx - 2;
        e.cols[2] = dx - 1;
    }
    return e;
}

__device__ __forceinline__ bool row_inner(int y, int z, int dy, int dz)
{
    return y >= 1 && y <= dy - 3 && z >= 1 && z <= dz - 3;
}

// ---------------------------------------------------------------- encode

template <typename T>
__device__ __forceinline__ uint32_t enc1(T tgt, T p, const LevelArgs& a, int b, uint32_t pos,
                                         uint16_t* __restrict__ codes, T* __restrict__ side, T& rec)
{
    const T d = tgt - p;
    T dec;
    const uint32_t q = quantize(d, a.eb, a.recip, dec);
    codes[a.code_off[b] + pos] = (uint16_t)q;
    if (!q) side[a.code_off[b] + pos] = d;
    rec = dec + p;  // decoded diff + prediction (F5:1381-1385 then F5:1077-1083)
    return q;
}

// high level: targets of the 2x2x2 cell of mid point (x,y,z) straight from the field
template <typename T>
__device__ __forceinline__ void high_enc_point(const T* __restrict__ full, int px, int py, const T* __restrict__ mid,
                                               const Geom& g, const LevelArgs& a, uint16_t* __restrict__ codes,
                                               T* __restrict__ side, int x, int y, int z, bool inner, uint32_t q[7])
{
    typedef typename V2<T>::type v2;
    T p[7];
    predict7(inner, mid, x, y, z, g.dx, g.dy, g.dz, p);
    const uint32_t pos = stream_pos(g, x, y, z);
    const size_t fy = (size_t)px, fz = (size_t)px * py;
    const size_t f = ((size_t)(2 * z) * py + 2 * y) * px + 2 * x;
    const v2 r00 = *(const v2*)(full + f);
    const v2 r01 = *(const v2*)(full + f + fy);
    const v2 r10 = *(const v2*)(full + f + fz);
    const v2 r11 = *(const v2*)(full + f + fz + fy);
    const T tg[7] = {r00.y, r01.x, r01.y, r10.x, r10.y, r11.x, r11.y};
    T rec;
#pragma unroll
    for (int b = 0; b < 7; ++b) q[b] = enc1(tg[b], p[b], a, b, pos, codes, side, rec);
}

// low level: targets from the field at stride 4, reconstruction written into the mid grid
template <typename T>
__device__ __forceinline__ void low_enc_point(const T* __restrict__ full, int px, int py, const T* __restrict__ base,
                                              const Geom& g, const LevelArgs& a, uint16_t* __restrict__ codes,
                                              T* __restrict__ side, T* __restrict__ mid, int x, int y, int z,
                                              bool inner, bool vec, uint32_t q[7])
{
    typedef typename V2<T>::type v2;
    const int lx = g.dx, ly = g.dy;
    const int mx = 2 * lx, my = 2 * ly;
    T p[7];
    predict7(inner, base, x, y, z, lx, ly, g.dz, p);
    const uint32_t pos = stream_pos(g, x, y, z);
    // low block k at (x,y,z) = mid[2z+pz][2y+py][2x+px] = full[4z+2pz][4y+2py][4x+2px]
    const size_t fy = (size_t)2 * px, fz = (size_t)2 * px * py;
    const size_t f = ((size_t)(4 * z) * py + 4 * y) * px + 4 * x;
    const T tg[7] = {full[f + 2], full[f + fy], full[f + fy + 2], full[f + fz], full[f + fz + 2],
                     full[f + fz + fy], full[f + fz + fy + 2]};
    T r[8];
    r[0] = base[((size_t)z * ly + y) * lx + x];
#pragma unroll
    for (int b = 0; b < 7; ++b) q[b] = enc1(tg[b], p[b], a, b, pos, codes, side, r[b + 1]);
    const size_t my_ = (size_t)mx, mz_ = (size_t)mx * my;
    const size_t m = ((size_t)(2 * z) * my + 2 * y) * mx + 2 * x;
    if (vec) {
        v2 w;
        w.x = r[0]; w.y = r[1]; *(v2*)(mid + m) = w;
        w.x = r[2]; w.y = r[3]; *(v2*)(mid + m + my_) = w;
        w.x = r[4]; w.y = r[5]; *(v2*)(mid + m + mz_) = w;
        w.x = r[6]; w.y = r[7]; *(v2*)(mid + m + mz_ + my_) = w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) mid[m + (k >> 2) * mz_ + ((k >> 1) & 1) * my_ + (k & 1)] = r[k];
    }
}

#define STZ_HIST_ADD7(q, ok)                                                    \
    do {                                                                        \
        _Pragma("unroll") for (int b_ = 0; b_ < 7; ++b_) hist_add(sh[b_], hist + b_ * 65536, q[b_], ok); \
    } while (0)

template <typename T>
__global__ __launch_bounds__(256) void k_high_encode(const T* __restrict__ full, int px, int py,
                                                     const T* __restrict__ mid, Geom g, LevelArgs a,
                                                     uint16_t* __restrict__ codes, T* __restrict__ side,
                                                     uint32_t* __restrict__ hist)
{
    __shared__ uint32_t sh[7][kHW];
    for (int k = threadIdx.x; k < 7 * kHW; k += 256) sh[k / kHW][k % kHW] = 0u;
    __syncthreads();
    const int mx = g.dx, my = g.dy, mz = g.dz;
    const int nxi = mx - 3;  // interior columns 1 .. mx-3
    const int ntx = (nxi + 63) >> 6, nty = (my + 3) >> 2;
    const uint64_t ntiles = nxi > 0 ? (uint64_t)ntx * nty * mz : 0;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    for (uint64_t t = xcd_block(); t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), z = (int)(t / ((uint64_t)ntx * nty));
        const int x = 1 + tx * 64 + lane, y = ty * 4 + row;
        const bool ok = x <= mx - 3 && y < my;
        uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
        if (ok) high_enc_point(full, px, py, mid, g, a, codes, side, x, y, z, row_inner(y, z, my, mz), q);
        STZ_HIST_ADD7(q, ok);
    }
    hist_flush(sh, 7, hist);
}

template <typename T>
__global__ __launch_bounds__(256) void k_high_encode_edge(const T* __restrict__ full, int px, int py,
                                                          const T* __restrict__ mid, Geom g, LevelArgs a,
                                                          uint16_t* __restrict__ codes, T* __restrict__ side,
                                                          uint32_t* __restrict__ hist)
{
    __shared__ uint32_t sh[7][kHW];
    for (int k = threadIdx.x; k < 7 * kHW; k += 256) sh[k / kHW][k % kHW] = 0u;
    __syncthreads();
    const EdgeMap e = edge_map(g.dx);
    const uint64_t n = (uint64_t)e.ncol * g.dy * g.dz;
    const uint64_t kend = (n + 255) / 256 * 256;
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < kend; k += (uint64_t)gridDim.x * 256ull) {
        const bool ok = k < n;
        uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
        if (ok) {
            const int c = (int)(k % e.ncol), y = (int)((k / e.ncol) % g.dy), z = (int)(k / ((uint64_t)e.ncol * g.dy));
            high_enc_point(full, px, py, mid, g, a, codes, side, e.cols[c], y, z, false, q);
        }
        STZ_HIST_ADD7(q, ok);
    }
    hist_flush(sh, 7, hist);
}

template <typename T>
__global__ __launch_bounds__(256) void k_low_encode(const T* __restrict__ full, int px, int py,
                                                    const T* __restrict__ base, Geom g, LevelArgs a,
                                                    uint16_t* __restrict__ codes, T* __restrict__ side,
                                                    uint32_t* __restrict__ hist, T* __restrict__ mid)
{
    __shared__ uint32_t sh[7][kHW];
    for (int k = threadIdx.x; k < 7 * kHW; k += 256) sh[k / kHW][k % kHW] = 0u;
    __syncthreads();
    const int lx = g.dx, ly = g.dy, lz = g.dz;
    const int nxi = lx - 3;
    const int ntx = (nxi + 63) >> 6, nty = (ly + 3) >> 2;
    const uint64_t ntiles = nxi > 0 ? (uint64_t)ntx * nty * lz : 0;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    for (uint64_t t = xcd_block(); t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), z = (int)(t / ((uint64_t)ntx * nty));
        const int x = 1 + tx * 64 + lane, y = ty * 4 + row;
        const bool ok = x <= lx - 3 && y < ly;
        uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
        if (ok) low_enc_point(full, px, py, base, g, a, codes, side, mid, x, y, z, row_inner(y, z, ly, lz), true, q);
        STZ_HIST_ADD7(q, ok);
    }
    hist_flush(sh, 7, hist);
}

template <typename T>
__global__ __launch_bounds__(256) void k_low_encode_edge(const T* __restrict__ full, int px, int py,
                                                         const T* __restrict__ base, Geom g, LevelArgs a,
                                                         uint16_t* __restrict__ codes, T* __restrict__ side,
                                                         uint32_t* __restrict__ hist, T* __restrict__ mid)
{
    __shared__ uint32_t sh[7][kHW];
    for (int k = threadIdx.x; k < 7 * kHW; k += 256) sh[k / kHW][k % kHW] = 0u;
    __syncthreads();
    const EdgeMap e = edge_map(g.dx);
    const uint64_t n = (uint64_t)e.ncol * g.dy * g.dz;
    const uint64_t kend = (n + 255) / 256 * 256;
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < kend; k += (uint64_t)gridDim.x * 256ull) {
        const bool ok = k < n;
        uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
        if (ok) {
            const int c = (int)(k % e.ncol), y = (int)((k / e.ncol) % g.dy), z = (int)(k / ((uint64_t)e.ncol * g.dy));
            low_enc_point(full, px, py, base, g, a, codes, side, mid, e.cols[c], y, z, false, false, q);
        }
        STZ_HIST_ADD7(q, ok);
    }
    hist_flush(sh, 7, hist);
}

// ---------------------------------------------------------------- decode

template <typename T>
__device__ __forceinline__ T dec1(T p, const LevelArgs& a, int b, uint32_t pos, const uint16_t* __restrict__ codes,
                                  const T* __restrict__ side)
{
    const uint32_t q = codes[a.code_off[b] + pos];
    const T v = q ? dequantize<T>(q, a.eb) : side[a.code_off[b] + pos];
    return v + p;  // F5:1080 de_sub = deData + pred
}

template <typename T>
__device__ __forceinline__ void low_dec_point(const T* __restrict__ base, const Geom& g, const LevelArgs& a,
                                              const uint16_t* __restrict__ codes, const T* __restrict__ side,
                                              T* __restrict__ mid, int x, int y, int z, bool inner, bool vec)
{
    typedef typename V2<T>::type v2;
    const int lx = g.dx, ly = g.dy;
    const int mx = 2 * lx, my = 2 * ly;
    T p[7];
    predict7(inner, base, x, y, z, lx, ly, g.dz, p);
    const uint32_t pos = stream_pos(g, x, y, z);
    T r[8];
    r[0] = base[((size_t)z * ly + y) * lx + x];
#pragma unroll
    for (int b = 0; b < 7; ++b) r[b + 1] = dec1(p[b], a, b, pos, codes, side);
    const size_t my_ = (size_t)mx, mz_ = (size_t)mx * my;
    const size_t m = ((size_t)(2 * z) * my + 2 * y) * mx + 2 * x;
    if (vec) {
        v2 w;
        w.x = r[0]; w.y = r[1]; *(v2*)(mid + m) = w;
        w.x = r[2]; w.y = r[3]; *(v2*)(mid + m + my_) = w;
        w.x = r[4]; w.y = r[5]; *(v2*)(mid + m + mz_) = w;
        w.x = r[6]; w.y = r[7]; *(v2*)(mid + m + mz_ + my_) = w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) mid[m + (k >> 2) * mz_ + ((k >> 1) & 1) * my_ + (k & 1)] = r[k];
    }
}

// the 2x2x2 cell of mid point (x,y,z) into the caller's (nx,ny,nz) field;
// vec = 2-wide stores (nx even and the output 2-element aligned)
template <typename T>
__device__ __forceinline__ void high_dec_point(const T* __restrict__ mid, const Geom& g, const LevelArgs& a,
                                               const uint16_t* __restrict__ codes, const T* __restrict__ side,
                                               T* __restrict__ out, int nx, int ny, int nz, int x, int y, int z,
                                               bool inner, bool vec)
{
    typedef typename V2<T>::type v2;
    const int X = 2 * x, Y = 2 * y, Z = 2 * z;
    if (X >= nx || Y >= ny || Z >= nz) return;  // cell entirely in the padding
    T p[7];
    predict7(inner, mid, x, y, z, g.dx, g.dy, g.dz, p);
    const uint32_t pos = stream_pos(g, x, y, z);
    T v[8];
    v[0] = mid[((size_t)z * g.dy + y) * g.dx + x];
#pragma unroll
    for (int b = 0; b < 7; ++b) v[b + 1] = dec1(p[b], a, b, pos, codes, side);
    const size_t oy = (size_t)nx, oz = (size_t)nx * ny;
    const size_t o = (size_t)Z * oz + (size_t)Y * oy + X;
    if (vec) {
        v2 w;
        w.x = v[0]; w.y = v[1]; *(v2*)(out + o) = w;
        if (Y + 1 < ny) { w.x = v[2]; w.y = v[3]; *(v2*)(out + o + oy) = w; }
        if (Z + 1 < nz) {
            w.x = v[4]; w.y = v[5]; *(v2*)(out + o + oz) = w;
            if (Y + 1 < ny) { w.x = v[6]; w.y = v[7]; *(v2*)(out + o + oz + oy) = w; }
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int kx = k & 1, ky = (k >> 1) & 1, kz = k >> 2;
            if (X + kx < nx && Y + ky < ny && Z + kz < nz) out[o + (size_t)kz * oz + (size_t)ky * oy + kx] = v[k];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_low_decode(const T* __restrict__ base, Geom g, LevelArgs a,
                                                    const uint16_t* __restrict__ codes,
                                                    const T* __restrict__ side, T* __restrict__ mid)
{
    const int lx = g.dx, ly = g.dy, lz = g.dz;
    const int nxi = lx - 3;
    const int ntx = (nxi + 63) >> 6, nty = (ly + 3) >> 2;
    const uint64_t ntiles = nxi > 0 ? (uint64_t)ntx * nty * lz : 0;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    for (uint64_t t = xcd_block(); t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), z = (int)(t / ((uint64_t)ntx * nty));
        const int x = 1 + tx * 64 + lane, y = ty * 4 + row;
        if (x > lx - 3 || y >= ly) continue;
        low_dec_point(base, g, a, codes, side, mid, x, y, z, row_inner(y, z, ly, lz), true);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_low_decode_edge(const T* __restrict__ base, Geom g, LevelArgs a,
                                                         const uint16_t* __restrict__ codes,
                                                         const T* __restrict__ side, T* __restrict__ mid)
{
    const EdgeMap e = edge_map(g.dx);
    const uint64_t n = (uint64_t)e.ncol * g.dy * g.dz;
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < n; k += (uint64_t)gridDim.x * 256ull) {
        const int c = (int)(k % e.ncol), y = (int)((k / e.ncol) % g.dy), z = (int)(k / ((uint64_t)e.ncol * g.dy));
        low_dec_point(base, g, a, codes, side, mid, e.cols[c], y, z, false, false);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_high_decode(const T* __restrict__ mid, Geom g, LevelArgs a,
                                                     const uint16_t* __restrict__ codes,
                                                     const T* __restrict__ side, T* __restrict__ out,
                                                     int nx, int ny, int nz, int vec)
{
    const int mx = g.dx, my = g.dy, mz = g.dz;
    const int nxi = mx - 3;
    const int ntx = (nxi + 63) >> 6, nty = (my + 3) >> 2;
    const uint64_t ntiles = nxi > 0 ? (uint64_t)ntx * nty * mz : 0;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    for (uint64_t t = xcd_block(); t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), z = (int)(t / ((uint64_t)ntx * nty));
        const int x = 1 + tx * 64 + lane, y = ty * 4 + row;
        if (x > mx - 3 || y >= my) continue;
        high_dec_point(mid, g, a, codes, side, out, nx, ny, nz, x, y, z, row_inner(y, z, my, mz), vec != 0);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_high_decode_edge(const T* __restrict__ mid, Geom g, LevelArgs a,
                                                          const uint16_t* __restrict__ codes,
                                                          const T* __restrict__ side, T* __restrict__ out,
                                                          int nx, int ny, int nz)
{
    const EdgeMap e = edge_map(g.dx);
    const uint64_t n = (uint64_t)e.ncol * g.dy * g.dz;
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < n; k += (uint64_t)gridDim.x * 256ull) {
        const int c = (int)(k % e.ncol), y = (int)((k / e.ncol) % g.dy), z = (int)(k / ((uint64_t)e.ncol * g.dy));
        high_dec_point(mid, g, a, codes, side, out, nx, ny, nz, e.cols[c], y, z, false, false);
    }
}

// Random access reads codes through the chunk map: the element of block
// point (x,y,z) is cmap[cm_off[b] + chunk] + its chunk-local row-major index
// (the stream order of F5:1884-1905).  Without a map it is the stream
// position in the full layout.
__device__ __forceinline__ uint64_t ra_index(const LevelArgs& a, const Geom& g, int b, int x, int y, int z)
{
    if (!a.cmap) return a.code_off[b] + stream_pos(g, x, y, z);
    const int C = g.chunk;
    const int cx = x / C, cy = y / C, cz = z / C;
    const uint32_t m = a.cmap[a.cm_off[b] + ((uint32_t)cz * g.ncy + cy) * g.ncx + cx];
    if (m == kNoChunk) {
        atomicOr(a.err, 2u);
        return 0;
    }
    const int wx = imin_(C, g.dx - cx * C), wy = imin_(C, g.dy - cy * C);
    return (uint64_t)m + ((uint32_t)(z - cz * C) * wy + (uint32_t)(y - cy * C)) * wx + (uint32_t)(x - cx * C);
}

// the job a workgroup serves (jobs sorted by blk0, jobs[0].blk0 == 0)
__device__ __forceinline__ uint32_t find_job(const BoxJob* __restrict__ jobs, uint32_t njobs, uint32_t blk)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (jobs[m].blk0 <= blk) lo = m; else hi = m;
    }
    return lo;
}

// Random access: reference values of sub-block 0 on boxes of the mid grid
// (F5:2508-2538, low depreprocess restricted to the regions + gather)
template <typename T>
__global__ __launch_bounds__(256) void k_low_decode_boxes(const T* __restrict__ base, Geom g, LevelArgs a,
                                                          const uint16_t* __restrict__ codes,
                                                          const T* __restrict__ side, T* __restrict__ mid,
                                                          int mx, int my, const BoxJob* __restrict__ jobs,
                                                          uint32_t njobs, uint32_t nblocks)
{
    const uint32_t j = find_job(jobs, njobs, blockIdx.x);
    const BoxJob J = jobs[j];
    const uint32_t nb = (j + 1 < njobs ? jobs[j + 1].blk0 : nblocks) - J.blk0;
    const int bx = J.x1 - J.x0 + 1, by = J.y1 - J.y0 + 1, bz = J.z1 - J.z0 + 1;
    const uint64_t n = (uint64_t)bx * by * bz;
    for (uint64_t k = (blockIdx.x - J.blk0) * 256ull + threadIdx.x; k < n; k += (uint64_t)nb * 256u) {
        const uint64_t t = k / (uint64_t)bx;
        const int X = J.x0 + (int)(k - t * bx), Y = J.y0 + (int)(t % (uint64_t)by), Z = J.z0 + (int)(t / (uint64_t)by);
        const int par = ((Z & 1) << 2) | ((Y & 1) << 1) | (X & 1);
        const int x = X >> 1, y = Y >> 1, z = Z >> 1;
        const int i = (z * g.dy + y) * g.dx + x;
        T v;
        if (par == 0) {
            v = base[i];
        } else {
            const int b = par - 1;
            const uint64_t e = ra_index(a, g, b, x, y, z);
            const uint32_t q = codes[e];
            const T r = q ? dequantize<T>(q, a.eb) : side[e];
            v = r + predict_rt(b, base, i, x, y, z, g.dx, g.dy, g.dz);
        }
        mid[((size_t)Z * my + Y) * mx + X] = v;
    }
}

// Random access: full-resolution values of query boxes (F5:2574-2605)
template <typename T>
__global__ __launch_bounds__(256) void k_high_decode_boxes(const T* __restrict__ mid, Geom g, LevelArgs a,
                                                           const uint16_t* __restrict__ codes,
                                                           const T* __restrict__ side,
                                                           const BoxJob* __restrict__ jobs, uint32_t njobs,
                                                           uint32_t nblocks)
{
    const uint32_t j = find_job(jobs, njobs, blockIdx.x);
    const BoxJob J = jobs[j];
    T* __restrict__ out = (T*)J.out;
    const uint32_t nb = (j + 1 < njobs ? jobs[j + 1].blk0 : nblocks) - J.blk0;
    const int bx = J.x1 - J.x0 + 1, by = J.y1 - J.y0 + 1, bz = J.z1 - J.z0 + 1;
    const uint64_t n = (uint64_t)bx * by * bz;
    for (uint64_t k = (blockIdx.x - J.blk0) * 256ull + threadIdx.x; k < n; k += (uint64_t)nb * 256u) {
        const uint64_t t = k / (uint64_t)bx;
        const int X = J.x0 + (int)(k - t * bx), Y = J.y0 + (int)(t % (uint64_t)by), Z = J.z0 + (int)(t / (uint64_t)by);
        const int par = ((Z & 1) << 2) | ((Y & 1) << 1) | (X & 1);
        const int x = X >> 1, y = Y >> 1, z = Z >> 1;
        const int i = (z * g.dy + y) * g.dx + x;
        T v;
        if (par == 0) {
            v = mid[i];
        } else {
            const int b = par - 1;
            const uint64_t e = ra_index(a, g, b, x, y, z);
            const uint32_t q = codes[e];
            const T r = q ? dequantize<T>(q, a.eb) : side[e];
            v = r + predict_rt(b, mid, i, x, y, z, g.dx, g.dy, g.dz);
        }
        out[k] = v;
    }
}

// ---------------------------------------------------------------- many fields
// Persistent multi-field kernel: each workgroup takes work items from the
// queue head (one atomic per item) until the queue is empty.  Items are
// handed out in increasing order, so a workgroup meets each field once;
// it flushes its LDS histograms when the field changes.
__device__ __forceinline__ uint32_t find_level_job(const LevelJob* __restrict__ jobs, uint32_t njobs, uint32_t it)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (jobs[m].item0 <= it) lo = m; else hi = m;
    }
    return lo;
}

template <typename T, int kMode, int kWaves>
__global__ __launch_bounds__(64 * kWaves) void k_level_multi(const LevelJob* __restrict__ jobs, uint32_t njobs,
                                                             uint32_t total, uint32_t* __restrict__ counter)
{
    constexpr bool kEnc = kMode == kLowEnc || kMode == kHighEnc;
    constexpr int kThreads = 64 * kWaves;
    __shared__ uint32_t sh[kEnc ? 7 : 1][kHW];
    __shared__ uint32_t item_s;
    if constexpr (kEnc) {
        for (int k = threadIdx.x; k < 7 * kHW; k += kThreads) sh[k / kHW][k % kHW] = 0u;
    }
    int cur = -1;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = item_s;
        __syncthreads();
        if (it >= total) break;  // every wave of the workgroup sees the same item
        const int j = (int)find_level_job(jobs, njobs, it);
        const LevelJob& J = jobs[j];
        if constexpr (kEnc) {
            if (j != cur && cur >= 0) {
                hist_flush(sh, 7, jobs[cur].hist);
                __syncthreads();
                for (int k = threadIdx.x; k < 7 * kHW; k += kThreads) sh[k / kHW][k % kHW] = 0u;
            }
            if (j != cur) __syncthreads();
            cur = j;
        }
        const Geom g = J.g;
        const uint32_t local = it - J.item0;
        int x, y, z;
        bool ok, inner, vec;
        if (local < J.ntiles) {
            const int nxi = g.dx - 3;
            const int ntx = (nxi + 63) >> 6, nty = (g.dy + kWaves - 1) / kWaves;
            const int tx = (int)(local % ntx), ty = (int)((local / ntx) % nty), tz = (int)(local / ((uint32_t)ntx * nty));
            x = 1 + tx * 64 + lane;
            y = ty * kWaves + row;
            z = tz;
            ok = x <= g.dx - 3 && y < g.dy;
            inner = row_inner(y, z, g.dy, g.dz);
            vec = true;
        } else {
            const EdgeMap e = edge_map(g.dx);
            const uint64_t k = (uint64_t)(local - J.ntiles) * kThreads + threadIdx.x;
            ok = k < (uint64_t)e.ncol * g.dy * g.dz;
            const int c = (int)(k % e.ncol);
            y = (int)((k / e.ncol) % g.dy);
            z = (int)(k / ((uint64_t)e.ncol * g.dy));
            x = e.cols[c];
            inner = false;
            vec = false;
        }
        if constexpr (kEnc) {
            uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
            if (ok) {
                if (kMode == kHighEnc)
                    high_enc_point((const T*)J.full, J.px, J.py, (const T*)J.ref, g, J.a, J.codes, (T*)J.side, x, y,
                                   z, inner, q);
                else
                    low_enc_point((const T*)J.full, J.px, J.py, (const T*)J.ref, g, J.a, J.codes, (T*)J.side,
                                  (T*)J.mid, x, y, z, inner, vec, q);
            }
            uint32_t* hist = J.hist;
            STZ_HIST_ADD7(q, ok);
        } else if (ok) {
            if (kMode == kHighDec)
                high_dec_point((const T*)J.ref, g, J.a, J.codes, (const T*)J.side, (T*)J.out, J.nx, J.ny, J.nz, x, y,
                               z, inner, vec && J.vec != 0);
            else
                low_dec_point((const T*)J.ref, g, J.a, J.codes, (const T*)J.side, (T*)J.mid, x, y, z, inner, vec);
        }
    }
    if constexpr (kEnc) {
        if (cur >= 0) hist_flush(sh, 7, jobs[cur].hist);
    }
}

inline unsigned tile_grid(int dx, int dy, int dz)
{
    const int nxi = dx - 3 > 0 ? dx - 3 : 1;
    const uint64_t nt = (uint64_t)((nxi + 63) / 64) * ((dy + 3) / 4) * dz;
    uint64_t g = nt < 2048 ? nt : 2048;
    if (g >= 64) g &= ~7ull;  // multiple of 8 for the XCD mapping
    return (unsigned)(g ? g : 1);
}

inline unsigned flat_grid(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : b > 8192 ? 8192 : b);
}

inline uint64_t edge_points(const Geom& g) { return (uint64_t)edge_map(g.dx).ncol * g.dy * g.dz; }

}  // namespace

template <typename T>
void launch_low_encode(hipStream_t s, const T* full, int px, int py, int pz, const T* base, Geom g, LevelArgs a,
                       uint16_t* codes, T* side, uint32_t* hist, T* mid)
{
    (void)pz;
    if (g.dx >= 4)
        k_low_encode<T><<<dim3(tile_grid(g.dx, g.dy, g.dz)), dim3(256), 0, s>>>(full, px, py, base, g, a, codes,
                                                                               side, hist, mid);
    k_low_encode_edge<T><<<dim3(flat_grid(edge_points(g))), dim3(256), 0, s>>>(full, px, py, base, g, a, codes, side,
                                                                             hist, mid);
}

template <typename T>
void launch_high_encode(hipStream_t s, const T* full, int px, int py, int pz, const T* mid, Geom g, LevelArgs a,
                        uint16_t* codes, T* side, uint32_t* hist)
{
    (void)pz;
    if (g.dx >= 4)
        k_high_encode<T><<<dim3(tile_grid(g.dx, g.dy, g.dz)), dim3(256), 0, s>>>(full, px, py, mid, g, a, codes,
                                                                                side, hist);
    k_high_encode_edge<T><<<dim3(flat_grid(edge_points(g))), dim3(256), 0, s>>>(full, px, py, mid, g, a, codes,
                                                                              side, hist);
}

template <typename T>
void launch_low_decode(hipStream_t s, const T* base, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                       T* mid)
{
    if (g.dx >= 4)
        k_low_decode<T><<<dim3(tile_grid(g.dx, g.dy, g.dz)), dim3(256), 0, s>>>(base, g, a, codes, side, mid);
    k_low_decode_edge<T><<<dim3(flat_grid(edge_points(g))), dim3(256), 0, s>>>(base, g, a, codes, side, mid);
}

template <typename T>
void launch_high_decode(hipStream_t s, const T* mid, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                        T* out, int nx, int ny, int nz)
{
    const int vec = ((nx & 1) == 0) && (((uintptr_t)out % (2 * sizeof(T))) == 0);
    if (g.dx >= 4)
        k_high_decode<T><<<dim3(tile_grid(g.dx, g.dy, g.dz)), dim3(256), 0, s>>>(mid, g, a, codes, side, out, nx,
                                                                                ny, nz, vec);
    k_high_decode_edge<T><<<dim3(flat_grid(edge_points(g))), dim3(256), 0, s>>>(mid, g, a, codes, side, out, nx,
                                                                             ny, nz);
}

template <typename T>
void launch_low_decode_boxes(hipStream_t s, const T* base, Geom g, LevelArgs a, const uint16_t* codes,
                             const T* side, T* mid, int mx, int my, const BoxJob* jobs, uint32_t njobs,
                             uint32_t nblocks)
{
    if (!njobs || !nblocks) return;
    k_low_decode_boxes<T><<<dim3(nblocks), dim3(256), 0, s>>>(base, g, a, codes, side, mid, mx, my, jobs, njobs,
                                                              nblocks);
}

template <typename T>
void launch_high_decode_boxes(hipStream_t s, const T* mid, Geom g, LevelArgs a, const uint16_t* codes,
                              const T* side, const BoxJob* jobs, uint32_t njobs, uint32_t nblocks)
{
    if (!njobs || !nblocks) return;
    k_high_decode_boxes<T><<<dim3(nblocks), dim3(256), 0, s>>>(mid, g, a, codes, side, jobs, njobs, nblocks);
}

uint32_t level_tiles(const Geom& g, int waves)
{
    const int nxi = g.dx - 3;
    if (nxi <= 0) return 0;
    return (uint32_t)(((nxi + 63) / 64) * ((g.dy + waves - 1) / waves) * g.dz);
}

uint32_t level_edge_groups(const Geom& g, int waves)
{
    const uint64_t n = edge_points(g);
    return (uint32_t)((n + 64u * waves - 1) / (64u * waves));
}

template <typename T, int kW>
void level_multi_w(hipStream_t s, int mode, const LevelJob* jobs, uint32_t njobs, uint32_t total, uint32_t* counter,
                   uint32_t grid)
{
    const dim3 G(grid), B(64 * kW);
    switch (mode) {
    case kLowEnc: k_level_multi<T, kLowEnc, kW><<<G, B, 0, s>>>(jobs, njobs, total, counter); break;
    case kHighEnc: k_level_multi<T, kHighEnc, kW><<<G, B, 0, s>>>(jobs, njobs, total, counter); break;
    case kLowDec: k_level_multi<T, kLowDec, kW><<<G, B, 0, s>>>(jobs, njobs, total, counter); break;
    default: k_level_multi<T, kHighDec, kW><<<G, B, 0, s>>>(jobs, njobs, total, counter); break;
    }
}

template <typename T>
void launch_level_multi(hipStream_t s, int mode, const LevelJob* jobs, uint32_t njobs, uint32_t total_items,
                        uint32_t* counter, int waves, uint32_t grid)
{
    if (!njobs || !total_items || !grid) return;
    if (waves == 8) level_multi_w<T, 8>(s, mode, jobs, njobs, total_items, counter, grid);
    else level_multi_w<T, 4>(s, mode, jobs, njobs, total_items, counter, grid);
}

#define STZ_INST(T)                                                                                              \
    template void launch_low_encode<T>(hipStream_t, const T*, int, int, int, const T*, Geom, LevelArgs,         \
                                       uint16_t*, T*, uint32_t*, T*);                                            \
    template void launch_high_encode<T>(hipStream_t, const T*, int, int, int, const T*, Geom, LevelArgs,        \
                                        uint16_t*, T*, uint32_t*);                                               \
    template void launch_low_decode<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, T*);  \
    template void launch_high_decode<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, T*,  \
                                        int, int, int);                                                          \
    template void launch_low_decode_boxes<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, \
                                             T*, int, int, const BoxJob*, uint32_t, uint32_t);                  \
    template void launch_high_decode_boxes<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*,         \
                                              const T*, const BoxJob*, uint32_t, uint32_t);                      \
    template void launch_level_multi<T>(hipStream_t, int, const LevelJob*, uint32_t, uint32_t, uint32_t*, int,  \
                                        uint32_t);
STZ_INST(float)
STZ_INST(double)

}  // namespace stz

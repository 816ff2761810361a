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

template <typename T>
struct Tile {
    int x, y, z;
    bool ok;
};

// ---------------------------------------------------------------- encode

template <int B, typename T>
__device__ __forceinline__ void enc_one(const T* __restrict__ ref, int i, int x, int y, int z, int dx, int dy,
                                        int dz, T tgt, const LevelArgs& a, uint32_t pos,
                                        uint16_t* __restrict__ codes, T* __restrict__ side, uint32_t& q, T& rec)
{
    const T p = predict<B>(ref, i, x, y, z, dx, dy, dz);
    const T d = tgt - p;
    T dec;
    q = quantize(d, a.eb, a.recip, dec);
    codes[a.code_off[B] + pos] = (uint16_t)q;
    if (!q) side[a.code_off[B] + pos] = d;
    rec = dec + p;  // decoded diff + prediction (F5:1381-1385 then F5:1077-1083)
}

template <typename T>
__global__ __launch_bounds__(256) void k_high_encode(const T* __restrict__ full, int px, int py,
                                                     const T* __restrict__ mid, Geom g, LevelArgs a,
                                                     uint16_t* __restrict__ codes, T* __restrict__ side,
                                                     uint32_t* __restrict__ hist)
{
    typedef typename V2<T>::type v2;
    __shared__ uint32_t sh[7][kHW];
    for (int k = threadIdx.x; k < 7 * kHW; k += 256) sh[k / kHW][k % kHW] = 0u;
    __syncthreads();
    const int mx = g.dx, my = g.dy, mz = g.dz;
    const int ntx = (mx + 63) >> 6, nty = (my + 3) >> 2;
    const uint64_t ntiles = (uint64_t)ntx * nty * mz;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const size_t fy = (size_t)px, fz = (size_t)px * py;
    for (uint64_t t = xcd_block(); t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), z = (int)(t / ((uint64_t)ntx * nty));
        const int x = tx * 64 + lane, y = ty * 4 + row;
        const bool ok = x < mx && y < my;
        uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0, q4 = 0, q5 = 0, q6 = 0;
        if (ok) {
            const int i = (z * my + y) * mx + x;
            const uint32_t pos = stream_pos(g, x, y, z);
            const size_t f = ((size_t)(2 * z) * py + 2 * y) * px + 2 * x;
            const v2 r00 = *(const v2*)(full + f);
            const v2 r01 = *(const v2*)(full + f + fy);
            const v2 r10 = *(const v2*)(full + f + fz);
            const v2 r11 = *(const v2*)(full + f + fz + fy);
            T rec;
            enc_one<0>(mid, i, x, y, z, mx, my, mz, r00.y, a, pos, codes, side, q0, rec);
            enc_one<1>(mid, i, x, y, z, mx, my, mz, r01.x, a, pos, codes, side, q1, rec);
            enc_one<2>(mid, i, x, y, z, mx, my, mz, r01.y, a, pos, codes, side, q2, rec);
            enc_one<3>(mid, i, x, y, z, mx, my, mz, r10.x, a, pos, codes, side, q3, rec);
            enc_one<4>(mid, i, x, y, z, mx, my, mz, r10.y, a, pos, codes, side, q4, rec);
            enc_one<5>(mid, i, x, y, z, mx, my, mz, r11.x, a, pos, codes, side, q5, rec);
            enc_one<6>(mid, i, x, y, z, mx, my, mz, r11.y, a, pos, codes, side, q6, rec);
        }
        hist_add(sh[0], hist + 0 * 65536, q0, ok);
        hist_add(sh[1], hist + 1 * 65536, q1, ok);
        hist_add(sh[2], hist + 2 * 65536, q2, ok);
        hist_add(sh[3], hist + 3 * 65536, q3, ok);
        hist_add(sh[4], hist + 4 * 65536, q4, ok);
        hist_add(sh[5], hist + 5 * 65536, q5, ok);
        hist_add(sh[6], hist + 6 * 65536, q6, ok);
    }
    hist_flush(sh, 7, hist);
}

template <typename T>
__global__ __launch_bounds__(256) void k_low_encode(const T* __restrict__ full, int px, int py,
                                                    const T* __restrict__ base, Geom g, LevelArgs a,
                                                    uint16_t* __restrict__ codes, T* __restrict__ side,
                                                    uint32_t* __restrict__ hist, T* __restrict__ mid)
{
    typedef typename V2<T>::type v2;
    __shared__ uint32_t sh[7][kHW];
    for (int k = threadIdx.x; k < 7 * kHW; k += 256) sh[k / kHW][k % kHW] = 0u;
    __syncthreads();
    const int lx = g.dx, ly = g.dy, lz = g.dz;
    const int mx = 2 * lx, my = 2 * ly;
    const int ntx = (lx + 63) >> 6, nty = (ly + 3) >> 2;
    const uint64_t ntiles = (uint64_t)ntx * nty * lz;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const size_t fy = (size_t)2 * px, fz = (size_t)2 * px * py;  // one mid step = two full steps
    const size_t my_ = (size_t)mx, mz_ = (size_t)mx * my;
    for (uint64_t t = xcd_block(); t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), z = (int)(t / ((uint64_t)ntx * nty));
        const int x = tx * 64 + lane, y = ty * 4 + row;
        const bool ok = x < lx && y < ly;
        uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0, q4 = 0, q5 = 0, q6 = 0;
        if (ok) {
            const int i = (z * ly + y) * lx + x;
            const uint32_t pos = stream_pos(g, x, y, z);
            // low block k at (x,y,z) = mid[2z+pz][2y+py][2x+px] = full[4z+2pz][4y+2py][4x+2px]
            const size_t f = ((size_t)(4 * z) * py + 4 * y) * px + 4 * x;
            const T t1 = full[f + 2], t2 = full[f + fy], t3 = full[f + fy + 2];
            const T t4 = full[f + fz], t5 = full[f + fz + 2], t6 = full[f + fz + fy], t7 = full[f + fz + fy + 2];
            T r1, r2, r3, r4, r5, r6, r7;
            enc_one<0>(base, i, x, y, z, lx, ly, lz, t1, a, pos, codes, side, q0, r1);
            enc_one<1>(base, i, x, y, z, lx, ly, lz, t2, a, pos, codes, side, q1, r2);
            enc_one<2>(base, i, x, y, z, lx, ly, lz, t3, a, pos, codes, side, q2, r3);
            enc_one<3>(base, i, x, y, z, lx, ly, lz, t4, a, pos, codes, side, q3, r4);
            enc_one<4>(base, i, x, y, z, lx, ly, lz, t5, a, pos, codes, side, q4, r5);
            enc_one<5>(base, i, x, y, z, lx, ly, lz, t6, a, pos, codes, side, q5, r6);
            enc_one<6>(base, i, x, y, z, lx, ly, lz, t7, a, pos, codes, side, q6, r7);
            // unpredictable points reconstruct as raw residual + prediction
            const size_t m = ((size_t)(2 * z) * my + 2 * y) * mx + 2 * x;
            v2 w;
            w.x = base[i]; w.y = r1; *(v2*)(mid + m) = w;
            w.x = r2; w.y = r3; *(v2*)(mid + m + my_) = w;
            w.x = r4; w.y = r5; *(v2*)(mid + m + mz_) = w;
            w.x = r6; w.y = r7; *(v2*)(mid + m + mz_ + my_) = w;
        }
        hist_add(sh[0], hist + 0 * 65536, q0, ok);
        hist_add(sh[1], hist + 1 * 65536, q1, ok);
        hist_add(sh[2], hist + 2 * 65536, q2, ok);
        hist_add(sh[3], hist + 3 * 65536, q3, ok);
        hist_add(sh[4], hist + 4 * 65536, q4, ok);
        hist_add(sh[5], hist + 5 * 65536, q5, ok);
        hist_add(sh[6], hist + 6 * 65536, q6, ok);
    }
    hist_flush(sh, 7, hist);
}

// ---------------------------------------------------------------- decode

template <int B, typename T>
__device__ __forceinline__ T dec_one(const T* __restrict__ ref, int i, int x, int y, int z, int dx, int dy,
                                     int dz, const LevelArgs& a, uint32_t pos,
                                     const uint16_t* __restrict__ codes, const T* __restrict__ side)
{
    const uint32_t q = codes[a.code_off[B] + pos];
    const T v = q ? dequantize<T>(q, a.eb) : side[a.code_off[B] + pos];
    return v + predict<B>(ref, i, x, y, z, dx, dy, dz);  // F5:1080 de_sub = deData + pred
}

template <typename T>
__global__ __launch_bounds__(256) void k_low_decode(const T* __restrict__ base, Geom g, LevelArgs a,
                                                    const uint16_t* __restrict__ codes,
                                                    const T* __restrict__ side, T* __restrict__ mid)
{
    typedef typename V2<T>::type v2;
    const int lx = g.dx, ly = g.dy, lz = g.dz;
    const int mx = 2 * lx, my = 2 * ly;
    const int ntx = (lx + 63) >> 6, nty = (ly + 3) >> 2;
    const uint64_t ntiles = (uint64_t)ntx * nty * lz;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const size_t my_ = (size_t)mx, mz_ = (size_t)mx * my;
    for (uint64_t t = xcd_block(); t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), z = (int)(t / ((uint64_t)ntx * nty));
        const int x = tx * 64 + lane, y = ty * 4 + row;
        if (x >= lx || y >= ly) continue;
        const int i = (z * ly + y) * lx + x;
        const uint32_t pos = stream_pos(g, x, y, z);
        const size_t m = ((size_t)(2 * z) * my + 2 * y) * mx + 2 * x;
        v2 w;
        w.x = base[i];
        w.y = dec_one<0>(base, i, x, y, z, lx, ly, lz, a, pos, codes, side);
        *(v2*)(mid + m) = w;
        w.x = dec_one<1>(base, i, x, y, z, lx, ly, lz, a, pos, codes, side);
        w.y = dec_one<2>(base, i, x, y, z, lx, ly, lz, a, pos, codes, side);
        *(v2*)(mid + m + my_) = w;
        w.x = dec_one<3>(base, i, x, y, z, lx, ly, lz, a, pos, codes, side);
        w.y = dec_one<4>(base, i, x, y, z, lx, ly, lz, a, pos, codes, side);
        *(v2*)(mid + m + mz_) = w;
        w.x = dec_one<5>(base, i, x, y, z, lx, ly, lz, a, pos, codes, side);
        w.y = dec_one<6>(base, i, x, y, z, lx, ly, lz, a, pos, codes, side);
        *(v2*)(mid + m + mz_ + my_) = w;
    }
}

// writes the 2x2x2 cell of mid point (x,y,z) into the caller's (nx,ny,nz)
// field; vec = 1 when nx is even and the output is 2-element aligned
template <typename T>
__global__ __launch_bounds__(256) void k_high_decode(const T* __restrict__ mid, Geom g, LevelArgs a,
                                                     const uint16_t* __restrict__ codes,
                                                     const T* __restrict__ side, T* __restrict__ out,
                                                     int nx, int ny, int nz, int vec)
{
    typedef typename V2<T>::type v2;
    const int mx = g.dx, my = g.dy, mz = g.dz;
    const int ntx = (mx + 63) >> 6, nty = (my + 3) >> 2;
    const uint64_t ntiles = (uint64_t)ntx * nty * mz;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const size_t oy = (size_t)nx, oz = (size_t)nx * ny;
    for (uint64_t t = xcd_block(); t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % ntx), ty = (int)((t / ntx) % nty), z = (int)(t / ((uint64_t)ntx * nty));
        const int x = tx * 64 + lane, y = ty * 4 + row;
        if (x >= mx || y >= my) continue;
        const int X = 2 * x, Y = 2 * y, Z = 2 * z;
        if (X >= nx || Y >= ny || Z >= nz) continue;  // cell entirely in the padding
        const int i = (z * my + y) * mx + x;
        const uint32_t pos = stream_pos(g, x, y, z);
        T v[8];
        v[0] = mid[i];
        v[1] = dec_one<0>(mid, i, x, y, z, mx, my, mz, a, pos, codes, side);
        v[2] = dec_one<1>(mid, i, x, y, z, mx, my, mz, a, pos, codes, side);
        v[3] = dec_one<2>(mid, i, x, y, z, mx, my, mz, a, pos, codes, side);
        v[4] = dec_one<3>(mid, i, x, y, z, mx, my, mz, a, pos, codes, side);
        v[5] = dec_one<4>(mid, i, x, y, z, mx, my, mz, a, pos, codes, side);
        v[6] = dec_one<5>(mid, i, x, y, z, mx, my, mz, a, pos, codes, side);
        v[7] = dec_one<6>(mid, i, x, y, z, mx, my, mz, a, pos, codes, side);
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
                if (X + kx < nx && Y + ky < ny && Z + kz < nz)
                    out[o + (size_t)kz * oz + (size_t)ky * oy + kx] = v[k];
            }
        }
    }
}

// ROI: reference values of sub-block 0 on a box of the mid grid
// (F5:2508-2538 low depreprocess restricted to the regions + gather)
template <typename T>
__global__ void k_low_decode_box(const T* __restrict__ base, Geom g, LevelArgs a,
                                 const uint16_t* __restrict__ codes, const T* __restrict__ side,
                                 T* __restrict__ mid, int mx, int my, int x0, int x1, int y0, int y1, int z0,
                                 int z1)
{
    const int bx = x1 - x0 + 1, by = y1 - y0 + 1, bz = z1 - z0 + 1;
    const uint64_t n = (uint64_t)bx * by * bz;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const int X = x0 + (int)(k % bx), Y = y0 + (int)((k / bx) % by), Z = z0 + (int)(k / ((uint64_t)bx * by));
        const int par = ((Z & 1) << 2) | ((Y & 1) << 1) | (X & 1);
        const int x = X >> 1, y = Y >> 1, z = Z >> 1;
        const int i = (z * g.dy + y) * g.dx + x;
        T v;
        if (par == 0) {
            v = base[i];
        } else {
            const int b = par - 1;
            const uint32_t pos = stream_pos(g, x, y, z);
            const uint32_t q = codes[a.code_off[b] + pos];
            const T r = q ? dequantize<T>(q, a.eb) : side[a.code_off[b] + pos];
            v = r + predict_rt(b, base, i, x, y, z, g.dx, g.dy, g.dz);
        }
        mid[((size_t)Z * my + Y) * mx + X] = v;
    }
}

// ROI: full-resolution values of the query box (F5:2574-2605)
template <typename T>
__global__ void k_high_decode_box(const T* __restrict__ mid, Geom g, LevelArgs a,
                                  const uint16_t* __restrict__ codes, const T* __restrict__ side,
                                  T* __restrict__ out, int x0, int x1, int y0, int y1, int z0, int z1)
{
    const int bx = x1 - x0 + 1, by = y1 - y0 + 1, bz = z1 - z0 + 1;
    const uint64_t n = (uint64_t)bx * by * bz;
    for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x) {
        const int X = x0 + (int)(k % bx), Y = y0 + (int)((k / bx) % by), Z = z0 + (int)(k / ((uint64_t)bx * by));
        const int par = ((Z & 1) << 2) | ((Y & 1) << 1) | (X & 1);
        const int x = X >> 1, y = Y >> 1, z = Z >> 1;
        const int i = (z * g.dy + y) * g.dx + x;
        T v;
        if (par == 0) {
            v = mid[i];
        } else {
            const int b = par - 1;
            const uint32_t pos = stream_pos(g, x, y, z);
            const uint32_t q = codes[a.code_off[b] + pos];
            const T r = q ? dequantize<T>(q, a.eb) : side[a.code_off[b] + pos];
            v = r + predict_rt(b, mid, i, x, y, z, g.dx, g.dy, g.dz);
        }
        out[k] = v;
    }
}

inline unsigned tile_grid(int dx, int dy, int dz)
{
    const uint64_t nt = (uint64_t)((dx + 63) / 64) * ((dy + 3) / 4) * dz;
    uint64_t g = nt < 2048 ? nt : 2048;
    if (g >= 64) g &= ~7ull;  // multiple of 8 for the XCD mapping
    return (unsigned)(g ? g : 1);
}

inline unsigned flat_grid(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : b > 8192 ? 8192 : b);
}

}  // namespace

template <typename T>
void launch_low_encode(hipStream_t s, const T* full, int px, int py, int pz, const T* base, Geom g, LevelArgs a,
                       uint16_t* codes, T* side, uint32_t* hist, T* mid)
{
    (void)pz;
    k_low_encode<T><<<dim3(tile_grid(g.dx, g.dy, g.dz)), dim3(256), 0, s>>>(full, px, py, base, g, a, codes, side,
                                                                           hist, mid);
}

template <typename T>
void launch_high_encode(hipStream_t s, const T* full, int px, int py, int pz, const T* mid, Geom g, LevelArgs a,
                        uint16_t* codes, T* side, uint32_t* hist)
{
    (void)pz;
    k_high_encode<T><<<dim3(tile_grid(g.dx, g.dy, g.dz)), dim3(256), 0, s>>>(full, px, py, mid, g, a, codes, side,
                                                                            hist);
}

template <typename T>
void launch_low_decode(hipStream_t s, const T* base, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                       T* mid)
{
    k_low_decode<T><<<dim3(tile_grid(g.dx, g.dy, g.dz)), dim3(256), 0, s>>>(base, g, a, codes, side, mid);
}

template <typename T>
void launch_high_decode(hipStream_t s, const T* mid, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                        T* out, int nx, int ny, int nz)
{
    const int vec = ((nx & 1) == 0) && (((uintptr_t)out % (2 * sizeof(T))) == 0);
    k_high_decode<T><<<dim3(tile_grid(g.dx, g.dy, g.dz)), dim3(256), 0, s>>>(mid, g, a, codes, side, out, nx, ny,
                                                                            nz, vec);
}

template <typename T>
void launch_low_decode_box(hipStream_t s, const T* base, Geom g, LevelArgs a, const uint16_t* codes,
                           const T* side, T* mid, int mx, int my, int x0, int x1, int y0, int y1, int z0, int z1)
{
    const uint64_t n = (uint64_t)(x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
    k_low_decode_box<T><<<dim3(flat_grid(n)), dim3(256), 0, s>>>(base, g, a, codes, side, mid, mx, my, x0, x1, y0,
                                                                y1, z0, z1);
}

template <typename T>
void launch_high_decode_box(hipStream_t s, const T* mid, Geom g, LevelArgs a, const uint16_t* codes,
                            const T* side, T* out, int x0, int x1, int y0, int y1, int z0, int z1)
{
    const uint64_t n = (uint64_t)(x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
    k_high_decode_box<T><<<dim3(flat_grid(n)), dim3(256), 0, s>>>(mid, g, a, codes, side, out, x0, x1, y0, y1,
                                                                 z0, z1);
}

#define STZ_INST(T)                                                                                              \
    template void launch_low_encode<T>(hipStream_t, const T*, int, int, int, const T*, Geom, LevelArgs,         \
                                       uint16_t*, T*, uint32_t*, T*);                                            \
    template void launch_high_encode<T>(hipStream_t, const T*, int, int, int, const T*, Geom, LevelArgs,        \
                                        uint16_t*, T*, uint32_t*);                                               \
    template void launch_low_decode<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, T*);  \
    template void launch_high_decode<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, T*,  \
                                        int, int, int);                                                          \
    template void launch_low_decode_box<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*,   \
                                           T*, int, int, int, int, int, int, int, int);                          \
    template void launch_high_decode_box<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*,  \
                                            T*, int, int, int, int, int, int);
STZ_INST(float)
STZ_INST(double)

}  // namespace stz

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

#include <map>
#include <mutex>

#include "device_math.h"
#include "kernels.h"

namespace stz {

namespace {

constexpr int kHW = 1024;                 // LDS histogram window
constexpr uint32_t kHLo = kRadius - kHW / 2;

template <typename T> struct V2;
template <> struct V2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct V2<double> { typedef double type __attribute__((ext_vector_type(2))); };

// index of the calling wave in its workgroup, as a wave-uniform (scalar) value
// so that row-dependent addresses stay in scalar registers
__device__ __forceinline__ int wave_row()
{
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// blockIdx -> logical block so that each XCD (round-robin dispatch over 8)
// walks a contiguous range of work items and keeps its halo rows in its own L2
__device__ __forceinline__ uint32_t xcd_block()
{
    const uint32_t G = gridDim.x;
    if (G % 8u) return blockIdx.x;
    return (blockIdx.x % 8u) * (G / 8u) + blockIdx.x / 8u;
}

__device__ __forceinline__ void hist_put(uint32_t* sh, uint32_t* gh, uint32_t q, uint32_t n)
{
    if (q - kHLo < (uint32_t)kHW) lds_add(&sh[q - kHLo], n);
    else glb_add(&gh[q], n);
}

// Histogram update by every active lane of the wave (`valid` says which
// lanes count a code).  The centre code (zero residual, by far the most
// frequent) is counted with one ballot and a single atomic; the other codes
// go in per lane.
__device__ __forceinline__ void hist_add(uint32_t* sh, uint32_t* gh, uint32_t q, bool valid)
{
    const bool centre = valid && q == (uint32_t)kRadius;
    const unsigned long long m = __ballot(centre);
    const unsigned long long act = __ballot(1);
    if (m && __lane_id() == (unsigned)(__ffsll((long long)act) - 1))
        lds_add(&sh[kRadius - kHLo], (uint32_t)__popcll(m));
    if (valid && !centre) hist_put(sh, gh, q, 1u);
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

// element k of the row at `base` relative to the lane's byte offset xo: a
// wave-uniform base plus a 32-bit lane offset, so the load is addressed as
// scalar base + vector offset + immediate with no 64-bit vector arithmetic
template <typename T>
__device__ __forceinline__ T ldx(const T* __restrict__ base, uint32_t xo, int k)
{
    return *(const T*)((const char*)base + xo + (ptrdiff_t)k * (ptrdiff_t)sizeof(T));
}

// Interior predictions of all 7 blocks of a point from an accessor:
// at(dy, dz, k) is the reconstructed value at (x + k, y + dy, z + dz), with
// dy, dz in -1..2 and k in -1..2 (requires 1 <= x <= dx-3 and the same on y
// and z).  Same expressions, in the same order, as predict<B>'s interior;
// the global-memory and LDS forms below both go through it.
// `opt`: F5's predictor options (kPred* bits; a constant 0 in the default
// kernels, so their code is F5's default build only).
template <typename T, typename At>
__device__ __forceinline__ void predict7_at(const At& at, T p[7], int opt)
{
    // A(ox, oy, oz): the value at that offset from the point
    auto A = [&](int ox, int oy, int oz) -> T { return at(oy, oz, ox); };
    const T c = at(0, 0, 0);
    p[0] = cub(at(0, 0, -1), c, at(0, 0, 1), at(0, 0, 2));
    p[1] = cub(at(-1, 0, 0), c, at(1, 0, 0), at(2, 0, 0));
    p[3] = cub(at(0, -1, 0), c, at(0, 1, 0), at(0, 2, 0));
    T p0, v0, p1, v1, p2, v2, p3, v3;
    if (opt & kPredBicubic12) {
        p[2] = bicubic12<T>([&](int o1, int o2) { return A(o1, o2, 0); });
        p[4] = bicubic12<T>([&](int o1, int o2) { return A(o1, 0, o2); });
        p[5] = bicubic12<T>([&](int o1, int o2) { return A(0, o1, o2); });
    } else if (opt & kPredPlaneAxis) {
        p[2] = plane_dyn<T>([&](int o1, int o2) { return A(o1, o2, 0); });
        p[4] = plane_dyn<T>([&](int o1, int o2) { return A(o1, 0, o2); });
        p[5] = plane_dyn<T>([&](int o1, int o2) { return A(0, o1, o2); });
    } else {
        pv(at(-1, 0, -1), c, at(1, 0, 1), at(2, 0, 2), p0, v0);
        pv(at(-1, 0, 2), at(0, 0, 1), at(1, 0, 0), at(2, 0, -1), p1, v1);
        p[2] = blend2(p0, v0, p1, v1, opt);
        pv(at(0, -1, -1), c, at(0, 1, 1), at(0, 2, 2), p0, v0);
        pv(at(0, -1, 2), at(0, 0, 1), at(0, 1, 0), at(0, 2, -1), p1, v1);
        p[4] = blend2(p0, v0, p1, v1, opt);
        pv(at(-1, -1, 0), c, at(1, 1, 0), at(2, 2, 0), p0, v0);
        pv(at(2, -1, 0), at(1, 0, 0), at(0, 1, 0), at(-1, 2, 0), p1, v1);
        p[5] = blend2(p0, v0, p1, v1, opt);
    }
    if (opt & kPredDiag16) {
        p[6] = diag16<T>(A);
        return;
    }
    pv(at(-1, -1, -1), c, at(1, 1, 1), at(2, 2, 2), p0, v0);
    pv(at(-1, -1, 2), at(0, 0, 1), at(1, 1, 0), at(2, 2, -1), p1, v1);
    pv(at(2, -1, -1), at(1, 0, 0), at(0, 1, 1), at(-1, 2, 2), p2, v2);
    pv(at(-1, 2, -1), at(0, 1, 0), at(1, 0, 1), at(2, -1, 2), p3, v3);
    if (opt & kPredCubeAxis) p[6] = cube7<T>(A, p0, v0, p1, v1, p2, v2, p3, v3);
    else p[6] = blend4(p0, v0, p1, v1, p2, v2, p3, v3, opt);
}

// global-memory form: `row` = block grid row of the point (wave-uniform in
// the tile kernels), xo = the point's byte offset in it
template <typename T>
__device__ __forceinline__ void predict7_interior(const T* __restrict__ row, uint32_t xo, int sx, int sxy, T p[7],
                                                  int opt)
{
    auto at = [&](int dy, int dz, int k) -> T { return ldx(row + (ptrdiff_t)dz * sxy + (ptrdiff_t)dy * sx, xo, k); };
    predict7_at<T>(at, p, opt);
}

// all 7 predictions at (x,y,z): the interior form when `inner` (wave-uniform), else checked
template <typename T>
__device__ __forceinline__ void predict7(bool inner, const T* __restrict__ r, int x, int y, int z, int dx, int dy,
                                         int dz, T p[7], int opt)
{
    if (inner) {
        predict7_interior(r + ((size_t)z * dy + y) * dx, (uint32_t)x * (uint32_t)sizeof(T), dx, dx * dy, p, opt);
    } else {
        const int i = (z * dy + y) * dx + x;
        p[0] = predict<0>(r, i, x, y, z, dx, dy, dz, opt);
        p[1] = predict<1>(r, i, x, y, z, dx, dy, dz, opt);
        p[2] = predict<2>(r, i, x, y, z, dx, dy, dz, opt);
        p[3] = predict<3>(r, i, x, y, z, dx, dy, dz, opt);
        p[4] = predict<4>(r, i, x, y, z, dx, dy, dz, opt);
        p[5] = predict<5>(r, i, x, y, z, dx, dy, dz, opt);
        p[6] = predict<6>(r, i, x, y, z, dx, dy, dz, opt);
    }
}

// Level kernels are instantiated per mode; a mode with kVarBit set is the
// instantiation for payloads with predictor options, and reads them from
// the level arguments (the default instantiations see a constant 0).
constexpr int kVarBit = 8;
template <int kMode>
__device__ __forceinline__ int mode_opt(const LevelArgs& a)
{
    return (kMode & kVarBit) ? a.opt : 0;
}

// Work split of a block grid: tiles cover x in [1, dx-3]; the edge columns
// x in {0, dx-2, dx-1} (every x when dx < 4) take the checked predictors.
struct EdgeMap {
    int ncol;  // edge columns
    int dx;
    // column c of the edge set (arithmetic, so no local array is indexed at run time)
    __host__ __device__ int col(int c) const { return dx < 4 ? c : (c == 0 ? 0 : dx - 3 + c); }
};

__host__ __device__ inline EdgeMap edge_map(int dx)
{
    EdgeMap e;
    e.ncol = dx < 4 ? dx : 3;
    e.dx = dx;
    return e;
}

__device__ __forceinline__ bool row_inner(int y, int z, int dy, int dz)
{
    return y >= 1 && y <= dy - 3 && z >= 1 && z <= dz - 3;
}

// the 2x2x2 cell r[k] (k = 4 kz + 2 ky + kx) at (X,Y,Z) of a grid with
// strides (1, sy, sz): row base + 32-bit byte offset; vec = 2-wide stores
template <typename T>
__device__ __forceinline__ void store_cell(T* __restrict__ dst, size_t sy, size_t sz, int X, int Y, int Z,
                                           const T r[8], bool vec)
{
    typedef typename V2<T>::type v2;
    T* row = dst + (size_t)Z * sz + (size_t)Y * sy;
    const uint32_t xo = (uint32_t)X * (uint32_t)sizeof(T);
    if (vec) {
        v2 w;
        w.x = r[0]; w.y = r[1]; *(v2*)((char*)row + xo) = w;
        w.x = r[2]; w.y = r[3]; *(v2*)((char*)(row + sy) + xo) = w;
        w.x = r[4]; w.y = r[5]; *(v2*)((char*)(row + sz) + xo) = w;
        w.x = r[6]; w.y = r[7]; *(v2*)((char*)(row + sz + sy) + xo) = w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            *(T*)((char*)(row + (k >> 2) * sz + ((k >> 1) & 1) * sy) + xo + (k & 1) * sizeof(T)) = r[k];
    }
}

// ---------------------------------------------------------------- encode

template <typename T>
__device__ __forceinline__ uint32_t enc1(T tgt, T p, const LevelArgs& a, int b, uint32_t pos,
                                         uint16_t* __restrict__ codes, T* __restrict__ side, T& rec)
{
    const T d = tgt - p;
    T dec;
    const uint32_t q = quantize(d, a.eb, a.recip, dec);
    // streams hold < 2^31 symbols, so the byte offset 2*pos fits 32 bits
    *(uint16_t*)((char*)(codes + a.code_off[b]) + 2u * pos) = (uint16_t)q;
    if (!q) side[a.code_off[b] + pos] = d;
    rec = dec + p;  // decoded diff + prediction (F5:1381-1385 then F5:1077-1083)
    return q;
}

// high level: targets of the 2x2x2 cell of mid point (x,y,z) straight from
// the field, predictions p
template <typename T>
__device__ __forceinline__ void high_enc_fin(const T p[7], const T* __restrict__ full, int px, int py, const Geom& g,
                                             const LevelArgs& a, uint16_t* __restrict__ codes, T* __restrict__ side,
                                             int x, int y, int z, uint32_t q[7])
{
    typedef typename V2<T>::type v2;
    const uint32_t pos = stream_pos(g, x, y, z);
    const size_t fy = (size_t)px, fz = (size_t)px * py;
    const T* f = full + ((size_t)(2 * z) * py + 2 * y) * px;  // row base (wave-uniform in tiles)
    const uint32_t fo = (uint32_t)(2 * x) * (uint32_t)sizeof(T);
    const v2 r00 = *(const v2*)((const char*)f + fo);
    const v2 r01 = *(const v2*)((const char*)(f + fy) + fo);
    const v2 r10 = *(const v2*)((const char*)(f + fz) + fo);
    const v2 r11 = *(const v2*)((const char*)(f + fz + fy) + fo);
    const T tg[7] = {r00.y, r01.x, r01.y, r10.x, r10.y, r11.x, r11.y};
    T rec;
#pragma unroll
    for (int b = 0; b < 7; ++b) q[b] = enc1(tg[b], p[b], a, b, pos, codes, side, rec);
}

template <typename T>
__device__ __forceinline__ void high_enc_point(const T* __restrict__ full, int px, int py, const T* __restrict__ mid,
                                               const Geom& g, const LevelArgs& a, uint16_t* __restrict__ codes,
                                               T* __restrict__ side, int x, int y, int z, bool inner, uint32_t q[7],
                                               int opt)
{
    T p[7];
    predict7(inner, mid, x, y, z, g.dx, g.dy, g.dz, p, opt);
    high_enc_fin(p, full, px, py, g, a, codes, side, x, y, z, q);
}

// low level: targets from the field at stride 4, predictions p, base value
// `centre`; the reconstruction goes into the mid grid
template <typename T>
__device__ __forceinline__ void low_enc_fin(const T p[7], T centre, const T* __restrict__ full, int px, int py,
                                            const Geom& g, const LevelArgs& a, uint16_t* __restrict__ codes,
                                            T* __restrict__ side, T* __restrict__ mid, int x, int y, int z, bool vec,
                                            uint32_t q[7])
{
    const int mx = 2 * g.dx, my = 2 * g.dy;
    const uint32_t pos = stream_pos(g, x, y, z);
    // low block k at (x,y,z) = mid[2z+pz][2y+py][2x+px] = full[4z+2pz][4y+2py][4x+2px]
    const size_t fy = (size_t)2 * px, fz = (size_t)2 * px * py;
    const T* f = full + ((size_t)(4 * z) * py + 4 * y) * px;  // row base (wave-uniform in tiles)
    const uint32_t fo = (uint32_t)(4 * x) * (uint32_t)sizeof(T);
    const T tg[7] = {ldx(f, fo, 2), ldx(f + fy, fo, 0), ldx(f + fy, fo, 2), ldx(f + fz, fo, 0),
                     ldx(f + fz, fo, 2), ldx(f + fz + fy, fo, 0), ldx(f + fz + fy, fo, 2)};
    T r[8];
    r[0] = centre;
#pragma unroll
    for (int b = 0; b < 7; ++b) q[b] = enc1(tg[b], p[b], a, b, pos, codes, side, r[b + 1]);
    store_cell(mid, (size_t)mx, (size_t)mx * my, 2 * x, 2 * y, 2 * z, r, vec);
}

template <typename T>
__device__ __forceinline__ void low_enc_point(const T* __restrict__ full, int px, int py, const T* __restrict__ base,
                                              const Geom& g, const LevelArgs& a, uint16_t* __restrict__ codes,
                                              T* __restrict__ side, T* __restrict__ mid, int x, int y, int z,
                                              bool inner, bool vec, uint32_t q[7], int opt)
{
    T p[7];
    predict7(inner, base, x, y, z, g.dx, g.dy, g.dz, p, opt);
    const T c = ldx(base + ((size_t)z * g.dy + y) * g.dx, (uint32_t)x * (uint32_t)sizeof(T), 0);
    low_enc_fin(p, c, full, px, py, g, a, codes, side, mid, x, y, z, vec, q);
}

#define STZ_HIST_ADD7(q, ok)                                                    \
    do {                                                                        \
        _Pragma("unroll") for (int b_ = 0; b_ < 7; ++b_) hist_add(sh[b_], hist + b_ * 65536, q[b_], ok); \
    } while (0)

// ---------------------------------------------------------------- decode

// the 7 block codes of one point
struct Codes7 {
    uint32_t q[7];
};

__device__ __forceinline__ Codes7 load_codes7(const LevelArgs& a, const uint16_t* __restrict__ codes, uint32_t pos)
{
    Codes7 c;
#pragma unroll
    for (int b = 0; b < 7; ++b) c.q[b] = *(const uint16_t*)((const char*)(codes + a.code_off[b]) + 2u * pos);
    return c;
}

// dequantized values of the kDqW codes around the centre, staged in LDS by
// the z-march decode (the same dequantize() per code, so the same values;
// F5:1531-1535 keeps such a table too)
constexpr int kDqW = 2048;
constexpr uint32_t kDqLo = (uint32_t)kRadius - kDqW / 2;

template <typename T>
__device__ __forceinline__ T dec1(T p, const LevelArgs& a, int b, uint32_t pos, uint32_t q, const T* __restrict__ side,
                                  const T* dq)
{
    if (dq) {
        const uint32_t r = q - kDqLo;  // symbol 0 (unpredictable) falls outside
        if (r < (uint32_t)kDqW) return dq[r] + p;
    }
    const T v = q ? dequantize<T>(q, a.eb) : side[a.code_off[b] + pos];
    return v + p;  // F5:1080 de_sub = deData + pred
}

template <typename T>
__device__ __forceinline__ void low_dec_fin(const T p[7], T centre, const Geom& g, const LevelArgs& a, Codes7 cq,
                                            const T* __restrict__ side, T* __restrict__ mid, int x, int y, int z,
                                            bool vec, const T* dq = nullptr)
{
    const int mx = 2 * g.dx, my = 2 * g.dy;
    const uint32_t pos = stream_pos(g, x, y, z);
    T r[8];
    r[0] = centre;
#pragma unroll
    for (int b = 0; b < 7; ++b) r[b + 1] = dec1(p[b], a, b, pos, cq.q[b], side, dq);
    store_cell(mid, (size_t)mx, (size_t)mx * my, 2 * x, 2 * y, 2 * z, r, vec);
}

template <typename T>
__device__ __forceinline__ void low_dec_point(const T* __restrict__ base, const Geom& g, const LevelArgs& a,
                                              const uint16_t* __restrict__ codes, const T* __restrict__ side,
                                              T* __restrict__ mid, int x, int y, int z, bool inner, bool vec,
                                              int opt)
{
    T p[7];
    predict7(inner, base, x, y, z, g.dx, g.dy, g.dz, p, opt);
    const T c = ldx(base + ((size_t)z * g.dy + y) * g.dx, (uint32_t)x * (uint32_t)sizeof(T), 0);
    low_dec_fin(p, c, g, a, load_codes7(a, codes, stream_pos(g, x, y, z)), side, mid, x, y, z, vec);
}

// the 2x2x2 cell of mid point (x,y,z) into the caller's (nx,ny,nz) field
// from predictions p and the mid value `centre`; vec = 2-wide stores (nx
// even and the output 2-element aligned)
template <typename T>
__device__ __forceinline__ void high_dec_fin(const T p[7], T centre, const Geom& g, const LevelArgs& a, Codes7 cq,
                                             const T* __restrict__ side, T* __restrict__ out, int nx, int ny, int nz,
                                             int x, int y, int z, bool vec, const T* dq = nullptr)
{
    typedef typename V2<T>::type v2;
    const int X = 2 * x, Y = 2 * y, Z = 2 * z;
    if (X >= nx || Y >= ny || Z >= nz) return;  // cell entirely in the padding
    const uint32_t pos = stream_pos(g, x, y, z);
    T v[8];
    v[0] = centre;
#pragma unroll
    for (int b = 0; b < 7; ++b) v[b + 1] = dec1(p[b], a, b, pos, cq.q[b], side, dq);
    const size_t oy = (size_t)nx, oz = (size_t)nx * ny;
    T* o = out + (size_t)Z * oz + (size_t)Y * oy;  // row base (wave-uniform in tiles)
    const uint32_t xo = (uint32_t)X * (uint32_t)sizeof(T);
    if (vec) {
        v2 w;
        w.x = v[0]; w.y = v[1]; *(v2*)((char*)o + xo) = w;
        if (Y + 1 < ny) { w.x = v[2]; w.y = v[3]; *(v2*)((char*)(o + oy) + xo) = w; }
        if (Z + 1 < nz) {
            w.x = v[4]; w.y = v[5]; *(v2*)((char*)(o + oz) + xo) = w;
            if (Y + 1 < ny) { w.x = v[6]; w.y = v[7]; *(v2*)((char*)(o + oz + oy) + xo) = w; }
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int kx = k & 1, ky = (k >> 1) & 1, kz = k >> 2;
            if (X + kx < nx && Y + ky < ny && Z + kz < nz)
                *(T*)((char*)(o + (size_t)kz * oz + (size_t)ky * oy) + xo + kx * sizeof(T)) = v[k];
        }
    }
}

template <typename T>
__device__ __forceinline__ void high_dec_point(const T* __restrict__ mid, const Geom& g, const LevelArgs& a,
                                               const uint16_t* __restrict__ codes, const T* __restrict__ side,
                                               T* __restrict__ out, int nx, int ny, int nz, int x, int y, int z,
                                               bool inner, bool vec, int opt)
{
    if (2 * x >= nx || 2 * y >= ny || 2 * z >= nz) return;
    T p[7];
    predict7(inner, mid, x, y, z, g.dx, g.dy, g.dz, p, opt);
    const T c = ldx(mid + ((size_t)z * g.dy + y) * g.dx, (uint32_t)x * (uint32_t)sizeof(T), 0);
    high_dec_fin(p, c, g, a, load_codes7(a, codes, stream_pos(g, x, y, z)), side, out, nx, ny, nz, x, y, z, vec);
}

// ---------------------------------------------------------------- z-march
//
// One launch per level and block grid.  Work items 0 .. nmarch-1 are column
// items: 64 interior columns (x0 = 1 + 64 tx ..) x kZR rows (y0 = kZR ty ..)
// over a segment of `seg` planes.  The workgroup walks the segment in z and
// keeps the predictor grid's planes z-1 .. z+2 of its column (rows y0-1 ..
// y0+kZR+1, columns x0-1 .. x0+65) in a 5-slot LDS ring: each step stages one
// new plane (coalesced, fetched into registers a step ahead) behind a single
// barrier, and every interior point reads its 34 neighbours from LDS.  Wave
// w owns rows y0 + kZRW w .. y0 + kZRW w + kZRW-1, so y and z are
// wave-uniform; rows and planes on the grid faces take the checked
// predictors from global memory.  The items after nmarch hold 256 points
// each of the edge columns (x = 0, dx-2, dx-1, or every x when dx < 4).  The
// grid is persistent (one resident wave of workgroups), walking the items in
// the XCD-aware order.
constexpr int kZR = 8;
constexpr int kZRW = kZR / 4;  // rows per wave
static_assert(kZR % 4 == 0, "four waves share the rows of a column item");
constexpr int kZRows = kZR + 3;
constexpr int kZCols = 67;
constexpr int kZPitch = 68;
constexpr int kZSlots = 5;
constexpr int kZPlane = kZRows * kZPitch;
constexpr int kZLoads = (kZRows * kZCols + 255) / 256;

template <typename T>
struct LevelK {
    const T* full;        // encode: the padded field (targets)
    int px, py;
    const T* ref;         // predictor grid: mid (high level) or base (low level)
    Geom g;               // block grid = predictor grid dims, stream chunking
    LevelArgs a;
    uint16_t* codes;      // encode: written; decode: read
    T* side;
    uint32_t* hist;       // encode: 7 x 65536 histogram bins
    T* dst;               // low level: the mid grid; high decode: the field
    int nx, ny, nz, vec;  // high decode output dims; 2-wide stores allowed
    int seg;              // planes per column item
};

struct ZItems {
    int ntx, nty;
    uint64_t nmarch, total;
};

__host__ __device__ inline ZItems z_items(const Geom& g, int seg)
{
    ZItems Z;
    const int nxi = g.dx - 3;
    Z.ntx = nxi > 0 ? (nxi + 63) >> 6 : 0;
    Z.nty = (g.dy + kZR - 1) / kZR;
    const int nseg = (g.dz + seg - 1) / seg;
    Z.nmarch = (uint64_t)Z.ntx * Z.nty * nseg;
    const uint64_t ne = (uint64_t)edge_map(g.dx).ncol * g.dy * g.dz;
    Z.total = Z.nmarch + (ne + 255) / 256;
    return Z;
}

// edge item t (after nmarch): the lane's point; ok = it exists
__device__ __forceinline__ void edge_point(const Geom& g, uint64_t first, uint64_t t, int& x, int& y, int& z,
                                           bool& ok)
{
    const EdgeMap e = edge_map(g.dx);
    const uint64_t k = (t - first) * 256u + threadIdx.x;
    ok = k < (uint64_t)e.ncol * g.dy * g.dz;
    const int c = (int)(k % e.ncol);
    y = (int)((k / e.ncol) % g.dy);
    z = (int)(k / ((uint64_t)e.ncol * g.dy));
    x = e.col(c);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : v > hi ? hi : v; }

// one point through the level's operation: predictions p and centre value
// given, or (checked form) evaluated from global memory
// decode modes take the point's codes in cq (the caller loads or prefetched
// them); encode modes ignore it
template <typename T, int kMode>
__device__ __forceinline__ void z_point(const LevelK<T>& K, const T p[7], T c, int x, int y, int z, bool vec,
                                        uint32_t q[7], Codes7 cq, const T* dq = nullptr)
{
    if constexpr ((kMode & 7) == kHighEnc) high_enc_fin(p, K.full, K.px, K.py, K.g, K.a, K.codes, K.side, x, y, z, q);
    else if constexpr ((kMode & 7) == kLowEnc)
        low_enc_fin(p, c, K.full, K.px, K.py, K.g, K.a, K.codes, K.side, K.dst, x, y, z, vec, q);
    else if constexpr ((kMode & 7) == kLowDec) low_dec_fin(p, c, K.g, K.a, cq, K.side, K.dst, x, y, z, vec, dq);
    else high_dec_fin(p, c, K.g, K.a, cq, K.side, K.dst, K.nx, K.ny, K.nz, x, y, z, vec && K.vec != 0, dq);
}

// the codes a decode mode needs at block point (x,y,z) (nothing for encode)
template <typename T, int kMode>
__device__ __forceinline__ Codes7 point_codes(const LevelK<T>& K, int x, int y, int z)
{
    if constexpr ((kMode & 7) == kLowDec || (kMode & 7) == kHighDec) return load_codes7(K.a, K.codes, stream_pos(K.g, x, y, z));
    else return Codes7{};
}

template <typename T, int kMode>
__device__ __forceinline__ void z_point_checked(const LevelK<T>& K, int x, int y, int z, uint32_t q[7])
{
    const Geom& g = K.g;
    if constexpr ((kMode & 7) == kHighDec) {
        if (2 * x >= K.nx || 2 * y >= K.ny || 2 * z >= K.nz) return;
    }
    T p[7];
    predict7(false, K.ref, x, y, z, g.dx, g.dy, g.dz, p, mode_opt<kMode>(K.a));
    const T c = K.ref[((size_t)z * g.dy + y) * g.dx + x];
    z_point<T, kMode>(K, p, c, x, y, z, false, q, point_codes<T, kMode>(K, x, y, z));
}

// Tile form of a level launch: work items 0 .. ntiles-1 are 64 x (4 kTW)
// tiles of one plane (wave w owns rows y0 + kTW w .., so y and z are
// wave-uniform; interior rows read their 34 neighbours straight from global
// memory, the kTW rows of a wave in one straight-line block so their loads
// overlap), the items after them the edge columns as above.  No LDS besides
// the encode histograms, so more workgroups share a CU than in the z-march
// form.  Rows per wave by level (measured on the 512^3 field): the high
// level's larger grid gains from three, the low level's from one.
template <int kMode>
constexpr int tile_rows()
{
    return ((kMode & 7) == kHighEnc || (kMode & 7) == kHighDec) ? 3 : 1;
}

template <int kTW>
__host__ __device__ inline void tile_items(const Geom& g, int& ntx, int& nty, uint64_t& ntiles, uint64_t& total)
{
    const int nxi = g.dx - 3;
    ntx = nxi > 0 ? (nxi + 63) >> 6 : 0;
    nty = (g.dy + 4 * kTW - 1) / (4 * kTW);
    ntiles = (uint64_t)ntx * nty * g.dz;
    const uint64_t ne = (uint64_t)edge_map(g.dx).ncol * g.dy * g.dz;
    total = ntiles + (ne + 255) / 256;
}

// interior point (predictions from global memory)
template <typename T, int kMode>
__device__ __forceinline__ void tile_point(const LevelK<T>& K, int x, int y, int z, uint32_t q[7])
{
    const Geom& g = K.g;
    if constexpr ((kMode & 7) == kHighDec) {
        if (2 * x >= K.nx || 2 * y >= K.ny || 2 * z >= K.nz) return;
    }
    T p[7];
    predict7(true, K.ref, x, y, z, g.dx, g.dy, g.dz, p, mode_opt<kMode>(K.a));
    const T c = ldx(K.ref + ((size_t)z * g.dy + y) * g.dx, (uint32_t)x * (uint32_t)sizeof(T), 0);
    z_point<T, kMode>(K, p, c, x, y, z, true, q, point_codes<T, kMode>(K, x, y, z));
}

template <typename T, int kMode>
__global__ __launch_bounds__(256) void k_level_tile(const LevelK<T> K)
{
    constexpr bool kEnc = (kMode & 7) == kLowEnc || (kMode & 7) == kHighEnc;
    __shared__ uint32_t sh[kEnc ? 7 : 1][kEnc ? kHW : 1];
    uint32_t* hist = K.hist;
    if constexpr (kEnc) {
        for (int k = threadIdx.x; k < 7 * kHW; k += 256) sh[k / kHW][k % kHW] = 0u;
        __syncthreads();
    }
    constexpr int kTW = tile_rows<kMode>();
    const Geom& g = K.g;
    int ntx, nty;
    uint64_t ntiles, total;
    tile_items<kTW>(g, ntx, nty, ntiles, total);
    const int lane = threadIdx.x & 63, row = wave_row();
    // item walk: the high level in the interleaved XCD order; the low level
    // with each XCD sweeping one contiguous range of items (z-planes), so
    // the base-grid planes a tile's stencil shares with the next ones stay
    // in that XCD's L2 (low encode 104 -> 94 us at 512^3; the high encode
    // measured slower that way, 416 -> 447 us)
    constexpr bool kSweep = (kMode & 7) == kLowEnc;
    const uint32_t G = gridDim.x;
    const bool sweep = kSweep && (G % 8u) == 0;
    const uint64_t per = sweep ? (total + 7) / 8 : total;
    const uint64_t base = sweep ? (uint64_t)(blockIdx.x % 8u) * per : 0;
    const uint64_t first = sweep ? blockIdx.x / 8u : xcd_block(), step = sweep ? G / 8u : G;
    Walk3 wk;
    wk.init(base + first, step, (uint32_t)(ntx > 0 ? ntx : 1), (uint32_t)nty);
    for (uint64_t j = first; j < per; j += step, wk.next()) {
        const uint64_t t = base + j;
        if (t >= total) break;
        if (t < ntiles) {
            const int tx = (int)wk.a, ty = (int)wk.b;
            const int z = (int)wk.c;
            const int x = 1 + tx * 64 + lane;
            const int y0 = ty * 4 * kTW + kTW * row;
            const bool okx = x <= g.dx - 3;
            uint32_t q[kTW][7] = {};
            if (y0 + kTW - 1 < g.dy && row_inner(y0, z, g.dy, g.dz) && row_inner(y0 + kTW - 1, z, g.dy, g.dz)) {
                if (okx) {
#pragma unroll
                    for (int r = 0; r < kTW; ++r) tile_point<T, kMode>(K, x, y0 + r, z, q[r]);
                }
#pragma unroll
                for (int r = 0; r < kTW; ++r)
                    if constexpr (kEnc) STZ_HIST_ADD7(q[r], okx);
            } else {
                for (int r = 0; r < kTW; ++r) {
                    const int y = y0 + r;
                    const bool ok = okx && y < g.dy;
                    if (ok) {
                        if (row_inner(y, z, g.dy, g.dz)) tile_point<T, kMode>(K, x, y, z, q[r]);
                        else z_point_checked<T, kMode>(K, x, y, z, q[r]);
                    }
                    if constexpr (kEnc) STZ_HIST_ADD7(q[r], ok);
                }
            }
        } else {
            int x, y, z;
            bool ok;
            uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
            edge_point(g, ntiles, t, x, y, z, ok);
            if (ok) z_point_checked<T, kMode>(K, x, y, z, q);
            if constexpr (kEnc) STZ_HIST_ADD7(q, ok);
        }
    }
    if constexpr (kEnc) hist_flush(sh, 7, hist);
}

// at most 96 VGPRs: five waves per SIMD (measured faster than four or six)
template <typename T, int kMode>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_level_z(const LevelK<T> K)
{
    constexpr bool kEnc = (kMode & 7) == kLowEnc || (kMode & 7) == kHighEnc;
    __shared__ uint32_t sh[kEnc ? 7 : 1][kEnc ? kHW : 1];
    __shared__ T ring[kZSlots * kZPlane];
    __shared__ T dq[kEnc ? 1 : kDqW];
    uint32_t* hist = K.hist;
    if constexpr (kEnc) {
        for (int k = threadIdx.x; k < 7 * kHW; k += 256) sh[k / kHW][k % kHW] = 0u;
        __syncthreads();
    } else {
        // read after the first ring barrier
        for (int k = threadIdx.x; k < kDqW; k += 256) dq[k] = dequantize<T>(kDqLo + (uint32_t)k, K.a.eb);
    }
    const Geom& g = K.g;
    const ZItems Z = z_items(g, K.seg);
    const int lane = threadIdx.x & 63, w = wave_row();
    const size_t sxy = (size_t)g.dx * g.dy;
    for (uint64_t t = xcd_block(); t < Z.total; t += gridDim.x) {
        if (t >= Z.nmarch) {
            int x, y, z;
            bool ok;
            edge_point(g, Z.nmarch, t, x, y, z, ok);
            uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
            if (ok) z_point_checked<T, kMode>(K, x, y, z, q);
            if constexpr (kEnc) STZ_HIST_ADD7(q, ok);
            continue;
        }
        const int tx = (int)(t % Z.ntx), ty = (int)((t / Z.ntx) % Z.nty);
        const int zb = (int)(t / ((uint64_t)Z.ntx * Z.nty)) * K.seg;
        const int ze = zb + K.seg < g.dz ? zb + K.seg : g.dz;
        const int x0 = 1 + 64 * tx, y0 = kZR * ty;
        // this thread's share of a staged plane: element e = threadIdx + 256 k
        // -> (row, column) of the ring, grid offset inside a plane (clamped;
        // clamped entries are never read by an interior point)
        int roff[kZLoads], loff[kZLoads];
#pragma unroll
        for (int k = 0; k < kZLoads; ++k) {
            const int e = threadIdx.x + 256 * k;
            const int r = e / kZCols, c = e - r * kZCols;
            const int gy = clampi(y0 - 1 + r, 0, g.dy - 1), gx = clampi(x0 - 1 + c, 0, g.dx - 1);
            roff[k] = gy * g.dx + gx;
            loff[k] = e < kZRows * kZCols ? r * kZPitch + c : -1;
        }
        auto fetch = [&](int pz, T v[kZLoads]) {
            const T* P = K.ref + (size_t)clampi(pz, 0, g.dz - 1) * sxy;
#pragma unroll
            for (int k = 0; k < kZLoads; ++k) v[k] = loff[k] >= 0 ? P[roff[k]] : T(0);
        };
        auto put = [&](int pz, const T v[kZLoads]) {
            T* S = ring + ((pz + kZSlots) % kZSlots) * kZPlane;
#pragma unroll
            for (int k = 0; k < kZLoads; ++k)
                if (loff[k] >= 0) S[loff[k]] = v[k];
        };
        __syncthreads();  // the previous item's readers are done with the ring
        // decode: the codes of a step's points are fetched a step ahead too
        Codes7 cq[kZRW] = {}, cn[kZRW] = {};
        auto fetch_codes = [&](int zz) {
#pragma unroll
            for (int r = 0; r < kZRW; ++r) {
                const int y = y0 + kZRW * w + r, x = x0 + lane;
                if (zz < ze && y < g.dy && x <= g.dx - 3) cn[r] = point_codes<T, kMode>(K, x, y, zz);
            }
        };
        T v[kZLoads];
        for (int pz = zb - 1; pz <= zb + 1; ++pz) {
            fetch(pz, v);
            put(pz, v);
        }
        fetch(zb + 2, v);
        if constexpr (!kEnc) fetch_codes(zb);
        for (int z = zb; z < ze; ++z) {
            put(z + 2, v);
            if (z + 1 < ze) fetch(z + 3, v);
            if constexpr (!kEnc) {
#pragma unroll
                for (int r = 0; r < kZRW; ++r) cq[r] = cn[r];
                fetch_codes(z + 1);
            }
            __syncthreads();
            const T* S[4];
#pragma unroll
            for (int d = 0; d < 4; ++d) S[d] = ring + ((z - 1 + d + kZSlots) % kZSlots) * kZPlane;
#pragma unroll
            for (int r = 0; r < kZRW; ++r) {
                const int y = y0 + kZRW * w + r;
                if (y >= g.dy) continue;  // wave-uniform
                const int x = x0 + lane;
                const bool ok = x <= g.dx - 3;
                uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
                if (row_inner(y, z, g.dy, g.dz)) {
                    if (ok) {
                        const int o = (y - y0 + 1) * kZPitch + lane + 1;
                        auto at = [&](int dy, int dz, int k) -> T { return S[dz + 1][o + dy * kZPitch + k]; };
                        T p[7];
                        predict7_at<T>(at, p, mode_opt<kMode>(K.a));
                        z_point<T, kMode>(K, p, S[1][o], x, y, z, true, q, cq[r], kEnc ? nullptr : dq);
                    }
                } else if (ok) {
                    z_point_checked<T, kMode>(K, x, y, z, q);
                }
                if constexpr (kEnc) STZ_HIST_ADD7(q, ok);
            }
        }
    }
    if constexpr (kEnc) hist_flush(sh, 7, hist);
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
template <typename T, bool kVar>
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
            v = r + predict_rt(b, base, i, x, y, z, g.dx, g.dy, g.dz, kVar ? a.opt : 0);
        }
        mid[((size_t)Z * my + Y) * mx + X] = v;
    }
}

// Random access: full-resolution values of query boxes (F5:2574-2605)
template <typename T, bool kVar>
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
            v = r + predict_rt(b, mid, i, x, y, z, g.dx, g.dy, g.dz, kVar ? a.opt : 0);
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

// one point of a level job: encode (codes, side values, histogram counts of
// every active lane) or decode; `ok` = the lane owns a point
template <typename T, int kMode>
__device__ __forceinline__ void level_point(const LevelJob& J, uint32_t (*sh)[kHW], int x, int y, int z, bool ok,
                                            bool inner, bool vec)
{
    constexpr bool kEnc = (kMode & 7) == kLowEnc || (kMode & 7) == kHighEnc;
    const Geom& g = J.g;
    if constexpr (kEnc) {
        uint32_t q[7] = {0, 0, 0, 0, 0, 0, 0};
        if (ok) {
            if ((kMode & 7) == kHighEnc)
                high_enc_point((const T*)J.full, J.px, J.py, (const T*)J.ref, g, J.a, J.codes, (T*)J.side, x, y, z,
                               inner, q, mode_opt<kMode>(J.a));
            else
                low_enc_point((const T*)J.full, J.px, J.py, (const T*)J.ref, g, J.a, J.codes, (T*)J.side,
                              (T*)J.mid, x, y, z, inner, vec, q, mode_opt<kMode>(J.a));
        }
        uint32_t* hist = J.hist;
        STZ_HIST_ADD7(q, ok);
    } else if (ok) {
        if ((kMode & 7) == kHighDec)
            high_dec_point((const T*)J.ref, g, J.a, J.codes, (const T*)J.side, (T*)J.out, J.nx, J.ny, J.nz, x, y, z,
                           inner, vec && J.vec != 0, mode_opt<kMode>(J.a));
        else
            low_dec_point((const T*)J.ref, g, J.a, J.codes, (const T*)J.side, (T*)J.mid, x, y, z, inner, vec,
                          mode_opt<kMode>(J.a));
    }
}

template <typename T, int kMode, int kWaves>
__global__ __launch_bounds__(64 * kWaves) void k_level_multi(const LevelJob* __restrict__ jobs, uint32_t njobs,
                                                             uint32_t total, uint32_t* __restrict__ counter)
{
    constexpr bool kEnc = (kMode & 7) == kLowEnc || (kMode & 7) == kHighEnc;
    constexpr int kThreads = 64 * kWaves;
    __shared__ uint32_t sh[kEnc ? 7 : 1][kHW];
    __shared__ uint32_t item_s;
    if constexpr (kEnc) {
        for (int k = threadIdx.x; k < 7 * kHW; k += kThreads) sh[k / kHW][k % kHW] = 0u;
    }
    int cur = -1;
    const int lane = threadIdx.x & 63, row = wave_row();
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = __builtin_amdgcn_readfirstlane(item_s);
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
        if (local < J.ntiles) {
            // tile: y and z are wave-uniform
            const int nxi = g.dx - 3;
            const int ntx = (nxi + 63) >> 6, nty = (g.dy + kWaves - 1) / kWaves;
            const int tx = (int)(local % ntx), ty = (int)((local / ntx) % nty), tz = (int)(local / ((uint32_t)ntx * nty));
            const int x = 1 + tx * 64 + lane, y = ty * kWaves + row;
            level_point<T, kMode>(J, sh, x, y, tz, x <= g.dx - 3 && y < g.dy, row_inner(y, tz, g.dy, g.dz), true);
        } else {
            const EdgeMap e = edge_map(g.dx);
            const uint64_t k = (uint64_t)(local - J.ntiles) * kThreads + threadIdx.x;
            const int c = (int)(k % e.ncol);
            const int y = (int)((k / e.ncol) % g.dy), z = (int)(k / ((uint64_t)e.ncol * g.dy));
            level_point<T, kMode>(J, sh, e.col(c), y, z, k < (uint64_t)e.ncol * g.dy * g.dz, false, false);
        }
    }
    if constexpr (kEnc) {
        if (cur >= 0) hist_flush(sh, 7, jobs[cur].hist);
    }
}

// Persistent grid of a level kernel: the workgroups that are resident at
// once on the device (CUs x occupancy, asked of the runtime once per
// kernel).
inline unsigned resident_slots(const void* kernel)
{
    static std::mutex mu;
    static std::map<const void*, unsigned> slots;
    std::lock_guard<std::mutex> lk(mu);
    auto it = slots.find(kernel);
    if (it == slots.end()) {
        int dev = 0, ncu = 0, occ = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) ncu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, 256, 0) != hipSuccess || occ < 1) occ = 1;
        it = slots.emplace(kernel, (unsigned)(ncu * occ)).first;
    }
    return it->second;
}

// Tile form: a persistent grid of resident workgroups, at most one per item,
// a multiple of 8 for the XCD order.
// at most `share` of the slots, at least one XCD-ordered round of 8
inline unsigned shared_slots(unsigned slots, float share)
{
    if (share >= 1.0f) return slots;
    const unsigned s = (unsigned)((float)slots * share);
    return s < 8u ? 8u : s;
}

template <typename T, int kMode>
void launch_level_tile(hipStream_t s, LevelK<T> K, float share)
{
    const unsigned slots = shared_slots(resident_slots((const void*)k_level_tile<T, kMode>), share);
    int ntx, nty;
    uint64_t ntiles, total;
    tile_items<tile_rows<kMode>()>(K.g, ntx, nty, ntiles, total);
    uint64_t G = total < slots ? total : slots;
    if (G >= 64) G &= ~7ull;
    if (!G) return;
    k_level_tile<T, kMode><<<dim3((unsigned)G), dim3(256), 0, s>>>(K);
}

template <typename T, int kMode>
void launch_level_z(hipStream_t s, LevelK<T> K, float share);

// level launch form per mode: encode is VALU-bound and keeps the higher
// occupancy of the tile form; decode is latency-bound and takes the z-march
constexpr bool kZMarchEnc = false;
constexpr bool kZMarchDec = true;

template <typename T, int kMode>
void launch_level(hipStream_t s, const LevelK<T>& K, float share = 1.0f)
{
    constexpr bool kEnc = (kMode & 7) == kLowEnc || (kMode & 7) == kHighEnc;
    if ((kEnc && kZMarchEnc) || (!kEnc && kZMarchDec)) launch_level_z<T, kMode>(s, K, share);
    else launch_level_tile<T, kMode>(s, K, share);
}

// Z-march: plane segments so that the column items fill about one resident
// wave of workgroups (4 to 32 planes per segment: the 3-plane halo is
// amortised, and a large grid still splits into many items), then a
// persistent grid of resident workgroups, a multiple of 8 for the XCD order.
// (One workgroup per item measured 30% slower at 256^3 per sub-band.)
template <typename T, int kMode>
void launch_level_z(hipStream_t s, LevelK<T> K, float share)
{
    const void* fn = (const void*)k_level_z<T, kMode>;
    const unsigned slots = resident_slots(fn);
    const Geom& g = K.g;
    const int nxi = g.dx - 3;
    const uint64_t ncol = (uint64_t)(nxi > 0 ? (nxi + 63) >> 6 : 0) * ((g.dy + kZR - 1) / kZR);
    uint64_t nseg = ncol ? slots / ncol : 1;
    if (nseg < 1) nseg = 1;
    int seg = (int)((g.dz + nseg - 1) / nseg);
    if (seg < 4) seg = 4;
    if (seg > 32) seg = 32;
    if (seg > g.dz) seg = g.dz;
    K.seg = seg;
    const ZItems Z = z_items(g, seg);
    const unsigned gs = shared_slots(slots, share);
    uint64_t G = Z.total < gs ? Z.total : gs;
    if (G >= 64) G &= ~7ull;
    if (!G) return;
    k_level_z<T, kMode><<<dim3((unsigned)G), dim3(256), 0, s>>>(K);
}

}  // namespace

template <typename T>
void launch_low_encode(hipStream_t s, const T* full, int px, int py, int pz, const T* base, Geom g, LevelArgs a,
                       uint16_t* codes, T* side, uint32_t* hist, T* mid)
{
    (void)pz;
    LevelK<T> K{};
    K.full = full; K.px = px; K.py = py; K.ref = base; K.g = g; K.a = a;
    K.codes = codes; K.side = side; K.hist = hist; K.dst = mid;
    if (a.opt) launch_level<T, kLowEnc | kVarBit>(s, K);
    else launch_level<T, kLowEnc>(s, K);
}

template <typename T>
void launch_high_encode(hipStream_t s, const T* full, int px, int py, int pz, const T* mid, Geom g, LevelArgs a,
                        uint16_t* codes, T* side, uint32_t* hist)
{
    (void)pz;
    LevelK<T> K{};
    K.full = full; K.px = px; K.py = py; K.ref = mid; K.g = g; K.a = a;
    K.codes = codes; K.side = side; K.hist = hist;
    if (a.opt) launch_level<T, kHighEnc | kVarBit>(s, K);
    else launch_level<T, kHighEnc>(s, K);
}

template <typename T>
void launch_low_decode(hipStream_t s, const T* base, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                       T* mid, float share)
{
    LevelK<T> K{};
    K.ref = base; K.g = g; K.a = a; K.codes = (uint16_t*)codes; K.side = (T*)side; K.dst = mid;
    if (a.opt) launch_level<T, kLowDec | kVarBit>(s, K, share);
    else launch_level<T, kLowDec>(s, K, share);
}

template <typename T>
void launch_high_decode(hipStream_t s, const T* mid, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                        T* out, int nx, int ny, int nz, float share)
{
    LevelK<T> K{};
    K.ref = mid; K.g = g; K.a = a; K.codes = (uint16_t*)codes; K.side = (T*)side; K.dst = out;
    K.nx = nx; K.ny = ny; K.nz = nz;
    K.vec = ((nx & 1) == 0) && (((uintptr_t)out % (2 * sizeof(T))) == 0);
    if (a.opt) launch_level<T, kHighDec | kVarBit>(s, K, share);
    else launch_level<T, kHighDec>(s, K, share);
}

template <typename T>
void launch_low_decode_boxes(hipStream_t s, const T* base, Geom g, LevelArgs a, const uint16_t* codes,
                             const T* side, T* mid, int mx, int my, const BoxJob* jobs, uint32_t njobs,
                             uint32_t nblocks)
{
    if (!njobs || !nblocks) return;
    if (a.opt)
        k_low_decode_boxes<T, true><<<dim3(nblocks), dim3(256), 0, s>>>(base, g, a, codes, side, mid, mx, my, jobs,
                                                                        njobs, nblocks);
    else
        k_low_decode_boxes<T, false><<<dim3(nblocks), dim3(256), 0, s>>>(base, g, a, codes, side, mid, mx, my, jobs,
                                                                         njobs, nblocks);
}

template <typename T>
void launch_high_decode_boxes(hipStream_t s, const T* mid, Geom g, LevelArgs a, const uint16_t* codes,
                              const T* side, const BoxJob* jobs, uint32_t njobs, uint32_t nblocks)
{
    if (!njobs || !nblocks) return;
    if (a.opt)
        k_high_decode_boxes<T, true><<<dim3(nblocks), dim3(256), 0, s>>>(mid, g, a, codes, side, jobs, njobs, nblocks);
    else
        k_high_decode_boxes<T, false><<<dim3(nblocks), dim3(256), 0, s>>>(mid, g, a, codes, side, jobs, njobs,
                                                                          nblocks);
}

uint32_t level_tiles(const Geom& g, int waves)
{
    const int nxi = g.dx - 3;
    if (nxi <= 0) return 0;
    return (uint32_t)(((nxi + 63) / 64) * ((g.dy + waves - 1) / waves) * g.dz);
}

uint32_t level_edge_groups(const Geom& g, int waves)
{
    const uint64_t n = (uint64_t)edge_map(g.dx).ncol * g.dy * g.dz;
    return (uint32_t)((n + 64u * waves - 1) / (64u * waves));
}

template <typename T, int kW, int kV>
void level_multi_w(hipStream_t s, int mode, const LevelJob* jobs, uint32_t njobs, uint32_t total, uint32_t* counter,
                   uint32_t grid)
{
    const dim3 G(grid), B(64 * kW);
    switch (mode) {
    case kLowEnc: k_level_multi<T, kLowEnc | kV, kW><<<G, B, 0, s>>>(jobs, njobs, total, counter); break;
    case kHighEnc: k_level_multi<T, kHighEnc | kV, kW><<<G, B, 0, s>>>(jobs, njobs, total, counter); break;
    case kLowDec: k_level_multi<T, kLowDec | kV, kW><<<G, B, 0, s>>>(jobs, njobs, total, counter); break;
    default: k_level_multi<T, kHighDec | kV, kW><<<G, B, 0, s>>>(jobs, njobs, total, counter); break;
    }
}

// var: some job has predictor options (each job reads its own)
template <typename T>
void launch_level_multi(hipStream_t s, int mode, const LevelJob* jobs, uint32_t njobs, uint32_t total_items,
                        uint32_t* counter, int waves, uint32_t grid, bool var)
{
    if (!njobs || !total_items || !grid) return;
    if (var) {
        if (waves == 8) level_multi_w<T, 8, kVarBit>(s, mode, jobs, njobs, total_items, counter, grid);
        else level_multi_w<T, 4, kVarBit>(s, mode, jobs, njobs, total_items, counter, grid);
    } else {
        if (waves == 8) level_multi_w<T, 8, 0>(s, mode, jobs, njobs, total_items, counter, grid);
        else level_multi_w<T, 4, 0>(s, mode, jobs, njobs, total_items, counter, grid);
    }
}

#define STZ_INST(T)                                                                                              \
    template void launch_low_encode<T>(hipStream_t, const T*, int, int, int, const T*, Geom, LevelArgs,         \
                                       uint16_t*, T*, uint32_t*, T*);                                            \
    template void launch_high_encode<T>(hipStream_t, const T*, int, int, int, const T*, Geom, LevelArgs,        \
                                        uint16_t*, T*, uint32_t*);                                               \
    template void launch_low_decode<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, T*,  \
                                       float);                                                                   \
    template void launch_high_decode<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, T*,  \
                                        int, int, int, float);                                                   \
    template void launch_low_decode_boxes<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, \
                                             T*, int, int, const BoxJob*, uint32_t, uint32_t);                  \
    template void launch_high_decode_boxes<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*,         \
                                              const T*, const BoxJob*, uint32_t, uint32_t);                      \
    template void launch_level_multi<T>(hipStream_t, int, const LevelJob*, uint32_t, uint32_t, uint32_t*, int,  \
                                        uint32_t, bool);
STZ_INST(float)
STZ_INST(double)

}  // namespace stz

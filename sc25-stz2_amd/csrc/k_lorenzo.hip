// k_lorenzo.hip -- Lorenzo base codec (stz_params.interp == 2) on gfx950.
//
// The alternative to interpolation that SZ3's ALGO_INTERP_LORENZO trials for
// the reference's base block (F5:1439-1449 through SZ3-ncp; sz.cpp:63-76
// turns it on): the first-order 3D Lorenzo predictor.  SZ3 sweeps the whole
// block in raster order, one point after another.  Here the base grid is cut
// into 16^3 tiles coded on their own (a neighbour outside the tile reads as
// zero), one workgroup per tile, and inside a tile the points are visited on
// the diagonal planes x + y + z = s: the seven neighbours of a point lie on
// earlier planes, so a plane is one data-parallel step and the raster result
// is reproduced bit for bit.  Lane (x, y) of the 256 owns the tile column at
// (x, y) and handles its point of each plane; the tile stays in LDS.
//
// Prediction ((((((a + b) + c) - ab) - ac) - bc) + abc) in T from the
// reconstructed neighbours, the residual quantized with the base bound
// (device_math.h quantize); an unpredictable point keeps its value.  Stream
// order: tiles in raster order (x fastest), raster order inside a tile, so
// the tile at (tx, ty, tz) starts at z0*by*bx + dz*y0*bx + dz*dy*x0.
//
// Lorenzo / regression (interp == 3; sz.cpp:63-76 runs SZ3 with both
// predictors on, blockSize 6): each tile is split into eight 8^3 sub-blocks
// (cropped at the grid's edge), and a sub-block takes the least-squares
// plane a*i + b*j + c*k + d of its original values (local coordinates,
// clamped to the field's range) when the plane's absolute error sum is
// below the Lorenzo estimate's (on the original values) plus 1.22*eb per
// point, Lorenzo otherwise.  Wave w of the four fits sub-blocks w and w+4:
// lane l = j + 8k sums its row over i in double, a butterfly over the wave
// (v += v[l ^ m], m = 32 .. 1) totals the sums, so the host restatement
// can follow the same order and get the same bits.  The tile's stream
// starts with 32 header symbols, four per sub-block: code 0 carrying a, b,
// c, d as T (unpredictables) for a plane, the zero residual for Lorenzo;
// tile t's symbols start kLorenzoHdr * t later than without them.
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "kernels.h"

namespace stz {

namespace {

constexpr int kT = kLorenzoTile;
constexpr int kHW = 1024;
constexpr uint32_t kHLo = kRadius - kHW / 2;

__device__ __forceinline__ void hist_put(uint32_t* sh, uint32_t* gh, uint32_t q, uint32_t n)
{
    if (q - kHLo < (uint32_t)kHW) lds_add(&sh[q - kHLo], n);
    else glb_add(&gh[q], n);
}

// one wave's codes into the LDS histogram: lanes holding the most common
// code add it once (two rounds), the rest one by one
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

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, 64);
    return v;
}

// Lorenzo prediction of tile-local (x, y, z) from the tile in S, neighbours
// outside the tile reading as zero
template <typename T>
__device__ __forceinline__ T lz_pred(const T* S, int x, int y, int z)
{
    const int i = (z * kT + y) * kT + x;
    const T zero = T(0);
    const T a = x ? S[i - 1] : zero, b = y ? S[i - kT] : zero, c = z ? S[i - kT * kT] : zero;
    const T ab = x && y ? S[i - 1 - kT] : zero, ac = x && z ? S[i - 1 - kT * kT] : zero;
    const T bc = y && z ? S[i - kT - kT * kT] : zero;
    const T abc = x && y && z ? S[i - 1 - kT - kT * kT] : zero;
    T p = a + b;
    p = p + c;
    p = p - ab;
    p = p - ac;
    p = p - bc;
    p = p + abc;
    return p;
}

template <typename T>
__device__ __forceinline__ T reg_pred(T ca, T cb, T cc, T cd, int i, int j, int k, T lo, T hi)
{
    T p = ca * (T)i;
    p = p + cb * (T)j;
    p = p + cc * (T)k;
    p = p + cd;
    p = p < lo ? lo : p;
    p = p > hi ? hi : p;
    return p;
}

// one wave: the plane fit of the sub-block at tile-local (sx, sy, sz) with
// extents nx, ny, nz (1..8) from the original values in S, and the choice
// (true: the plane); every lane returns the same
template <typename T>
__device__ bool reg_choose(const T* S, int sx, int sy, int sz, int nx, int ny, int nz, double eb, T lo, T hi,
                           T& ca, T& cb, T& cc, T& cd)
{
    const int l = (int)__lane_id(), j = l & 7, k = l >> 3;
    const bool row = j < ny && k < nz;
    const double xm = 0.5 * (nx - 1), ym = 0.5 * (ny - 1), zm = 0.5 * (nz - 1);
    const T* R = S + ((sz + k) * kT + (sy + j)) * kT + sx;
    double r = 0.0, rx = 0.0;
    if (row)
        for (int i = 0; i < nx; ++i) {
            const double f = (double)R[i];
            r = r + f;
            rx = rx + ((double)i - xm) * f;
        }
    const double S0 = wave_sum(r), Sx = wave_sum(rx), Sy = wave_sum(((double)j - ym) * r),
                 Sz = wave_sum(((double)k - zm) * r);
    const double n = (double)nx * ny * nz;
    const double a = nx > 1 ? Sx / ((double)ny * nz * ((double)((nx - 1) * nx * (nx + 1)) / 12.0)) : 0.0;
    const double b = ny > 1 ? Sy / ((double)nx * nz * ((double)((ny - 1) * ny * (ny + 1)) / 12.0)) : 0.0;
    const double c = nz > 1 ? Sz / ((double)nx * ny * ((double)((nz - 1) * nz * (nz + 1)) / 12.0)) : 0.0;
    double d = S0 / n;
    d = d - a * xm;
    d = d - b * ym;
    d = d - c * zm;
    ca = (T)a; cb = (T)b; cc = (T)c; cd = (T)d;
    double e1 = 0.0, e2 = 0.0;
    if (row)
        for (int i = 0; i < nx; ++i) {
            const double f = (double)R[i];
            e1 = e1 + fabs(f - (double)reg_pred<T>(ca, cb, cc, cd, i, j, k, lo, hi));
            e2 = e2 + fabs(f - (double)lz_pred<T>(S, sx + i, sy + j, sz + k));
        }
    const double ER = wave_sum(e1), EL = wave_sum(e2);
    return ER < EL + (1.22 * eb) * n;
}

template <typename T, bool kEnc>
__device__ __forceinline__ void lorenzo_tile(const LorenzoJob& J, uint32_t tile)
{
    __shared__ T S[kT * kT * kT];
    __shared__ uint32_t sh[kEnc ? kHW : 1];
    __shared__ T Cf[8][4];
    __shared__ int Rg[8];
    const uint32_t ntx = (uint32_t)(J.bx + kT - 1) / kT, nty = (uint32_t)(J.by + kT - 1) / kT;
    const uint32_t txy = tile / ntx;
    const int x0 = (int)(tile - txy * ntx) * kT;
    const int y0 = (int)(txy % nty) * kT;
    const int z0 = (int)(txy / nty) * kT;
    const int dx = min(kT, J.bx - x0), dy = min(kT, J.by - y0), dz = min(kT, J.bz - z0);
    // the tile's first symbol (its header's, with reg) and first point
    const uint64_t hdr = (uint64_t)z0 * J.by * J.bx + (uint64_t)dz * y0 * J.bx + (uint64_t)dz * dy * x0 +
                         (J.reg ? (uint64_t)kLorenzoHdr * tile : 0);
    const uint64_t off = hdr + (J.reg ? kLorenzoHdr : 0);
    const size_t sy = (size_t)J.bx, sz = (size_t)J.bx * J.by;
    T* __restrict__ v = (T*)J.v;
    T* __restrict__ side = (T*)J.side;
    const T lo = (T)J.lo, hi = (T)J.hi;
    const int npts = dx * dy * dz;
    if (threadIdx.x < 8) Rg[threadIdx.x] = 0;
    if (kEnc) {
        for (int k = threadIdx.x; k < kHW; k += 256) sh[k] = 0u;
        for (int k = threadIdx.x; k < npts; k += 256) {
            const int x = k % dx, t = k / dx, y = t % dy, z = t / dy;
            S[(z * kT + y) * kT + x] = v[(size_t)(x0 + x) + (size_t)(y0 + y) * sy + (size_t)(z0 + z) * sz];
        }
    }
    __syncthreads();
    if (J.reg) {
        if (kEnc) {
            // wave w fits sub-blocks w and w + 4 and writes their header symbols
            const int w = threadIdx.x >> 6, l = (int)__lane_id();
            for (int b = w; b < 8; b += 4) {
                const int sx = (b & 1) * 8, sy0 = ((b >> 1) & 1) * 8, sz0 = (b >> 2) * 8;
                const int nx = min(8, dx - sx), ny = min(8, dy - sy0), nz = min(8, dz - sz0);
                bool rg = false;
                T ca = T(0), cb = T(0), cc = T(0), cd = T(0);
                if (nx > 0 && ny > 0 && nz > 0) rg = reg_choose<T>(S, sx, sy0, sz0, nx, ny, nz, J.eb, lo, hi, ca, cb, cc, cd);
                if (l < 4) {
                    const uint64_t pos = hdr + 4 * b + l;
                    J.codes[pos] = rg ? (uint16_t)0 : (uint16_t)kRadius;
                    if (rg) side[pos] = l == 0 ? ca : l == 1 ? cb : l == 2 ? cc : cd;
                }
                if (l == 0) {
                    Rg[b] = rg;
                    Cf[b][0] = ca; Cf[b][1] = cb; Cf[b][2] = cc; Cf[b][3] = cd;
                }
                hist_add(sh, J.hist, rg ? 0u : (uint32_t)kRadius, l < 4);
            }
        } else if (threadIdx.x < 32) {
            const int b = threadIdx.x >> 2, e = threadIdx.x & 3;
            const uint32_t q = J.codes[hdr + threadIdx.x];
            Cf[b][e] = q == 0 ? side[hdr + threadIdx.x] : T(0);
            if (e == 0) Rg[b] = q == 0;
        }
        __syncthreads();
    }
    const int x = threadIdx.x & (kT - 1), y = threadIdx.x / kT;
    const bool col = x < dx && y < dy;
    const int nplanes = dx + dy + dz - 2;
    for (int s = 0; s < nplanes; ++s) {
        const int z = s - x - y;
        const bool on = col && z >= 0 && z < dz;
        uint32_t q = 0;
        if (on) {
            const int i = (z * kT + y) * kT + x;
            const int b = ((z >> 3) << 2) | ((y >> 3) << 1) | (x >> 3);
            const T p = Rg[b] ? reg_pred<T>(Cf[b][0], Cf[b][1], Cf[b][2], Cf[b][3], x & 7, y & 7, z & 7, lo, hi)
                              : lz_pred<T>(S, x, y, z);
            const uint64_t pos = off + (uint64_t)(z * dy + y) * dx + x;
            if (kEnc) {
                const T orig = S[i];
                T dec;
                q = quantize(orig - p, J.eb, J.recip, dec);
                J.codes[pos] = (uint16_t)q;
                if (q) S[i] = p + dec;
                else side[pos] = orig;
            } else {
                q = J.codes[pos];
                S[i] = q ? p + dequantize<T>(q, J.eb) : side[pos];
            }
        }
        if (kEnc) hist_add(sh, J.hist, q, on);
        __syncthreads();
    }
    for (int k = threadIdx.x; k < npts; k += 256) {
        const int xx = k % dx, t = k / dx, yy = t % dy, zz = t / dy;
        v[(size_t)(x0 + xx) + (size_t)(y0 + yy) * sy + (size_t)(z0 + zz) * sz] = S[(zz * kT + yy) * kT + xx];
    }
    if (kEnc)
        for (int k = threadIdx.x; k < kHW; k += 256)
            if (sh[k]) atomicAdd(&J.hist[kHLo + k], sh[k]);
}

template <typename T, bool kEnc>
__global__ __launch_bounds__(256) void k_lorenzo(LorenzoJob J)
{
    lorenzo_tile<T, kEnc>(J, blockIdx.x);
}

// many fields: workgroups wg0 .. wg0 + ntiles - 1 code field j's tiles
template <typename T, bool kEnc>
__global__ __launch_bounds__(256) void k_lorenzo_multi(const LorenzoJob* __restrict__ jobs, uint32_t njobs)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (jobs[m].wg0 <= blockIdx.x) lo = m; else hi = m;
    }
    lorenzo_tile<T, kEnc>(jobs[lo], blockIdx.x - jobs[lo].wg0);
}

}  // namespace

uint32_t lorenzo_tiles(int bx, int by, int bz)
{
    return (uint32_t)((bx + kT - 1) / kT) * (uint32_t)((by + kT - 1) / kT) * (uint32_t)((bz + kT - 1) / kT);
}

template <typename T>
void launch_lorenzo(hipStream_t s, const LorenzoJob& j, bool encode)
{
    const uint32_t nt = lorenzo_tiles(j.bx, j.by, j.bz);
    if (!nt) return;
    if (encode) k_lorenzo<T, true><<<dim3(nt), dim3(256), 0, s>>>(j);
    else k_lorenzo<T, false><<<dim3(nt), dim3(256), 0, s>>>(j);
}

template <typename T>
void launch_lorenzo_multi(hipStream_t s, const LorenzoJob* jobs, uint32_t njobs, uint32_t wgs, bool encode)
{
    if (!njobs || !wgs) return;
    if (encode) k_lorenzo_multi<T, true><<<dim3(wgs), dim3(256), 0, s>>>(jobs, njobs);
    else k_lorenzo_multi<T, false><<<dim3(wgs), dim3(256), 0, s>>>(jobs, njobs);
}

template void launch_lorenzo<float>(hipStream_t, const LorenzoJob&, bool);
template void launch_lorenzo<double>(hipStream_t, const LorenzoJob&, bool);
template void launch_lorenzo_multi<float>(hipStream_t, const LorenzoJob*, uint32_t, uint32_t, bool);
template void launch_lorenzo_multi<double>(hipStream_t, const LorenzoJob*, uint32_t, uint32_t, bool);

}  // namespace stz

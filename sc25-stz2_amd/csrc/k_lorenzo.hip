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

template <typename T, bool kEnc>
__device__ __forceinline__ void lorenzo_tile(const LorenzoJob& J, uint32_t tile)
{
    __shared__ T S[kT * kT * kT];
    __shared__ uint32_t sh[kEnc ? kHW : 1];
    const uint32_t ntx = (uint32_t)(J.bx + kT - 1) / kT, nty = (uint32_t)(J.by + kT - 1) / kT;
    const uint32_t txy = tile / ntx;
    const int x0 = (int)(tile - txy * ntx) * kT;
    const int y0 = (int)(txy % nty) * kT;
    const int z0 = (int)(txy / nty) * kT;
    const int dx = min(kT, J.bx - x0), dy = min(kT, J.by - y0), dz = min(kT, J.bz - z0);
    const uint64_t off = (uint64_t)z0 * J.by * J.bx + (uint64_t)dz * y0 * J.bx + (uint64_t)dz * dy * x0;
    const size_t sy = (size_t)J.bx, sz = (size_t)J.bx * J.by;
    T* __restrict__ v = (T*)J.v;
    T* __restrict__ side = (T*)J.side;
    const int npts = dx * dy * dz;
    if (kEnc) {
        for (int k = threadIdx.x; k < kHW; k += 256) sh[k] = 0u;
        for (int k = threadIdx.x; k < npts; k += 256) {
            const int x = k % dx, t = k / dx, y = t % dy, z = t / dy;
            S[(z * kT + y) * kT + x] = v[(size_t)(x0 + x) + (size_t)(y0 + y) * sy + (size_t)(z0 + z) * sz];
        }
    }
    __syncthreads();
    const int x = threadIdx.x & (kT - 1), y = threadIdx.x / kT;
    const bool col = x < dx && y < dy;
    const int nplanes = dx + dy + dz - 2;
    const T zero = T(0);
    for (int s = 0; s < nplanes; ++s) {
        const int z = s - x - y;
        const bool on = col && z >= 0 && z < dz;
        uint32_t q = 0;
        if (on) {
            const int i = (z * kT + y) * kT + x;
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

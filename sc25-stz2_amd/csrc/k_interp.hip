// k_interp.hip -- base codec passes on gfx950.
//
// The reference compresses its 1/4-resolution base block with SZ3-ncp's
// interpolation compressor (SZ_compress4De, F5:1439-1449), which is not
// available.  This project's base codec is a dyadic interpolation scheme whose
// points within one (level, axis) pass depend only on earlier passes, so each
// pass is one data-parallel launch: predict from reconstructed neighbours at
// +-s / +-3s along the pass axis, quantize, reconstruct in place.
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "kernels.h"

namespace stz {

namespace {

constexpr int kHW = 1024;
constexpr int kFusedDecThreads = 1024;
constexpr uint32_t kHLo = kRadius - kHW / 2;
// A one-field encode pass: at most kPassEncWgs workgroups (each flushes its
// LDS histogram with atomics on the same few bins) of kPassEncThreads
// threads, so the CUs stay occupied (512^3 field, base passes per step:
// 256-thread x 512 workgroups 108 us, 1024 x 512 93 us, 1024 x 256 91 us)
constexpr int kPassEncThreads = 1024;
constexpr unsigned kPassEncWgs = 256;

template <typename T>
__device__ __forceinline__ T interp_pred(const T* __restrict__ v, long c, long st, int pos, int s, int dim,
                                         int interp)
{
    const bool has_r1 = pos + s < dim, has_l3 = pos >= 3 * s, has_r3 = pos + 3 * s < dim;
    const T b = v[c - (long)s * st];
    if (interp == 1) {
        if (has_r1) {
            const T cc = v[c + (long)s * st];
            if (has_l3 && has_r3) return cub(v[c - 3L * s * st], b, cc, v[c + 3L * s * st]);
            if (has_l3) return T(-0.125f) * v[c - 3L * s * st] + T(0.75f) * b + T(0.375f) * cc;
            if (has_r3) return T(0.375f) * b + T(0.75f) * cc - T(0.125f) * v[c + 3L * s * st];
            return T(0.5f) * (b + cc);
        }
        if (has_l3) return T(1.5f) * b - T(0.5f) * v[c - 3L * s * st];
        return b;
    }
    if (has_r1) return T(0.5f) * (b + v[c + (long)s * st]);
    if (has_l3) return T(1.5f) * b - T(0.5f) * v[c - 3L * s * st];
    return b;
}

struct PassPoint {
    long c;
    int pos_axis;
    long st;
    int dim;
};

__device__ __forceinline__ PassPoint pass_point(const InterpPass& p, uint64_t k64)
{
    const uint32_t k = (uint32_t)k64;  // a pass has < 2^31 points
    const uint32_t c0 = (uint32_t)p.cnt[0], c1 = (uint32_t)p.cnt[1];
    const uint32_t t = k / c0;
    const int ix = (int)(k - t * c0);
    const int iz = (int)(t / c1);
    const int iy = (int)(t - (uint32_t)iz * c1);
    const int cx = p.start[0] + ix * p.step[0];
    const int cy = p.start[1] + iy * p.step[1];
    const int cz = p.start[2] + iz * p.step[2];
    PassPoint r;
    r.c = cx + (long)cy * p.bx + (long)cz * p.bx * p.by;
    r.pos_axis = p.axis == 0 ? cx : p.axis == 1 ? cy : cz;
    r.st = p.axis == 0 ? 1L : p.axis == 1 ? (long)p.bx : (long)p.bx * p.by;
    r.dim = p.axis == 0 ? p.bx : p.axis == 1 ? p.by : p.bz;
    return r;
}

__device__ __forceinline__ void hist_put(uint32_t* sh, uint32_t* gh, uint32_t q, uint32_t n)
{
    if (q - kHLo < (uint32_t)kHW) lds_add(&sh[q - kHLo], n);
    else glb_add(&gh[q], n);
}

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

template <typename T, int kThreads>
__device__ __forceinline__ void interp_encode_wg(const InterpPass& p, T* __restrict__ v, uint16_t* __restrict__ codes,
                                                 T* __restrict__ side, uint32_t* __restrict__ hist, uint32_t wg,
                                                 uint32_t nwg)
{
    __shared__ uint32_t sh[kHW];
    for (int k = threadIdx.x; k < kHW; k += kThreads) sh[k] = 0u;
    __syncthreads();
    const uint64_t stride = (uint64_t)nwg * kThreads;
    const uint64_t kend = (p.n + 255u) / 256u * 256u;  // every lane of a wave runs the same trip count
    for (uint64_t k = (uint64_t)wg * kThreads + threadIdx.x; k < kend; k += stride) {
        const bool ok = k < p.n;
        uint32_t q = 0;
        if (ok) {
            T dec;
            if (p.axis < 0) {  // anchor: predicted by zero
                const long c = (long)p.start[0] + (long)p.start[1] * p.bx + (long)p.start[2] * p.bx * p.by;
                q = quantize(v[c], p.eb, p.recip, dec);
                codes[p.pos + k] = (uint16_t)q;
                if (q) v[c] = dec; else side[p.pos + k] = v[c];
            } else {
                const PassPoint pp = pass_point(p, k);
                const T pred = interp_pred(v, pp.c, pp.st, pp.pos_axis, p.s, pp.dim, p.interp);
                const T orig = v[pp.c];
                q = quantize(orig - pred, p.eb, p.recip, dec);
                codes[p.pos + k] = (uint16_t)q;
                if (q) v[pp.c] = pred + dec; else side[p.pos + k] = orig;
            }
        }
        hist_add(sh, hist, q, ok);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kHW; k += kThreads)
        if (sh[k]) atomicAdd(&hist[kHLo + k], sh[k]);
}

template <typename T>
__global__ __launch_bounds__(kPassEncThreads) void k_interp_encode(InterpPass p, T* __restrict__ v,
                                                       uint16_t* __restrict__ codes, T* __restrict__ side,
                                                       uint32_t* __restrict__ hist)
{
    interp_encode_wg<T, kPassEncThreads>(p, v, codes, side, hist, blockIdx.x, gridDim.x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_interp_encode_multi(const PassEncJob* __restrict__ jobs, uint32_t njobs)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (jobs[m].wg0 <= blockIdx.x) lo = m; else hi = m;
    }
    const PassEncJob& J = jobs[lo];
    interp_encode_wg<T, 256>(J.p, (T*)J.v, J.codes, (T*)J.side, J.hist, blockIdx.x - J.wg0, J.nwg);
}

template <typename T>
__global__ __launch_bounds__(256) void k_interp_decode(InterpPass p, T* __restrict__ v,
                                                       const uint16_t* __restrict__ codes,
                                                       const T* __restrict__ side)
{
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < p.n; k += stride) {
        const uint32_t q = codes[p.pos + k];
        if (p.axis < 0) {
            const long c = (long)p.start[0] + (long)p.start[1] * p.bx + (long)p.start[2] * p.bx * p.by;
            v[c] = q ? dequantize<T>(q, p.eb) : side[p.pos + k];
        } else {
            const PassPoint pp = pass_point(p, k);
            const T pred = interp_pred(v, pp.c, pp.st, pp.pos_axis, p.s, pp.dim, p.interp);
            v[pp.c] = q ? pred + dequantize<T>(q, p.eb) : side[p.pos + k];
        }
    }
}

// Coarse passes back to back in one workgroup of 1024 lanes, on the
// stride-S0 lattice held in LDS (its predicates in lattice units equal the
// base-grid ones: pos + s < dim  <=>  pos/S0 + s/S0 < (dim-1)/S0 + 1).
// The barrier between passes orders them.
template <typename T>
__device__ __forceinline__ void lattice_io(const InterpPassList& l, T* __restrict__ L, T* __restrict__ v, bool load)
{
    const uint32_t n = (uint32_t)l.lx * l.ly * l.lz;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        const uint32_t t = k / (uint32_t)l.lx;
        const int x = (int)(k - t * (uint32_t)l.lx);
        const int z = (int)(t / (uint32_t)l.ly);
        const int y = (int)(t - (uint32_t)z * (uint32_t)l.ly);
        const size_t g = ((size_t)(z * l.S0) * l.by + (size_t)y * l.S0) * l.bx + (size_t)x * l.S0;
        if (load) L[k] = v[g];
        else v[g] = L[k];
    }
}

template <typename T>
__device__ __forceinline__ void interp_encode_fused_wg(const InterpPassList& l, T* __restrict__ v,
                                                       uint16_t* __restrict__ codes, T* __restrict__ side,
                                                       uint32_t* __restrict__ hist)
{
    __shared__ uint32_t sh[kHW];
    __shared__ T L[kLatticeBytes / sizeof(T)];
    for (int k = threadIdx.x; k < kHW; k += 1024) sh[k] = 0u;
    lattice_io(l, L, v, true);
    __syncthreads();
    for (int pi = 0; pi < l.n; ++pi) {
        const InterpPass& p = l.p[pi];
        const uint64_t kend = (p.n + 1023u) / 1024u * 1024u;
        for (uint64_t k = threadIdx.x; k < kend; k += 1024u) {
            const bool ok = k < p.n;
            uint32_t q = 0;
            if (ok) {
                T dec;
                if (p.axis < 0) {
                    q = quantize(L[0], p.eb, p.recip, dec);
                    codes[p.pos + k] = (uint16_t)q;
                    if (q) L[0] = dec; else side[p.pos + k] = L[0];
                } else {
                    const PassPoint pp = pass_point(p, k);
                    const T pred = interp_pred(L, pp.c, pp.st, pp.pos_axis, p.s, pp.dim, p.interp);
                    const T orig = L[pp.c];
                    q = quantize(orig - pred, p.eb, p.recip, dec);
                    codes[p.pos + k] = (uint16_t)q;
                    if (q) L[pp.c] = pred + dec; else side[p.pos + k] = orig;
                }
            }
            hist_add(sh, hist, q, ok);
        }
        __syncthreads();
    }
    lattice_io(l, L, v, false);
    for (int k = threadIdx.x; k < kHW; k += 1024)
        if (sh[k]) atomicAdd(&hist[kHLo + k], sh[k]);
}

template <typename T>
__global__ __launch_bounds__(1024) void k_interp_encode_fused(InterpPassList l, T* __restrict__ v,
                                                              uint16_t* __restrict__ codes, T* __restrict__ side,
                                                              uint32_t* __restrict__ hist)
{
    interp_encode_fused_wg(l, v, codes, side, hist);
}

template <typename T>
__global__ __launch_bounds__(1024) void k_interp_encode_fused_multi(const FusedEncJob* __restrict__ jobs)
{
    const FusedEncJob& J = jobs[blockIdx.x];
    interp_encode_fused_wg(J.l, (T*)J.v, J.codes, (T*)J.side, J.hist);
}

template <typename T>
__device__ __forceinline__ void interp_decode_fused_wg(const InterpPassList& l, T* __restrict__ v,
                                                       const uint16_t* __restrict__ codes,
                                                       const T* __restrict__ side, T* L)
{
    for (int pi = 0; pi < l.n; ++pi) {
        const InterpPass& p = l.p[pi];
        for (uint64_t k = threadIdx.x; k < p.n; k += blockDim.x) {
            const uint32_t q = codes[p.pos + k];
            if (p.axis < 0) {
                L[0] = q ? dequantize<T>(q, p.eb) : side[p.pos + k];
            } else {
                const PassPoint pp = pass_point(p, k);
                const T pred = interp_pred(L, pp.c, pp.st, pp.pos_axis, p.s, pp.dim, p.interp);
                L[pp.c] = q ? pred + dequantize<T>(q, p.eb) : side[p.pos + k];
            }
        }
        __syncthreads();
    }
    lattice_io(l, L, v, false);
}

template <typename T>
__global__ __launch_bounds__(kFusedDecThreads) void k_interp_decode_fused(InterpPassList l, T* __restrict__ v,
                                                              const uint16_t* __restrict__ codes,
                                                              const T* __restrict__ side)
{
    __shared__ T L[kFusedPoints];
    interp_decode_fused_wg(l, v, codes, side, L);
}

template <typename T>
__global__ __launch_bounds__(kFusedDecThreads) void k_interp_decode_fused_multi(const FusedJob* __restrict__ jobs)
{
    __shared__ T L[kLatticeBytes / sizeof(T)];
    const FusedJob& J = jobs[blockIdx.x];
    interp_decode_fused_wg(J.l, (T*)J.v, J.codes, (const T*)J.side, L);
}

template <typename T>
__global__ __launch_bounds__(256) void k_interp_decode_multi(const PassJob* __restrict__ jobs, uint32_t njobs)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (jobs[m].wg0 <= blockIdx.x) lo = m; else hi = m;
    }
    const PassJob& J = jobs[lo];
    const InterpPass& p = J.p;
    T* __restrict__ v = (T*)J.v;
    const T* __restrict__ side = (const T*)J.side;
    const uint64_t stride = (uint64_t)J.nwg * 256u;
    for (uint64_t k = (blockIdx.x - J.wg0) * 256ull + threadIdx.x; k < p.n; k += stride) {
        const uint32_t q = J.codes[p.pos + k];
        if (p.axis < 0) {
            const long c = (long)p.start[0] + (long)p.start[1] * p.bx + (long)p.start[2] * p.bx * p.by;
            v[c] = q ? dequantize<T>(q, p.eb) : side[p.pos + k];
        } else {
            const PassPoint pp = pass_point(p, k);
            const T pred = interp_pred(v, pp.c, pp.st, pp.pos_axis, p.s, pp.dim, p.interp);
            v[pp.c] = q ? pred + dequantize<T>(q, p.eb) : side[p.pos + k];
        }
    }
}

inline unsigned pass_grid(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : b > 4096 ? 4096 : b);
}

}  // namespace

template <typename T>
void launch_interp_encode(hipStream_t s, const InterpPass& p, T* v, uint16_t* codes, T* side, uint32_t* hist)
{
    if (!p.n) return;
    const uint64_t g = (p.n + kPassEncThreads - 1) / kPassEncThreads;
    k_interp_encode<T><<<dim3((unsigned)(g < kPassEncWgs ? g : kPassEncWgs)), dim3(kPassEncThreads), 0, s>>>(
        p, v, codes, side, hist);
}

template <typename T>
void launch_interp_decode(hipStream_t s, const InterpPass& p, T* v, const uint16_t* codes, const T* side)
{
    if (!p.n) return;
    k_interp_decode<T><<<dim3(pass_grid(p.n)), dim3(256), 0, s>>>(p, v, codes, side);
}

template <typename T>
void launch_interp_encode_fused(hipStream_t s, const InterpPassList& l, T* v, uint16_t* codes, T* side, uint32_t* hist)
{
    if (l.n <= 0) return;
    k_interp_encode_fused<T><<<dim3(1), dim3(1024), 0, s>>>(l, v, codes, side, hist);
}

template <typename T>
void launch_interp_decode_fused(hipStream_t s, const InterpPassList& l, T* v, const uint16_t* codes, const T* side)
{
    if (l.n <= 0) return;
    k_interp_decode_fused<T><<<dim3(1), dim3(kFusedDecThreads), 0, s>>>(l, v, codes, side);
}

template <typename T>
void launch_interp_decode_fused_multi(hipStream_t s, const FusedJob* jobs, uint32_t njobs)
{
    if (!njobs) return;
    k_interp_decode_fused_multi<T><<<dim3(njobs), dim3(kFusedDecThreads), 0, s>>>(jobs);
}

template <typename T>
void launch_interp_decode_multi(hipStream_t s, const PassJob* jobs, uint32_t njobs, uint32_t wgs)
{
    if (!njobs || !wgs) return;
    k_interp_decode_multi<T><<<dim3(wgs), dim3(256), 0, s>>>(jobs, njobs);
}

template <typename T>
void launch_interp_encode_fused_multi(hipStream_t s, const FusedEncJob* jobs, uint32_t njobs)
{
    if (!njobs) return;
    k_interp_encode_fused_multi<T><<<dim3(njobs), dim3(1024), 0, s>>>(jobs);
}

template <typename T>
void launch_interp_encode_multi(hipStream_t s, const PassEncJob* jobs, uint32_t njobs, uint32_t wgs)
{
    if (!njobs || !wgs) return;
    k_interp_encode_multi<T><<<dim3(wgs), dim3(256), 0, s>>>(jobs, njobs);
}

template void launch_interp_encode_fused_multi<float>(hipStream_t, const FusedEncJob*, uint32_t);
template void launch_interp_encode_fused_multi<double>(hipStream_t, const FusedEncJob*, uint32_t);
template void launch_interp_encode_multi<float>(hipStream_t, const PassEncJob*, uint32_t, uint32_t);
template void launch_interp_encode_multi<double>(hipStream_t, const PassEncJob*, uint32_t, uint32_t);
template void launch_interp_decode_fused_multi<float>(hipStream_t, const FusedJob*, uint32_t);
template void launch_interp_decode_fused_multi<double>(hipStream_t, const FusedJob*, uint32_t);
template void launch_interp_decode_multi<float>(hipStream_t, const PassJob*, uint32_t, uint32_t);
template void launch_interp_decode_multi<double>(hipStream_t, const PassJob*, uint32_t, uint32_t);
template void launch_interp_encode_fused<float>(hipStream_t, const InterpPassList&, float*, uint16_t*, float*, uint32_t*);
template void launch_interp_encode_fused<double>(hipStream_t, const InterpPassList&, double*, uint16_t*, double*, uint32_t*);
template void launch_interp_decode_fused<float>(hipStream_t, const InterpPassList&, float*, const uint16_t*, const float*);
template void launch_interp_decode_fused<double>(hipStream_t, const InterpPassList&, double*, const uint16_t*, const double*);
template void launch_interp_encode<float>(hipStream_t, const InterpPass&, float*, uint16_t*, float*, uint32_t*);
template void launch_interp_encode<double>(hipStream_t, const InterpPass&, double*, uint16_t*, double*, uint32_t*);
template void launch_interp_decode<float>(hipStream_t, const InterpPass&, float*, const uint16_t*, const float*);
template void launch_interp_decode<double>(hipStream_t, const InterpPass&, double*, const uint16_t*, const double*);

}  // namespace stz

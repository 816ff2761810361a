// k_stage.hip -- kernels of the per-stream stage entry points (gfx950):
// one residual block's fused predict + quantize, and the moves between a
// block's row-major order and a marked stream's chunk-major order.
//
// F5 = src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp
//   fused residual encode   preprocess_block_zslab (F5:1161-1175) feeding
//                           SZ_compress's quantizer (F5:1345-1397)
//   stream order            encode_marked_fast's chunk order (row-major chunk
//                           ids F5:1578-1582, chunk-local row-major F5:1884-1905)
//   marked-chunk placement  SZ_decompress_nopred_marked_chunks (F5:1870-1905)
#include <hip/hip_runtime.h>

#include <mutex>

#include "device_math.h"
#include "kernels.h"

namespace stz {

// Device scratch of the allocation-free residual encode: the block's
// residuals, which the ordered unpredictable compaction reads back.  Every
// GPU holds its own copy for the life of the process; its address on a
// device comes from hipGetSymbolAddress on that device (the name itself is a
// host-side stand-in, never a device address).  All calls on one GPU share
// it, so calls that overlap on one GPU race.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_dev[kMaxDevices] = {};
}  // namespace

int scratch_setup(int device, void** ptr)
{
    if (device < 0 || device >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[device]) {
        // the caller has made `device` current
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return -1;
        scratch_dev[device] = p;
    }
    *ptr = scratch_dev[device];
    return 0;
}

namespace {

inline unsigned grid_of(uint64_t n, unsigned cap = 16384)
{
    uint64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    return (unsigned)(b > cap ? cap : b);
}

// diff = sub - pred(ref) (F5:1161-1175), code and decoded value from the
// quantizer (F5:1345-1397); rec = decoded diff + pred, the value the low
// level writes back (F5:1381-1385, then F5:1077-1083); diff kept for the
// ordered unpredictable list
template <int B, typename T>
__global__ void k_residual_encode(const T* __restrict__ sub, const T* __restrict__ ref, int dx, int dy, int dz,
                                  double eb, double recip, uint16_t* __restrict__ codes, T* __restrict__ diff,
                                  T* __restrict__ rec)
{
    const uint64_t n = (uint64_t)dx * dy * dz;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % dx), y = (int)((i / dx) % dy), z = (int)(i / ((uint64_t)dx * dy));
        const T p = predict<B>(ref, (int)i, x, y, z, dx, dy, dz);
        const T d = sub[i] - p;
        T dec;
        codes[i] = (uint16_t)quantize(d, eb, recip, dec);
        diff[i] = d;
        if (rec) rec[i] = dec + p;
    }
}

// row-major block codes -> stream order
__global__ void k_codes_to_stream(const uint16_t* __restrict__ codes, Geom g, uint16_t* __restrict__ out)
{
    const uint64_t n = (uint64_t)g.dx * g.dy * g.dz;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % g.dx), y = (int)((i / g.dx) % g.dy), z = (int)(i / ((uint64_t)g.dx * g.dy));
        out[stream_pos(g, x, y, z)] = codes[i];
    }
}

__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// One workgroup of 1024 walks the block in row-major order: the k-th
// unpredictable value (element order, as SZ_compress collects them) goes to
// the stream position of the k-th point whose code is 0.
template <typename T>
__global__ __launch_bounds__(1024) void k_unpred_to_stream(const uint16_t* __restrict__ codes,
                                                           const T* __restrict__ unpred, Geom g,
                                                           T* __restrict__ side)
{
    __shared__ uint32_t ws[16];
    const uint64_t n = (uint64_t)g.dx * g.dy * g.dz;
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    uint64_t carry = 0;
    for (uint64_t base = 0; base < n; base += 1024) {
        const uint64_t i = base + threadIdx.x;
        const bool f = i < n && codes[i] == 0;
        const uint32_t inc = wave_scan_incl(f ? 1u : 0u);
        if (lane == 63) ws[wid] = inc;
        __syncthreads();
        if (wid == 0) {
            uint32_t t = lane < 16 ? ws[lane] : 0u;
            t = wave_scan_incl(t);
            if (lane < 16) ws[lane] = t;
        }
        __syncthreads();
        const uint32_t r = inc - (f ? 1u : 0u) + (wid ? ws[wid - 1] : 0u);
        const uint32_t tot = ws[15];
        __syncthreads();
        if (f) {
            const int x = (int)(i % g.dx), y = (int)((i / g.dx) % g.dy), z = (int)(i / ((uint64_t)g.dx * g.dy));
            side[stream_pos(g, x, y, z)] = unpred[carry + r];
        }
        carry += tot;
    }
}

// symbol histogram of n codes (65536 bins); the centre code counted once
// per wave with a ballot
__global__ void k_hist_codes(const uint16_t* __restrict__ codes, uint64_t n, uint32_t* __restrict__ hist)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t first = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    // every lane of a wave runs the same trip count (ballots below)
    const uint64_t wave0 = first - __lane_id();
    for (uint64_t i0 = wave0; i0 < n; i0 += stride) {
        const uint64_t i = i0 + __lane_id();
        const bool ok = i < n;
        const uint32_t q = ok ? codes[i] : 0u;
        const bool centre = ok && q == (uint32_t)kRadius;
        const unsigned long long m = __ballot(centre);
        if (m && __lane_id() == 0) atomicAdd(&hist[kRadius], (uint32_t)__popcll(m));
        if (ok && !centre) atomicAdd(&hist[q], 1u);
    }
}

// Marked-chunk placement: workgroup k takes listed chunk list[2k] whose
// decoded symbols start at element list[2k+1] of the compact code / side
// buffers (chunk-local row-major) and writes the recovered values
// (F5:1531-1535 LUT values; unpredictable points their stored value) and,
// when asked, the codes at the chunk's points of the row-major block.
template <typename T>
__global__ __launch_bounds__(256) void k_chunks_to_block(const uint32_t* __restrict__ list, Geom g,
                                                         const uint16_t* __restrict__ codes,
                                                         const T* __restrict__ side, double eb,
                                                         T* __restrict__ out, uint16_t* __restrict__ codes_out)
{
    const uint32_t c = list[2 * blockIdx.x], p0 = list[2 * blockIdx.x + 1];
    const int C = g.chunk;
    const int cx = (int)(c % g.ncx), cy = (int)((c / g.ncx) % g.ncy), cz = (int)(c / ((uint32_t)g.ncx * g.ncy));
    const int wx = imin_(C, g.dx - cx * C), wy = imin_(C, g.dy - cy * C), wz = imin_(C, g.dz - cz * C);
    const uint32_t vol = (uint32_t)wx * wy * wz;
    for (uint32_t j = threadIdx.x; j < vol; j += blockDim.x) {
        const int lx = (int)(j % wx), ly = (int)((j / wx) % wy), lz = (int)(j / ((uint32_t)wx * wy));
        const uint64_t idx = ((uint64_t)(cz * C + lz) * g.dy + (cy * C + ly)) * g.dx + (cx * C + lx);
        const uint32_t q = codes[p0 + j];
        out[idx] = q ? dequantize<T>(q, eb) : side[p0 + j];
        if (codes_out) codes_out[idx] = (uint16_t)q;
    }
}

}  // namespace

template <typename T>
void launch_residual_encode(hipStream_t s, int block, const T* sub, const T* ref, int dx, int dy, int dz, double eb,
                            uint16_t* codes, T* diff, T* rec)
{
    const dim3 gr(grid_of((uint64_t)dx * dy * dz)), b(256);
    const double recip = 1.0 / eb;
    switch (block) {
    case 0: k_residual_encode<0, T><<<gr, b, 0, s>>>(sub, ref, dx, dy, dz, eb, recip, codes, diff, rec); break;
    case 1: k_residual_encode<1, T><<<gr, b, 0, s>>>(sub, ref, dx, dy, dz, eb, recip, codes, diff, rec); break;
    case 2: k_residual_encode<2, T><<<gr, b, 0, s>>>(sub, ref, dx, dy, dz, eb, recip, codes, diff, rec); break;
    case 3: k_residual_encode<3, T><<<gr, b, 0, s>>>(sub, ref, dx, dy, dz, eb, recip, codes, diff, rec); break;
    case 4: k_residual_encode<4, T><<<gr, b, 0, s>>>(sub, ref, dx, dy, dz, eb, recip, codes, diff, rec); break;
    case 5: k_residual_encode<5, T><<<gr, b, 0, s>>>(sub, ref, dx, dy, dz, eb, recip, codes, diff, rec); break;
    default: k_residual_encode<6, T><<<gr, b, 0, s>>>(sub, ref, dx, dy, dz, eb, recip, codes, diff, rec); break;
    }
}

template <typename T>
void launch_residual_encode_noalloc(hipStream_t s, void* scratch_ptr, int block, const T* sub, const T* ref, int dx,
                                    int dy, int dz, double eb, uint16_t* codes, T* unpred, uint64_t* count, T* rec)
{
    T* diff = (T*)scratch_ptr;   // the per-device address, never the array's name
    launch_residual_encode<T>(s, block, sub, ref, dx, dy, dz, eb, codes, diff, rec);
    launch_compact_unpred<T>(s, codes, diff, (uint64_t)dx * dy * dz, unpred, count);
}

template <typename T>
void launch_block_to_stream(hipStream_t s, const uint16_t* codes, const T* unpred, uint64_t n_unpred, const Geom& g,
                            uint16_t* codes_s, T* side_s)
{
    const uint64_t n = (uint64_t)g.dx * g.dy * g.dz;
    k_codes_to_stream<<<dim3(grid_of(n)), dim3(256), 0, s>>>(codes, g, codes_s);
    if (n_unpred) k_unpred_to_stream<T><<<dim3(1), dim3(1024), 0, s>>>(codes, unpred, g, side_s);
}

void launch_hist_codes(hipStream_t s, const uint16_t* codes, uint64_t n, uint32_t* hist)
{
    k_hist_codes<<<dim3(grid_of(n, 2048)), dim3(256), 0, s>>>(codes, n, hist);
}

template <typename T>
void launch_chunks_to_block(hipStream_t s, const uint32_t* list, uint32_t nlist, const Geom& g, const uint16_t* codes,
                            const T* side, double eb, T* out, uint16_t* codes_out)
{
    if (nlist) k_chunks_to_block<T><<<dim3(nlist), dim3(256), 0, s>>>(list, g, codes, side, eb, out, codes_out);
}

#define STZ_INST(T)                                                                                               \
    template void launch_residual_encode<T>(hipStream_t, int, const T*, const T*, int, int, int, double, uint16_t*, \
                                            T*, T*);                                                              \
    template void launch_residual_encode_noalloc<T>(hipStream_t, void*, int, const T*, const T*, int, int, int,   \
                                                    double, uint16_t*, T*, uint64_t*, T*);                       \
    template void launch_block_to_stream<T>(hipStream_t, const uint16_t*, const T*, uint64_t, const Geom&,          \
                                            uint16_t*, T*);                                                       \
    template void launch_chunks_to_block<T>(hipStream_t, const uint32_t*, uint32_t, const Geom&, const uint16_t*,   \
                                            const T*, double, T*, uint16_t*);
STZ_INST(float)
STZ_INST(double)

}  // namespace stz

// k_field.hip -- whole-field kernels: value range, padding, parity gathers and
// the single-kernel stage entry points (split / merge / residual / quantizer).
#include <hip/hip_runtime.h>

#include "device_math.h"
#include "kernels.h"

namespace stz {

namespace {

// computeRange (F5:92-104) in double; NaN never wins a comparison
template <typename T>
__device__ __forceinline__ void mm_upd(double v, double& mn, double& mx)
{
    if (v < mn) mn = v;
    if (v > mx) mx = v;
}

template <typename T>
__device__ __forceinline__ void block_minmax(double mn, double mx, double* smn, double* smx)
{
    smn[threadIdx.x] = mn;
    smx[threadIdx.x] = mx;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            if (smn[threadIdx.x + st] < smn[threadIdx.x]) smn[threadIdx.x] = smn[threadIdx.x + st];
            if (smx[threadIdx.x + st] > smx[threadIdx.x]) smx[threadIdx.x] = smx[threadIdx.x + st];
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_minmax(const T* __restrict__ d, uint64_t n, double* __restrict__ partial)
{
    __shared__ double smn[256], smx[256];
    double mn = __builtin_inf(), mx = -__builtin_inf();
    constexpr int V = 16 / sizeof(T);  // 16-byte vectors
    typedef T vec __attribute__((ext_vector_type(V)));
    const uint64_t nv = ((uintptr_t)d % 16 == 0) ? n / V : 0;
    const vec* dv = (const vec*)d;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < nv; i += (uint64_t)gridDim.x * 256ull) {
        const vec v = dv[i];
#pragma unroll
        for (int k = 0; k < V; ++k) mm_upd<T>((double)v[k], mn, mx);
    }
    for (uint64_t i = nv * V + blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull)
        mm_upd<T>((double)d[i], mn, mx);
    block_minmax<T>(mn, mx, smn, smx);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = smn[0];
        partial[2 * blockIdx.x + 1] = smx[0];
    }
}

__global__ __launch_bounds__(256) void k_minmax_final(const double* __restrict__ partial, int nb, double* __restrict__ out)
{
    __shared__ double smn[256], smx[256];
    double mn = __builtin_inf(), mx = -__builtin_inf();
    for (int b = threadIdx.x; b < nb; b += 256) {
        if (partial[2 * b] < mn) mn = partial[2 * b];
        if (partial[2 * b + 1] > mx) mx = partial[2 * b + 1];
    }
    block_minmax<double>(mn, mx, smn, smx);
    if (threadIdx.x == 0) {
        out[0] = smn[0];
        out[1] = smx[0];
    }
}

// edge-replicate padding to multiples of 4
template <typename T>
__global__ void k_pad(const T* __restrict__ in, int nx, int ny, int nz, T* __restrict__ out, int px, int py, int pz)
{
    const uint64_t n = (uint64_t)px * py * pz;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % px), y = (int)((i / px) % py), z = (int)(i / ((uint64_t)px * py));
        const int sx = x < nx ? x : nx - 1, sy = y < ny ? y : ny - 1, sz = z < nz ? z : nz - 1;
        out[i] = in[((uint64_t)sz * ny + sy) * nx + sx];
    }
}

// base-codec input: low block 0 = full[4z][4y][4x] (F5:2102-2109 two splits)
template <typename T>
__global__ void k_gather_low0(const T* __restrict__ full, int px, int py, int pz, T* __restrict__ v)
{
    const int lx = px / 4, ly = py / 4, lz = pz / 4;
    const uint64_t n = (uint64_t)lx * ly * lz;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % lx), y = (int)((i / lx) % ly), z = (int)(i / ((uint64_t)lx * ly));
        v[i] = full[((uint64_t)(4 * z) * py + 4 * y) * px + 4 * x];
    }
}

struct Ptr8 { void* p[8]; };

// slice_full_data (F5:107-122)
template <typename T>
__global__ void k_split(const T* __restrict__ full, int nx, int ny, int nz, Ptr8 sub)
{
    const uint64_t n = (uint64_t)nx * ny * nz;
    const int hx = nx / 2, hy = ny / 2;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((uint64_t)nx * ny));
        const int k = ((z & 1) << 2) | ((y & 1) << 1) | (x & 1);
        ((T*)sub.p[k])[((uint64_t)(z / 2) * hy + y / 2) * hx + x / 2] = full[i];
    }
}

// merge_sub_blocks_to_full (F5:68-90)
template <typename T>
__global__ void k_merge(Ptr8 sub, int sx, int sy, int sz, T* __restrict__ full)
{
    const int fx = 2 * sx, fy = 2 * sy;
    const uint64_t n = (uint64_t)fx * fy * 2 * sz;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % fx), y = (int)((i / fx) % fy), z = (int)(i / ((uint64_t)fx * fy));
        const int k = ((z & 1) << 2) | ((y & 1) << 1) | (x & 1);
        full[i] = ((const T*)sub.p[k])[((uint64_t)(z / 2) * sy + y / 2) * sx + x / 2];
    }
}

// sign > 0: out = sub - pred (F5:1005-1047); sign < 0: out = sub + pred (F5:1050-1091)
template <int B, typename T>
__global__ void k_residual(const T* __restrict__ sub, const T* __restrict__ ref, int dx, int dy, int dz,
                           T* __restrict__ out, int sign)
{
    const uint64_t n = (uint64_t)dx * dy * dz;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % dx), y = (int)((i / dx) % dy), z = (int)(i / ((uint64_t)dx * dy));
        const T p = predict<B>(ref, (int)i, x, y, z, dx, dy, dz);
        out[i] = sign > 0 ? sub[i] - p : sub[i] + p;
    }
}

template <typename T>
__global__ void k_quantize(const T* __restrict__ d, uint64_t n, double eb, double recip, uint16_t* __restrict__ codes)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        T dec;
        codes[i] = (uint16_t)quantize(d[i], eb, recip, dec);
    }
}

template <typename T>
__global__ void k_dequantize(const uint16_t* __restrict__ codes, const T* __restrict__ side, uint64_t n,
                             double eb, T* __restrict__ out)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = codes[i];
        out[i] = q ? dequantize<T>(q, eb) : side[i];
    }
}

// strided crop: out[z][y][x] = in[(sz + z*step)][(sy + y*step)][(sx + x*step)]
template <typename T>
__global__ void k_crop(const T* __restrict__ in, int px, int py, int ox, int oy, int oz, int sx, int sy,
                       int sz, int step, T* __restrict__ out)
{
    const uint64_t n = (uint64_t)ox * oy * oz;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % ox), y = (int)((i / ox) % oy), z = (int)(i / ((uint64_t)ox * oy));
        out[i] = in[((uint64_t)(sz + z * step) * py + (sy + y * step)) * px + (sx + x * step)];
    }
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// one workgroup, element order: exclusive rank of each flagged element
__device__ __forceinline__ uint32_t block_rank(bool f, uint32_t* ws, uint32_t& total)
{
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(f ? 1u : 0u);
    if (lane == 63) ws[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        uint32_t t = lane < 16 ? ws[lane] : 0u;
        t = wave_incl_scan(t);
        if (lane < 16) ws[lane] = t;
    }
    __syncthreads();
    const uint32_t r = inc - (f ? 1u : 0u) + (wid ? ws[wid - 1] : 0u);
    total = ws[15];
    __syncthreads();
    return r;
}

// unpredictable values (code 0) in element order (F5:1250 unpred.push_back)
template <typename T>
__global__ __launch_bounds__(1024) void k_compact_unpred(const uint16_t* __restrict__ codes, const T* __restrict__ data,
                                                         uint64_t n, T* __restrict__ out, uint64_t* __restrict__ count)
{
    __shared__ uint32_t ws[16];
    uint64_t carry = 0;
    for (uint64_t base = 0; base < n; base += 1024) {
        const uint64_t i = base + threadIdx.x;
        const bool f = i < n && codes[i] == 0;
        uint32_t tot;
        const uint32_t r = block_rank(f, ws, tot);
        if (f) out[carry + r] = data[i];
        carry += tot;
    }
    if (threadIdx.x == 0) *count = carry;
}

// decode_marked_transform's recover with the unpredictable list consumed in element order
template <typename T>
__global__ __launch_bounds__(1024) void k_dequant_list(const uint16_t* __restrict__ codes, const T* __restrict__ unpred,
                                                       uint64_t n, double eb, T* __restrict__ out)
{
    __shared__ uint32_t ws[16];
    uint64_t carry = 0;
    for (uint64_t base = 0; base < n; base += 1024) {
        const uint64_t i = base + threadIdx.x;
        const uint32_t q = i < n ? codes[i] : 1u;
        uint32_t tot;
        const uint32_t r = block_rank(q == 0u, ws, tot);
        if (i < n) out[i] = q ? dequantize<T>(q, eb) : unpred[carry + r];
        carry += tot;
    }
}

// ---------------------------------------------------------------- many fields

// the job whose range of one kind holds workgroup b: the last job with
// wg0 <= b (jobs with no workgroups of that kind share the next one's wg0)
template <typename F>
__device__ __forceinline__ uint32_t find_pro_job(const ProJob* __restrict__ jobs, uint32_t n, uint32_t b, F wg0)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (wg0(jobs[m]) <= b) lo = m; else hi = m;
    }
    return lo;
}

// k_pad per field, a workgroup per range of output rows
template <typename T>
__global__ __launch_bounds__(256) void k_pad_multi(const ProJob* __restrict__ jobs, uint32_t n)
{
    const ProJob& J = jobs[find_pro_job(jobs, n, blockIdx.x, [](const ProJob& q) { return q.pad_wg0; })];
    const uint32_t rows = (uint32_t)J.py * (uint32_t)J.pz, per = (rows + J.pad_nwg - 1) / J.pad_nwg;
    const uint32_t r0 = (blockIdx.x - J.pad_wg0) * per, r1 = min(rows, r0 + per);
    const T* in = (const T*)J.in;
    T* out = (T*)J.padded;
    for (uint32_t r = r0; r < r1; ++r) {
        const int z = (int)(r / (uint32_t)J.py), y = (int)(r % (uint32_t)J.py);
        const int sz = z < J.nz ? z : J.nz - 1, sy = y < J.ny ? y : J.ny - 1;
        const T* src = in + ((uint64_t)sz * J.ny + sy) * J.nx;
        T* dst = out + (uint64_t)r * J.px;
        for (int x = threadIdx.x; x < J.px; x += 256) dst[x] = src[x < J.nx ? x : J.nx - 1];
    }
}

// k_minmax per field: the job's workgroups stride over its values
template <typename T>
__global__ __launch_bounds__(256) void k_minmax_multi(const ProJob* __restrict__ jobs, uint32_t n)
{
    __shared__ double smn[256], smx[256];
    const ProJob& J = jobs[find_pro_job(jobs, n, blockIdx.x, [](const ProJob& q) { return q.mm_wg0; })];
    const uint32_t local = blockIdx.x - J.mm_wg0;
    const T* d = (const T*)J.in;
    double mn = __builtin_inf(), mx = -__builtin_inf();
    constexpr int V = 16 / sizeof(T);
    typedef T vec __attribute__((ext_vector_type(V)));
    const uint64_t nv = ((uintptr_t)d % 16 == 0) ? J.nin / V : 0;
    const vec* dv = (const vec*)d;
    const uint64_t stride = (uint64_t)J.mm_nwg * 256ull;
    for (uint64_t i = local * 256ull + threadIdx.x; i < nv; i += stride) {
        const vec v = dv[i];
#pragma unroll
        for (int k = 0; k < V; ++k) mm_upd<T>((double)v[k], mn, mx);
    }
    for (uint64_t i = nv * V + local * 256ull + threadIdx.x; i < J.nin; i += stride) mm_upd<T>((double)d[i], mn, mx);
    block_minmax<T>(mn, mx, smn, smx);
    if (threadIdx.x == 0) {
        J.partial[2 * local] = smn[0];
        J.partial[2 * local + 1] = smx[0];
    }
}

// one workgroup per field: its partials -> mm_out[2 j], mm_out[2 j + 1]
__global__ __launch_bounds__(256) void k_minmax_final_multi(const ProJob* __restrict__ jobs, double* __restrict__ out)
{
    __shared__ double smn[256], smx[256];
    const ProJob& J = jobs[blockIdx.x];
    double mn = __builtin_inf(), mx = -__builtin_inf();
    for (uint32_t b = threadIdx.x; b < J.mm_nwg; b += 256) {
        if (J.partial[2 * b] < mn) mn = J.partial[2 * b];
        if (J.partial[2 * b + 1] > mx) mx = J.partial[2 * b + 1];
    }
    block_minmax<double>(mn, mx, smn, smx);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = smn[0];
        out[2 * blockIdx.x + 1] = smx[0];
    }
}

// k_gather_low0 per field
template <typename T>
__global__ __launch_bounds__(256) void k_gather_multi(const ProJob* __restrict__ jobs, uint32_t n)
{
    const ProJob& J = jobs[find_pro_job(jobs, n, blockIdx.x, [](const ProJob& q) { return q.ga_wg0; })];
    const uint32_t lx = (uint32_t)J.px / 4, ly = (uint32_t)J.py / 4, lz = (uint32_t)J.pz / 4;
    const uint64_t nb = (uint64_t)lx * ly * lz;
    const T* full = (const T*)J.full;
    T* v = (T*)J.vbase;
    const uint64_t stride = (uint64_t)J.ga_nwg * 256ull;
    for (uint64_t i = (uint64_t)(blockIdx.x - J.ga_wg0) * 256ull + threadIdx.x; i < nb; i += stride) {
        const uint32_t x = (uint32_t)(i % lx), t = (uint32_t)(i / lx), y = t % ly, z = t / ly;
        v[i] = full[((uint64_t)(4 * z) * J.py + 4 * y) * J.px + 4 * x];
    }
}

inline unsigned grid_for(uint64_t n, unsigned cap = 16384)
{
    uint64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    return (unsigned)(b > cap ? cap : b);
}

}  // namespace

template <typename T>
void launch_minmax(hipStream_t s, const T* d, uint64_t n, double* partial, double* out)
{
    const unsigned nb = grid_for(n, 2048);
    k_minmax<T><<<dim3(nb), dim3(256), 0, s>>>(d, n, partial);
    k_minmax_final<<<dim3(1), dim3(256), 0, s>>>(partial, (int)nb, out);
}

template <typename T>
void launch_pad(hipStream_t s, const T* in, int nx, int ny, int nz, T* out, int px, int py, int pz)
{
    k_pad<T><<<dim3(grid_for((uint64_t)px * py * pz)), dim3(256), 0, s>>>(in, nx, ny, nz, out, px, py, pz);
}

template <typename T>
void launch_gather_low0(hipStream_t s, const T* full, int px, int py, int pz, T* v)
{
    const uint64_t n = (uint64_t)(px / 4) * (py / 4) * (pz / 4);
    k_gather_low0<T><<<dim3(grid_for(n)), dim3(256), 0, s>>>(full, px, py, pz, v);
}

template <typename T>
void launch_prologue_multi(hipStream_t s, const ProJob* jobs, uint32_t n, uint32_t pad_wgs, uint32_t mm_wgs,
                           uint32_t ga_wgs, double* mm_out)
{
    if (!n) return;
    if (pad_wgs) k_pad_multi<T><<<dim3(pad_wgs), dim3(256), 0, s>>>(jobs, n);
    if (mm_wgs) {
        k_minmax_multi<T><<<dim3(mm_wgs), dim3(256), 0, s>>>(jobs, n);
        k_minmax_final_multi<<<dim3(n), dim3(256), 0, s>>>(jobs, mm_out);
    }
    if (ga_wgs) k_gather_multi<T><<<dim3(ga_wgs), dim3(256), 0, s>>>(jobs, n);
}

template <typename T>
void launch_split(hipStream_t s, const T* full, int nx, int ny, int nz, T* const* sub)
{
    Ptr8 p;
    for (int k = 0; k < 8; ++k) p.p[k] = sub[k];
    k_split<T><<<dim3(grid_for((uint64_t)nx * ny * nz)), dim3(256), 0, s>>>(full, nx, ny, nz, p);
}

template <typename T>
void launch_merge(hipStream_t s, const T* const* sub, int sx, int sy, int sz, T* full)
{
    Ptr8 p;
    for (int k = 0; k < 8; ++k) p.p[k] = (void*)sub[k];
    k_merge<T><<<dim3(grid_for((uint64_t)8 * sx * sy * sz)), dim3(256), 0, s>>>(p, sx, sy, sz, full);
}

template <typename T>
void launch_residual_predict(hipStream_t s, int block, const T* sub, const T* ref, int dx, int dy, int dz,
                             T* out, int sign)
{
    const dim3 g(grid_for((uint64_t)dx * dy * dz)), b(256);
    switch (block) {
    case 0: k_residual<0, T><<<g, b, 0, s>>>(sub, ref, dx, dy, dz, out, sign); break;
    case 1: k_residual<1, T><<<g, b, 0, s>>>(sub, ref, dx, dy, dz, out, sign); break;
    case 2: k_residual<2, T><<<g, b, 0, s>>>(sub, ref, dx, dy, dz, out, sign); break;
    case 3: k_residual<3, T><<<g, b, 0, s>>>(sub, ref, dx, dy, dz, out, sign); break;
    case 4: k_residual<4, T><<<g, b, 0, s>>>(sub, ref, dx, dy, dz, out, sign); break;
    case 5: k_residual<5, T><<<g, b, 0, s>>>(sub, ref, dx, dy, dz, out, sign); break;
    default: k_residual<6, T><<<g, b, 0, s>>>(sub, ref, dx, dy, dz, out, sign); break;
    }
}

template <typename T>
void launch_quantize(hipStream_t s, const T* d, uint64_t n, double eb, uint16_t* codes)
{
    k_quantize<T><<<dim3(grid_for(n)), dim3(256), 0, s>>>(d, n, eb, 1.0 / eb, codes);
}

template <typename T>
void launch_dequantize(hipStream_t s, const uint16_t* codes, const T* side, uint64_t n, double eb, T* out)
{
    k_dequantize<T><<<dim3(grid_for(n)), dim3(256), 0, s>>>(codes, side, n, eb, out);
}

template <typename T>
void launch_crop(hipStream_t s, const T* in, int px, int py, int ox, int oy, int oz, int sx, int sy, int sz,
                 int step, T* out)
{
    k_crop<T><<<dim3(grid_for((uint64_t)ox * oy * oz)), dim3(256), 0, s>>>(in, px, py, ox, oy, oz, sx, sy, sz,
                                                                           step, out);
}

template <typename T>
void launch_compact_unpred(hipStream_t s, const uint16_t* codes, const T* data, uint64_t n, T* out, uint64_t* count)
{
    k_compact_unpred<T><<<dim3(1), dim3(1024), 0, s>>>(codes, data, n, out, count);
}

template <typename T>
void launch_dequant_list(hipStream_t s, const uint16_t* codes, const T* unpred, uint64_t n, double eb, T* out)
{
    k_dequant_list<T><<<dim3(1), dim3(1024), 0, s>>>(codes, unpred, n, eb, out);
}

#define STZ_INST(T)                                                                                    \
    template void launch_compact_unpred<T>(hipStream_t, const uint16_t*, const T*, uint64_t, T*, uint64_t*); \
    template void launch_dequant_list<T>(hipStream_t, const uint16_t*, const T*, uint64_t, double, T*); \
    template void launch_minmax<T>(hipStream_t, const T*, uint64_t, double*, double*);                \
    template void launch_pad<T>(hipStream_t, const T*, int, int, int, T*, int, int, int);             \
    template void launch_gather_low0<T>(hipStream_t, const T*, int, int, int, T*);                    \
    template void launch_prologue_multi<T>(hipStream_t, const ProJob*, uint32_t, uint32_t, uint32_t, uint32_t, \
                                           double*);                                                   \
    template void launch_split<T>(hipStream_t, const T*, int, int, int, T* const*);                   \
    template void launch_merge<T>(hipStream_t, const T* const*, int, int, int, T*);                   \
    template void launch_residual_predict<T>(hipStream_t, int, const T*, const T*, int, int, int, T*, \
                                             int);                                                     \
    template void launch_quantize<T>(hipStream_t, const T*, uint64_t, double, uint16_t*);             \
    template void launch_dequantize<T>(hipStream_t, const uint16_t*, const T*, uint64_t, double, T*); \
    template void launch_crop<T>(hipStream_t, const T*, int, int, int, int, int, int, int, int, int, T*);
STZ_INST(float)
STZ_INST(double)

}  // namespace stz

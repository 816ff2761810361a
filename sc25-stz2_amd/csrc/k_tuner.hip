// k_tuner.hip -- the learned tuner: per-sub-block field statistics and a
// 16 -> 64 -> 64 -> 208 MLP (ReLU) whose forward pass, backward pass and Adam
// update run here.  Matrix products are bf16 MFMA (v_mfma_f32_16x16x16_bf16)
// with fp32 accumulation; weights are kept in fp32 (master copy) and rounded
// to bf16 (round to nearest even) when a workgroup stages them in LDS.
//
// MFMA operand layout (wave64, checked on gfx950): lane l holds
//   A[l%16][4*(l/16) + i], B[4*(l/16) + i][l%16], D[4*(l/16) + i][l%16], i = 0..3.
// A wave owns 16 samples (rows); a workgroup is 4 waves = 64 samples.
#include "kernels.h"

namespace stz {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = kMlpIn, kHid = kMlpHid, kOut = kMlpOut;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kHid, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kHid * kOut;
static_assert(oB3 + kOut == kMlpParams, "parameter layout");

__device__ __forceinline__ uint16_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x4 ld4(const uint16_t* p)
{
    const uint2 v = *(const uint2*)p;  // 8-byte aligned LDS read
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFFu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFFu);
    r[3] = (short)(v.y >> 16);
    return r;
}
// four bf16 at a stride (transposed operand reads)
__device__ __forceinline__ bf16x4 ld4s(const uint16_t* p, int stride)
{
    bf16x4 r;
    r[0] = (short)p[0];
    r[1] = (short)p[stride];
    r[2] = (short)p[2 * stride];
    r[3] = (short)p[3 * stride];
    return r;
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// waves per MLP workgroup (16 sample rows each): two keep the weights and
// both waves' activations within the 160 KiB of LDS at 208 outputs
constexpr int kMlpWaves = 2;
constexpr int kMlpThreads = 64 * kMlpWaves;

struct MlpLds {
    uint16_t w1t[kHid * kIn];   // W1^T: [n][k]
    uint16_t w2t[kHid * kHid];  // W2^T
    uint16_t w2[kHid * kHid];   // W2: [k][n]
    uint16_t w3t[kOut * kHid];  // W3^T
    uint16_t w3[kHid * kOut];   // W3
    float b1[kHid], b2[kHid], b3[kOut];
    uint16_t x[kMlpWaves][16 * kIn];
    uint16_t h1[kMlpWaves][16 * kHid];
    uint16_t h2[kMlpWaves][16 * kHid];
    float y[kMlpWaves][16 * kOut];
    uint16_t dy[kMlpWaves][16 * kOut];
    uint16_t dh2[kMlpWaves][16 * kHid];
    uint16_t dh1[kMlpWaves][16 * kHid];
};

__device__ void stage_weights(MlpLds& S, const float* __restrict__ P)
{
    for (int e = threadIdx.x; e < kIn * kHid; e += blockDim.x) {
        const int k = e / kHid, n = e % kHid;
        S.w1t[n * kIn + k] = bf16_rne(P[oW1 + e]);
    }
    for (int e = threadIdx.x; e < kHid * kHid; e += blockDim.x) {
        const int k = e / kHid, n = e % kHid;
        const uint16_t h = bf16_rne(P[oW2 + e]);
        S.w2[e] = h;
        S.w2t[n * kHid + k] = h;
    }
    for (int e = threadIdx.x; e < kHid * kOut; e += blockDim.x) {
        const int k = e / kOut, n = e % kOut;
        const uint16_t h = bf16_rne(P[oW3 + e]);
        S.w3[e] = h;
        S.w3t[n * kHid + k] = h;
    }
    for (int e = threadIdx.x; e < kHid; e += blockDim.x) {
        S.b1[e] = P[oB1 + e];
        S.b2[e] = P[oB2 + e];
    }
    for (int e = threadIdx.x; e < kOut; e += blockDim.x) S.b3[e] = P[oB3 + e];
}

// rows r0 .. r0+15 of X (zero past n) -> logits in S.y[wave]
__device__ void forward16(MlpLds& S, const float* __restrict__ X, uint32_t r0, uint32_t n)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c16 = lane & 15, q = lane >> 4;
    for (int e = lane; e < 16 * kIn; e += 64) {
        const uint32_t r = r0 + (uint32_t)(e / kIn);
        S.x[wv][e] = r < n ? bf16_rne(X[(size_t)r * kIn + e % kIn]) : (uint16_t)0;
    }
    __syncthreads();
    {
        const bf16x4 a = ld4(&S.x[wv][c16 * kIn + 4 * q]);
        for (int t = 0; t < kHid / 16; ++t) {
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            d = mfma(a, ld4(&S.w1t[(16 * t + c16) * kIn + 4 * q]), d);
            for (int i = 0; i < 4; ++i) {
                const float h = d[i] + S.b1[16 * t + c16];
                S.h1[wv][(4 * q + i) * kHid + 16 * t + c16] = bf16_rne(h > 0.f ? h : 0.f);
            }
        }
    }
    __syncthreads();
    for (int t = 0; t < kHid / 16; ++t) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < kHid / 16; ++s)
            d = mfma(ld4(&S.h1[wv][c16 * kHid + 16 * s + 4 * q]), ld4(&S.w2t[(16 * t + c16) * kHid + 16 * s + 4 * q]), d);
        for (int i = 0; i < 4; ++i) {
            const float h = d[i] + S.b2[16 * t + c16];
            S.h2[wv][(4 * q + i) * kHid + 16 * t + c16] = bf16_rne(h > 0.f ? h : 0.f);
        }
    }
    __syncthreads();
    for (int t = 0; t < kOut / 16; ++t) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < kHid / 16; ++s)
            d = mfma(ld4(&S.h2[wv][c16 * kHid + 16 * s + 4 * q]), ld4(&S.w3t[(16 * t + c16) * kHid + 16 * s + 4 * q]), d);
        for (int i = 0; i < 4; ++i) S.y[wv][(4 * q + i) * kOut + 16 * t + c16] = d[i] + S.b3[16 * t + c16];
    }
    __syncthreads();
}

__global__ __launch_bounds__(kMlpThreads) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                     uint32_t n, float* __restrict__ Y)
{
    __shared__ MlpLds S;
    stage_weights(S, P);
    __syncthreads();
    const uint32_t r0 = blockIdx.x * (16u * kMlpWaves) + (threadIdx.x >> 6) * 16u;
    forward16(S, X, r0, n);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int e = lane; e < 16 * kOut; e += 64) {
        const uint32_t r = r0 + (uint32_t)(e / kOut);
        if (r < n) Y[(size_t)r * kOut + e % kOut] = S.y[wv][e];
    }
}

// Forward, soft-target cross entropy (mean over the n samples) and backward;
// gradients and loss are accumulated with fp32 atomics into G and *loss.
__global__ __launch_bounds__(kMlpThreads) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                  const float* __restrict__ T, uint32_t n, float* __restrict__ G,
                                                  float* __restrict__ loss)
{
    __shared__ MlpLds S;
    stage_weights(S, P);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c16 = lane & 15, q = lane >> 4;
    const uint32_t r0 = blockIdx.x * (16u * kMlpWaves) + (uint32_t)wv * 16u;
    forward16(S, X, r0, n);
    const float inv_n = 1.0f / (float)n;
    // softmax + loss + dY, one lane per row
    if (lane < 16) {
        const uint32_t r = r0 + (uint32_t)lane;
        const float* y = &S.y[wv][lane * kOut];
        if (r < n) {
            float mx = y[0];
            for (int c = 1; c < kOut; ++c) mx = fmaxf(mx, y[c]);
            float se = 0.f;
            for (int c = 0; c < kOut; ++c) se += __expf(y[c] - mx);
            const float lse = mx + __logf(se);
            float l = 0.f;
            for (int c = 0; c < kOut; ++c) {
                const float t = T[(size_t)r * kOut + c];
                const float p = __expf(y[c] - lse);
                l -= t * (y[c] - lse);
                S.dy[wv][lane * kOut + c] = bf16_rne((p - t) * inv_n);
            }
            atomicAdd(loss, l * inv_n);
        } else {
            for (int c = 0; c < kOut; ++c) S.dy[wv][lane * kOut + c] = 0;
        }
    }
    __syncthreads();
    // dW3 = H2^T dY, db3
    for (int mt = 0; mt < kHid / 16; ++mt)
        for (int nt = 0; nt < kOut / 16; ++nt) {
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            d = mfma(ld4s(&S.h2[wv][(4 * q) * kHid + 16 * mt + c16], kHid),
                     ld4s(&S.dy[wv][(4 * q) * kOut + 16 * nt + c16], kOut), d);
            for (int i = 0; i < 4; ++i) atomicAdd(&G[oW3 + (16 * mt + 4 * q + i) * kOut + 16 * nt + c16], d[i]);
        }
    for (int c = lane; c < kOut; c += 64) {
        float s = 0.f;
        for (int r = 0; r < 16; ++r) s += bf16_f(S.dy[wv][r * kOut + c]);
        atomicAdd(&G[oB3 + c], s);
    }
    // dH2 = dY W3^T, masked by ReLU
    for (int nt = 0; nt < kHid / 16; ++nt) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < kOut / 16; ++ks)
            d = mfma(ld4(&S.dy[wv][c16 * kOut + 16 * ks + 4 * q]), ld4(&S.w3[(16 * nt + c16) * kOut + 16 * ks + 4 * q]), d);
        for (int i = 0; i < 4; ++i) {
            const int e = (4 * q + i) * kHid + 16 * nt + c16;
            S.dh2[wv][e] = S.h2[wv][e] != 0 ? bf16_rne(d[i]) : (uint16_t)0;
        }
    }
    __syncthreads();
    // dW2 = H1^T dH2, db2
    for (int mt = 0; mt < kHid / 16; ++mt)
        for (int nt = 0; nt < kHid / 16; ++nt) {
            f32x4 d = {0.f, 0.f, 0.f, 0.f};
            d = mfma(ld4s(&S.h1[wv][(4 * q) * kHid + 16 * mt + c16], kHid),
                     ld4s(&S.dh2[wv][(4 * q) * kHid + 16 * nt + c16], kHid), d);
            for (int i = 0; i < 4; ++i) atomicAdd(&G[oW2 + (16 * mt + 4 * q + i) * kHid + 16 * nt + c16], d[i]);
        }
    {
        float s = 0.f;
        for (int r = 0; r < 16; ++r) s += bf16_f(S.dh2[wv][r * kHid + lane]);
        atomicAdd(&G[oB2 + lane], s);
    }
    // dH1 = dH2 W2^T, masked by ReLU
    for (int nt = 0; nt < kHid / 16; ++nt) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < kHid / 16; ++ks)
            d = mfma(ld4(&S.dh2[wv][c16 * kHid + 16 * ks + 4 * q]), ld4(&S.w2[(16 * nt + c16) * kHid + 16 * ks + 4 * q]), d);
        for (int i = 0; i < 4; ++i) {
            const int e = (4 * q + i) * kHid + 16 * nt + c16;
            S.dh1[wv][e] = S.h1[wv][e] != 0 ? bf16_rne(d[i]) : (uint16_t)0;
        }
    }
    __syncthreads();
    // dW1 = X^T dH1, db1
    for (int nt = 0; nt < kHid / 16; ++nt) {
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = mfma(ld4s(&S.x[wv][(4 * q) * kIn + c16], kIn), ld4s(&S.dh1[wv][(4 * q) * kHid + 16 * nt + c16], kHid), d);
        for (int i = 0; i < 4; ++i) atomicAdd(&G[oW1 + (4 * q + i) * kHid + 16 * nt + c16], d[i]);
    }
    {
        float s = 0.f;
        for (int r = 0; r < 16; ++r) s += bf16_f(S.dh1[wv][r * kHid + lane]);
        atomicAdd(&G[oB1 + lane], s);
    }
}

// Adam (bias-corrected) on the fp32 master weights; clears the gradient
__global__ void k_adam(float* __restrict__ P, float* __restrict__ G, float* __restrict__ M, float* __restrict__ V,
                       uint32_t n, float lr, float b1, float b2, float eps, float c1, float c2, float wd)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    G[i] = 0.f;
    const float m = b1 * M[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    const float mh = m / c1, vh = v / c2;
    P[i] -= lr * (mh / (sqrtf(vh) + eps) + wd * P[i]);
}

// ---------------------------------------------------------------- statistics
// One workgroup per 64^3 sub-block; samples on a stride-2 lattice.  Sums per
// block (kFeatSums doubles): n, sum v, sum v^2, sum v^4, min, max,
// sum|lin4_a| (3), sum|cub4_a| (3), sum|lin1_a| (3), #(max|lin1| <= eb),
// #(max|cub4| <= eb), sum|d1|.  lin4/cub4: residual of the stride-4 linear /
// cubic (-1, 9, 9, -1)/16 interpolation along axis a (the base codec's
// finest level); lin1: stride-1 linear residual (the high level); d1: first
// difference.  Points whose stencil leaves the field are skipped.
// Each 64^3 block is split into kStatParts z-slabs of its sample lattice,
// one 256-lane workgroup each (about 80 VGPRs: six workgroups per CU), whose
// partial sums k_block_stats_combine adds up.

template <typename T>
__global__ __launch_bounds__(256) void k_block_stats(const T* __restrict__ f, int nx, int ny, int nz, int bx,
                                                     int by, double eb, double* __restrict__ part)
{
    __shared__ double red[4][kFeatSums];
    const int b = blockIdx.x / kStatParts, q = blockIdx.x % kStatParts;
    const int cx = b % bx, cy = (b / bx) % by, cz = b / (bx * by);
    const int x0 = cx * 64, y0 = cy * 64, z0 = cz * 64;
    const int x1 = min(x0 + 64, nx), y1 = min(y0 + 64, ny), z1 = min(z0 + 64, nz);
    const int sx = (x1 - x0 + 1) / 2, sy = (y1 - y0 + 1) / 2, sz = (z1 - z0 + 1) / 2;
    const int ns = sx * sy * sz;
    // this part's samples: lattice planes [q sz / P, (q + 1) sz / P)
    const int k0 = sx * sy * (q * sz / kStatParts), k1 = sx * sy * ((q + 1) * sz / kStatParts);
    double acc[kFeatSums];
    for (int k = 0; k < kFeatSums; ++k) acc[k] = 0.0;
    acc[4] = 1e300;
    acc[5] = -1e300;
    const size_t st[3] = {1, (size_t)nx, (size_t)nx * ny};
    const int dim[3] = {nx, ny, nz};
    for (int k = k0 + threadIdx.x; k < k1 && k < ns; k += 256) {
        const int x = x0 + 2 * (k % sx), y = y0 + 2 * ((k / sx) % sy), z = z0 + 2 * (k / (sx * sy));
        const int c[3] = {x, y, z};
        const size_t i = ((size_t)z * ny + y) * nx + x;
        const double v = (double)f[i];
        acc[4] = fmin(acc[4], v);
        acc[5] = fmax(acc[5], v);
        bool ok = true;
        for (int a = 0; a < 3; ++a) ok = ok && c[a] >= 12 && c[a] + 12 < dim[a];
        if (!ok) continue;
        acc[0] += 1.0;
        acc[1] += v;
        acc[2] += v * v;
        acc[3] += v * v * v * v;
        double ml1 = 0, mc4 = 0, d1 = 0;
        for (int a = 0; a < 3; ++a) {
            const size_t s = st[a];
            const double m4 = (double)f[i - 4 * s], p4 = (double)f[i + 4 * s];
            const double m12 = (double)f[i - 12 * s], p12 = (double)f[i + 12 * s];
            const double l4 = fabs(v - 0.5 * (m4 + p4));
            const double c4 = fabs(v - (-0.0625 * m12 + 0.5625 * m4 + 0.5625 * p4 - 0.0625 * p12));
            const double l1 = fabs(v - 0.5 * ((double)f[i - s] + (double)f[i + s]));
            acc[6 + a] += l4;
            acc[9 + a] += c4;
            acc[12 + a] += l1;
            ml1 = fmax(ml1, l1);
            mc4 = fmax(mc4, c4);
            d1 += fabs((double)f[i + s] - v);
        }
        acc[15] += ml1 <= eb ? 1.0 : 0.0;
        acc[16] += mc4 <= eb ? 1.0 : 0.0;
        acc[17] += d1 / 3.0;
    }
    // a shuffle reduction per wave, then one pass over the four waves
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kFeatSums; ++k) {
        double v = acc[k];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const double o = __shfl_xor(v, d, 64);
            v = k == 4 ? fmin(v, o) : k == 5 ? fmax(v, o) : v + o;
        }
        if (lane == 0) red[wv][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kFeatSums) {
        const int k = threadIdx.x;
        double r = red[0][k];
        for (int w = 1; w < 4; ++w) {
            const double o = red[w][k];
            r = k == 4 ? fmin(r, o) : k == 5 ? fmax(r, o) : r + o;
        }
        part[(size_t)blockIdx.x * kFeatSums + k] = r;
    }
}

__global__ void k_block_stats_combine(const double* __restrict__ part, uint32_t nb, double* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * kFeatSums) return;
    const uint32_t b = i / kFeatSums, k = i % kFeatSums;
    double r = part[((size_t)b * kStatParts) * kFeatSums + k];
    for (int q = 1; q < kStatParts; ++q) {
        const double o = part[((size_t)b * kStatParts + q) * kFeatSums + k];
        r = k == 4 ? fmin(r, o) : k == 5 ? fmax(r, o) : r + o;
    }
    out[i] = r;
}

}  // namespace

void launch_mlp_forward(hipStream_t s, const float* P, const float* X, uint32_t n, float* Y)
{
    if (!n) return;
    k_mlp_forward<<<dim3((n + 16 * kMlpWaves - 1) / (16 * kMlpWaves)), dim3(kMlpThreads), 0, s>>>(P, X, n, Y);
}

void launch_mlp_grad(hipStream_t s, const float* P, const float* X, const float* T, uint32_t n, float* G, float* loss)
{
    if (!n) return;
    k_mlp_grad<<<dim3((n + 16 * kMlpWaves - 1) / (16 * kMlpWaves)), dim3(kMlpThreads), 0, s>>>(P, X, T, n, G, loss);
}

void launch_adam(hipStream_t s, float* P, float* G, float* M, float* V, uint32_t n, float lr, float b1, float b2,
                 float eps, int step, float wd)
{
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    k_adam<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(P, G, M, V, n, lr, b1, b2, eps, c1, c2, wd);
}

template <typename T>
void launch_block_stats(hipStream_t s, const T* f, int nx, int ny, int nz, double eb, double* out,
                        double* part)
{
    const int bx = (nx + 63) / 64, by = (ny + 63) / 64, bz = (nz + 63) / 64;
    const uint32_t nb = (uint32_t)(bx * by * bz);
    k_block_stats<T><<<dim3(nb * kStatParts), dim3(256), 0, s>>>(f, nx, ny, nz, bx, by, eb, part);
    k_block_stats_combine<<<dim3((nb * kFeatSums + 255) / 256), dim3(256), 0, s>>>(part, nb, out);
}
template void launch_block_stats<float>(hipStream_t, const float*, int, int, int, double, double*, double*);
template void launch_block_stats<double>(hipStream_t, const double*, int, int, int, double, double*, double*);

}  // namespace stz

// k_staged.hip -- the residual levels under dependency.cpp's staged
// inverse-variance predictors (option bit kPredStaged).
//
// dependency.cpp:594-636 (low level) and 681-704 (high level) predict the
// blocks in three stages: blocks 0, 1, 3 (one offset axis) from the coarser
// grid by F5's single-axis rules, then blocks 2, 4, 5 from the reconstructed
// blocks 0, 1, 3, each interpolated along its missing axis and blended by
// inverse variance, then block 6 from the reconstructed 2, 4, 5
// (dependency.cpp:162-292).  A stage reads reconstructions of the stage
// before at other points, so each stage is a launch of its own:
//   stage 1 writes blocks 0, 1, 3, stage 2 blocks 2, 4, 5, each into a dense
//   block-grid array of the scratch `rec` (6 arrays of g.dx x g.dy x g.dz,
//   block b at rec + b n), so the next stage's stencils read whole rows;
//   stage 3 computes block 6 and writes every 2x2x2 cell of the level's 2x
//   grid (coarser grid at parity 0, block b at the parity of sub-block b + 1)
//   as 2-wide row stores, cropped to the output's dims; the high level's
//   encode writes no cells (nothing reads them).
// Work items are 64-point rows of one plane, one per wave (y and z
// wave-uniform, so the stencils along y and z read whole rows and take their
// edge rules as uniform branches; along x each lane selects its rule).
// Encode quantizes as it predicts (F5's quantizer at the level's bound) and
// writes codes, unpredictable residuals and histograms like the other level
// kernels; decode reads them back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

#include "device_math.h"
#include "kernels.h"

namespace stz {

namespace {

constexpr int kSHW = 1024;  // LDS histogram window per block (the codes around the centre)
constexpr uint32_t kSLo = (uint32_t)kRadius - kSHW / 2;
constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;

template <typename T>
struct StagedK {
    const T* full;  // encode: the padded field (targets)
    int px, py;
    int tstep;      // target stride in the field: 2 (high level) or 4 (low level)
    const T* ref;   // the coarser grid, dims g.dx, g.dy, g.dz
    T* rec;         // blocks 0..5 reconstructed, dense, block b at rec + b n
    Geom g;         // block grid and stream chunking
    LevelArgs a;
    uint16_t* codes;  // encode: written; decode: read
    T* side;
    uint32_t* hist;   // encode: 7 x 65536 bins
    T* out;           // stage 3: the level's 2x grid (row strides ox, ox * oy), or null
    int ox, oy, oz;   // its dims; points at or past them are not written
    int vec;          // 2-wide stores allowed (ox even, out 2-element aligned)
};

// Histogram update by the active lanes of the wave: the centre code (zero
// residual, the most frequent) with one ballot and one atomic, the others
// per lane; codes outside the LDS window go straight to the global bins.
__device__ __forceinline__ void hist_add(uint32_t* sh, uint32_t* gh, uint32_t q)
{
    const bool centre = q == (uint32_t)kRadius;
    const unsigned long long m = __ballot(centre);
    const unsigned long long act = __ballot(1);
    if (m && __lane_id() == (unsigned)(__ffsll((long long)act) - 1)) lds_add(&sh[kRadius - kSLo], (uint32_t)__popcll(m));
    if (!centre) {
        if (q - kSLo < (uint32_t)kSHW) lds_add(&sh[q - kSLo], 1u);
        else glb_add(&gh[q], 1u);
    }
}

// F5's single-axis rule (predict<0/1/3>, F5:690-748; dependency.cpp's
// interp_axis, dependency.cpp:90-118, evaluates the same) at coordinate c of
// n along the axis from the values at c-1 .. c+2 (those past the grid are
// clamped copies and never chosen)
template <typename T>
__device__ __forceinline__ T rule_f5(int c, int n, T a, T b, T cc, T d)
{
    if (c == n - 1) return b;
    if (c == n - 2 || c == 0) return T(0.5f) * (b + cc);
    return cub(a, b, cc, d);
}

// dependency.cpp:162-239 interp_axis_with_variance at coordinate c of n from
// the reconstructed values at c-1 .. c+2 (clamped as above); |b - a| squared
// is (b - a) squared bit for bit
template <typename T>
__device__ __forceinline__ void rule_var(int c, int n, T a, T b, T cc, T d, T& p, T& v)
{
    const T eps = T(1e-12f);
    if (c == n - 1) {
        p = b;
        v = eps;
    } else if (c == n - 2 || c == 0) {
        const T e = cc - b;
        p = T(0.5f) * (b + cc);
        v = e * e + eps;
    } else {
        const T c0 = a - T(2.0f) * b + cc;
        const T c1 = b - T(2.0f) * cc + d;
        p = T(-0.0625f) * a + T(0.5625f) * b + T(0.5625f) * cc + T(-0.0625f) * d;
        v = T(0.5f) * (c0 * c0 + c1 * c1) + eps;
    }
}

// dependency.cpp:241-248 blend2 (blocks 2, 4, 5) and 250-256 blend3 (block 6)
template <typename T>
__device__ __forceinline__ T blend2s(T p0, T v0, T p1, T v1)
{
    return (p0 * v1 + p1 * v0) / (v0 + v1);
}

template <typename T>
__device__ __forceinline__ T blend3s(T p0, T v0, T p1, T v1, T p2, T v2)
{
    const T w0 = v1 * v2, w1 = v0 * v2, w2 = v0 * v1;
    return (w0 * p0 + w1 * p1 + w2 * p2) / (w0 + w1 + w2);
}

// One point: wave-uniform (y, z), the lane's x, the dense index
// i = (z dy + y) dx + x and the clamped neighbour offsets along each axis
// (steps -1, +1, +2; the rules never pick a clamped one)
struct Pt {
    int x, y, z;
    size_t i;
    int xo[3];        // along x
    ptrdiff_t yo[3];  // along y (elements)
    ptrdiff_t zo[3];  // along z (elements)
};

__device__ __forceinline__ Pt make_pt(const Geom& g, int x, int y, int z)
{
    Pt p;
    p.x = x;
    p.y = y;
    p.z = z;
    const ptrdiff_t sx = g.dx, sxy = (ptrdiff_t)g.dx * g.dy;
    p.i = ((size_t)z * g.dy + y) * g.dx + x;
    const int ks[3] = {-1, 1, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        p.xo[k] = min(max(x + ks[k], 0), g.dx - 1) - x;
        p.yo[k] = (ptrdiff_t)(min(max(y + ks[k], 0), g.dy - 1) - y) * sx;
        p.zo[k] = (ptrdiff_t)(min(max(z + ks[k], 0), g.dz - 1) - z) * sxy;
    }
    return p;
}

// the values at c-1, c, c+1, c+2 along axis A (0 x, 1 y, 2 z) of array r
template <int A, typename T>
__device__ __forceinline__ void taps(const T* __restrict__ r, const Pt& p, T& a, T& b, T& c, T& d)
{
    const T* q = r + p.i;
    if constexpr (A == 0) {
        a = q[p.xo[0]]; b = q[0]; c = q[p.xo[1]]; d = q[p.xo[2]];
    } else if constexpr (A == 1) {
        a = q[p.yo[0]]; b = q[0]; c = q[p.yo[1]]; d = q[p.yo[2]];
    } else {
        a = q[p.zo[0]]; b = q[0]; c = q[p.zo[1]]; d = q[p.zo[2]];
    }
}

template <int A>
__device__ __forceinline__ int coord(const Pt& p)
{
    return A == 0 ? p.x : A == 1 ? p.y : p.z;
}

template <int A>
__device__ __forceinline__ int extent(const Geom& g)
{
    return A == 0 ? g.dx : A == 1 ? g.dy : g.dz;
}

// F5's rule along axis A of the coarser grid
template <int A, typename T>
__device__ __forceinline__ T pred_f5(const T* __restrict__ r, const Pt& p, const Geom& g)
{
    T a, b, c, d;
    taps<A>(r, p, a, b, c, d);
    return rule_f5<T>(coord<A>(p), extent<A>(g), a, b, c, d);
}

// the variance rule along axis A of reconstructed block array r
template <int A, typename T>
__device__ __forceinline__ void pred_var(const T* __restrict__ r, const Pt& p, const Geom& g, T& pv, T& vv)
{
    T a, b, c, d;
    taps<A>(r, p, a, b, c, d);
    rule_var<T>(coord<A>(p), extent<A>(g), a, b, c, d, pv, vv);
}

// the input of block b (parity qx, qy, qz) at p: its target (encode) or
// its code (decode)
template <typename T, bool kEnc>
using In = typename std::conditional<kEnc, T, uint32_t>::type;

template <typename T, bool kEnc>
__device__ __forceinline__ In<T, kEnc> load_in(const StagedK<T>& K, int b, int qx, int qy, int qz, const Pt& p)
{
    if constexpr (kEnc) {
        const int ts = K.tstep, hs = K.tstep / 2;
        return K.full[((size_t)(ts * p.z + hs * qz) * K.py + (size_t)(ts * p.y + hs * qy)) * K.px +
                      (size_t)(ts * p.x + hs * qx)];
    } else {
        return K.codes[K.a.code_off[b] + stream_pos(K.g, p.x, p.y, p.z)];
    }
}

// block b at p from its prediction and input: encode quantizes the target
// against the prediction and writes the code, decode dequantizes the code;
// returns the reconstruction
template <typename T, bool kEnc>
__device__ __forceinline__ T recon(const StagedK<T>& K, int b, T pr, In<T, kEnc> v, const Pt& p, uint32_t* shb)
{
    if constexpr (kEnc) {
        const uint64_t e = K.a.code_off[b] + stream_pos(K.g, p.x, p.y, p.z);
        const T d = v - pr;
        T dec;
        const uint32_t q = quantize(d, K.a.eb, K.a.recip, dec);
        K.codes[e] = (uint16_t)q;
        if (!q) K.side[e] = d;
        hist_add(shb, K.hist + (size_t)b * 65536, q);
        return dec + pr;  // dependency.cpp:600-614: the decoded residual + the prediction
    } else {
        (void)shb;
        if (v) return dequantize<T>(v, K.a.eb) + pr;
        return K.side[K.a.code_off[b] + stream_pos(K.g, p.x, p.y, p.z)] + pr;
    }
}

// the four taps of a stencil, loaded ahead of their use
template <typename T>
struct Taps {
    T a, b, c, d;
};

template <int A, typename T>
__device__ __forceinline__ Taps<T> load_taps(const T* __restrict__ r, const Pt& p)
{
    Taps<T> t;
    taps<A>(r, p, t.a, t.b, t.c, t.d);
    return t;
}

template <int A, typename T>
__device__ __forceinline__ T use_f5(const Taps<T>& t, const Pt& p, const Geom& g)
{
    return rule_f5<T>(coord<A>(p), extent<A>(g), t.a, t.b, t.c, t.d);
}

template <int A, typename T>
__device__ __forceinline__ void use_var(const Taps<T>& t, const Pt& p, const Geom& g, T& pv, T& vv)
{
    rule_var<T>(coord<A>(p), extent<A>(g), t.a, t.b, t.c, t.d, pv, vv);
}

// items: kWaves x kRows rows of 64 points of one plane
constexpr int kRows = 2;  // rows per wave: their loads are all issued before any is used

__host__ __device__ inline uint64_t row_items(const Geom& g, int& ntx, int& nty)
{
    ntx = (g.dx + 63) / 64;
    nty = (g.dy + kWaves * kRows - 1) / (kWaves * kRows);
    return (uint64_t)ntx * nty * g.dz;
}

// index of the calling wave in its workgroup, as a wave-uniform value
__device__ __forceinline__ int wave_id()
{
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

template <typename T, bool kEnc, int kStage>
__global__ __launch_bounds__(kThreads) void k_staged(const StagedK<T> K)
{
    constexpr int kNB = kStage == 2 ? 1 : 3;
    __shared__ uint32_t sh[kEnc ? kNB : 1][kEnc ? kSHW : 1];
    if constexpr (kEnc) {
        for (int k = threadIdx.x; k < kNB * kSHW; k += kThreads) sh[k / kSHW][k % kSHW] = 0u;
        __syncthreads();
    }
    const Geom& g = K.g;
    const size_t n = (size_t)g.dx * g.dy * g.dz;
    int ntx, nty;
    const uint64_t total = row_items(g, ntx, nty);
    const int lane = threadIdx.x & 63, w = wave_id();
    // each XCD (workgroups dispatched round-robin over 8) sweeps one
    // contiguous range of items, so the rows a stencil shares with the next
    // items stay in that XCD's L2
    const uint32_t G = gridDim.x;
    const bool sweep = (G % 8u) == 0;
    const uint64_t per = sweep ? (total + 7) / 8 : total;
    const uint64_t base = sweep ? (uint64_t)(blockIdx.x % 8u) * per : 0;
    const uint64_t first = sweep ? blockIdx.x / 8u : blockIdx.x, step = sweep ? G / 8u : G;
    T* const R = K.rec;
    uint32_t* const sh0 = sh[0];
    uint32_t* const sh1 = sh[kEnc && kNB > 1 ? 1 : 0];
    uint32_t* const sh2 = sh[kEnc && kNB > 2 ? 2 : 0];
    Walk3 it;
    it.init(base + first, step, (uint32_t)ntx, (uint32_t)nty);
    for (uint64_t j = first; j < per; j += step, it.next()) {
        const uint64_t t = base + j;
        if (t >= total) break;
        const int z = (int)it.c;
        const int x = 64 * (int)it.a + lane, y0 = (kWaves * (int)it.b + w) * kRows;
        if (y0 >= g.dy) continue;  // wave-uniform
        const int nr = y0 + kRows <= g.dy ? kRows : g.dy - y0;
        if (x >= g.dx) continue;
        Pt P[kRows];
#pragma unroll
        for (int r = 0; r < kRows; ++r) P[r] = make_pt(g, x, r < nr ? y0 + r : y0, z);
        if constexpr (kStage == 0) {
            // blocks 0 (x), 1 (y), 3 (z) from the coarser grid
            Taps<T> t0[kRows], t1[kRows], t3[kRows];
            In<T, kEnc> v0[kRows], v1[kRows], v3[kRows];
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                t0[r] = load_taps<0>(K.ref, P[r]);
                t1[r] = load_taps<1>(K.ref, P[r]);
                t3[r] = load_taps<2>(K.ref, P[r]);
                v0[r] = load_in<T, kEnc>(K, 0, 1, 0, 0, P[r]);
                v1[r] = load_in<T, kEnc>(K, 1, 0, 1, 0, P[r]);
                v3[r] = load_in<T, kEnc>(K, 3, 0, 0, 1, P[r]);
            }
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                if (r >= nr) break;  // wave-uniform
                const Pt& p = P[r];
                R[0 * n + p.i] = recon<T, kEnc>(K, 0, use_f5<0>(t0[r], p, g), v0[r], p, sh0);
                R[1 * n + p.i] = recon<T, kEnc>(K, 1, use_f5<1>(t1[r], p, g), v1[r], p, sh1);
                R[3 * n + p.i] = recon<T, kEnc>(K, 3, use_f5<2>(t3[r], p, g), v3[r], p, sh2);
            }
        } else if constexpr (kStage == 1) {
            // block 2 (xy): block 0 along y, block 1 along x; block 4 (xz):
            // block 0 along z, block 3 along x; block 5 (yz): block 1 along
            // z, block 3 along y
            Taps<T> a2[kRows], b2[kRows], a4[kRows], b4[kRows], a5[kRows], b5[kRows];
            In<T, kEnc> v2[kRows], v4[kRows], v5[kRows];
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                a2[r] = load_taps<1>(R + 0 * n, P[r]);
                b2[r] = load_taps<0>(R + 1 * n, P[r]);
                a4[r] = load_taps<2>(R + 0 * n, P[r]);
                b4[r] = load_taps<0>(R + 3 * n, P[r]);
                a5[r] = load_taps<2>(R + 1 * n, P[r]);
                b5[r] = load_taps<1>(R + 3 * n, P[r]);
                v2[r] = load_in<T, kEnc>(K, 2, 1, 1, 0, P[r]);
                v4[r] = load_in<T, kEnc>(K, 4, 1, 0, 1, P[r]);
                v5[r] = load_in<T, kEnc>(K, 5, 0, 1, 1, P[r]);
            }
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                if (r >= nr) break;  // wave-uniform
                const Pt& p = P[r];
                T p0, q0, p1, q1;
                use_var<1>(a2[r], p, g, p0, q0);
                use_var<0>(b2[r], p, g, p1, q1);
                R[2 * n + p.i] = recon<T, kEnc>(K, 2, blend2s(p0, q0, p1, q1), v2[r], p, sh0);
                use_var<2>(a4[r], p, g, p0, q0);
                use_var<0>(b4[r], p, g, p1, q1);
                R[4 * n + p.i] = recon<T, kEnc>(K, 4, blend2s(p0, q0, p1, q1), v4[r], p, sh1);
                use_var<2>(a5[r], p, g, p0, q0);
                use_var<1>(b5[r], p, g, p1, q1);
                R[5 * n + p.i] = recon<T, kEnc>(K, 5, blend2s(p0, q0, p1, q1), v5[r], p, sh2);
            }
        } else {
            // block 6 (xyz): block 2 along z, block 4 along y, block 5 along x;
            // then the point's 2x2x2 cell
            Taps<T> a6[kRows], b6[kRows], c6[kRows];
            In<T, kEnc> v6[kRows];
            T cr[kRows], c0[kRows], c1[kRows], c3[kRows];
            const bool cells = K.out != nullptr;
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                a6[r] = load_taps<2>(R + 2 * n, P[r]);
                b6[r] = load_taps<1>(R + 4 * n, P[r]);
                c6[r] = load_taps<0>(R + 5 * n, P[r]);
                v6[r] = load_in<T, kEnc>(K, 6, 1, 1, 1, P[r]);
                if (cells) {
                    cr[r] = K.ref[P[r].i];
                    c0[r] = R[P[r].i];
                    c1[r] = R[n + P[r].i];
                    c3[r] = R[3 * n + P[r].i];
                }
            }
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                if (r >= nr) break;  // wave-uniform
                const Pt& p = P[r];
                T p0, q0, p1, q1, p2, q2;
                use_var<2>(a6[r], p, g, p0, q0);
                use_var<1>(b6[r], p, g, p1, q1);
                use_var<0>(c6[r], p, g, p2, q2);
                const T r6 = recon<T, kEnc>(K, 6, blend3s(p0, q0, p1, q1, p2, q2), v6[r], p, sh0);
                if (!cells) continue;
                // rows (qy, qz) of (x parity 0, x parity 1): blocks at
                // parities 0, (1,0,0) | (0,1,0), (1,1,0) | (0,0,1), (1,0,1) |
                // (0,1,1), (1,1,1); the taps' centres are blocks 2, 4, 5 here
                const int X = 2 * x, Y = 2 * p.y, Z = 2 * z;
                if (X >= K.ox || Y >= K.oy || Z >= K.oz) continue;
                const T v[8] = {cr[r], c0[r], c1[r], a6[r].b, c3[r], b6[r].b, c6[r].b, r6};
                typedef typename std::conditional<sizeof(T) == 4, float2, double2>::type P2;
                const size_t sy = (size_t)K.ox, sz = (size_t)K.ox * K.oy;
                T* o = K.out + (size_t)Z * sz + (size_t)Y * sy + X;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int qy = c & 1, qz = c >> 1;
                    if (Y + qy >= K.oy || Z + qz >= K.oz) continue;
                    T* op = o + (size_t)qz * sz + (size_t)qy * sy;
                    if (K.vec) {
                        P2 w2;
                        w2.x = v[2 * c];
                        w2.y = v[2 * c + 1];
                        *(P2*)op = w2;
                    } else {
                        op[0] = v[2 * c];
                        if (X + 1 < K.ox) op[1] = v[2 * c + 1];
                    }
                }
            }
        }
    }
    if constexpr (kEnc) {
        __syncthreads();
        for (int k = threadIdx.x; k < kNB * kSHW; k += kThreads) {
            const uint32_t v = sh[k / kSHW][k % kSHW];
            const int j = k / kSHW;  // the stage's j-th block: {0, 1, 3}, {2, 4, 5}, {6}
            const int blk = kStage == 0 ? (j == 2 ? 3 : j) : kStage == 1 ? (j == 0 ? 2 : j + 3) : 6;
            if (v) glb_add(&K.hist[(size_t)blk * 65536 + kSLo + (k % kSHW)], v);
        }
    }
}

// workgroups resident on the device at once for a kernel
unsigned resident(const void* kernel)
{
    static std::mutex mu;
    static std::map<const void*, unsigned> slots;
    std::lock_guard<std::mutex> lk(mu);
    auto it = slots.find(kernel);
    if (it == slots.end()) {
        int dev = 0, ncu = 0, occ = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) ncu = 1;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kThreads, 0) != hipSuccess || occ < 1) occ = 1;
        it = slots.emplace(kernel, (unsigned)(ncu * occ)).first;
    }
    return it->second;
}

// a persistent grid of resident workgroups, at most one per item, a multiple
// of 8 for the XCD sweep
template <typename T, bool kEnc, int kStage>
void launch_stage(hipStream_t s, const StagedK<T>& K)
{
    int ntx, nty;
    const uint64_t items = row_items(K.g, ntx, nty);
    if (!items) return;
    const unsigned slots = resident((const void*)k_staged<T, kEnc, kStage>);
    uint64_t G = items < slots ? items : slots;
    if (G >= 64) G &= ~7ull;
    k_staged<T, kEnc, kStage><<<dim3((unsigned)G), dim3(kThreads), 0, s>>>(K);
}

template <typename T, bool kEnc>
void launch_stages(hipStream_t s, const StagedK<T>& K)
{
    if (K.g.dx <= 0 || K.g.dy <= 0 || K.g.dz <= 0) return;
    launch_stage<T, kEnc, 0>(s, K);
    launch_stage<T, kEnc, 1>(s, K);
    launch_stage<T, kEnc, 2>(s, K);
}

}  // namespace

template <typename T>
void launch_staged_encode(hipStream_t s, const T* full, int px, int py, int tstep, const T* ref, Geom g, LevelArgs a,
                          uint16_t* codes, T* side, uint32_t* hist, T* out, T* rec)
{
    StagedK<T> K{};
    K.full = full;
    K.px = px;
    K.py = py;
    K.tstep = tstep;
    K.ref = ref;
    K.rec = rec;
    K.g = g;
    K.a = a;
    K.codes = codes;
    K.side = side;
    K.hist = hist;
    K.out = out;
    K.ox = 2 * g.dx;
    K.oy = 2 * g.dy;
    K.oz = 2 * g.dz;
    K.vec = out && ((uintptr_t)out % (2 * sizeof(T))) == 0;
    launch_stages<T, true>(s, K);
}

template <typename T>
void launch_staged_decode(hipStream_t s, const T* ref, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                          T* out, int nx, int ny, int nz, T* rec)
{
    StagedK<T> K{};
    K.ref = ref;
    K.rec = rec;
    K.g = g;
    K.a = a;
    K.codes = const_cast<uint16_t*>(codes);
    K.side = const_cast<T*>(side);
    K.out = out;
    K.ox = nx;
    K.oy = ny;
    K.oz = nz;
    K.vec = (nx % 2) == 0 && ((uintptr_t)out % (2 * sizeof(T))) == 0;
    launch_stages<T, false>(s, K);
}

#define STZ_INST_STAGED(T)                                                                                     \
    template void launch_staged_encode<T>(hipStream_t, const T*, int, int, int, const T*, Geom, LevelArgs,      \
                                          uint16_t*, T*, uint32_t*, T*, T*);                                    \
    template void launch_staged_decode<T>(hipStream_t, const T*, Geom, LevelArgs, const uint16_t*, const T*, T*, \
                                          int, int, int, T*);
STZ_INST_STAGED(float)
STZ_INST_STAGED(double)

}  // namespace stz

// layout.h -- payload layout and block/chunk geometry of the STZ device path.
//
// Payload = file header (512 B) + 15 streams (stream 0: base, 1..7: low
// blocks 0..6, 8..14: high blocks 0..6).  A residual stream holds the codes of
// one parity block in chunk-major order (row-major chunk ids, F5:1577-1582;
// chunk-local row-major inside a chunk, the order F5:1884-1905 reads them
// back), cut into sub-chunks of 1024 symbols that decode in parallel.  The
// per-chunk bit offsets are the reference's "markers" (F5:50-55); they also
// drive random access.  Unpredictable values are kept per chunk, in stream
// order, so a marked-chunk decode can place them (the reference keeps them in
// global row-major order, F5:1201-1203, which a chunk decode cannot consume).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stz {

constexpr uint32_t kMagic = 0x325A5453u;        // "STZ2"
constexpr uint32_t kStreamMagic = 0x4D525453u;  // "STRM"
constexpr uint32_t kVersion = 1;
constexpr int kRadius = 32768;      // quantbinCnt/2 (F5:1334)
constexpr int kLmax = 24;           // Huffman length limit
constexpr int kSub = 1024;          // symbols per decode sub-chunk
constexpr int kBaseChunk = 32768;   // symbols per base-stream chunk
constexpr int kStreams = 15;
constexpr int kLutBits = 12;        // decode table window (bits)

// Predictor options as run-time bits (stz_params.pred, the file header's
// pred).  F5's build options: PRED_WPOW2 squares the inverse-variance
// weights (F5:177-203); PRED_PLANE_AXIS adds two axis-pair candidates to
// blocks 2, 4 and 5 (F5:579-607, F5:627-633); PRED_CUBE_AXIS adds three to
// block 6 (F5:609-622, F5:675-686).  stz2.cpp's static predictors replace
// the dynamic ones: the 12-point truncated bicubic for blocks 2, 4, 5
// (stz2.cpp:122-155, 186-199, 222-235) and 16 points for block 6
// (stz2.cpp:271-293).  PRED_STAGED: dependency.cpp's staged inverse-variance
// predictors (dependency.cpp:162-292): blocks 2, 4, 5 from the reconstructed
// blocks 0, 1, 3 and block 6 from the reconstructed 2, 4, 5; alone (no other
// bit with it).  0 is F5's default build.
constexpr int kPredWpow2 = 1, kPredPlaneAxis = 2, kPredCubeAxis = 4, kPredBicubic12 = 8, kPredDiag16 = 16;
constexpr int kPredStaged = 32;
constexpr int kPredAll = 63;
// a valid predictor option set: known bits, the staged predictors alone
__host__ __device__ inline bool pred_valid(uint64_t p)
{
    return p <= (uint64_t)kPredAll && (!(p & kPredStaged) || p == (uint64_t)kPredStaged);
}

struct FileHeader {
    uint32_t magic, version, dtype, eb_mode;
    uint64_t nx, ny, nz;
    uint64_t px, py, pz;
    double eb_user, eb_unit;
    double mult_base, mult_low, mult_high;
    double eb_base, eb_low, eb_high;
    double vmin, vmax;
    uint32_t radius, chunk_high, chunk_low, sub_size;
    uint32_t interp, order, lmax, n_streams;
    double alpha, beta;
    uint64_t total_bytes;
    uint64_t pred;  // predictor options (kPred* bits), 0 = F5's default build
    uint64_t reserved[2];
    uint64_t stream_off[kStreams];
    uint64_t stream_bytes[kStreams];
    uint8_t pad[48];
};
static_assert(sizeof(FileHeader) == 512, "file header");

struct StreamHeader {
    uint32_t magic, kind, block, dtype;  // kind: 0 base, 1 low, 2 high
    uint64_t n_symbols;
    uint32_t dx, dy, dz, chunk;
    uint32_t n_chunks, n_sub, n_used, lmax;
    uint64_t n_unpred, total_bits;
    double eb;
    uint32_t radius, sub_size;
    uint64_t off_codebook, off_chunk_bit, off_chunk_unpred, off_chunk_sub;
    uint64_t off_sub_bits, off_sub_unpred, off_unpred, off_bits;
    uint64_t stream_bytes;
};
static_assert(sizeof(StreamHeader) == 160, "stream header");

// Geometry of one stream's block grid, passed to kernels by value.
struct Geom {
    int dx, dy, dz;
    int chunk;  // C for residual streams, 0 for the base stream (linear chunks)
    int ncx, ncy, ncz;
    uint32_t n_chunks;
    int cshift;  // log2(C) when C is a power of two, else -1
};

__host__ __device__ inline int imin_(int a, int b) { return a < b ? a : b; }

// stream position of block point (x,y,z): chunk start + chunk-local row-major
// index.  Chunks before (cz,cy,cx) are full along the axes they precede on.
__host__ __device__ inline uint32_t stream_pos(const Geom& g, int x, int y, int z)
{
    const int C = g.chunk;
    int cx, cy, cz;
    if (g.cshift >= 0) {
        cx = x >> g.cshift; cy = y >> g.cshift; cz = z >> g.cshift;
    } else {
        cx = x / C; cy = y / C; cz = z / C;
    }
    const int wx = imin_(C, g.dx - cx * C), wy = imin_(C, g.dy - cy * C), wz = imin_(C, g.dz - cz * C);
    const uint32_t start = (uint32_t)cz * C * g.dx * g.dy + (uint32_t)cy * C * g.dx * wz +
                           (uint32_t)cx * C * wy * wz;
    return start + ((uint32_t)(z - cz * C) * wy + (uint32_t)(y - cy * C)) * wx + (uint32_t)(x - cx * C);
}

// first stream position of chunk c
__host__ __device__ inline uint64_t chunk_start(const Geom& g, uint32_t c)
{
    if (g.chunk <= 0) return (uint64_t)c * kBaseChunk;
    const int C = g.chunk;
    const int cx = (int)(c % g.ncx), cy = (int)((c / g.ncx) % g.ncy), cz = (int)(c / ((uint32_t)g.ncx * g.ncy));
    const int wy = imin_(C, g.dy - cy * C), wz = imin_(C, g.dz - cz * C);
    return (uint64_t)cz * C * g.dx * g.dy + (uint64_t)cy * C * g.dx * wz + (uint64_t)cx * C * wy * wz;
}

__host__ __device__ inline uint64_t chunk_volume(const Geom& g, uint32_t c, uint64_t n)
{
    if (g.chunk <= 0) {
        const uint64_t p = (uint64_t)c * kBaseChunk;
        return n - p < (uint64_t)kBaseChunk ? n - p : (uint64_t)kBaseChunk;
    }
    const int C = g.chunk;
    const int cx = (int)(c % g.ncx), cy = (int)((c / g.ncx) % g.ncy), cz = (int)(c / ((uint32_t)g.ncx * g.ncy));
    return (uint64_t)imin_(C, g.dx - cx * C) * imin_(C, g.dy - cy * C) * imin_(C, g.dz - cz * C);
}

// extra: symbols a linearly chunked stream holds beyond its grid points (the
// tiles' header symbols of a Lorenzo/regression base)
inline Geom make_geom(int dx, int dy, int dz, int chunk, uint64_t extra = 0)
{
    Geom g{};
    g.dx = dx; g.dy = dy; g.dz = dz; g.chunk = chunk;
    g.cshift = -1;
    for (int k = 0; k < 31 && chunk > 0; ++k)
        if ((1 << k) == chunk) g.cshift = k;
    if (chunk > 0) {
        g.ncx = (dx + chunk - 1) / chunk;
        g.ncy = (dy + chunk - 1) / chunk;
        g.ncz = (dz + chunk - 1) / chunk;
        g.n_chunks = (uint32_t)g.ncx * g.ncy * g.ncz;
    } else {
        const uint64_t n = (uint64_t)dx * dy * dz + extra;
        g.n_chunks = (uint32_t)((n + kBaseChunk - 1) / kBaseChunk);
    }
    return g;
}

}  // namespace stz

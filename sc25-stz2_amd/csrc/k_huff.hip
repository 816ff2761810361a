// k_huff.hip -- entropy-coding kernels of the STZ device path (gfx950).
//
// Replaces SZ3-ncp's HuffmanEncoder<int> as the reference calls it:
// preprocess_encode_fast (histogram + tree, F5:1271 / F5:1402),
// encode_marked_fast (chunk-marked bit packing, F5:1292 / F5:1412),
// decode_marked_transform / decode_marked_chunks_fast (F5:1536, F5:1872).
// The tree itself is built on the host from the compacted histogram; all
// per-symbol work runs here:
//   histogram   fused into the quantize kernels (LDS bins; the centre code
//               counted per wave with one ballot)
//   encode      per-sub-chunk code-length sums -> per-stream chunk scan ->
//               wavefront bit-pack (16 symbols per lane, LDS word assembly,
//               atomics only on the two words a sub-chunk may share)
//   decode      one lane per 1024-symbol sub-chunk, 12-bit LDS lookup table
//               giving up to three symbols per lookup, canonical fallback
//               for longer codes
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.h"

namespace stz {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint64_t t = (uint64_t)__shfl_up((long long)v, d, kWave);
        if (lane >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, d, kWave);
    return v;
}

__device__ __forceinline__ int find_stream_by_chunk(const StreamSet& ss, uint32_t gc)
{
    int s = 0;
    for (int k = 1; k < (int)ss.nstreams; ++k)
        if (gc >= ss.s[k].chunk_base) s = k;
    return s;
}

// The encoder table entries of a stream's used symbols (at most kEncWin of
// them, those around the centre when the range is wider; nearly all codes
// fall there) staged in LDS per workgroup; the rest read global memory.
constexpr int kEncWin = (int)kEncWindow;

__device__ __forceinline__ void stage_enc_window(const uint32_t* __restrict__ es, uint32_t* tab, uint32_t lo,
                                                 uint32_t n)
{
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) tab[k] = es[lo + k];
}

// a position past the sub-chunk's end: encoder entry 0 (no bits), not a zero code
constexpr uint32_t kNoCode = 0xFFFFFFFFu;

// symbols of codes longer than the decode table window held in LDS by a
// staged decode workgroup (2 KiB; two workgroups still share a CU)
constexpr int kLongSyms = 1024;
// the same for the unstaged form, whose 16 KiB table leaves room
constexpr int kLongSyms1 = 2048;

// A lane's 16 codes of sub-chunk j (positions pstart + 1024 j + 16 lane ..
// +15): two 16-byte loads for a whole aligned sub-chunk, else code by code
// with kNoCode past its end
__device__ __forceinline__ void load16(const uint16_t* __restrict__ cs, uint64_t pstart, uint64_t pend, bool al8,
                                       uint32_t j, int lane, uint32_t (&q)[16])
{
    const uint64_t p0 = pstart + (uint64_t)j * kSub;
    const uint64_t p1 = p0 + kSub < pend ? p0 + kSub : pend;
    const uint64_t pl = p0 + (uint64_t)lane * 16;
    if (al8 && p1 - p0 == kSub) {
        const uint4 a = *(const uint4*)(cs + pl);
        const uint4 b = *(const uint4*)(cs + pl + 8);
        const uint32_t w8[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            q[2 * k] = w8[k] & 0xFFFFu;
            q[2 * k + 1] = w8[k] >> 16;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) q[k] = pl + k < p1 ? (uint32_t)cs[pl + k] : kNoCode;
    }
}

// encoder entries of a lane's 16 codes: every code reads the LDS window
// (clamped index), and only a lane holding a code outside it takes the
// branch to global memory -- one branch per 16 codes instead of one each
__device__ __forceinline__ void lookup16(const uint32_t* tab, const uint32_t* __restrict__ es, uint32_t lo,
                                         uint32_t n, const uint32_t (&q)[16], uint32_t (&e)[16])
{
    uint32_t outm = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint32_t r = q[k] - lo;
        const bool inw = r < n;
        e[k] = tab[inw ? r : 0u];
        outm |= inw ? 0u : (1u << k);
    }
    if (outm) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if ((outm >> k) & 1u) e[k] = q[k] <= 0xFFFFu ? es[q[k]] : 0u;
    }
}

// orders one wave's LDS accesses around this point (a wave's LDS operations
// execute in issue order, so only the compiler has to be held back)
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_zero_bits(StreamSet ss, uint8_t* __restrict__ payload)
{
    const StreamDesc& d = ss.s[blockIdx.y];
    uint32_t* w = (uint32_t*)(payload + d.off_bits);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < d.bits_words;
         i += (uint64_t)gridDim.x * blockDim.x)
        w[i] = 0u;
}

// one workgroup per stream: nonzero histogram bins -> (symbol << 32 | freq), symbol order
__device__ __forceinline__ void hist_compact_body(const uint32_t* __restrict__ hist, uint64_t* __restrict__ out,
                                                  uint32_t* __restrict__ counts, int s)
{
    __shared__ uint32_t wsum[16];
    const int tid = threadIdx.x, lane = __lane_id(), wid = tid >> 6;
    const uint32_t* h = hist + (size_t)s * 65536 + tid * 64;
    uint32_t v[64];
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < 64; k += 4) {
        const uint4 q = *(const uint4*)(h + k);
        v[k] = q.x; v[k + 1] = q.y; v[k + 2] = q.z; v[k + 3] = q.w;
    }
#pragma unroll
    for (int k = 0; k < 64; ++k) cnt += v[k] != 0u;
    const uint32_t inc = wave_incl_scan_u32(cnt);
    if (lane == kWave - 1) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        uint32_t t = lane < 16 ? wsum[lane] : 0u;
        t = wave_incl_scan_u32(t);
        if (lane < 16) wsum[lane] = t;
    }
    __syncthreads();
    uint32_t idx = inc - cnt + (wid ? wsum[wid - 1] : 0u);
    uint64_t* o = out + (size_t)s * 65536;
#pragma unroll
    for (int k = 0; k < 64; ++k)
        if (v[k]) o[idx++] = ((uint64_t)(tid * 64 + k) << 32) | v[k];
    if (tid == 0) counts[s] = wsum[15];
}

__global__ __launch_bounds__(1024) void k_hist_compact(const uint32_t* __restrict__ hist,
                                                       uint64_t* __restrict__ out,
                                                       uint32_t* __restrict__ counts)
{
    hist_compact_body(hist, out, counts, blockIdx.x);
}

__global__ void k_enc_scatter(const uint32_t* __restrict__ list, uint32_t n, uint32_t* __restrict__ enc)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = list[2 * i], val = list[2 * i + 1];
    enc[(size_t)(key >> 16) * 65536 + (key & 0xFFFFu)] = val;
}

// per chunk: bits and unpredictable count of each sub-chunk, in-chunk prefixes, chunk totals
__device__ __forceinline__ void huff_len_body(const StreamSet& ss, const uint16_t* __restrict__ codes,
                                              const uint32_t* __restrict__ enc, uint8_t* __restrict__ payload,
                                              uint32_t* __restrict__ sub_pre_bits, uint32_t* __restrict__ sub_pre_un,
                                              uint64_t* __restrict__ chunk_tot_bits,
                                              uint64_t* __restrict__ chunk_tot_un, uint32_t gc)
{
    __shared__ uint32_t lb[256], lu[256];
    __shared__ uint32_t tab[kEncWin];
    const int s = find_stream_by_chunk(ss, gc);
    const StreamDesc& d = ss.s[s];
    const uint32_t c = gc - d.chunk_base;
    const uint32_t* csub = (const uint32_t*)(payload + d.off_chunk_sub);
    const uint32_t sb = csub[c], nsub = csub[c + 1] - sb;
    const uint64_t pstart = chunk_start(d.g, c);
    const uint64_t pend = pstart + chunk_volume(d.g, c, d.n_symbols);
    const uint16_t* cs = codes + d.code_off;
    const uint32_t* es = enc + (size_t)s * 65536;
    stage_enc_window(es, tab, d.enc_lo, d.enc_n);
    __syncthreads();
    uint16_t* sbits = (uint16_t*)(payload + d.off_sub_bits);
    uint16_t* sun = (uint16_t*)(payload + d.off_sub_unpred);
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    const bool al8 = ((d.code_off + pstart) & 7) == 0;  // sub-chunks start 1024 codes apart
    for (uint32_t j = wid; j < nsub; j += 4) {
        // the lane's 16 codes first (two 16-byte loads for a whole aligned
        // sub-chunk), then 16 independent table reads
        uint32_t q[16], e[16];
        load16(cs, pstart, pend, al8, j, lane, q);
        lookup16(tab, es, d.enc_lo, d.enc_n, q, e);
        uint32_t bits = 0, un = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            bits += e[k] >> 24;
            un += q[k] == 0u ? 1u : 0u;
        }
        bits = wave_sum_u32(bits);
        un = wave_sum_u32(un);
        if (lane == 0) {
            lb[j] = bits;
            lu[j] = un;
            sbits[sb + j] = (uint16_t)bits;
            sun[sb + j] = (uint16_t)un;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ab = 0, au = 0;
        for (uint32_t j = 0; j < nsub; ++j) {
            sub_pre_bits[d.sub_base + sb + j] = ab;
            sub_pre_un[d.sub_base + sb + j] = au;
            ab += lb[j];
            au += lu[j];
        }
        chunk_tot_bits[gc] = ab;
        chunk_tot_un[gc] = au;
    }
}

__global__ __launch_bounds__(256) void k_huff_len(StreamSet ss, const uint16_t* __restrict__ codes,
                                                  const uint32_t* __restrict__ enc,
                                                  uint8_t* __restrict__ payload,
                                                  uint32_t* __restrict__ sub_pre_bits,
                                                  uint32_t* __restrict__ sub_pre_un,
                                                  uint64_t* __restrict__ chunk_tot_bits,
                                                  uint64_t* __restrict__ chunk_tot_un)
{
    huff_len_body(ss, codes, enc, payload, sub_pre_bits, sub_pre_un, chunk_tot_bits, chunk_tot_un, blockIdx.x);
}

// per stream: exclusive scan of chunk totals -> marker table (chunk bit offsets)
__device__ __forceinline__ void huff_scan_body(const StreamSet& ss, const uint64_t* __restrict__ tb,
                                               const uint64_t* __restrict__ tu, uint8_t* __restrict__ payload,
                                               int sidx)
{
    __shared__ uint64_t wsb[16], wsu[16];
    const StreamDesc& d = ss.s[sidx];
    const uint32_t n = d.n_chunks;
    const int tid = threadIdx.x, lane = __lane_id(), wid = tid >> 6;
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t c0 = tid * per < n ? tid * per : n;
    const uint32_t c1 = c0 + per < n ? c0 + per : n;
    uint64_t sbits = 0, sun = 0;
    for (uint32_t c = c0; c < c1; ++c) {
        sbits += tb[d.chunk_base + c];
        sun += tu[d.chunk_base + c];
    }
    const uint64_t ib = wave_incl_scan_u64(sbits), iu = wave_incl_scan_u64(sun);
    if (lane == kWave - 1) { wsb[wid] = ib; wsu[wid] = iu; }
    __syncthreads();
    if (wid == 0) {
        uint64_t a = lane < 16 ? wsb[lane] : 0, b = lane < 16 ? wsu[lane] : 0;
        a = wave_incl_scan_u64(a);
        b = wave_incl_scan_u64(b);
        if (lane < 16) { wsb[lane] = a; wsu[lane] = b; }
    }
    __syncthreads();
    uint64_t ab = ib - sbits + (wid ? wsb[wid - 1] : 0), au = iu - sun + (wid ? wsu[wid - 1] : 0);
    uint64_t* cbit = (uint64_t*)(payload + d.off_chunk_bit);
    uint64_t* cun = (uint64_t*)(payload + d.off_chunk_unpred);
    for (uint32_t c = c0; c < c1; ++c) {
        cbit[c] = ab;
        cun[c] = au;
        ab += tb[d.chunk_base + c];
        au += tu[d.chunk_base + c];
    }
    if (tid == 0) {
        cbit[n] = wsb[15];
        cun[n] = wsu[15];
    }
}

__global__ __launch_bounds__(1024) void k_huff_scan(StreamSet ss, const uint64_t* __restrict__ tb,
                                                    const uint64_t* __restrict__ tu,
                                                    uint8_t* __restrict__ payload)
{
    huff_scan_body(ss, tb, tu, payload, blockIdx.x);
}

// per chunk: wavefront bit-pack of each sub-chunk + its unpredictable values
__device__ __forceinline__ void huff_pack_body(const StreamSet& ss, const uint16_t* __restrict__ codes,
                                               const uint8_t* __restrict__ side, const uint32_t* __restrict__ enc,
                                               const uint32_t* __restrict__ sub_pre_bits,
                                               const uint32_t* __restrict__ sub_pre_un,
                                               uint8_t* __restrict__ payload, uint32_t gc)
{
    constexpr int kWords = (31 + kSub * kLmax + 31) / 32 + 2;  // 772
    __shared__ uint32_t wbuf[4][kWords];
    __shared__ uint32_t tab[kEncWin];
    const int s = find_stream_by_chunk(ss, gc);
    const StreamDesc& d = ss.s[s];
    const uint32_t c = gc - d.chunk_base;
    const uint32_t* csub = (const uint32_t*)(payload + d.off_chunk_sub);
    const uint32_t sb = csub[c], nsub = csub[c + 1] - sb;
    const uint64_t cbit0 = ((const uint64_t*)(payload + d.off_chunk_bit))[c];
    const uint64_t cun0 = ((const uint64_t*)(payload + d.off_chunk_unpred))[c];
    const uint64_t pstart = chunk_start(d.g, c);
    const uint64_t pend = pstart + chunk_volume(d.g, c, d.n_symbols);
    const uint16_t* cs = codes + d.code_off;
    const uint32_t* es = enc + (size_t)s * 65536;
    stage_enc_window(es, tab, d.enc_lo, d.enc_n);
    __syncthreads();
    uint32_t* bw = (uint32_t*)(payload + d.off_bits);
    uint8_t* un_out = payload + d.off_unpred;
    const uint8_t* side_s = side + d.code_off * d.tsize;
    const int lane = __lane_id(), wid = threadIdx.x >> 6;
    uint32_t* wb = wbuf[wid];
    const bool al8 = ((d.code_off + pstart) & 7) == 0;  // sub-chunks start 1024 codes apart
    // each wave packs its own sub-chunks in its own word buffer: no
    // workgroup barriers between the rounds
    for (uint32_t j = wid; j < nsub; j += 4) {
        const uint32_t gs = d.sub_base + sb + j;
        const uint64_t bit0 = cbit0 + sub_pre_bits[gs];
        const uint64_t u0 = cun0 + sub_pre_un[gs];
        const uint64_t pl = pstart + (uint64_t)j * kSub + (uint64_t)lane * 16;
        uint32_t q[16], e[16];
        load16(cs, pstart, pend, al8, j, lane, q);
        lookup16(tab, es, d.enc_lo, d.enc_n, q, e);
        uint32_t lbits = 0, lzero = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            lbits += e[k] >> 24;
            lzero += q[k] == 0u ? 1u : 0u;
        }
        const uint32_t ib = wave_incl_scan_u32(lbits);
        const uint32_t iu = wave_incl_scan_u32(lzero);
        const uint32_t tot = __shfl((int)ib, kWave - 1, kWave);
        const uint32_t off = (uint32_t)(bit0 & 31u);
        const uint32_t nW = (off + tot + 31) >> 5;
        for (uint32_t w = lane; w < nW; w += kWave) wb[w] = 0u;
        wave_lds_sync();
        // the lane's bits go to [pos, pos + lbits) of the word buffer, MSB first
        const uint32_t pos = off + ib - lbits;
        const uint32_t o = pos & 31u;
        if (o + lbits <= 64u) {
            // they span at most two words: assemble one 64-bit value
            uint64_t acc = 0;
#pragma unroll
            for (int k = 0; k < 16; ++k) acc = (acc << (e[k] >> 24)) | (e[k] & 0xFFFFFFu);
            if (lbits) {
                const uint64_t v = acc << (64u - lbits);
                const uint32_t w0 = pos >> 5;
                atomicOr(&wb[w0], (uint32_t)(v >> (32u + o)));
                if (o + lbits > 32u) atomicOr(&wb[w0 + 1], (uint32_t)(v >> o));
            }
        } else {
            // whole words as they fill; the first may be shared with the previous lane
            uint32_t wcur = pos >> 5;
            uint64_t acc = 0;
            int nacc = (int)o;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const int l = (int)(e[k] >> 24);
                acc = (acc << l) | (e[k] & 0xFFFFFFu);
                nacc += l;
                if (nacc >= 32) {
                    atomicOr(&wb[wcur], (uint32_t)(acc >> (nacc - 32)));
                    wcur++;
                    nacc -= 32;
                    acc &= nacc ? ((1ull << nacc) - 1ull) : 0ull;
                }
            }
            if (nacc > 0) atomicOr(&wb[wcur], (uint32_t)(acc << (32 - nacc)));
        }
        // unpredictable values of this lane, in stream order
        if (lzero) {
            uint64_t u = u0 + iu - lzero;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                if (q[k] == 0u) {
                    if (d.tsize == 4)
                        ((uint32_t*)un_out)[u] = ((const uint32_t*)side_s)[pl + k];
                    else
                        ((uint64_t*)un_out)[u] = ((const uint64_t*)side_s)[pl + k];
                    u++;
                }
            }
        }
        wave_lds_sync();
        const uint64_t W0 = bit0 >> 5;
        for (uint32_t w = lane; w < nW; w += kWave) {
            const uint32_t v = wb[w];
            if (w == 0 || w == nW - 1) {
                if (v) atomicOr(&bw[W0 + w], v);
            } else {
                bw[W0 + w] = v;
            }
        }
        wave_lds_sync();  // the next round clears the buffer
    }
}

__global__ __launch_bounds__(256) void k_huff_pack(StreamSet ss, const uint16_t* __restrict__ codes,
                                                   const uint8_t* __restrict__ side,
                                                   const uint32_t* __restrict__ enc,
                                                   const uint32_t* __restrict__ sub_pre_bits,
                                                   const uint32_t* __restrict__ sub_pre_un,
                                                   uint8_t* __restrict__ payload)
{
    huff_pack_body(ss, codes, side, enc, sub_pre_bits, sub_pre_un, payload, blockIdx.x);
}

// ---------------------------------------------------------------- many fields (encode)
__device__ __forceinline__ uint32_t find_enc_job(const EncJob* __restrict__ jobs, uint32_t njobs, uint32_t gc)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (jobs[m].chunk0 <= gc) lo = m; else hi = m;
    }
    return lo;
}

// streams s0 .. s0+ns-1 of every field
__global__ __launch_bounds__(1024) void k_hist_compact_multi(const EncJob* __restrict__ jobs, int s0, int ns)
{
    const EncJob& J = jobs[blockIdx.x / ns];
    hist_compact_body(J.hist, J.compact, J.counts, s0 + (int)(blockIdx.x % ns));
}

__global__ void k_enc_scatter_multi(const uint32_t* __restrict__ list, uint32_t n, const EncJob* __restrict__ jobs)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t f = list[3 * i], key = list[3 * i + 1], val = list[3 * i + 2];
    jobs[f].enc[(size_t)(key >> 16) * 65536 + (key & 0xFFFFu)] = val;
}

__global__ __launch_bounds__(256) void k_huff_len_multi(const EncJob* __restrict__ jobs, uint32_t njobs)
{
    const EncJob& J = jobs[find_enc_job(jobs, njobs, blockIdx.x)];
    huff_len_body(J.ss, J.codes, J.enc, J.payload, J.sub_pre_bits, J.sub_pre_un, J.chunk_tot_bits, J.chunk_tot_un,
                  blockIdx.x - J.chunk0);
}

__global__ __launch_bounds__(1024) void k_huff_scan_multi(const EncJob* __restrict__ jobs)
{
    const EncJob& J = jobs[blockIdx.x / kStreams];
    huff_scan_body(J.ss, J.chunk_tot_bits, J.chunk_tot_un, J.payload, (int)(blockIdx.x % kStreams));
}

__global__ __launch_bounds__(256) void k_huff_pack_multi(const EncJob* __restrict__ jobs, uint32_t njobs)
{
    const EncJob& J = jobs[find_enc_job(jobs, njobs, blockIdx.x)];
    huff_pack_body(J.ss, J.codes, (const uint8_t*)J.side, J.enc, J.sub_pre_bits, J.sub_pre_un, J.payload,
                   blockIdx.x - J.chunk0);
}

// one workgroup per segment: src[off .. off+bytes) -> dst
__global__ __launch_bounds__(256) void k_copy_segments(const uint8_t* __restrict__ src, const Segment* __restrict__ segs)
{
    const Segment g = segs[blockIdx.x];
    const uint8_t* a = src + g.src_off;
    uint8_t* b = (uint8_t*)g.dst;
    if (((g.src_off | (uint64_t)(uintptr_t)b | g.bytes) & 3u) == 0) {
        const uint32_t* a4 = (const uint32_t*)a;
        uint32_t* b4 = (uint32_t*)b;
        for (uint64_t k = threadIdx.x; k < g.bytes / 4; k += 256) b4[k] = a4[k];
    } else {
        for (uint64_t k = threadIdx.x; k < g.bytes; k += 256) b[k] = a[k];
    }
}

__device__ __forceinline__ int find_stream_by_wg(const StreamSet& ss, uint32_t wg)
{
    int s = -1;
    for (int k = 0; k < (int)ss.nstreams; ++k) {
        const uint32_t nwg = (ss.s[k].item_count + ss.s[k].ipw - 1) / ss.s[k].ipw;
        if (ss.s[k].item_count && wg >= ss.s[k].wg_base && wg < ss.s[k].wg_base + nwg) s = k;
    }
    return s;
}

// expand one chunk into its 1024-symbol decode work items: one wave per
// chunk, its lanes over the sub-chunks (64 per round), bit and
// unpredictable offsets by wave prefix sums with the running total carried
// between rounds
__device__ __forceinline__ void dec_prep_chunk(const StreamDesc& d, uint32_t s, uint32_t c, uint32_t ib,
                                               uint64_t pstart, const uint8_t* __restrict__ payload,
                                               DecItem* __restrict__ items)
{
    const int lane = __lane_id();
    const uint32_t* csub = (const uint32_t*)(payload + d.off_chunk_sub);
    const uint16_t* sbits = (const uint16_t*)(payload + d.off_sub_bits);
    const uint16_t* sun = (const uint16_t*)(payload + d.off_sub_unpred);
    uint64_t bit = ((const uint64_t*)(payload + d.off_chunk_bit))[c];
    uint64_t u = ((const uint64_t*)(payload + d.off_chunk_unpred))[c];
    const uint32_t sb = csub[c], nsub = csub[c + 1] - sb;
    const uint64_t vol = chunk_volume(d.g, c, d.n_symbols);
    for (uint32_t j0 = 0; j0 < nsub; j0 += kWave) {
        const uint32_t j = j0 + (uint32_t)lane;
        const bool ok = j < nsub;
        const uint32_t b = ok ? (uint32_t)sbits[sb + j] : 0u, n = ok ? (uint32_t)sun[sb + j] : 0u;
        const uint32_t ibits = wave_incl_scan_u32(b), iun = wave_incl_scan_u32(n);
        if (ok) {
            DecItem it;
            it.bit0 = bit + (ibits - b);
            it.u0 = u + (iun - n);
            it.p0 = (uint32_t)(pstart + (uint64_t)j * kSub);
            const uint64_t rem = vol - (uint64_t)j * kSub;
            it.cnt = (uint32_t)(rem < (uint64_t)kSub ? rem : (uint64_t)kSub);
            it.stream = s;
            it.bits = b;
            items[ib + j] = it;
        }
        bit += (uint32_t)__shfl((int)ibits, kWave - 1, kWave);
        u += (uint32_t)__shfl((int)iun, kWave - 1, kWave);
    }
}

// one 64-lane workgroup per chunk of the set, or per listed chunk
__global__ __launch_bounds__(64) void k_dec_prep(StreamSet ss, const uint8_t* __restrict__ payload,
                                                 const uint32_t* __restrict__ list, uint32_t nlist,
                                                 DecItem* __restrict__ items, uint32_t* __restrict__ cmap)
{
    const uint32_t i = blockIdx.x;
    uint32_t s, c, ib;
    uint64_t pstart;
    if (list) {
        if (i >= nlist) return;
        s = list[4 * i];
        c = list[4 * i + 1];
        ib = ss.s[s].item_base + list[4 * i + 2];
        pstart = list[4 * i + 3];
        if (threadIdx.x == 0) cmap[ss.s[s].cm_base + c] = (uint32_t)pstart;
    } else {
        if (i >= ss.total_chunks) return;
        s = (uint32_t)find_stream_by_chunk(ss, i);
        c = i - ss.s[s].chunk_base;
        ib = ss.s[s].item_base + ((const uint32_t*)(payload + ss.s[s].off_chunk_sub))[c];
        pstart = chunk_start(ss.s[s].g, c);
    }
    dec_prep_chunk(ss.s[s], s, c, ib, pstart, payload, items);
}

// Decode tables: one thread per kLutBits-bit window of one stream
// (blockIdx.y).  Codes are found by the canonical test in increasing length
// -- codes of length l are first_code[l] .. first_code[l] + count[l] - 1 --
// which names the unique code that prefixes the window.  Up to three codes
// are chained while they fit in the window; symbol 0 only ever stands alone.
__device__ __forceinline__ void build_dectab(const DecHead& h, const uint16_t* __restrict__ sorted, DecTab& t)
{
    if (blockIdx.x == 0 && threadIdx.x < kLmax + 2) {
        t.first_code[threadIdx.x] = h.first_code[threadIdx.x];
        t.count[threadIdx.x] = h.count[threadIdx.x];
        t.first_index[threadIdx.x] = h.first_index[threadIdx.x];
        if (threadIdx.x == 0) t.sorted_off = h.sorted_off;
    }
    const uint32_t w = blockIdx.x * 256u + threadIdx.x;
    if (w >= (1u << kLutBits)) return;
    const uint16_t* srt = sorted + h.sorted_off;
    // first code of the `avail` leading bits of the kLutBits-bit window `win`
    auto code_at = [&](uint32_t win, uint32_t avail, uint32_t& sym, uint32_t& len) -> bool {
        for (uint32_t l = 1; l <= avail; ++l) {
            const uint32_t cd = win >> (kLutBits - l);
            if (cd - h.first_code[l] < h.count[l]) {
                sym = srt[h.first_index[l] + cd - h.first_code[l]];
                len = l;
                return true;
            }
        }
        return false;
    };
    uint32_t s1, l1;
    if (!code_at(w, kLutBits, s1, l1)) {
        t.lut[w] = 0ull;  // n = 0: longer than the window
        t.lut1[w] = 0xFFFFFFFFu;
        return;
    }
    t.lut1[w] = (l1 << 16) | s1;
    uint64_t e = s1;
    uint32_t cnt = 1, used = l1;
    while (s1 != 0 && cnt < 3 && used < (uint32_t)kLutBits) {
        uint32_t sk, lk;
        if (!code_at((w << used) & ((1u << kLutBits) - 1u), kLutBits - used, sk, lk) || sk == 0) break;
        e |= (uint64_t)sk << (16 * cnt);
        ++cnt;
        used += lk;
    }
    t.lut[w] = e | (uint64_t)cnt << kDecNShift | (uint64_t)l1 << kDecL1Shift | (uint64_t)used << kDecLtShift;
}

__global__ __launch_bounds__(256) void k_build_dectabs(const DecHead* __restrict__ heads,
                                                       const uint16_t* __restrict__ sorted, DecTab* __restrict__ tabs)
{
    build_dectab(heads[blockIdx.y], sorted, tabs[blockIdx.y]);
}

// many stream sets: set blockIdx.z, stream blockIdx.y
__global__ __launch_bounds__(256) void k_build_dectabs_multi(const DecTabJob* __restrict__ jobs)
{
    const DecTabJob& J = jobs[blockIdx.z];
    build_dectab(J.heads[blockIdx.y], J.sorted, J.tabs[blockIdx.y]);
}

// Symbol at canonical index idx of a code longer than the table window:
// from the LDS copy of the first ln of them (lsrt, indexed from llo) or, when
// not all of them fit (kLdsLong false), from global memory past it.  An index
// outside both (a corrupt bitstream) sets `bad`.
template <bool kLdsLong>
__device__ __forceinline__ uint32_t long_symbol(uint32_t idx, const uint16_t* __restrict__ srt, const uint16_t* lsrt,
                                                uint32_t llo, uint32_t ln, uint32_t& bad)
{
    const uint32_t r = idx - llo;
    if constexpr (kLdsLong) {
        bad |= r < ln ? 0u : 1u;
        return (uint32_t)lsrt[r < ln ? r : 0u];
    } else {
        return r < ln ? (uint32_t)lsrt[r] : (uint32_t)srt[idx];
    }
}

// Sequential Huffman decode of one lane's sub-chunk.  Bits come from `W`
// (the workgroup's LDS stage or global memory; word w of the stream is
// W[w - wbase]); each refill word is fetched one refill ahead.  One table
// lookup takes up to three symbols (the codes that lie wholly in the next
// kLutBits bits).  Codes leave in 8-byte groups once the destination is
// 8-byte aligned (`al` = element offset of co[0]); the first symbols up to
// that alignment and the last ones go one at a time.
// kLdsLong: every symbol of a code longer than the window is in the LDS copy
// `lsrt` (the workgroup checked), so that path reads LDS only.  Otherwise a
// long code may read global memory, and on this GPU a global load waits for
// every store the wave issued before it (the vector memory counter retires
// loads and stores in order): with a long code every few steps in some lane
// of a wave, the whole loop then runs at the speed of store completions.
template <bool kLdsBits, bool kLdsLong>
__device__ __forceinline__ uint32_t decode_run(const uint32_t* __restrict__ W, uint64_t wbase, uint64_t nwords,
                                               const DecItem& it, const uint64_t* __restrict__ lut,
                                               const uint32_t* __restrict__ fc, const uint32_t* __restrict__ cn,
                                               const uint32_t* __restrict__ fi, const uint16_t* __restrict__ srt,
                                               uint16_t* __restrict__ co, uint8_t* __restrict__ so,
                                               const uint8_t* __restrict__ un, uint32_t tsize, uint64_t n_unpred,
                                               uint64_t al, const uint16_t* lsrt, uint32_t llo, uint32_t ln)
{
    const uint64_t wpos0 = (it.bit0 >> 5) - wbase;
    uint64_t wpos = wpos0;
    const int skip = (int)(it.bit0 & 31u);
    uint32_t bad = 0;
    if (wpos + 2 > nwords) return 1u;
    uint64_t buf = (((uint64_t)W[wpos] << 32) | W[wpos + 1]) << skip;
    int nb = 64 - skip;
    wpos += 2;
    // the look-ahead word is read unconditionally (index clamped; a word
    // past the end is never consumed without setting `bad`), so the load
    // lands in its register and is only waited for at the next refill
    const uint64_t wlast = nwords - 1;
    uint32_t nextw = W[wpos < wlast ? wpos : wlast];
    uint64_t u = it.u0;
    const uint32_t cnt = it.cnt;
    auto refill = [&]() {
        if (nb < 32) {
            buf |= (uint64_t)nextw << (32 - nb);
            bad |= wpos >= nwords ? 1u : 0u;
            ++wpos;
            nb += 32;
            nextw = W[wpos < wlast ? wpos : wlast];
        }
    };
    // a code longer than the table window: canonical search on the top 32 bits
    auto slow = [&](int& l) -> uint32_t {
        const uint32_t win = (uint32_t)(buf >> 32);
        for (l = kLutBits + 1; l <= kLmax; ++l)
            if ((win >> (32 - l)) - fc[l] < cn[l]) break;
        if (l > kLmax) {
            bad = 1;
            l = 1;
            return 0u;
        }
        // the long codes' symbols: the workgroup's LDS copy of their first ln
        const uint32_t idx = fi[l] + (win >> (32 - l)) - fc[l];
        return long_symbol<kLdsLong>(idx, srt, lsrt, llo, ln, bad);
    };
    // symbol 0 at stream position `pos` of the sub-chunk: the next raw value
    auto unpred = [&](uint32_t pos) {
        if (u < n_unpred) {
            if (tsize == 4) ((uint32_t*)so)[pos] = ((const uint32_t*)un)[u];
            else ((uint64_t*)so)[pos] = ((const uint64_t*)un)[u];
        } else {
            bad = 1;
        }
        u++;
    };
    auto single = [&](uint32_t pos) -> uint32_t {
        refill();
        const uint64_t e = lut[buf >> (64 - kLutBits)];
        uint32_t sym;
        int l;
        if ((e >> kDecNShift) & 3u) {
            sym = (uint32_t)e & 0xFFFFu;
            l = (int)((e >> kDecL1Shift) & 15u);
        } else {
            sym = slow(l);
        }
        buf <<= l;
        nb -= l;
        if (sym == 0u) unpred(pos);
        return sym;
    };
    uint32_t i = 0;
    const uint32_t head = (uint32_t)((4u - (uint32_t)(al & 3u)) & 3u);
    for (; i < head && i < cnt; ++i) co[i] = (uint16_t)single(i);
    // co + i is 8-byte aligned; `lo` holds pc decoded codes not yet stored
    uint64_t lo = 0;
    uint32_t pc = 0;
    while (i + pc + 3 <= cnt) {
        if constexpr (!kLdsBits) {
            // bits in global memory: a read is a long round trip, so only
            // when the buffer runs low, one word ahead
            refill();
        } else {
            // bits in LDS: refill without a branch -- the word after the
            // look-ahead one is read every step (beside the table read) and
            // taken by selects
            const bool need = nb < 32;
            const uint64_t wn = wpos + 1;
            const uint32_t w2 = W[wn < wlast ? wn : wlast];
            buf |= need ? (uint64_t)nextw << ((32 - nb) & 63) : 0ull;
            bad |= (need && wpos >= nwords) ? 1u : 0u;
            wpos += need ? 1u : 0u;
            nb += need ? 32 : 0;
            nextw = need ? w2 : nextw;
        }
        const uint64_t e = lut[buf >> (64 - kLutBits)];
        uint32_t n = (uint32_t)(e >> kDecNShift) & 3u;
        int l = (int)((e >> kDecLtShift) & 15u);
        uint64_t v = e & ((1ull << (16 * n)) - 1ull);
        // the two rare cases under one branch: a code longer than the
        // window (n = 0), and symbol 0, which always comes alone (v = 0)
        if (v == 0ull) {
            if (n == 0) {
                v = slow(l);
                n = 1;
            }
            if (v == 0ull) unpred(i + pc);
        }
        buf <<= l;
        nb -= l;
        const uint64_t spill = pc ? v >> (64 - 16 * pc) : 0ull;
        lo |= v << (16 * pc);
        pc += n;
        if (pc >= 4) {
            *(uint2*)(co + i) = make_uint2((uint32_t)lo, (uint32_t)(lo >> 32));
            i += 4;
            pc -= 4;
            lo = spill;
        }
    }
    for (uint32_t k = 0; k < pc; ++k) co[i + k] = (uint16_t)(lo >> (16 * k));
    i += pc;
    for (; i < cnt; ++i) co[i] = (uint16_t)single(i);
    // the sub-chunk uses exactly its recorded bit count (else corrupt):
    // bits used = words loaded into buf x 32 - the start offset - bits left
    const int64_t used = (int64_t)(wpos - wpos0) * 32 - skip - nb;
    bad |= used != (int64_t)it.bits ? 1u : 0u;
    return bad;
}

// Canonical-code limits of the lengths above the table window, uniform over
// a workgroup (scalar registers): lim[j] = first_code[l] + count[l] of length
// l = kLutBits + 1 + j, left-aligned in 32 bits (64-bit values, since a
// complete code's last limit is 2^32).  Canonical codes of increasing length
// have increasing left-aligned values, so the length of the code that
// prefixes a 32-bit window w is kLutBits + 1 + (the number of limits w is at
// or above): twelve compares, no branches.  The symbol's place in the LDS
// copy of the long codes' symbols is then loff[j] + (w >> (32 - l)), loff[j]
// = first_index[l] - first_code[l] - (first long index), from an LDS table.
constexpr int kLong = kLmax - kLutBits;  // 12 lengths
struct LongLim {
    uint64_t lim[kLong];
};

__device__ __forceinline__ LongLim long_limits(const uint32_t* fc, const uint32_t* cn)
{
    LongLim L;
#pragma unroll
    for (int j = 0; j < kLong; ++j) {
        const int l = kLutBits + 1 + j;
        const uint64_t v = ((uint64_t)fc[l] + cn[l]) << (32 - l);
        // the builtin returns int: cast both halves back before widening
        // (a sign-extended low half would set every bit of the high one)
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
        L.lim[j] = ((uint64_t)hi << 32) | lo;
    }
    return L;
}

// Sequential Huffman decode of one lane's sub-chunk with its bits in the
// workgroup's LDS stage (word w of the stream is W[w - wbase]; the stage has
// kStageSlack readable words past the staged ones, so the reads below need no
// clamp).  The reader holds two words A, B and k, the bits of A already
// consumed (1..32): the stream's next 32 bits are alignbit(A, B, 32 - k), one
// instruction, and 32-bit arithmetic throughout.  Every step reads, beside
// its table entry, the word two after A (both reads issued before either is
// waited on); when the step's codes carry k past 32, B moves to A and that
// word's predecessor to B.  One table lookup takes up to three symbols (the
// codes that lie wholly in the next kLutBits bits).  A longer code finds its
// length by the compares above and its symbol in the LDS copy (all of them
// are there: the caller checked), so the loop reads LDS only; a global load
// would wait for every store the wave issued before it (the vector memory
// counter retires loads and stores in order).  Codes leave in 8-byte groups
// once the destination is 8-byte aligned (`al` = element offset of co[0]);
// the first symbols up to that alignment and the last ones go one at a time.
// An unpredictable value (symbol 0) is the only global read, and rare.
constexpr int kStageSlack = 4;

__device__ __forceinline__ uint32_t decode_run_lds(const uint32_t* W, uint64_t wbase, const DecItem& it,
                                                   const uint64_t* lut, const LongLim LL, const uint32_t* loff,
                                                   const uint16_t* lsrt, uint32_t ln, uint16_t* __restrict__ co, uint8_t* __restrict__ so,
                                                   const uint8_t* __restrict__ un, uint32_t tsize, uint64_t n_unpred,
                                                   uint64_t al)
{
    // bit offset of the sub-chunk in the stage (>= 0: the stage starts at
    // the workgroup's first word); A = stage word a, k its consumed bits
    const int64_t rel0 = (int64_t)it.bit0 - (int64_t)(wbase << 5);
    int32_t a = (int32_t)((rel0 + 31) >> 5) - 1;  // -1 when rel0 = 0 (k = 32, A unused)
    int32_t k = (int32_t)(rel0 - 32 * (int64_t)a);
    const int32_t a0 = a, k0 = k;
    // W[-1] is the word before the stage (the caller's layout keeps it
    // readable); it only ever stands in for an A that is never read
    uint32_t A = W[a], B = W[a + 1], N = W[a + 2];
    uint32_t bad = 0;
    uint64_t u = it.u0;
    const uint32_t cnt = it.cnt;
    auto consume = [&](int32_t l, uint32_t NN) {
        k += l;
        const bool t = k > 32;
        k -= t ? 32 : 0;
        a += t ? 1 : 0;
        A = t ? B : A;
        B = t ? N : B;
        N = t ? NN : N;
    };
    // a code longer than the table window
    auto slow = [&](uint32_t win, int32_t& l) -> uint32_t {
        uint32_t j = 0;
#pragma unroll
        for (int q = 0; q < kLong; ++q) j += (uint64_t)win >= LL.lim[q] ? 1u : 0u;
        if (j >= (uint32_t)kLong) {  // no code prefixes the window
            bad = 1;
            l = 1;
            return 0u;
        }
        l = kLutBits + 1 + (int32_t)j;
        const uint32_t r = loff[j] + (win >> (32 - l));
        bad |= r < ln ? 0u : 1u;
        return (uint32_t)lsrt[r < ln ? r : 0u];
    };
    auto unpred = [&](uint32_t pos) {
        if (u < n_unpred) {
            if (tsize == 4) ((uint32_t*)so)[pos] = ((const uint32_t*)un)[u];
            else ((uint64_t*)so)[pos] = ((const uint64_t*)un)[u];
        } else {
            bad = 1;
        }
        u++;
    };
    // one symbol (the first code of the window)
    auto single = [&](uint32_t pos) -> uint32_t {
        const uint32_t NN = W[a + 3];
        const uint32_t win = __builtin_amdgcn_alignbit(A, B, (uint32_t)(32 - k));
        const uint64_t e = lut[win >> (32 - kLutBits)];
        uint32_t sym;
        int32_t l;
        if ((e >> kDecNShift) & 3u) {
            sym = (uint32_t)e & 0xFFFFu;
            l = (int32_t)((e >> kDecL1Shift) & 15u);
        } else {
            sym = slow(win, l);
        }
        consume(l, NN);
        if (sym == 0u) unpred(pos);
        return sym;
    };
    uint32_t i = 0;
    const uint32_t head = (uint32_t)((8u - (uint32_t)(al & 7u)) & 7u);
    for (; i < head && i < cnt; ++i) co[i] = (uint16_t)single(i);
    // co + i is 16-byte aligned.  Codes leave in 16-byte pairs of 4-code
    // groups: `pv` holds the group at co + i when `hp`, `lo` the pc codes
    // after it.  A lane's stores land on lines of their own (one sub-chunk
    // each), so every store instruction addresses 64 lines; halving their
    // number halves that address work, which stalled the lanes' chains
    // (8-byte stores took a quarter of the decode's time).
    uint64_t lo = 0, pv = 0;
    uint32_t pc = 0, hp = 0;
    while (i + hp + pc + 3 <= cnt) {
        const uint32_t NN = W[a + 3];
        const uint32_t win = __builtin_amdgcn_alignbit(A, B, (uint32_t)(32 - k));
        const uint64_t e = lut[win >> (32 - kLutBits)];
        // the table leaves unused symbol fields zero: the low 48 bits are
        // the entry's symbols, 16 bits each
        uint64_t v = e & 0xFFFFFFFFFFFFull;
        uint32_t n = (uint32_t)(e >> kDecNShift) & 3u;
        int32_t l = (int32_t)((e >> kDecLtShift) & 15u);
        // the two rare cases under one branch: a code longer than the
        // window (n = 0), and symbol 0, which always comes alone (v = 0)
        if (v == 0ull) {
            if (n == 0) {
                v = slow(win, l);
                n = 1;
            }
            if (v == 0ull) unpred(i + hp + pc);
        }
        consume(l, NN);
        const uint64_t spill = pc ? v >> (64 - 16 * pc) : 0ull;
        lo |= v << (16 * pc);
        pc += n;
        if (pc >= 4) {
            if (hp) {
                *(uint4*)(co + i) = make_uint4((uint32_t)pv, (uint32_t)(pv >> 32), (uint32_t)lo, (uint32_t)(lo >> 32));
                i += 8;
                hp = 0;
            } else {
                pv = lo;
                hp = 4;
            }
            pc -= 4;
            lo = spill;
        }
    }
    if (hp) {
        *(uint2*)(co + i) = make_uint2((uint32_t)pv, (uint32_t)(pv >> 32));
        i += 4;
    }
    for (uint32_t q = 0; q < pc; ++q) co[i + q] = (uint16_t)(lo >> (16 * q));
    i += pc;
    for (; i < cnt; ++i) co[i] = (uint16_t)single(i);
    // the sub-chunk uses exactly its recorded bit count (else corrupt)
    const int64_t used = 32 * (int64_t)(a - a0) + (k - k0);
    bad |= used != (int64_t)it.bits ? 1u : 0u;
    return bad;
}

// Single-symbol form for dense payloads (a kLutBits window rarely holds two
// codes): one table entry (len << 16 | symbol, 0xFFFFFFFF for a longer code)
// per step from a 16 KiB table, bits from global memory fetched one refill
// ahead.  Codes leave in 64-byte bursts (32 codes held in registers) when the
// destination is 64-byte aligned, in 8-byte groups otherwise.  kLdsLong as
// for decode_run.
template <int kGroup, bool kLdsLong>
__device__ __forceinline__ uint32_t decode_run1(const uint32_t* __restrict__ W, uint64_t wbase, uint64_t nwords,
                                               const DecItem& it, const uint32_t* __restrict__ lut,
                                               const uint32_t* __restrict__ fc, const uint32_t* __restrict__ cn,
                                               const uint32_t* __restrict__ fi, const uint16_t* __restrict__ srt,
                                               uint16_t* __restrict__ co, uint8_t* __restrict__ so,
                                               const uint8_t* __restrict__ un, uint32_t tsize, uint64_t n_unpred,
                                               const uint16_t* lsrt, uint32_t llo, uint32_t ln)
{
    const uint64_t wpos0 = (it.bit0 >> 5) - wbase;
    uint64_t wpos = wpos0;
    const int skip = (int)(it.bit0 & 31u);
    uint32_t bad = 0;
    if (wpos + 2 > nwords) return 1u;
    uint64_t buf = (((uint64_t)W[wpos] << 32) | W[wpos + 1]) << skip;
    int nb = 64 - skip;
    wpos += 2;
    uint32_t nextw = wpos < nwords ? W[wpos] : 0u;
    uint64_t u = it.u0;
    const uint32_t cnt = it.cnt;
    uint32_t i = 0;
    auto step = [&]() -> uint32_t {
        if (nb < 32) {
            buf |= (uint64_t)nextw << (32 - nb);
            bad |= wpos >= nwords ? 1u : 0u;
            ++wpos;
            nb += 32;
            nextw = wpos < nwords ? W[wpos] : 0u;
        }
        const uint32_t e = lut[(uint32_t)(buf >> (64 - kLutBits))];
        uint32_t sym;
        int l;
        if (e != 0xFFFFFFFFu) {
            sym = e & 0xFFFFu;
            l = (int)(e >> 16);
        } else {
            const uint32_t win = (uint32_t)(buf >> 32);
            for (l = kLutBits + 1; l <= kLmax; ++l)
                if ((win >> (32 - l)) - fc[l] < cn[l]) break;
            if (l > kLmax) {
                bad = 1;
                l = 1;
                sym = 0;
            } else {
                const uint32_t idx = fi[l] + (win >> (32 - l)) - fc[l];
                sym = long_symbol<kLdsLong>(idx, srt, lsrt, llo, ln, bad);
            }
        }
        buf <<= l;
        nb -= l;
        if (sym == 0u) {
            if (u < n_unpred) {
                if (tsize == 4) ((uint32_t*)so)[i] = ((const uint32_t*)un)[u];
                else ((uint64_t*)so)[i] = ((const uint64_t*)un)[u];
            } else {
                bad = 1;
            }
            u++;
        }
        return sym;
    };
    if (kGroup == 32) {
        // 32 codes -> 4 x 16-byte stores
        for (; i + 32 <= cnt;) {
            uint32_t r[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const uint32_t a = step();
                ++i;
                const uint32_t b = step();
                ++i;
                r[k] = a | (b << 16);
            }
            uint4* d = (uint4*)(co + i - 32);
            d[0] = make_uint4(r[0], r[1], r[2], r[3]);
            d[1] = make_uint4(r[4], r[5], r[6], r[7]);
            d[2] = make_uint4(r[8], r[9], r[10], r[11]);
            d[3] = make_uint4(r[12], r[13], r[14], r[15]);
        }
    } else if (kGroup == 4) {
        for (; i + 4 <= cnt;) {
            const uint32_t a = step();
            ++i;
            const uint32_t b = step();
            ++i;
            const uint32_t c = step();
            ++i;
            const uint32_t d = step();
            ++i;
            *(uint2*)(co + i - 4) = make_uint2(a | (b << 16), c | (d << 16));
        }
    }
    for (; i < cnt; ++i) co[i] = (uint16_t)step();
    // the sub-chunk uses exactly its recorded bit count (else corrupt):
    // bits used = words loaded into buf x 32 - the start offset - bits left
    const int64_t used = (int64_t)(wpos - wpos0) * 32 - skip - nb;
    bad |= used != (int64_t)it.bits ? 1u : 0u;
    return bad;
}

// One workgroup = `ipw` consecutive sub-chunks of one stream, one per lane.
// Their bits form one contiguous range.  kStage: the workgroup first copies
// that range (up to kDecStageWords words of it) into LDS with coalesced
// loads, so the lanes' data-dependent refills are LDS reads (low-entropy
// payloads); lanes past the staged words read global memory.  Without the
// stage (dense payloads) every refill reads global memory, and the smaller
// LDS footprint lets more workgroups share a CU.
template <int kForm>
__device__ __forceinline__ void huff_decode_wg(const StreamSet& ss, const uint8_t* __restrict__ payload,
                                               const DecTab* __restrict__ tabs, const uint16_t* __restrict__ sorted,
                                               const DecItem* __restrict__ items, uint16_t* __restrict__ codes,
                                               uint8_t* __restrict__ side, uint32_t* __restrict__ err,
                                               uint32_t wgid, uint64_t* lut, uint32_t* fc, uint32_t* cn,
                                               uint32_t* fi, uint32_t* stage, uint16_t* lsrt, uint32_t* loff)
{
    const int s = find_stream_by_wg(ss, wgid);
    if (s < 0) return;  // uniform across the workgroup
    constexpr bool kStage = kForm == kDecStaged;
    const StreamDesc& d = ss.s[s];
    const DecTab* t = tabs + s;
    uint32_t* lut1 = (uint32_t*)lut;  // the single-symbol table of the dense form
    if constexpr (kStage)
        for (int k = threadIdx.x; k < (1 << kLutBits); k += 256) lut[k] = t->lut[k];
    else
        for (int k = threadIdx.x; k < (1 << kLutBits); k += 256) lut1[k] = t->lut1[k];
    // the symbols of the codes longer than the table window (the tail of
    // the canonical order), the first kLongSyms (kLongSyms1) of them, in LDS:
    // the long-code path then waits on no global load
    constexpr uint32_t kCap = kStage ? kLongSyms : kLongSyms1;
    const uint32_t llo = t->first_index[kLutBits + 1];
    const uint32_t lhi = t->first_index[kLmax] + t->count[kLmax];
    const uint32_t ln = lhi > llo ? (lhi - llo < kCap ? lhi - llo : kCap) : 0u;
    const uint16_t* s0 = sorted + t->sorted_off + llo;
    for (uint32_t k = threadIdx.x; k < ln; k += 256) lsrt[k] = s0[k];
    if (threadIdx.x < kLmax + 2) {
        fc[threadIdx.x] = t->first_code[threadIdx.x];
        cn[threadIdx.x] = t->count[threadIdx.x];
        fi[threadIdx.x] = t->first_index[threadIdx.x];
    }
    if (threadIdx.x < kLong) {
        const int l = kLutBits + 1 + (int)threadIdx.x;
        loff[threadIdx.x] = t->first_index[l] - t->first_code[l] - llo;
    }
    const uint32_t first = d.item_base + (wgid - d.wg_base) * d.ipw;
    const uint32_t end = d.item_base + d.item_count;
    const uint32_t lim = first + d.ipw < end ? first + d.ipw : end;
    const uint32_t last = lim - 1u;
    const uint32_t* bw = (const uint32_t*)(payload + d.off_bits);
    // items of a stream are in increasing bit order, so the range is [first.bit0, last.bit0 + last.bits)
    const uint64_t w0 = items[first].bit0 >> 5;
    const uint64_t w1r = ((items[last].bit0 + items[last].bits + 31) >> 5) + 2;
    const uint64_t w1a = w1r < d.bits_words ? w1r : d.bits_words;
    // the stage holds the first kDecStageWords words of the range; lanes
    // whose bits lie past it read global memory
    const uint64_t w1 = w1a > w0 + kDecStageWords ? w0 + kDecStageWords : w1a;
    const bool use_lds = kStage && w1 > w0;
    // staged word w0 + j sits at sw[j]; the stage array keeps one readable
    // word before sw and kStageSlack after the largest stage
    uint32_t* sw = stage + (kStage ? 1 : 0);
    if (use_lds) {
        const uint64_t nw = w1 - w0;
        const uint64_t a0 = (w0 + 3) & ~3ull;  // 16-byte aligned body
        const uint64_t h = a0 - w0 < nw ? a0 - w0 : nw;
        if (threadIdx.x < h) sw[threadIdx.x] = bw[w0 + threadIdx.x];
        const uint64_t nb4 = (nw - h) / 4;
        const uint4* src = (const uint4*)(bw + a0);
        for (uint64_t k = threadIdx.x; k < nb4; k += 256) {
            const uint4 v = src[k];
            uint32_t* dst = sw + h + 4 * k;
            dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; dst[3] = v.w;
        }
        for (uint64_t k = h + 4 * nb4 + threadIdx.x; k < nw; k += 256) sw[k] = bw[w0 + k];
    }
    __syncthreads();
    const uint32_t item = first + threadIdx.x;
    if (item >= lim) return;
    const DecItem it = items[item];
    const uint16_t* srt = sorted + t->sorted_off;
    uint16_t* co = codes + d.code_off + it.p0;
    uint8_t* so = side + (d.code_off + it.p0) * d.tsize;
    const uint8_t* un = payload + d.off_unpred;
    const uint64_t al = (d.code_off + it.p0);
    const bool in_stage = use_lds && it.bit0 >= (w0 << 5) && ((it.bit0 + it.bits + 31) >> 5) + 2 <= w1;
    // two calls rather than one through a pointer chosen per lane: each
    // sees the address space of its bits (LDS reads for the stage, global
    // loads otherwise) instead of generic flat loads
    // every long-code symbol in the LDS copy (uniform over the workgroup):
    // the instantiations whose long-code path reads LDS only
    const bool lds_long = lhi <= llo || lhi - llo <= kCap;
    auto run = [&](auto long_tag) -> uint32_t {
        constexpr bool kL = decltype(long_tag)::value;
        if constexpr (!kStage) {
            if ((al & 31u) == 0)
                return decode_run1<32, kL>(bw, 0, d.bits_words, it, lut1, fc, cn, fi, srt, co, so, un, d.tsize,
                                           d.n_unpred, lsrt, llo, ln);
            if ((al & 3u) == 0)
                return decode_run1<4, kL>(bw, 0, d.bits_words, it, lut1, fc, cn, fi, srt, co, so, un, d.tsize,
                                          d.n_unpred, lsrt, llo, ln);
            return decode_run1<1, kL>(bw, 0, d.bits_words, it, lut1, fc, cn, fi, srt, co, so, un, d.tsize,
                                      d.n_unpred, lsrt, llo, ln);
        } else if (in_stage) {
            if constexpr (kL)
                return decode_run_lds(sw, w0, it, lut, long_limits(fc, cn), loff, lsrt, ln, co, so, un, d.tsize,
                                      d.n_unpred, al);
            else
                return decode_run<true, false>(sw, w0, w1 - w0, it, lut, fc, cn, fi, srt, co, so, un, d.tsize,
                                               d.n_unpred, al, lsrt, llo, ln);
        } else {
            return decode_run<false, kL>(bw, 0, d.bits_words, it, lut, fc, cn, fi, srt, co, so, un, d.tsize,
                                         d.n_unpred, al, lsrt, llo, ln);
        }
    };
    const uint32_t bad = lds_long ? run(std::true_type{}) : run(std::false_type{});
    if (bad) atomicOr(err, 1u);
}


template <int kForm>
__global__ __launch_bounds__(256) void k_huff_decode(StreamSet ss, const uint8_t* __restrict__ payload,
                                                     const DecTab* __restrict__ tabs,
                                                     const uint16_t* __restrict__ sorted,
                                                     const DecItem* __restrict__ items,
                                                     uint16_t* __restrict__ codes,
                                                     uint8_t* __restrict__ side,
                                                     uint32_t* __restrict__ err)
{
    constexpr bool kStage = kForm == kDecStaged;
    __shared__ uint64_t lut[kStage ? (1 << kLutBits) : (1 << kLutBits) / 2];
    __shared__ uint32_t fc[kLmax + 2], cn[kLmax + 2], fi[kLmax + 2], loff[kLong];
    __shared__ uint32_t stage[kStage ? kDecStageWords + 1 + kStageSlack : 1];
    __shared__ uint16_t lsrt[kStage ? kLongSyms : kLongSyms1];
    huff_decode_wg<kForm>(ss, payload, tabs, sorted, items, codes, side, err, blockIdx.x, lut, fc, cn, fi, stage,
                           lsrt, loff);
}

// many fields: workgroups wg0 .. of job j decode field j
__device__ __forceinline__ uint32_t find_dec_job(const DecJob* __restrict__ jobs, uint32_t njobs, uint32_t wg)
{
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (jobs[m].wg0 <= wg) lo = m; else hi = m;
    }
    return lo;
}

template <int kForm>
__global__ __launch_bounds__(256) void k_huff_decode_multi(const DecJob* __restrict__ jobs, uint32_t njobs)
{
    constexpr bool kStage = kForm == kDecStaged;
    __shared__ uint64_t lut[kStage ? (1 << kLutBits) : (1 << kLutBits) / 2];
    __shared__ uint32_t fc[kLmax + 2], cn[kLmax + 2], fi[kLmax + 2], loff[kLong];
    __shared__ uint32_t stage[kStage ? kDecStageWords + 1 + kStageSlack : 1];
    __shared__ uint16_t lsrt[kStage ? kLongSyms : kLongSyms1];
    const DecJob& J = jobs[find_dec_job(jobs, njobs, blockIdx.x)];
    huff_decode_wg<kForm>(J.ss, J.payload, J.tabs, J.sorted, J.items, J.codes, (uint8_t*)J.side, J.err,
                           blockIdx.x - J.wg0, lut, fc, cn, fi, stage, lsrt, loff);
}

// many fields, every chunk: thread -> (field, chunk)
__global__ __launch_bounds__(64) void k_dec_prep_multi(const DecJob* __restrict__ jobs, uint32_t njobs,
                                                       uint32_t total)
{
    const uint32_t i = blockIdx.x;
    if (i >= total) return;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (jobs[m].chunk0 <= i) lo = m; else hi = m;
    }
    const DecJob& J = jobs[lo];
    const StreamSet& ss = J.ss;
    const uint32_t gc = i - J.chunk0;
    const uint32_t s = (uint32_t)find_stream_by_chunk(ss, gc);
    const uint32_t c = gc - ss.s[s].chunk_base;
    const StreamDesc& d = ss.s[s];
    const uint32_t sb = ((const uint32_t*)(J.payload + d.off_chunk_sub))[c];
    dec_prep_chunk(d, s, c, d.item_base + sb, chunk_start(d.g, c), J.payload, J.items);
}

}  // namespace

void launch_fill_u32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v)
{
    if (!n) return;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    k_fill_u32<<<dim3((unsigned)blocks), dim3(256), 0, s>>>(p, n, v);
}

void launch_zero_bits(hipStream_t s, const StreamSet& ss, uint8_t* payload)
{
    uint64_t mx = 0;
    for (uint32_t k = 0; k < ss.nstreams; ++k) mx = ss.s[k].bits_words > mx ? ss.s[k].bits_words : mx;
    uint64_t bx = (mx + 255) / 256;
    if (bx > 2048) bx = 2048;
    if (bx == 0) return;
    k_zero_bits<<<dim3((unsigned)bx, ss.nstreams), dim3(256), 0, s>>>(ss, payload);
}

void launch_hist_compact(hipStream_t s, const uint32_t* hist, int nstreams, uint64_t* out, uint32_t* counts)
{
    k_hist_compact<<<dim3(nstreams), dim3(1024), 0, s>>>(hist, out, counts);
}

void launch_enc_scatter(hipStream_t s, const uint32_t* list, uint32_t n, uint32_t* enc)
{
    if (!n) return;
    k_enc_scatter<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(list, n, enc);
}

void launch_huff_len(hipStream_t s, const StreamSet& ss, const uint16_t* codes, const uint32_t* enc,
                     uint8_t* payload, uint32_t* sub_pre_bits, uint32_t* sub_pre_un,
                     uint64_t* chunk_tot_bits, uint64_t* chunk_tot_un)
{
    k_huff_len<<<dim3(ss.total_chunks), dim3(256), 0, s>>>(ss, codes, enc, payload, sub_pre_bits,
                                                            sub_pre_un, chunk_tot_bits, chunk_tot_un);
}

void launch_huff_scan(hipStream_t s, const StreamSet& ss, const uint64_t* tb, const uint64_t* tu,
                      uint8_t* payload)
{
    k_huff_scan<<<dim3(ss.nstreams), dim3(1024), 0, s>>>(ss, tb, tu, payload);
}

void launch_huff_pack(hipStream_t s, const StreamSet& ss, const uint16_t* codes, const void* side,
                      const uint32_t* enc, const uint32_t* sub_pre_bits, const uint32_t* sub_pre_un,
                      uint8_t* payload)
{
    k_huff_pack<<<dim3(ss.total_chunks), dim3(256), 0, s>>>(ss, codes, (const uint8_t*)side, enc,
                                                             sub_pre_bits, sub_pre_un, payload);
}

void launch_dec_prep(hipStream_t s, const StreamSet& ss, const uint8_t* payload, const uint32_t* list,
                     uint32_t nlist, DecItem* items, uint32_t* cmap)
{
    const uint32_t n = list ? nlist : ss.total_chunks;
    if (!n) return;
    k_dec_prep<<<dim3(n), dim3(64), 0, s>>>(ss, payload, list, nlist, items, cmap);
}

void launch_hist_compact_multi(hipStream_t s, const EncJob* jobs, uint32_t njobs, int s0, int ns)
{
    if (njobs && ns > 0) k_hist_compact_multi<<<dim3(njobs * (unsigned)ns), dim3(1024), 0, s>>>(jobs, s0, ns);
}

void launch_enc_scatter_multi(hipStream_t s, const uint32_t* list, uint32_t n, const EncJob* jobs)
{
    if (n) k_enc_scatter_multi<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(list, n, jobs);
}

void launch_huff_encode_multi(hipStream_t s, const EncJob* jobs, uint32_t njobs, uint32_t total_chunks)
{
    if (!njobs || !total_chunks) return;
    k_huff_len_multi<<<dim3(total_chunks), dim3(256), 0, s>>>(jobs, njobs);
    k_huff_scan_multi<<<dim3(njobs * kStreams), dim3(1024), 0, s>>>(jobs);
    k_huff_pack_multi<<<dim3(total_chunks), dim3(256), 0, s>>>(jobs, njobs);
}

void launch_copy_segments(hipStream_t s, const uint8_t* src, const Segment* segs, uint32_t n)
{
    if (n) k_copy_segments<<<dim3(n), dim3(256), 0, s>>>(src, segs);
}

void launch_huff_decode_multi(hipStream_t s, const DecJob* jobs, uint32_t njobs, uint32_t wgs, int form)
{
    if (!njobs || !wgs) return;
    if (form == kDecStaged) k_huff_decode_multi<kDecStaged><<<dim3(wgs), dim3(256), 0, s>>>(jobs, njobs);
    else k_huff_decode_multi<kDecDense><<<dim3(wgs), dim3(256), 0, s>>>(jobs, njobs);
}

void launch_dec_prep_multi(hipStream_t s, const DecJob* jobs, uint32_t njobs, uint32_t total_chunks)
{
    if (!njobs || !total_chunks) return;
    k_dec_prep_multi<<<dim3(total_chunks), dim3(64), 0, s>>>(jobs, njobs, total_chunks);
}

void launch_build_dectabs(hipStream_t s, const DecHead* heads, const uint16_t* sorted, DecTab* tabs, int nstreams)
{
    if (nstreams > 0)
        k_build_dectabs<<<dim3((1u << kLutBits) / 256u, (unsigned)nstreams), dim3(256), 0, s>>>(heads, sorted, tabs);
}

void launch_build_dectabs_multi(hipStream_t s, const DecTabJob* jobs, uint32_t njobs, int nstreams)
{
    if (njobs && nstreams > 0)
        k_build_dectabs_multi<<<dim3((1u << kLutBits) / 256u, (unsigned)nstreams, njobs), dim3(256), 0, s>>>(jobs);
}

void launch_huff_decode(hipStream_t s, const StreamSet& ss, const uint8_t* payload, const DecTab* tabs,
                        const uint16_t* sorted, const DecItem* items, uint16_t* codes, void* side,
                        uint32_t* err, int form)
{
    if (!ss.total_wgs) return;
    if (form == kDecStaged)
        k_huff_decode<kDecStaged><<<dim3(ss.total_wgs), dim3(256), 0, s>>>(ss, payload, tabs, sorted, items, codes,
                                                                           (uint8_t*)side, err);
    else
        k_huff_decode<kDecDense><<<dim3(ss.total_wgs), dim3(256), 0, s>>>(ss, payload, tabs, sorted, items, codes,
                                                                          (uint8_t*)side, err);
}

}  // namespace stz

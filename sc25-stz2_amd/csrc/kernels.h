// kernels.h -- host-side launchers of the gfx950 kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace stz {

// Per-stream view used by the entropy-coding kernels.  Offsets *_off are byte
// offsets into the payload buffer; code_off is the element offset of the
// stream inside the codes / side arrays.
// encoder-table entries an encode workgroup stages in LDS (per stream)
constexpr uint32_t kEncWindow = 2048;

struct StreamDesc {
    uint64_t code_off;
    uint64_t n_symbols;
    uint64_t off_chunk_bit, off_chunk_unpred, off_chunk_sub;
    uint64_t off_sub_bits, off_sub_unpred, off_unpred, off_bits;
    uint64_t bits_words;     // words reserved for the bitstream (incl. padding)
    Geom g;
    uint32_t n_chunks, n_sub;
    uint32_t chunk_base;     // first global chunk id of this stream
    uint32_t sub_base;       // first global sub-chunk id of this stream
    uint32_t item_base;      // first decode work item of this stream (256-aligned)
    uint32_t item_count;     // decode work items of this stream
    uint32_t tsize;          // bytes per raw value (4 or 8)
    uint32_t enc_lo;         // encode: the symbols enc_lo .. enc_lo + enc_n - 1 have
    uint64_t n_unpred;       // unpredictable values stored in the stream
    uint32_t ipw;            // decode work items per workgroup (<= 256)
    uint32_t wg_base;        // first decode workgroup of this stream
    uint32_t cm_base;        // first entry of this stream in a chunk map
    uint32_t enc_n;          //   their encoder-table entries staged in LDS (<= 2048)
};

struct StreamSet {
    StreamDesc s[kStreams];
    uint32_t total_chunks;
    uint32_t total_sub;
    uint32_t total_items;    // padded to ipw per stream
    uint32_t nstreams;
    uint32_t total_wgs;      // decode workgroups
    uint32_t pad_[3];
};

// Multi-symbol decode table entry for one kLutBits-bit window: the codes
// that lie wholly inside the window, up to three, the first always; symbol 0
// (unpredictable) only ever as the single symbol of an entry.
//   bits  0-47  symbols 1..3, 16 bits each (unused slots zero)
//   bits 48-49  n, the number of symbols; 0: the first code is longer than
//               the window (canonical search)
//   bits 50-53  length of the first code
//   bits 54-57  total length of the n codes
constexpr int kDecNShift = 48, kDecL1Shift = 50, kDecLtShift = 54;
constexpr int kDecStageWords = 11264;  // 44 KiB LDS bitstream stage per decode workgroup

// canonical decode description of one stream: codes of length l are
// first_code[l] .. first_code[l] + count[l] - 1, their symbols
// sorted[sorted_off + first_index[l] ..] in code order
struct DecHead {
    uint32_t first_code[kLmax + 2];
    uint32_t count[kLmax + 2];
    uint32_t first_index[kLmax + 2];
    uint32_t sorted_off;
    uint32_t pad_[3];
};

// decode table of one stream (device memory)
struct DecTab {
    uint64_t lut[1 << kLutBits];       // multi-symbol entries (above)
    uint32_t lut1[1 << kLutBits];      // the first code only: (len << 16) | symbol, 0xFFFFFFFF if longer
    uint32_t first_code[kLmax + 2];
    uint32_t count[kLmax + 2];
    uint32_t first_index[kLmax + 2];
    uint32_t sorted_off;               // element offset of the sorted-symbol list
    uint32_t pad_[3];
};

struct DecItem {
    uint64_t bit0;     // absolute bit offset in the stream's bitstream
    uint64_t u0;       // index of the first unpredictable value
    uint32_t p0;       // stream position of the first symbol
    uint32_t cnt;      // symbols
    uint32_t stream;
    uint32_t bits;     // encoded bits of the sub-chunk
};

// ---- k_huff.hip ----
// one field of a multi-field decode launch (every chunk of every stream)
struct DecJob {
    StreamSet ss;
    const uint8_t* payload;
    const DecTab* tabs;
    const uint16_t* sorted;
    DecItem* items;
    uint16_t* codes;
    void* side;
    uint32_t* err;
    uint32_t wg0;      // first workgroup of this field in its decode launch
    uint32_t chunk0;   // first chunk of this field in the prep launch
};
void launch_dec_prep_multi(hipStream_t s, const DecJob* jobs, uint32_t njobs, uint32_t total_chunks);
// one field of a multi-field encode launch
struct EncJob {
    StreamSet ss;
    const uint16_t* codes;
    const void* side;
    uint32_t* enc;             // dense per-stream encoder tables (15 x 65536)
    uint8_t* payload;
    uint32_t* sub_pre_bits;
    uint32_t* sub_pre_un;
    uint64_t* chunk_tot_bits;
    uint64_t* chunk_tot_un;
    const uint32_t* hist;      // 15 x 65536 histograms
    uint64_t* compact;         // 15 x 65536 (symbol << 32 | freq)
    uint32_t* counts;          // symbols used per stream
    uint32_t chunk0;           // first chunk of this field in the per-chunk launches
};
// compaction of the histograms of streams s0 .. s0+ns-1 of every job
void launch_hist_compact_multi(hipStream_t s, const EncJob* jobs, uint32_t njobs, int s0, int ns);
// list: (field, stream << 16 | symbol, code | len << 24) triples
void launch_enc_scatter_multi(hipStream_t s, const uint32_t* list, uint32_t n, const EncJob* jobs);
// huff_len, huff_scan, huff_pack of every field (chunk prefix in chunk0)
void launch_huff_encode_multi(hipStream_t s, const EncJob* jobs, uint32_t njobs, uint32_t total_chunks);
struct Segment {
    uint64_t src_off;
    void* dst;
    uint64_t bytes;
};
void launch_copy_segments(hipStream_t s, const uint8_t* src, const Segment* segs, uint32_t n);
// Huffman decode forms: kDecStaged, the multi-symbol table with the
// workgroup's bits staged in LDS; kDecDense, the single-symbol table with
// bits from global memory
enum DecForm { kDecStaged = 0, kDecDense = 1 };
void launch_huff_decode_multi(hipStream_t s, const DecJob* jobs, uint32_t njobs, uint32_t wgs, int form);

void launch_fill_u32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v);
void launch_hist_compact(hipStream_t s, const uint32_t* hist, int nstreams, uint64_t* out,
                         uint32_t* counts);
void launch_enc_scatter(hipStream_t s, const uint32_t* list, uint32_t n, uint32_t* enc);
void launch_huff_len(hipStream_t s, const StreamSet& ss, const uint16_t* codes, const uint32_t* enc,
                     uint8_t* payload, uint32_t* sub_pre_bits, uint32_t* sub_pre_un,
                     uint64_t* chunk_tot_bits, uint64_t* chunk_tot_un);
void launch_huff_scan(hipStream_t s, const StreamSet& ss, const uint64_t* chunk_tot_bits,
                      const uint64_t* chunk_tot_un, uint8_t* payload);
void launch_huff_pack(hipStream_t s, const StreamSet& ss, const uint16_t* codes, const void* side,
                      const uint32_t* enc, const uint32_t* sub_pre_bits, const uint32_t* sub_pre_un,
                      uint8_t* payload);
void launch_zero_bits(hipStream_t s, const StreamSet& ss, uint8_t* payload);
// chunk list: nullptr = every chunk of every stream at its stream position;
// else (stream, chunk, first item relative to the stream's item_base,
// out_pos) quads: the chunk's symbols go to codes[out_pos ..] and
// cmap[cm_base + chunk] = out_pos
void launch_dec_prep(hipStream_t s, const StreamSet& ss, const uint8_t* payload,
                     const uint32_t* list, uint32_t nlist, DecItem* items, uint32_t* cmap);
// the decode tables of nstreams streams from their canonical descriptions
void launch_build_dectabs(hipStream_t s, const DecHead* heads, const uint16_t* sorted, DecTab* tabs, int nstreams);
// the tables of many stream sets in one launch: job j's nstreams streams
struct DecTabJob {
    const DecHead* heads;
    const uint16_t* sorted;
    DecTab* tabs;
};
void launch_build_dectabs_multi(hipStream_t s, const DecTabJob* jobs, uint32_t njobs, int nstreams);
void launch_huff_decode(hipStream_t s, const StreamSet& ss, const uint8_t* payload,
                        const DecTab* tabs, const uint16_t* sorted, const DecItem* items,
                        uint16_t* codes, void* side, uint32_t* err, int form);

// ---- k_field.hip ----
template <typename T>
void launch_minmax(hipStream_t s, const T* d, uint64_t n, double* partial, double* out);
template <typename T>
void launch_pad(hipStream_t s, const T* in, int nx, int ny, int nz, T* out, int px, int py, int pz);
template <typename T>
void launch_gather_low0(hipStream_t s, const T* full, int px, int py, int pz, T* v);

// The first steps of a batch compress for many fields in one launch each:
// padding to multiples of 4, the value range (partials, then one workgroup
// per field writing min and max to mm_out + 2 j), the base-grid gather.
// Job j's workgroups of each kind are [kind_wg0, kind_wg0 + kind_nwg), the
// ranges back to back in job order (a kind's total is the last job's
// wg0 + nwg).
struct ProJob {
    const void* in;       // the caller's field (nx x ny x nz)
    void* padded;         // its padded copy (px x py x pz), or null
    const void* full;     // the padded field (padded, or in)
    void* vbase;          // base grid out: full[4z][4y][4x]
    double* partial;      // 2 x mm_nwg value-range partials
    int nx, ny, nz, px, py, pz;
    uint64_t nin;
    uint32_t pad_wg0, pad_nwg, mm_wg0, mm_nwg, ga_wg0, ga_nwg;
};
template <typename T>
void launch_prologue_multi(hipStream_t s, const ProJob* jobs, uint32_t n, uint32_t pad_wgs, uint32_t mm_wgs,
                           uint32_t ga_wgs, double* mm_out);
template <typename T>
void launch_split(hipStream_t s, const T* full, int nx, int ny, int nz, T* const* sub);
template <typename T>
void launch_merge(hipStream_t s, const T* const* sub, int sx, int sy, int sz, T* full);
template <typename T>
void launch_residual_predict(hipStream_t s, int block, const T* sub, const T* ref, int dx, int dy,
                             int dz, T* out, int sign);
template <typename T>
void launch_quantize(hipStream_t s, const T* d, uint64_t n, double eb, uint16_t* codes);
template <typename T>
void launch_dequantize(hipStream_t s, const uint16_t* codes, const T* side, uint64_t n, double eb,
                       T* out);
template <typename T>
void launch_crop(hipStream_t s, const T* in, int px, int py, int ox, int oy, int oz, int sx, int sy,
                 int sz, int step, T* out);

// ---- k_stage.hip (per-stream stage entry points) ----
// codes / diff (the residual, for the unpredictable list) / rec (decoded
// value, may be null) of one block's points in row-major order
template <typename T>
void launch_residual_encode(hipStream_t s, int block, const T* sub, const T* ref, int dx, int dy, int dz, double eb,
                            uint16_t* codes, T* diff, T* rec);
// row-major codes and element-order unpredictables -> stream order
template <typename T>
void launch_block_to_stream(hipStream_t s, const uint16_t* codes, const T* unpred, uint64_t n_unpred, const Geom& g,
                            uint16_t* codes_s, T* side_s);
void launch_hist_codes(hipStream_t s, const uint16_t* codes, uint64_t n, uint32_t* hist);
// the allocation-free residual encode: residuals in the per-device scratch
// array (kScratchBytes), its address from scratch_setup (called once per
// device with that device current; 0 on success)
constexpr size_t kScratchBytes = 32u << 20;
int scratch_setup(int device, void** ptr);
template <typename T>
void launch_residual_encode_noalloc(hipStream_t s, void* scratch_ptr, int block, const T* sub, const T* ref, int dx,
                                    int dy, int dz, double eb, uint16_t* codes, T* unpred, uint64_t* count, T* rec);
template <typename T>
void launch_compact_unpred(hipStream_t s, const uint16_t* codes, const T* data, uint64_t n, T* out, uint64_t* count);
// list: (chunk, first element in codes/side) pairs
template <typename T>
void launch_chunks_to_block(hipStream_t s, const uint32_t* list, uint32_t nlist, const Geom& g, const uint16_t* codes,
                            const T* side, double eb, T* out, uint16_t* codes_out);

// ---- k_residual.hip ----
struct LevelArgs {
    double eb, recip;
    uint64_t code_off[7];   // element offsets of the 7 block streams in codes/side
    // random access into compact decode buffers: cmap[cm_off[b] + chunk] is the
    // element offset of the chunk's first symbol, kNoChunk when it was not
    // decoded (then *err is set and the point reads element 0)
    const uint32_t* cmap;
    uint32_t cm_off[7];
    uint32_t* err;
    int opt;                // predictor options (kPred* bits of device_math.h), 0 = F5's default build
};
constexpr uint32_t kNoChunk = 0xFFFFFFFFu;

// ---- k_staged.hip (dependency.cpp's staged predictors, option bit kPredStaged) ----
// One residual level in three launches (blocks 0/1/3, 2/4/5, 6).  rec:
// scratch for the reconstructions of blocks 0..5, 6 x g.dx x g.dy x g.dz
// values.  Encode: targets from the padded field at stride tstep (2 high
// level, 4 low level); codes, unpredictable residuals and the 7 streams'
// histograms (hist: 7 x 65536) as the other level kernels; out (or null):
// the level's 2x grid (2 g.dx x 2 g.dy x 2 g.dz) with ref at parity 0 and
// every block's reconstruction.  Decode: the same 2x grid into out with dims
// nx x ny x nz (points past them are not written: the padded grid, or the
// field it crops to).
template <typename T>
void launch_staged_encode(hipStream_t s, const T* full, int px, int py, int tstep, const T* ref, Geom g, LevelArgs a,
                          uint16_t* codes, T* side, uint32_t* hist, T* out, T* rec);
template <typename T>
void launch_staged_decode(hipStream_t s, const T* ref, Geom g, LevelArgs a, const uint16_t* codes, const T* side,
                          T* out, int nx, int ny, int nz, T* rec);


// one box of a batched random-access decode (inclusive bounds); workgroups
// blk0 .. next job's blk0 - 1 work on it
struct BoxJob {
    int x0, x1, y0, y1, z0, z1;
    uint32_t blk0;
    uint32_t pad_;
    void* out;              // dense output of a full-resolution box
};
template <typename T>
void launch_low_encode(hipStream_t s, const T* full, int px, int py, int pz, const T* base,
                       Geom g, LevelArgs a, uint16_t* codes, T* side, uint32_t* hist, T* mid);
template <typename T>
void launch_high_encode(hipStream_t s, const T* full, int px, int py, int pz, const T* mid, Geom g,
                        LevelArgs a, uint16_t* codes, T* side, uint32_t* hist);
// share: the persistent grid takes at most that fraction of the resident
// workgroup slots (the rest stay free for other streams' launches)
template <typename T>
void launch_low_decode(hipStream_t s, const T* base, Geom g, LevelArgs a, const uint16_t* codes,
                       const T* side, T* mid, float share = 1.0f);
template <typename T>
void launch_high_decode(hipStream_t s, const T* mid, Geom g, LevelArgs a, const uint16_t* codes,
                        const T* side, T* out, int nx, int ny, int nz, float share = 1.0f);
// random access: mid values on boxes of the mid grid -> mid array (full mid dims)
template <typename T>
void launch_low_decode_boxes(hipStream_t s, const T* base, Geom g, LevelArgs a, const uint16_t* codes,
                             const T* side, T* mid, int mx, int my, const BoxJob* jobs, uint32_t njobs,
                             uint32_t nblocks);
// random access: full-resolution boxes -> each job's dense output
template <typename T>
void launch_high_decode_boxes(hipStream_t s, const T* mid, Geom g, LevelArgs a, const uint16_t* codes,
                              const T* side, const BoxJob* jobs, uint32_t njobs, uint32_t nblocks);

// One field's level in a multi-field persistent launch: work items
// item0 .. item0 + ntiles - 1 are its main 64 x waves tiles, then nedge
// groups of one edge point per thread.
enum LevelMode { kLowEnc = 0, kHighEnc = 1, kLowDec = 2, kHighDec = 3 };
struct LevelJob {
    const void* full;     // encode: padded field
    const void* ref;      // base grid (low) or mid grid (high)
    void* mid;            // low level: mid grid written
    void* out;            // high decode: the caller's field
    uint16_t* codes;
    void* side;
    uint32_t* hist;       // encode: this field's 7 level histograms
    int px, py;           // padded dims of `full`
    int nx, ny, nz;       // high decode: output dims
    int vec;              // high decode: 2-wide output stores allowed
    Geom g;
    LevelArgs a;
    uint32_t item0, ntiles, nedge, pad_;
};
// all jobs of one mode; counter: a zeroed device word (the work queue head)
// var: some job's level arguments carry predictor options
template <typename T>
void launch_level_multi(hipStream_t s, int mode, const LevelJob* jobs, uint32_t njobs, uint32_t total_items,
                        uint32_t* counter, int waves, uint32_t grid, bool var);
// work items of one field's level for a given number of waves per block
uint32_t level_tiles(const Geom& g, int waves);
uint32_t level_edge_groups(const Geom& g, int waves);

// ---- k_lorenzo.hip ----
// Lorenzo base codec (interp == 2) and Lorenzo / regression base (interp ==
// 3): one workgroup per kLorenzoTile^3 tile of the base grid.  Encode
// reconstructs v in place and writes codes, the raw values of unpredictable
// points (side, stream positions) and the base stream's histogram; decode
// writes v from codes and side.  With reg set every tile's stream starts
// with kLorenzoHdr header symbols (four per 8^3 sub-block: its plane's
// coefficients as unpredictables, or four zero residuals for Lorenzo).
constexpr int kLorenzoTile = 16;
constexpr int kLorenzoHdr = 32;
struct LorenzoJob {
    void* v;
    uint16_t* codes;
    void* side;
    uint32_t* hist;   // encode only
    int bx, by, bz;
    int reg;          // 1: Lorenzo / regression per sub-block (interp 3)
    double eb, recip;
    double lo, hi;    // regression predictions are clamped to [lo, hi] (the field's range)
    uint32_t wg0;     // multi-field launch: the field's first workgroup
};
uint32_t lorenzo_tiles(int bx, int by, int bz);
// base stream symbols beyond the grid's points: the tiles' header symbols
inline uint64_t base_extra_symbols(int interp, int bx, int by, int bz)
{
    return interp == 3 ? (uint64_t)kLorenzoHdr * lorenzo_tiles(bx, by, bz) : 0;
}
template <typename T>
void launch_lorenzo(hipStream_t s, const LorenzoJob& j, bool encode);
template <typename T>
void launch_lorenzo_multi(hipStream_t s, const LorenzoJob* jobs, uint32_t njobs, uint32_t wgs, bool encode);

// ---- k_interp.hip ----
struct InterpPass {
    int start[3], step[3], cnt[3];
    int axis;        // -1: anchor (pred 0)
    int s;           // stride along the axis
    int interp;
    int bx, by, bz;
    double eb, recip;
    uint64_t pos;    // first code position of the pass in the base stream
    uint64_t n;      // points in the pass
};
constexpr int kMaxFusedPasses = 24;
// Coarse passes of the base codec in lattice units: every pass with stride
// s >= S0 lives on the stride-S0 sub-grid, which one workgroup keeps in LDS.
struct InterpPassList {
    int n;
    int S0;                    // lattice step in base-grid points
    int bx, by, bz;            // base grid dims
    int lx, ly, lz;            // lattice dims
    InterpPass p[kMaxFusedPasses];  // start/step/s/dims in lattice units
};
constexpr int kLatticeBytes = 128 * 1024;
// The fused kernel is one workgroup, so its lattice is also held to
// kFusedPoints (16^3): the finer levels run as passes over the whole chip
// (at 128^3, 23 us for the 16^3 lattice plus three stride-4 passes, where the
// 32^3 lattice took 86 us in one workgroup).
constexpr int kFusedPoints = 4096;
// An encode pass of many fields flushes one LDS histogram per workgroup
// with global atomics on the same bins: at most this many workgroups per
// field and pass.
constexpr unsigned kPassEncMaxWgs = 512;
// many fields: the fused coarse passes one workgroup per field; a fine pass
// of every field in one launch (workgroups wg0 .. wg0 + nwg - 1 per field)
struct FusedJob {
    InterpPassList l;
    void* v;
    const uint16_t* codes;
    const void* side;
};
struct PassJob {
    InterpPass p;
    void* v;
    const uint16_t* codes;
    const void* side;
    uint32_t wg0, nwg;
};
struct FusedEncJob {
    InterpPassList l;
    void* v;
    uint16_t* codes;
    void* side;
    uint32_t* hist;
};
struct PassEncJob {
    InterpPass p;
    void* v;
    uint16_t* codes;
    void* side;
    uint32_t* hist;
    uint32_t wg0, nwg;
};
template <typename T>
void launch_interp_encode_fused_multi(hipStream_t s, const FusedEncJob* jobs, uint32_t njobs);
template <typename T>
void launch_interp_encode_multi(hipStream_t s, const PassEncJob* jobs, uint32_t njobs, uint32_t wgs);
template <typename T>
void launch_interp_decode_fused_multi(hipStream_t s, const FusedJob* jobs, uint32_t njobs);
template <typename T>
void launch_interp_decode_multi(hipStream_t s, const PassJob* jobs, uint32_t njobs, uint32_t wgs);
template <typename T>
void launch_interp_encode_fused(hipStream_t s, const InterpPassList& l, T* v, uint16_t* codes, T* side,
                                uint32_t* hist);
template <typename T>
void launch_interp_decode_fused(hipStream_t s, const InterpPassList& l, T* v, const uint16_t* codes,
                                const T* side);
template <typename T>
void launch_interp_encode(hipStream_t s, const InterpPass& p, T* v, uint16_t* codes, T* side,
                          uint32_t* hist);
template <typename T>
void launch_interp_decode(hipStream_t s, const InterpPass& p, T* v, const uint16_t* codes,
                          const T* side);

// ---- k_tuner.hip ----
constexpr int kMlpIn = 16, kMlpHid = 64, kMlpOut = 208;
constexpr int kMlpParams = kMlpIn * kMlpHid + kMlpHid + kMlpHid * kMlpHid + kMlpHid + kMlpHid * kMlpOut + kMlpOut;
constexpr int kFeatSums = 18;   // per-sub-block sums of k_block_stats
void launch_mlp_forward(hipStream_t s, const float* P, const float* X, uint32_t n, float* Y);
void launch_mlp_grad(hipStream_t s, const float* P, const float* X, const float* T, uint32_t n, float* G,
                     float* loss);
void launch_adam(hipStream_t s, float* P, float* G, float* M, float* V, uint32_t n, float lr, float b1, float b2,
                 float eps, int step, float wd);
// 64^3 blocks' statistics are summed in this many partial workgroups each
constexpr int kStatParts = 4;
template <typename T>
void launch_block_stats(hipStream_t s, const T* f, int nx, int ny, int nz, double eb, double* out,
                        double* part);  // part: kStatParts x kFeatSums doubles per block

}  // namespace stz

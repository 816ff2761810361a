/*
 * stz.h -- C ABI of the MI355X-native STZ compress/decompress hot path.
 *
 * The reference (FabioGrosso/SC25-STZ2) has no library API: its boundary is a
 * set of internal C++ functions of the canonical program
 *   F5 = src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp
 * driven by run_typed<T>() (F5:1954-3038).  Each entry point below names the
 * reference function(s) it replaces.  All entry points are plain C: pointers,
 * sizes and integer status codes; no C++ or torch types cross the boundary.
 *
 * Memory: every field / output pointer may be host memory or HIP device
 * memory (detected with hipPointerGetAttributes).  Compressed streams
 * returned to the caller are host memory owned by the caller (stz_free).
 *
 * Threading: a context owns one HIP stream and its memory pool; calls on one
 * context are serialised by the caller, calls on different contexts are
 * independent (no process-global state, unlike F5's HuffmanEncoder forced
 * marker, F5:1264).  The library never prints or exits.
 */
#ifndef STZ_H
#define STZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define STZ_OK 0
#define STZ_E_ARG -1     /* bad argument (F5:1957-2070 usage errors) */
#define STZ_E_DIMS -2    /* dims unsupported (F5:1976-2006) */
#define STZ_E_STREAM -3  /* corrupt / wrong stream (F5:1477-1485, F5:1855-1861) */
#define STZ_E_OOM -4     /* allocation failed */
#define STZ_E_HIP -5     /* HIP runtime error */
#define STZ_E_ZSTD -6    /* host zstd failure */
#define STZ_E_NODEV -7   /* no usable GPU */
#define STZ_E_SCRATCH -8 /* input larger than the allocation-free entry point's scratch */

typedef enum { STZ_F32 = 0, STZ_F64 = 1 } stz_dtype; /* F5:3057-3076 -f / -d */

/* STZ_EB_ABS / STZ_EB_REL: the bound E = eb (REL: eb times the value range
 * of the finite values, in double) is a strict maximum of |decoded - input|
 * at every finite point.  The levels share E - 4u(M + E), u the unit
 * roundoff of the field's type and M its largest finite |value|: that much
 * covers the rounding of the reconstruction in the field's type.  A bound
 * below 4u(M + E) (a few ulps of the data) leaves 0: every value is stored,
 * and the error is then that rounding, at most 3.25u(M + E).
 * STZ_EB_REF: eb has the meaning of F5's argv[1]: level bounds are
 * mult_* x eb (F5:2139, F5:2201, F5:2259), and the decoded error is at most
 * the level bound plus the same rounding, as in the reference. */
typedef enum { STZ_EB_ABS = 0, STZ_EB_REL = 1, STZ_EB_REF = 2 } stz_eb_mode;

typedef struct {
    double mult_base, mult_low, mult_high; /* F5:2127-2129: 0.7 / 2.7 / 8.2 */
    int chunk_high, chunk_low;             /* F5:34-35 marker chunks: 32 / 16 */
    int interp;                            /* base codec: 0 linear, 1 cubic interpolation, 2 Lorenzo,
                                              3 Lorenzo or linear regression per 8^3 sub-block */
    int order;                             /* base codec axis order 0..5 */
    double alpha, beta;                    /* base codec level eb: eb/min(alpha^(l-1),beta) */
    int zstd_level;                        /* host zstd level (F5:3048: SZ3 default 3) */
    int tuner;                             /* 0 fixed params, 1 learned, 2 exhaustive trial */
    int pred;                              /* STZ_PRED_* bits: F5's predictor build options */
} stz_params;

/* Predictor options as run-time bits of stz_params.pred (0 = F5's default
 * build): F5's compile-time options, stz2.cpp's static predictors and
 * dependency.cpp's staged ones.
 * They change the interior predictions of the plane blocks (2, 4, 5) and
 * the cube block (6), and are written to the stream header, which the
 * decoder follows. */
#define STZ_PRED_WPOW2 1      /* inverse-square variance weights (F5:177-203) */
#define STZ_PRED_PLANE_AXIS 2 /* two axis-pair candidates more for blocks 2, 4, 5 (F5:579-607, F5:627-633) */
#define STZ_PRED_CUBE_AXIS 4  /* three axis-pair candidates more for block 6 (F5:609-622, F5:675-686) */
#define STZ_PRED_BICUBIC12 8  /* blocks 2, 4, 5: 12-point truncated bicubic (stz2.cpp:122-155, 186-199, 222-235) */
#define STZ_PRED_DIAG16 16    /* block 6: 16-point static predictor (stz2.cpp:271-293) */
/* dependency.cpp's staged inverse-variance predictors (dependency.cpp:162-292,
 * driven as at dependency.cpp:594-704): blocks 2, 4, 5 from the
 * reconstructed blocks 0, 1, 3, block 6 from the reconstructed 2, 4, 5.
 * Alone (no other bit with it).  Each level then runs in three stages, and
 * random access decodes the levels in full (a staged block depends on
 * reconstructed neighbours of other blocks, past the planner's footprint). */
#define STZ_PRED_STAGED 32

typedef struct {
    size_t nx, ny, nz;        /* caller's dims, x fastest */
    size_t px, py, pz;        /* padded dims (multiples of 4) */
    stz_dtype dtype;
    double eb_base, eb_low, eb_high;
    size_t payload_bytes;     /* device payload (before zstd) */
    int interp, order;
    double mult_base, mult_low, mult_high, alpha, beta;
    int pred;                 /* STZ_PRED_* bits */
    int tuner;                /* device streams: 0 parameters given, 1 the learned tuner's
                                 choice, 2 the exhaustive trial's (asked for, or the
                                 learned tuner's low-confidence fallback); payloads: 0 */
    float tune_confidence;    /* the model's confidence when it chose (0 if not asked) */
} stz_info;

typedef struct stz_ctx stz_ctx;
typedef struct stz_dstream stz_dstream; /* compressed stream resident in HBM */

/* ---- context ----------------------------------------------------------- */
int stz_create(stz_ctx** ctx, int device);
/* stz_create with flags.  STZ_CTX_OWN_QUEUE: the context's streams get
 * hardware queues of their own (streams created with a CU mask over every
 * CU), so its work never waits behind another stream's on a shared queue
 * whatever GPU_MAX_HW_QUEUES says (HIP otherwise spreads streams over that
 * many queues, 4 by default, and streams on one queue run in order).  For
 * serving contexts whose batches must overlap; each such stream holds one
 * of the GPU's hardware queue slots while it lives.
 * No counterpart in the reference, which runs one query at a time. */
#define STZ_CTX_OWN_QUEUE 1u
int stz_create_ex(stz_ctx** ctx, int device, unsigned flags);
void stz_destroy(stz_ctx* ctx);
/* run subsequent work on a caller-owned hipStream_t (NULL: the library's own) */
int stz_set_stream(stz_ctx* ctx, void* hip_stream);
void* stz_get_stream(stz_ctx* ctx);
/* Serving: this context's persistent full-level decode kernels (a
 * full-resolution or 1/2-resolution decode) take at most `share` (0 < share
 * <= 1, default 1) of the GPU's resident workgroup slots, so launches from
 * other contexts' streams find free CUs meanwhile, and its full decodes keep
 * their Huffman stage on one stream.  Results are unchanged.
 * No counterpart in the reference, which runs one query at a time. */
int stz_set_grid_share(stz_ctx* ctx, float share);
/* per-kernel timing with HIP events on the context stream */
int stz_profile(stz_ctx* ctx, int enable);
/* JSON {"kernel": [launches, total_ms], ...} of everything since enable */
int stz_profile_report(stz_ctx* ctx, char* buf, size_t cap);
void stz_default_params(stz_params* p);
const char* stz_strerror(int code);
void stz_free(void* p);

/* ---- whole path ---------------------------------------------------------
 * stz_compress replaces run_typed's compress half: slice_full_data(_parity)
 * (F5:2102-2109), SZ_compress4De (F5:1439, base), preprocess_block_zslab +
 * SZ_compress (F5:2182-2212, low), merge_sub_blocks_to_full (F5:2230),
 * fused preprocess + quantize_range_nodec + SZ_compress_finish (F5:2252-2311,
 * high).  The stream is a zstd container of the device payload; the input is
 * never modified (unlike F5:1312). */
int stz_compress(stz_ctx* ctx, stz_dtype dtype, const void* field, size_t nx, size_t ny,
                 size_t nz, double eb, stz_eb_mode mode, const stz_params* params,
                 void** stream, size_t* bytes);
/* replaces SZ_decompress4De + SZ_decompress_separated x14 + depreprocess /
 * merge (F5:2339-2451) */
int stz_decompress(stz_ctx* ctx, const void* stream, size_t bytes, void* out);
/* 3D box, inclusive corners normalised like make_query (F5:1637-1644); a == b
 * on an axis makes a slice.  Replaces the --auto-query path (F5:2460-2605)
 * built on SZ_decompress_nopred_marked_chunks (F5:1841-1911). */
int stz_decompress_roi(stz_ctx* ctx, const void* stream, size_t bytes, int x0, int y0, int z0,
                       int x1, int y1, int z1, void* out, size_t out_dims[3]);
/* axis-aligned 2D slice (axis 0=x,1=y,2=z); the disabled z-slice test of
 * F5:2678-2929 generalised to any axis */
int stz_decompress_slice(stz_ctx* ctx, const void* stream, size_t bytes, int axis, int index,
                         void* out, size_t out_dims[3]);
/* progressive resolution: 0 = 1/8 (base codec stride-2 points), 1 = 1/4
 * (base, F5:2339), 2 = 1/2 (F5:2392-2394), 3 = full */
int stz_decompress_level(stz_ctx* ctx, const void* stream, size_t bytes, int level, void* out,
                         size_t out_dims[3]);
int stz_query_info(const void* stream, size_t bytes, stz_info* info);
/* Host-only validation of a container or bare payload: every zstd frame,
 * header, codebook and chunk table is checked as a decode would check it
 * before its first kernel launch.  STZ_OK or STZ_E_STREAM; needs no GPU.
 * Replaces the nullptr returns of SZ_decompress_nopred_marked_chunks on a
 * bad stream (F5:1855-1861) with an up-front check. */
int stz_stream_check(const void* stream, size_t bytes);

/* ---- device-resident streams (benchmark, serving) ---------------------- */
int stz_compress_device(stz_ctx* ctx, stz_dtype dtype, const void* field, size_t nx, size_t ny,
                        size_t nz, double eb, stz_eb_mode mode, const stz_params* params,
                        stz_dstream** out);
int stz_dstream_upload(stz_ctx* ctx, const void* stream, size_t bytes, stz_dstream** out);
int stz_dstream_download(stz_ctx* ctx, const stz_dstream* ds, int zstd_level, void** stream,
                         size_t* bytes);
int stz_dstream_info(const stz_dstream* ds, stz_info* info);
int stz_dstream_decompress(stz_ctx* ctx, stz_dstream* ds, void* out);
int stz_dstream_decompress_roi(stz_ctx* ctx, stz_dstream* ds, int x0, int y0, int z0, int x1,
                               int y1, int z1, void* out, size_t out_dims[3]);
int stz_dstream_decompress_level(stz_ctx* ctx, stz_dstream* ds, int level, void* out,
                                 size_t out_dims[3]);
/* ---- random access in batches (serving) --------------------------------
 * A request is a 3D box (a = x0 y0 z0 x1 y1 z1, inclusive, normalised like
 * make_query F5:1637-1644), an axis slice (a = axis, index) or a level
 * (a = level as in stz_decompress_level).  One call decodes the union of the
 * requests' chunks once (SZ_decompress_nopred_marked_chunks, F5:1841-1911,
 * over all of them) and runs every request's region decode in one launch per
 * level; results are bit-identical to the requests run one by one. */
#define STZ_REQ_BOX 0
#define STZ_REQ_SLICE 1
#define STZ_REQ_LEVEL 2
typedef struct {
    int kind;
    int a[6];
} stz_request;
/* decode tables, and (keep_base != 0) the decoded base level kept on the
 * device so random-access requests skip it (F5:2339 decodes it per query) */
int stz_dstream_prepare(stz_ctx* ctx, stz_dstream* ds, int keep_base);
/* outs[i]: host or device buffer of the request's size; out_dims: 3 per request */
int stz_dstream_decompress_batch(stz_ctx* ctx, stz_dstream* ds, int n, const stz_request* reqs,
                                 void* const* outs, size_t* out_dims);
/* ---- many fields in one call ---------------------------------------------
 * Uneven fields of one dtype.  The residual levels of all fields run as one
 * persistent launch per level fed by a device work queue (fields under
 * STZ_PRED_STAGED: their own three stage launches per level); its launch
 * geometry (waves per workgroup, workgroups per CU) is timed per dtype,
 * level and size class the first time and kept in the context.  Streams and
 * decoded fields are bit-identical to one call per field.  With params.tuner
 * = 1 every field gets its own choice; the fields below the confidence
 * threshold take the exhaustive trial together.  A new set of streams
 * decodes eagerly; a second call naming the same streams and outputs (device
 * memory) captures the decode as a HIP graph, replayed from then on.
 * Replaces per-field runs of the compress / decompress halves of run_typed
 * (F5:2102-2451). */
int stz_compress_batch(stz_ctx* ctx, stz_dtype dtype, int n, const void* const* fields,
                       const size_t* dims /* 3 per field: nx ny nz */, double eb, stz_eb_mode mode,
                       const stz_params* params, stz_dstream** out);
int stz_decompress_batch(stz_ctx* ctx, int n, stz_dstream* const* ds, void* const* outs);
/* enable (1) / disable (0) the geometry timing (-1: leave); JSON of the
 * choices and their trial times into report when given */
int stz_autotune(stz_ctx* ctx, int enable, char* report, size_t cap);

/* ---- learned tuner -------------------------------------------------------
 * Chooses the residual predictor set (F5's; F5's PRED_PLANE_AXIS +
 * PRED_CUBE_AXIS; stz2.cpp's BICUBIC12 + DIAG16; dependency.cpp's staged
 * inverse-variance predictors, STZ_PRED_STAGED), the base codec
 * (interpolator and axis order, or Lorenzo / regression, interp 3) and the
 * level-eb scaling (the tuning targets of F5:2120-2129 and SZ3's
 * sample-and-trial between interpolation and Lorenzo / regression, F5:1439,
 * sz.cpp:63-76) among STZ_TUNE_CONFIGS configurations: 4 predictor sets x
 * (2 interpolators x 6 orders x 4 scalings + Lorenzo / regression x 3
 * scalings), index 51 * set + j, j = 24 * interp + 4 * order + scaling
 * (j < 48) or 48 + scaling (Lorenzo / regression).  The model has STZ_MLP_OUTPUTS outputs,
 * the first STZ_TUNE_CONFIGS of which are the configurations.
 * The choice lands in the stream header; decoding and the error bound are
 * unchanged.  params.tuner = 1 compresses with the model's choice (falling
 * back to the exhaustive trial below the model's confidence threshold),
 * params.tuner = 2 with the trial's.  Tuning needs an ABS or REL bound. */
#define STZ_TUNE_CONFIGS 204
#define STZ_MLP_OUTPUTS 208
#define STZ_TUNE_FEATURES 16
#define STZ_MLP_PARAMS 18768 /* 16x64 + 64 + 64x64 + 64 + 64x208 + 208 fp32 */
int stz_tune_config(int index, stz_params* p);
/* 16 statistics per 64^3 sub-block (x fastest block order) */
int stz_tune_features(stz_ctx* ctx, stz_dtype dtype, const void* field, size_t nx, size_t ny, size_t nz,
                      double eb_abs, float* feats, size_t cap_blocks, size_t* n_blocks);
/* host fp32 parameters; confidence below min_confidence falls back to the trial */
int stz_tune_set_model(stz_ctx* ctx, const float* params, size_t n, float min_confidence);
/* the model's configuration: fills pred / interp / order / mult_* / alpha / beta */
int stz_tune(stz_ctx* ctx, stz_dtype dtype, const void* field, size_t nx, size_t ny, size_t nz, double eb,
             stz_eb_mode mode, stz_params* inout, float* confidence, int* config);
/* exhaustive trial: payload bytes of every configuration (sizes may be NULL) */
int stz_tune_trial(stz_ctx* ctx, stz_dtype dtype, const void* field, size_t nx, size_t ny, size_t nz,
                   double eb, stz_eb_mode mode, const stz_params* base, uint64_t* sizes, int* best);
/* MLP kernels (bf16 MFMA, fp32 accumulation; device pointers): logits of n
 * feature rows; mean soft-target cross entropy and its gradient (overwrites
 * grad and loss); one Adam step (beta 0.9 / 0.999, eps 1e-8) that clears grad */
int stz_mlp_forward(stz_ctx* ctx, const float* params, const float* x, int n, float* logits);
int stz_mlp_grad(stz_ctx* ctx, const float* params, const float* x, const float* target, int n,
                 float* grad, float* loss);
int stz_mlp_adam(stz_ctx* ctx, float* params, float* grad, float* m, float* v, int step, float lr,
                 float weight_decay);
/* device payload pointer and size (pre-zstd layout) */
int stz_dstream_payload(const stz_dstream* ds, const void** d_payload, size_t* bytes);
void stz_dstream_free(stz_dstream* ds);

/* ---- stage entry points (parity tests of single kernels) ----------------- */
/* slice_full_data (F5:107-122): sub[k] gets dims (nx/2, ny/2, nz/2) */
int stz_split(stz_ctx* ctx, stz_dtype dtype, const void* full, size_t nx, size_t ny, size_t nz,
              void* const sub[8]);
/* merge_sub_blocks_to_full (F5:68-90) */
int stz_merge(stz_ctx* ctx, stz_dtype dtype, const void* const sub[8], size_t sx, size_t sy,
              size_t sz, void* full);
/* preprocess_block_zslab over all slabs (F5:1161-1175): diff = sub - pred(ref) */
int stz_residual_predict(stz_ctx* ctx, stz_dtype dtype, int block, const void* sub,
                         const void* ref, size_t dx, size_t dy, size_t dz, void* diff);
/* depreprocess_block_zslab (F5:1177-1191): out = de + pred(ref) */
int stz_residual_reconstruct(stz_ctx* ctx, stz_dtype dtype, int block, const void* de,
                             const void* ref, size_t dx, size_t dy, size_t dz, void* out);
/* quantize_range_nodec (F5:1204-1253) with radius 32768; codes are uint16;
 * unpredictable values in element order */
int stz_quantize(stz_ctx* ctx, stz_dtype dtype, const void* data, size_t n, double eb,
                 uint16_t* codes, void* unpred, size_t* n_unpred);
/* decode_marked_transform's recover LUT (F5:1531-1544) */
int stz_dequantize(stz_ctx* ctx, stz_dtype dtype, const uint16_t* codes, size_t n, double eb,
                   const void* unpred, void* out);
/* Fused predict + quantize of one residual block (the two halves of the
 * reference's per-block compress: preprocess_block_zslab, F5:1161-1175, then
 * SZ_compress's quantizer, F5:1314-1397): codes[i] (row-major, 0 =
 * unpredictable) of diff = sub - pred(ref) under bound eb; the unpredictable
 * residuals in element order into unpred (capacity dx*dy*dz values), their
 * count in n_unpred.  rec (may be NULL) receives every point's decoded value
 * dec + pred: the closed-loop reconstruction the low level writes back
 * (writeDec, F5:1381-1385, then F5:1077-1083).  The level only sets eb
 * (mult_low / mult_high x the unit, F5:2201, F5:2259), so it is not an
 * argument here. */
int stz_residual_encode(stz_ctx* ctx, stz_dtype dtype, int block, const void* sub, const void* ref, size_t dx,
                        size_t dy, size_t dz, double eb, uint16_t* codes, void* unpred, size_t* n_unpred,
                        void* rec);
/* Its inverse: out = dequantize(codes, unpred) + pred(ref), the recover of
 * SZ_decompress_separated (F5:1531-1562) followed by depreprocess_block_zslab
 * (F5:1177-1191); unpred in element order (NULL when no code is 0). */
int stz_residual_decode(stz_ctx* ctx, stz_dtype dtype, int block, const uint16_t* codes, const void* unpred,
                        const void* ref, size_t dx, size_t dy, size_t dz, double eb, void* out);
/* SZ_compress_finish (F5:1258-1307): one (dx,dy,dz) block's row-major codes
 * plus its unpredictable values (element order, n_unpred = the number of zero
 * codes) -> one chunk-marked stream of `chunk`^3 chunks: a stream header, the
 * canonical codebook, the per-chunk bit offsets (the markers, F5:50-55) and
 * unpredictable counts, the per-sub-chunk tables and the bitstream (this
 * project's stream format, as inside a payload).  Histogram, codebook and
 * marked encode run on the device.  zstd is not applied (stz_compress's
 * container does that per stream).  *stream is host memory (stz_free). */
int stz_huff_encode_chunks(stz_ctx* ctx, stz_dtype dtype, const uint16_t* codes, const void* unpred,
                           size_t n_unpred, size_t dx, size_t dy, size_t dz, int chunk, double eb, void** stream,
                           size_t* bytes);
/* SZ_decompress_nopred_marked_chunks (F5:1841-1911): decode only the chunks
 * chunk_ids[0 .. n_ids) (row-major ids (cz*ncy + cy)*ncx + cx, F5:1578-1582)
 * of a stream from stz_huff_encode_chunks and write their recovered values
 * (and, when codes_out is not NULL, their codes) at their points of the
 * row-major block `out` (dx*dy*dz values of the stream's dtype); other points
 * keep what they held.  chunk_ids NULL or n_ids < 0 decodes every chunk
 * (SZ_decompress_separated, F5:1466-1576).  decoded_symbols and unpred_zeros
 * count what the reference counts (F5:1870-1905); unlike the reference,
 * which consumes its row-major unpredictable list in chunk order, every
 * unpredictable point gets its own stored value. */
int stz_huff_decode_chunks(stz_ctx* ctx, const void* stream, size_t bytes, const int64_t* chunk_ids, int64_t n_ids,
                           void* out, uint16_t* codes_out, size_t* decoded_symbols, size_t* unpred_zeros);
/* Allocation-free variant of stz_residual_encode.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It allocates nothing (no hipMalloc / hipFree, no host staging, no
 * synchronisation): the residuals the unpredictable compaction reads live in
 * a per-GPU static device array of STZ_NOALLOC_SCRATCH_BYTES, so a block may
 * hold at most that many bytes of values (8 Mi f32 or 4 Mi f64 points; larger
 * is STZ_E_SCRATCH).  Every pointer is device memory of the context's device
 * (STZ_E_ARG otherwise); the unpredictable count goes to the device word
 * *n_unpred.  The work is enqueued on the context's stream (stz_get_stream)
 * and the outputs are ready when it completes.  All calls on one GPU share
 * the scratch, which keeps the last call's contents: calls that overlap on
 * one GPU (two contexts' streams, or two graph replays) race. */
#define STZ_NOALLOC_SCRATCH_BYTES (32u << 20)
int stz_residual_encode_noalloc(stz_ctx* ctx, stz_dtype dtype, int block, const void* sub, const void* ref,
                                size_t dx, size_t dy, size_t dz, double eb, uint16_t* codes, void* unpred,
                                uint64_t* n_unpred, void* rec);
/* Huffman code lengths (HuffmanEncoder::preprocess_encode_fast, F5:1271) --
 * this project's canonical length-limited code */
int stz_huff_lengths(const uint64_t* freq, int nsym, int lmax, uint8_t* len);
/* build_residual_plan / build_low_residual_plan (F5:1689-1830) for full dims
 * (fx,fy,fz): query[7] = x0 x1 y0 y1 z0 z1 is_slice; per block b (7 rows of
 * 8 ints): valid x0 x1 y0 y1 z0 z1 nchunks, chunk ids at chunks[b*cap ..] */
int stz_plan_query(int fx, int fy, int fz, int chunk_high, int chunk_low, int ax, int ay, int az,
                   int bx, int by, int bz, int* query, int* high_plan, int64_t* high_chunks,
                   int* low_plan, int64_t* low_chunks, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif

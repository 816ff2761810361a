/*
 * stz_oracle.h -- CPU oracle for the STZ compress/decompress hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the parity tests,
 * the smoke test and the CPU-baseline leg of the benchmark compare against.
 * The product (the HIP library under sc25-stz2_amd/) never links, loads or
 * calls it.
 *
 * It restates, in plain C11 + OpenMP, the canonical reference variant
 *   src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp
 * (cited below as F5:<line>) for the parts the reference defines itself:
 * parity split/merge, the 7 residual predictors, the linear quantizer and its
 * dequantization LUT, the ROI/slice access planners.  The parts that live in
 * the un-vendored SZ3-ncp dependency (interpolation base codec, Huffman coder,
 * stream layout) are this project's own design and are restated here from the
 * design, byte for byte as the GPU library writes them.
 */
#ifndef STZ_ORACLE_H
#define STZ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_F32 = 0, ORA_F64 = 1 };
/* base codec (ora_params.interp): interpolation 0 linear / 1 cubic; 2 Lorenzo
 * over independent ORA_LZ_TILE^3 tiles; 3 Lorenzo or linear regression per
 * 8^3 sub-block of those tiles, ORA_LZ_HDR header symbols per tile */
enum { ORA_BASE_LINEAR = 0, ORA_BASE_CUBIC = 1, ORA_BASE_LORENZO = 2, ORA_BASE_LZREG = 3, ORA_LZ_TILE = 16,
       ORA_LZ_HDR = 32 };
enum { ORA_EB_ABS = 0, ORA_EB_REL = 1, ORA_EB_REF = 2 };

typedef struct {
    double mult_base, mult_low, mult_high;   /* F5:2127-2129 defaults 0.7/2.7/8.2 */
    int chunk_high, chunk_low;               /* F5:34-35 defaults 32/16 */
    int interp;                              /* base codec: 0 linear, 1 cubic interpolation, 2 Lorenzo, 3 Lorenzo/regression */
    int order;                               /* base codec axis order 0..5 */
    double alpha, beta;                      /* base codec level eb scaling */
    int pred;                                /* ORA_PRED_* bits: F5's compile-time predictor options */
} ora_params;

/* Predictor options, as run-time bits: F5's build options PRED_WPOW2
 * (F5:177-203), PRED_PLANE_AXIS (F5:579-607, F5:627-633), PRED_CUBE_AXIS
 * (F5:609-622, F5:675-686); and the static predictors of stz2.cpp for the
 * plane blocks 2/4/5 (12-point truncated bicubic, stz2.cpp:122-155,
 * 186-199, 222-235) and the cube block 6 (16 points, stz2.cpp:271-293),
 * which replace the dynamic ones of those blocks; and dependency.cpp's
 * staged inverse-variance predictors (dependency.cpp:162-292, 594-704):
 * blocks 2, 4, 5 from the reconstructed blocks 0, 1, 3 and block 6 from the
 * reconstructed 2, 4, 5, alone (no other bit with it) */
enum { ORA_PRED_WPOW2 = 1, ORA_PRED_PLANE_AXIS = 2, ORA_PRED_CUBE_AXIS = 4, ORA_PRED_BICUBIC12 = 8,
       ORA_PRED_DIAG16 = 16, ORA_PRED_STAGED = 32 };
#define ORA_PRED_ALL 63

void ora_default_params(ora_params* p);

/* ---- whole path ------------------------------------------------------- */
/* Compress a field (x fastest) into the device payload layout (pre-zstd).
 * Returns payload bytes (>0) and a malloc'd buffer in *payload, or <0 on error. */
long long ora_compress(int dtype, const void* field, size_t nx, size_t ny, size_t nz,
                       double eb, int eb_mode, const ora_params* p, uint8_t** payload);
int ora_decompress(const uint8_t* payload, size_t bytes, void* out);
/* inclusive corners, normalised like F5:1622-1644 (a == b on an axis => slice) */
int ora_decompress_roi(const uint8_t* payload, size_t bytes, int x0, int y0, int z0,
                       int x1, int y1, int z1, void* out, size_t* out_dims /*3*/);
/* level 0: 1/8, 1: 1/4 (base), 2: 1/2 (mid), 3: full */
int ora_decompress_level(const uint8_t* payload, size_t bytes, int level, void* out,
                         size_t* out_dims /*3*/);
void ora_free(void* p);

/* ---- stage functions (parity tests of single kernels) ------------------ */
void ora_split_f32(const float* full, float* sub[8], int nx, int ny, int nz);
void ora_split_f64(const double* full, double* sub[8], int nx, int ny, int nz);
void ora_merge_f32(float* const sub[8], float* full, int sx, int sy, int sz);
void ora_merge_f64(double* const sub[8], double* full, int sx, int sy, int sz);
/* diff = sub - pred(ref) for residual block 0..6 over the whole block */
void ora_preprocess_f32(int block, const float* sub, const float* ref, float* diff,
                        int dx, int dy, int dz);
void ora_preprocess_f64(int block, const double* sub, const double* ref, double* diff,
                        int dx, int dy, int dz);
/* the same under F5's predictor options (pv: ORA_PRED_* bits) */
void ora_preprocess_pv_f32(int block, int pv, const float* sub, const float* ref, float* diff,
                           int dx, int dy, int dz);
void ora_preprocess_pv_f64(int block, int pv, const double* sub, const double* ref, double* diff,
                           int dx, int dy, int dz);
void ora_depreprocess_pv_f32(int block, int pv, const float* de, const float* ref, float* out,
                             int dx, int dy, int dz);
void ora_depreprocess_pv_f64(int block, int pv, const double* de, const double* ref, double* out,
                             int dx, int dy, int dz);
/* out = de + pred(ref) */
void ora_depreprocess_f32(int block, const float* de, const float* ref, float* out,
                          int dx, int dy, int dz);
void ora_depreprocess_f64(int block, const double* de, const double* ref, double* out,
                          int dx, int dy, int dz);
/* linear quantizer; writes dec (decompressed value or raw d) when dec != NULL,
 * appends unpredictable values in element order; returns the unpred count */
size_t ora_quantize_f32(const float* d, size_t n, double eb, int radius, uint16_t* codes,
                        float* unpred, float* dec);
size_t ora_quantize_f64(const double* d, size_t n, double eb, int radius, uint16_t* codes,
                        double* unpred, double* dec);
void ora_dequantize_f32(const uint16_t* codes, size_t n, double eb, int radius,
                        const float* unpred, float* out);
void ora_dequantize_f64(const uint16_t* codes, size_t n, double eb, int radius,
                        const double* unpred, double* out);
/* canonical, length-limited Huffman code lengths for a 65536-symbol histogram */
int ora_huff_lengths(const uint64_t* freq, int nsym, int lmax, uint8_t* len);
/* base codec on one grid (in-place reconstruction), codes in stream order;
 * [lo, hi] clamps the regression predictions of interp 3 (the field's range) */
size_t ora_interp_compress_f32(float* v, int bx, int by, int bz, double eb, int interp,
                               int order, double alpha, double beta, double lo, double hi,
                               uint16_t* codes, float* unpred);
size_t ora_interp_compress_f64(double* v, int bx, int by, int bz, double eb, int interp,
                               int order, double alpha, double beta, double lo, double hi,
                               uint16_t* codes, double* unpred);
/* its decode: v (zero-filled by the caller) from codes and the unpredictables */
void ora_interp_decompress_f32(float* v, int bx, int by, int bz, double eb, int interp, int order,
                               double alpha, double beta, double lo, double hi, const uint16_t* codes,
                               const float* unpred);
void ora_interp_decompress_f64(double* v, int bx, int by, int bz, double eb, int interp, int order,
                               double alpha, double beta, double lo, double hi, const uint16_t* codes,
                               const double* unpred);

/* one residual level with dependency.cpp's staged predictors: ref the
 * coarser grid (dx, dy, dz), g its 2x grid (filled: ref at parity 0, each
 * block's reconstruction at its parity); encode with tgt7 (the 7 target
 * blocks concatenated, grid order) into q7 / dv7 (codes, decoded residuals),
 * or decode res7 (the residuals) when tgt7 is NULL */
void ora_staged_level_f32(const float* ref, int dx, int dy, int dz, double eb, float* g, const float* tgt7,
                          uint16_t* q7, float* dv7, const float* res7);
void ora_staged_level_f64(const double* ref, int dx, int dy, int dz, double eb, double* g, const double* tgt7,
                          uint16_t* q7, double* dv7, const double* res7);

/* ROI planner (F5:1602-1830).  Fills per-block plans for the high and low
 * levels: for each of 7 blocks, [valid, x0,x1,y0,y1,z0,z1, nchunks] and the
 * row-major chunk ids into chunk_out (capacity per block = cap).  Returns 0. */
int ora_plan_query(int fx, int fy, int fz, int chunk_high, int chunk_low,
                   int ax, int ay, int az, int bx, int by, int bz,
                   int* query /*7: x0 x1 y0 y1 z0 z1 is_slice*/,
                   int* high_plan /*7x8*/, int64_t* high_chunks, int* low_plan /*7x8*/,
                   int64_t* low_chunks, int64_t cap);

/* payload inspection helpers */
int ora_payload_info(const uint8_t* payload, size_t bytes, double* eb_levels /*3*/,
                     int64_t* stream_bytes /*15*/, int64_t* n_unpred /*15*/);

#ifdef __cplusplus
}
#endif
#endif

/*
 * stz_format.h -- byte layout of the STZ device payload, as restated by the
 * oracle.  TEST INFRASTRUCTURE: the GPU library carries its own definition of
 * the same layout; the parity tests compare the two byte for byte.
 *
 * The reference stores each residual stream as an SZ3-ncp NOPRED blob
 * (quantizer section + Huffman tree + marked bitstream + zstd, F5:1258-1307).
 * SZ3-ncp is not vendored, so the layout below is this project's own; it keeps
 * the reference's information content: per-stream eb/radius (F5:1277-1289),
 * unpredictable values, a Huffman codebook, a per-chunk bit-offset table (the
 * reference's "markers", F5:50-55) and the bitstream.
 */
#ifndef STZ_FORMAT_H
#define STZ_FORMAT_H
#include <stdint.h>

#define STZF_MAGIC 0x325A5453u        /* "STZ2" */
#define STZF_STREAM_MAGIC 0x4D525453u /* "STRM" */
#define STZF_VERSION 1u
#define STZF_RADIUS 32768             /* quantbinCnt/2, SZ3 default 65536 bins */
#define STZF_LMAX 24                  /* Huffman length limit */
#define STZF_SUB 1024                 /* symbols per decode sub-chunk */
#define STZF_BASE_CHUNK 32768         /* symbols per base-stream chunk */
#define STZF_NSTREAMS 15              /* 1 base + 7 low + 7 high */

typedef struct {
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
    uint64_t pred;          /* predictor options (ORA_PRED_* bits), 0 = F5's default build */
    uint64_t reserved[2];
    uint64_t stream_off[STZF_NSTREAMS];
    uint64_t stream_bytes[STZF_NSTREAMS];
    uint8_t pad[48];
} stzf_file_header; /* 512 bytes */

typedef struct {
    uint32_t magic, kind, block, dtype; /* kind: 0 base, 1 low, 2 high */
    uint64_t n_symbols;
    uint32_t dx, dy, dz, chunk;
    uint32_t n_chunks, n_sub, n_used, lmax;
    uint64_t n_unpred, total_bits;
    double eb;
    uint32_t radius, sub_size;
    uint64_t off_codebook, off_chunk_bit, off_chunk_unpred, off_chunk_sub;
    uint64_t off_sub_bits, off_sub_unpred, off_unpred, off_bits;
    uint64_t stream_bytes;
} stzf_stream_header; /* 160 bytes */

#ifdef __cplusplus
static_assert(sizeof(stzf_file_header) == 512, "file header size");
static_assert(sizeof(stzf_stream_header) == 160, "stream header size");
#else
_Static_assert(sizeof(stzf_file_header) == 512, "file header size");
_Static_assert(sizeof(stzf_stream_header) == 160, "stream header size");
#endif

#endif

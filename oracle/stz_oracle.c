/*
 * stz_oracle.c -- CPU oracle for the STZ hot path.  TEST INFRASTRUCTURE ONLY:
 * the parity tests, smoke() and the benchmark's cpu_baseline leg use it as the
 * checker; the product library never links or calls it.
 *
 * Restates F5 = src/zero_all_line_dynamic_random_access_optfast3_optcr_lb_fastdec5.cpp
 * (split/merge, predictors, quantizer, planners, pipeline order) and, for
 * what F5 delegates to the un-vendored SZ3-ncp (base codec, Huffman, stream
 * layout), this project's own design, byte-identical with the GPU library.
 *
 * Pinning (see tests/): predictors, quantizer, split/merge and planners are
 * checked bit-for-bit against F5's own source lines compiled by
 * oracle/build_ref.sh (when /root/reference is present) and against the
 * golden vectors that script wrote to tests/golden/; Huffman against brute
 * force optimality and Kraft equality; the base codec against closed-form
 * polynomial reproduction and the error bound.
 */
#define _GNU_SOURCE
#include "stz_oracle.h"
#include "stz_format.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

#define OCAT_(a, b) a##b
#define OCAT(a, b) OCAT_(a, b)

/* F5:3046-3047 sets x86 FTZ+DAZ process-wide; MXCSR is per thread, so every
 * parallel region sets it on entry and restores the caller's state on exit. */
static inline unsigned ora_ftz_enter(void)
{
    unsigned m = _mm_getcsr();
    _mm_setcsr(m | 0x8040u);
    return m;
}
static inline void ora_ftz_leave(unsigned m) { _mm_setcsr(m); }

static inline uint64_t align8(uint64_t v) { return (v + 7) & ~(uint64_t)7; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------ */
/* geometry                                                                  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int dx, dy, dz;
    int chunk;               /* residual: C; base: 0 (linear chunks of STZF_BASE_CHUNK) */
    int ncx, ncy, ncz;
    uint32_t n_chunks, n_sub;
    uint64_t n;
    uint32_t* chunk_sub;     /* [n_chunks+1] first sub-chunk of each chunk */
    uint64_t* chunk_start;   /* [n_chunks+1] stream position of each chunk */
} ora_geom;

typedef struct {
    size_t nx, ny, nz, px, py, pz;
    int chunk_high, chunk_low;
    ora_geom geom[STZF_NSTREAMS];
} ora_layout;

static inline uint64_t ora_stream_pos(const ora_geom* g, int x, int y, int z)
{
    const int C = g->chunk;
    const int cx = x / C, cy = y / C, cz = z / C;
    const int wx = imin(C, g->dx - cx * C), wy = imin(C, g->dy - cy * C), wz = imin(C, g->dz - cz * C);
    const uint64_t start = (uint64_t)cz * C * g->dx * g->dy + (uint64_t)cy * C * g->dx * wz +
                           (uint64_t)cx * C * wy * wz;
    return start + ((uint64_t)(z - cz * C) * wy + (uint64_t)(y - cy * C)) * wx + (uint64_t)(x - cx * C);
}

/* extra: symbols the stream holds beyond its grid points (the header
 * symbols of a Lorenzo/regression base) */
static void geom_init(ora_geom* g, int dx, int dy, int dz, int chunk, uint64_t extra)
{
    memset(g, 0, sizeof(*g));
    g->dx = dx; g->dy = dy; g->dz = dz; g->chunk = chunk;
    g->n = (uint64_t)dx * dy * dz + extra;
    if (chunk > 0) {
        g->ncx = (dx + chunk - 1) / chunk;
        g->ncy = (dy + chunk - 1) / chunk;
        g->ncz = (dz + chunk - 1) / chunk;
        g->n_chunks = (uint32_t)g->ncx * g->ncy * g->ncz;
    } else {
        g->n_chunks = (uint32_t)((g->n + STZF_BASE_CHUNK - 1) / STZF_BASE_CHUNK);
    }
    g->chunk_sub = (uint32_t*)malloc((g->n_chunks + 1) * sizeof(uint32_t));
    g->chunk_start = (uint64_t*)malloc((g->n_chunks + 1) * sizeof(uint64_t));
    uint32_t sub = 0;
    uint64_t pos = 0;
    for (uint32_t c = 0; c < g->n_chunks; ++c) {
        uint64_t vol;
        if (chunk > 0) {
            const int cx = (int)(c % g->ncx), cy = (int)((c / g->ncx) % g->ncy), cz = (int)(c / ((uint32_t)g->ncx * g->ncy));
            vol = (uint64_t)imin(chunk, dx - cx * chunk) * imin(chunk, dy - cy * chunk) * imin(chunk, dz - cz * chunk);
        } else {
            vol = g->n - pos < STZF_BASE_CHUNK ? g->n - pos : STZF_BASE_CHUNK;
        }
        g->chunk_sub[c] = sub;
        g->chunk_start[c] = pos;
        sub += (uint32_t)((vol + STZF_SUB - 1) / STZF_SUB);
        pos += vol;
    }
    g->chunk_sub[g->n_chunks] = sub;
    g->chunk_start[g->n_chunks] = pos;
    g->n_sub = sub;
}

static void layout_free(ora_layout* L)
{
    for (int s = 0; s < STZF_NSTREAMS; ++s) {
        free(L->geom[s].chunk_sub);
        free(L->geom[s].chunk_start);
    }
}

static size_t pad4(size_t n) { size_t p = (n + 3) & ~(size_t)3; return p < 4 ? 4 : p; }

/* header symbols of a Lorenzo/regression base grid: ORA_LZ_HDR per tile */
static uint64_t ora_base_extra(int interp, int bx, int by, int bz)
{
    if (interp != ORA_BASE_LZREG) return 0;
    const uint64_t t = (uint64_t)((bx + ORA_LZ_TILE - 1) / ORA_LZ_TILE) * ((by + ORA_LZ_TILE - 1) / ORA_LZ_TILE) *
                       ((bz + ORA_LZ_TILE - 1) / ORA_LZ_TILE);
    return t * ORA_LZ_HDR;
}

static int ora_layout_dims(ora_layout* L, size_t nx, size_t ny, size_t nz, size_t px, size_t py,
                           size_t pz, int chunk_high, int chunk_low, int interp)
{
    memset(L, 0, sizeof(*L));
    L->nx = nx; L->ny = ny; L->nz = nz;
    L->px = px; L->py = py; L->pz = pz;
    L->chunk_high = chunk_high; L->chunk_low = chunk_low;
    const int lx = (int)(px / 4), ly = (int)(py / 4), lz = (int)(pz / 4);
    geom_init(&L->geom[0], lx, ly, lz, 0, ora_base_extra(interp, lx, ly, lz));
    for (int b = 0; b < 7; ++b) {
        geom_init(&L->geom[1 + b], lx, ly, lz, chunk_low, 0);
        geom_init(&L->geom[8 + b], (int)(px / 2), (int)(py / 2), (int)(pz / 2), chunk_high, 0);
    }
    return 0;
}

static int ora_make_layout(ora_layout* L, size_t nx, size_t ny, size_t nz, const ora_params* p)
{
    if (nx == 0 || ny == 0 || nz == 0) return -1;
    if (p->chunk_high <= 0 || p->chunk_low <= 0) return -1;
    return ora_layout_dims(L, nx, ny, nz, pad4(nx), pad4(ny), pad4(nz), p->chunk_high, p->chunk_low, p->interp);
}

/* axis orders of the base codec: axes listed first-to-last inside a level (0=x,1=y,2=z) */
static void ora_axis_order(int order, int ord[3])
{
    static const int tab[6][3] = {{2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 0, 2}, {0, 2, 1}, {0, 1, 2}};
    const int o = (order >= 0 && order < 6) ? order : 0;
    ord[0] = tab[o][0]; ord[1] = tab[o][1]; ord[2] = tab[o][2];
}

static int ora_interp_levels(int bx, int by, int bz)
{
    const int m = imax(bx, imax(by, bz));
    int L = 0;
    while ((1 << L) < m) L++;
    return L;
}

/* level l (1 = finest) error-bound divisor: min(alpha^(l-1), beta) */
static double ora_level_div(double alpha, double beta, int l)
{
    const double a = pow(alpha, (double)(l - 1));
    return a < beta ? a : beta;
}

void ora_default_params(ora_params* p)
{
    p->mult_base = 0.7;  /* F5:2127 */
    p->mult_low = 2.7;   /* F5:2128 */
    p->mult_high = 8.2;  /* F5:2129 */
    p->chunk_high = 32;  /* F5:34 roi_high_dim */
    p->chunk_low = 16;   /* F5:35 roi_low_dim */
    p->interp = 1;
    p->order = 0;
    p->alpha = 1.0;
    p->beta = 1.0;
    p->pred = 0;         /* F5's default build: no PRED_* option */
}

/* ABS/REL: the requested bound E is the maximum pointwise error.  The
 * reconstruction's two roundings in T (residual and value) and the rounded
 * difference the quantizer checks add at most 3.25 u (M + E), u the unit
 * roundoff of T and M = max |value|, so the levels share E - 4 u (M + E)
 * (0 when the bound is below that), and the level unit is that over
 * max(mult).  A Lorenzo prediction sums seven reconstructed values
 * (|prediction| <= 7 (M + E)), so the Lorenzo bases (interp 2 and 3, whose
 * regression predictions are clamped to the value range) take
 * E - 10 u (M + E).  REF keeps F5's CLI meaning (argv[1] times the level
 * multiplier, F5:2139, F5:2201, F5:2259). */
static void ora_level_ebs(double eb, int mode, double vmin, double vmax, int dtype, const ora_params* p,
                          double ebl[3], double* unit)
{
    if (mode == ORA_EB_REF) {
        *unit = eb;
    } else {
        const double E = (mode == ORA_EB_REL) ? eb * (vmax - vmin) : eb;
        const double M = fabs(vmin) > fabs(vmax) ? fabs(vmin) : fabs(vmax);
        const double u = dtype == ORA_F32 ? 0x1p-24 : 0x1p-53;
        double Es = E - (p->interp >= ORA_BASE_LORENZO ? 10.0 : 4.0) * u * (M + E);
        if (!(Es > 0)) Es = 0;
        double mm = p->mult_base;
        if (p->mult_low > mm) mm = p->mult_low;
        if (p->mult_high > mm) mm = p->mult_high;
        *unit = Es / mm;
    }
    ebl[0] = p->mult_base * *unit;
    ebl[1] = p->mult_low * *unit;
    ebl[2] = p->mult_high * *unit;
}

/* ------------------------------------------------------------------------ */
/* Huffman: two-queue construction with deterministic ties, JPEG Annex K.3   */
/* length limiting, DEFLATE-style canonical codes, MSB-first packing.        */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t f; int sym; } leaf_t;

static int leaf_cmp(const void* a, const void* b)
{
    const leaf_t* x = (const leaf_t*)a;
    const leaf_t* y = (const leaf_t*)b;
    if (x->f != y->f) return x->f < y->f ? -1 : 1;
    return x->sym - y->sym;
}

int ora_huff_lengths(const uint64_t* freq, int nsym, int lmax, uint8_t* len)
{
    memset(len, 0, (size_t)nsym);
    int n = 0;
    for (int s = 0; s < nsym; ++s) n += freq[s] > 0;
    if (n == 0) return 0;
    leaf_t* lv = (leaf_t*)malloc((size_t)n * sizeof(leaf_t));
    n = 0;
    for (int s = 0; s < nsym; ++s)
        if (freq[s]) { lv[n].f = freq[s]; lv[n].sym = s; n++; }
    if (n == 1) { len[lv[0].sym] = 1; free(lv); return 1; }
    qsort(lv, (size_t)n, sizeof(leaf_t), leaf_cmp);
    /* nodes 0..n-1 leaves, n..2n-2 internal in creation order */
    uint64_t* w = (uint64_t*)malloc((size_t)(2 * n) * sizeof(uint64_t));
    int* parent = (int*)malloc((size_t)(2 * n) * sizeof(int));
    for (int i = 0; i < n; ++i) w[i] = lv[i].f;
    int li = 0, ii = n, next = n;
    for (int k = 0; k < n - 1; ++k) {
        int pick[2];
        for (int t = 0; t < 2; ++t) {
            /* ties prefer the leaf queue */
            if (li < n && (ii >= next || w[li] <= w[ii])) pick[t] = li++;
            else pick[t] = ii++;
        }
        w[next] = w[pick[0]] + w[pick[1]];
        parent[pick[0]] = next;
        parent[pick[1]] = next;
        next++;
    }
    int* depth = (int*)malloc((size_t)(2 * n) * sizeof(int));
    depth[next - 1] = 0;
    for (int k = next - 2; k >= n; --k) depth[k] = depth[parent[k]] + 1;
    int maxlen = 0;
    for (int i = 0; i < n; ++i) {
        depth[i] = depth[parent[i]] + 1;
        if (depth[i] > maxlen) maxlen = depth[i];
    }
    if (maxlen <= lmax) {
        for (int i = 0; i < n; ++i) len[lv[i].sym] = (uint8_t)depth[i];
    } else {
        /* ITU-T T.81 Annex K.3 (Figure K.3) on the per-length counts */
        long* bits = (long*)calloc((size_t)maxlen + 2, sizeof(long));
        for (int i = 0; i < n; ++i) bits[depth[i]]++;
        for (int i = maxlen; i > lmax; --i) {
            while (bits[i] > 0) {
                int j = i - 2;
                while (bits[j] == 0) j--;
                bits[i] -= 2;
                bits[i - 1] += 1;
                bits[j + 1] += 2;
                bits[j] -= 1;
            }
        }
        /* hand the new lengths out in (old length, symbol) order */
        int* cnt = (int*)calloc((size_t)maxlen + 2, sizeof(int));
        for (int i = 0; i < n; ++i) cnt[depth[i]]++;
        int* start = (int*)calloc((size_t)maxlen + 2, sizeof(int));
        for (int d = 1; d <= maxlen; ++d) start[d] = start[d - 1] + cnt[d - 1];
        int* dep = (int*)malloc((size_t)nsym * sizeof(int));
        for (int s = 0; s < nsym; ++s) dep[s] = 0;
        for (int i = 0; i < n; ++i) dep[lv[i].sym] = depth[i];
        int* order = (int*)malloc((size_t)n * sizeof(int));
        for (int s = 0; s < nsym; ++s)
            if (dep[s] > 0) order[start[dep[s]]++] = s;
        int k = 0;
        for (int l = 1; l <= lmax; ++l)
            for (long c = 0; c < bits[l]; ++c) len[order[k++]] = (uint8_t)l;
        free(order); free(dep); free(start); free(cnt);
        free(bits);
    }
    free(depth); free(parent); free(w); free(lv);
    return n;
}

typedef struct {
    int n_used;
    uint8_t len[65536];
    uint32_t code[65536];
    /* canonical decode */
    uint32_t first_code[STZF_LMAX + 2], count[STZF_LMAX + 2], first_index[STZF_LMAX + 2];
    uint16_t sorted[65536];
    int lmax_used;
} huff_t;

/* DEFLATE-style canonical assignment over (length, symbol) order */
static void huff_canonical(huff_t* h)
{
    memset(h->count, 0, sizeof(h->count));
    h->n_used = 0;
    h->lmax_used = 0;
    for (int s = 0; s < 65536; ++s)
        if (h->len[s]) { h->count[h->len[s]]++; h->n_used++; if (h->len[s] > h->lmax_used) h->lmax_used = h->len[s]; }
    uint32_t code = 0;
    uint32_t idx = 0;
    uint32_t next_code[STZF_LMAX + 2];
    for (int l = 1; l <= STZF_LMAX; ++l) {
        code = (code + (l > 1 ? h->count[l - 1] : 0)) << 1;
        if (l == 1) code = 0;
        h->first_code[l] = code;
        next_code[l] = code;
        h->first_index[l] = idx;
        idx += h->count[l];
    }
    uint32_t fill[STZF_LMAX + 2];
    for (int l = 1; l <= STZF_LMAX; ++l) fill[l] = h->first_index[l];
    for (int s = 0; s < 65536; ++s)
        if (h->len[s]) {
            const int l = h->len[s];
            h->code[s] = next_code[l]++;
            h->sorted[fill[l]++] = (uint16_t)s;
        }
}

/* ------------------------------------------------------------------------ */
/* payload writer                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint16_t* codes;     /* stream order */
    void* unpred;        /* raw values of code-0 positions, stream order */
    size_t n_unpred;
} ora_stream_in;

static void ora_fill_header(stzf_file_header* fh, const ora_layout* L, int dtype, double eb,
                            int eb_mode, double unit, const double ebl[3], double vmin,
                            double vmax, const ora_params* p)
{
    memset(fh, 0, sizeof(*fh));
    fh->magic = STZF_MAGIC;
    fh->version = STZF_VERSION;
    fh->dtype = (uint32_t)dtype;
    fh->eb_mode = (uint32_t)eb_mode;
    fh->nx = L->nx; fh->ny = L->ny; fh->nz = L->nz;
    fh->px = L->px; fh->py = L->py; fh->pz = L->pz;
    fh->eb_user = eb;
    fh->eb_unit = unit;
    fh->mult_base = p->mult_base; fh->mult_low = p->mult_low; fh->mult_high = p->mult_high;
    fh->eb_base = ebl[0]; fh->eb_low = ebl[1]; fh->eb_high = ebl[2];
    fh->vmin = vmin; fh->vmax = vmax;
    fh->radius = STZF_RADIUS;
    fh->chunk_high = (uint32_t)L->chunk_high;
    fh->chunk_low = (uint32_t)L->chunk_low;
    fh->sub_size = STZF_SUB;
    fh->interp = (uint32_t)p->interp;
    fh->pred = (uint64_t)p->pred;
    fh->order = (uint32_t)p->order;
    fh->lmax = STZF_LMAX;
    fh->n_streams = STZF_NSTREAMS;
    fh->alpha = p->alpha;
    fh->beta = p->beta;
}

typedef struct {
    stzf_stream_header sh;
    huff_t* h;
    uint32_t* sub_bits;
    uint32_t* sub_un;
} stream_plan_t;

static long long ora_write_payload(stzf_file_header* fh, const ora_layout* L,
                                   const ora_stream_in* sin, size_t tsize, uint8_t** out)
{
    stream_plan_t sp[STZF_NSTREAMS];
    memset(sp, 0, sizeof(sp));
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < STZF_NSTREAMS; ++s) {
        const ora_geom* g = &L->geom[s];
        stream_plan_t* P = &sp[s];
        P->h = (huff_t*)calloc(1, sizeof(huff_t));
        uint64_t* freq = (uint64_t*)calloc(65536, sizeof(uint64_t));
        for (uint64_t i = 0; i < g->n; ++i) freq[sin[s].codes[i]]++;
        ora_huff_lengths(freq, 65536, STZF_LMAX, P->h->len);
        huff_canonical(P->h);
        uint64_t total_bits = 0;
        for (int q = 0; q < 65536; ++q) total_bits += freq[q] * P->h->len[q];
        free(freq);
        P->sub_bits = (uint32_t*)calloc(g->n_sub + 1, sizeof(uint32_t));
        P->sub_un = (uint32_t*)calloc(g->n_sub + 1, sizeof(uint32_t));
        for (uint32_t c = 0; c < g->n_chunks; ++c) {
            const uint64_t p0 = g->chunk_start[c], p1 = g->chunk_start[c + 1];
            for (uint64_t p = p0; p < p1; ++p) {
                const uint32_t k = g->chunk_sub[c] + (uint32_t)((p - p0) / STZF_SUB);
                P->sub_bits[k] += P->h->len[sin[s].codes[p]];
                P->sub_un[k] += sin[s].codes[p] == 0;
            }
        }
        stzf_stream_header* sh = &P->sh;
        memset(sh, 0, sizeof(*sh));
        sh->magic = STZF_STREAM_MAGIC;
        sh->kind = s == 0 ? 0u : s < 8 ? 1u : 2u;
        sh->block = s == 0 ? 0u : (uint32_t)((s - 1) % 7);
        sh->dtype = fh->dtype;
        sh->n_symbols = g->n;
        sh->dx = (uint32_t)g->dx; sh->dy = (uint32_t)g->dy; sh->dz = (uint32_t)g->dz;
        sh->chunk = g->chunk > 0 ? (uint32_t)g->chunk : STZF_BASE_CHUNK;
        sh->n_chunks = g->n_chunks;
        sh->n_sub = g->n_sub;
        sh->n_used = (uint32_t)P->h->n_used;
        sh->lmax = STZF_LMAX;
        sh->n_unpred = sin[s].n_unpred;
        sh->total_bits = total_bits;
        sh->eb = s == 0 ? fh->eb_base : s < 8 ? fh->eb_low : fh->eb_high;
        sh->radius = STZF_RADIUS;
        sh->sub_size = STZF_SUB;
        uint64_t off = sizeof(stzf_stream_header);
        sh->off_codebook = off;     off += align8(4ull * sh->n_used);
        sh->off_chunk_bit = off;    off += 8ull * (sh->n_chunks + 1);
        sh->off_chunk_unpred = off; off += 8ull * (sh->n_chunks + 1);
        sh->off_chunk_sub = off;    off += align8(4ull * (sh->n_chunks + 1));
        sh->off_sub_bits = off;     off += align8(2ull * sh->n_sub);
        sh->off_sub_unpred = off;   off += align8(2ull * sh->n_sub);
        sh->off_unpred = off;       off += align8(tsize * sh->n_unpred);
        sh->off_bits = off;         off += align8(4ull * ((total_bits + 31) / 32 + 2));
        sh->stream_bytes = off;
    }
    uint64_t total = sizeof(stzf_file_header);
    for (int s = 0; s < STZF_NSTREAMS; ++s) {
        fh->stream_off[s] = total;
        fh->stream_bytes[s] = sp[s].sh.stream_bytes;
        total += sp[s].sh.stream_bytes;
    }
    fh->total_bytes = total;
    uint8_t* buf = (uint8_t*)calloc(total, 1);
    memcpy(buf, fh, sizeof(*fh));
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < STZF_NSTREAMS; ++s) {
        const ora_geom* g = &L->geom[s];
        stream_plan_t* P = &sp[s];
        const stzf_stream_header* sh = &P->sh;
        uint8_t* base = buf + fh->stream_off[s];
        memcpy(base, sh, sizeof(*sh));
        uint32_t* cb = (uint32_t*)(base + sh->off_codebook);
        uint32_t k = 0;
        for (int q = 0; q < 65536; ++q)
            if (P->h->len[q]) cb[k++] = (uint32_t)q | ((uint32_t)P->h->len[q] << 16);
        uint64_t* cbit = (uint64_t*)(base + sh->off_chunk_bit);
        uint64_t* cun = (uint64_t*)(base + sh->off_chunk_unpred);
        uint32_t* csub = (uint32_t*)(base + sh->off_chunk_sub);
        uint16_t* sbits = (uint16_t*)(base + sh->off_sub_bits);
        uint16_t* sun = (uint16_t*)(base + sh->off_sub_unpred);
        uint64_t bacc = 0, uacc = 0;
        for (uint32_t c = 0; c < g->n_chunks; ++c) {
            cbit[c] = bacc;
            cun[c] = uacc;
            csub[c] = g->chunk_sub[c];
            for (uint32_t j = g->chunk_sub[c]; j < g->chunk_sub[c + 1]; ++j) {
                sbits[j] = (uint16_t)P->sub_bits[j];
                sun[j] = (uint16_t)P->sub_un[j];
                bacc += P->sub_bits[j];
                uacc += P->sub_un[j];
            }
        }
        cbit[g->n_chunks] = bacc;
        cun[g->n_chunks] = uacc;
        csub[g->n_chunks] = g->chunk_sub[g->n_chunks];
        if (sh->n_unpred) memcpy(base + sh->off_unpred, sin[s].unpred, tsize * sh->n_unpred);
        /* MSB-first bit packing into 32-bit words */
        uint32_t* words = (uint32_t*)(base + sh->off_bits);
        uint64_t acc = 0;
        int nacc = 0;
        uint64_t wi = 0;
        for (uint64_t p = 0; p < g->n; ++p) {
            const uint16_t q = sin[s].codes[p];
            const int l = P->h->len[q];
            acc = (acc << l) | P->h->code[q];
            nacc += l;
            while (nacc >= 32) {
                words[wi++] = (uint32_t)(acc >> (nacc - 32));
                nacc -= 32;
                acc &= (nacc ? ((1ull << nacc) - 1) : 0);
            }
        }
        if (nacc) words[wi++] = (uint32_t)(acc << (32 - nacc));
        free(P->sub_bits);
        free(P->sub_un);
        free(P->h);
    }
    *out = buf;
    return (long long)total;
}

/* ------------------------------------------------------------------------ */
/* payload reader                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    const uint8_t* base;
    size_t bytes;
    const stzf_file_header* fh;
    const stzf_stream_header* sh[STZF_NSTREAMS];
    ora_layout lay;
} ora_view;

static int ora_open_view(ora_view* v, const uint8_t* p, size_t bytes)
{
    memset(v, 0, sizeof(*v));
    if (bytes < sizeof(stzf_file_header)) return -1;
    const stzf_file_header* fh = (const stzf_file_header*)p;
    if (fh->magic != STZF_MAGIC || fh->version != STZF_VERSION || fh->n_streams != STZF_NSTREAMS ||
        fh->pred > ORA_PRED_ALL || ((fh->pred & ORA_PRED_STAGED) && fh->pred != ORA_PRED_STAGED) ||
        fh->interp > ORA_BASE_LZREG || fh->order > 5 || !(fh->alpha >= 1.0) ||
        !(fh->beta >= 1.0))
        return -1;
    if (fh->total_bytes > bytes) return -1;
    v->base = p;
    v->bytes = bytes;
    v->fh = fh;
    for (int s = 0; s < STZF_NSTREAMS; ++s) {
        if (fh->stream_off[s] + sizeof(stzf_stream_header) > bytes ||
            fh->stream_off[s] + fh->stream_bytes[s] > bytes)
            return -1;
        v->sh[s] = (const stzf_stream_header*)(p + fh->stream_off[s]);
        if (v->sh[s]->magic != STZF_STREAM_MAGIC) return -1;
    }
    ora_layout_dims(&v->lay, fh->nx, fh->ny, fh->nz, fh->px, fh->py, fh->pz, (int)fh->chunk_high,
                    (int)fh->chunk_low, (int)fh->interp);
    for (int s = 0; s < STZF_NSTREAMS; ++s)
        if (v->sh[s]->n_symbols != v->lay.geom[s].n || v->sh[s]->n_chunks != v->lay.geom[s].n_chunks) {
            layout_free(&v->lay);
            return -1;
        }
    return 0;
}

static inline uint32_t peek32(const uint32_t* w, uint64_t bit)
{
    const uint64_t wi = bit >> 5;
    const int o = (int)(bit & 31);
    const uint64_t two = ((uint64_t)w[wi] << 32) | w[wi + 1];
    return (uint32_t)(two >> (32 - o));
}

/* decode chunks (list == NULL: all) of stream s into codes[pos] / side[pos] */
static int ora_decode_stream(const ora_view* vw, int s, const int64_t* list, int64_t nlist,
                             uint16_t* codes, uint8_t* side, size_t tsize)
{
    const stzf_stream_header* sh = vw->sh[s];
    const uint8_t* base = vw->base + vw->fh->stream_off[s];
    const ora_geom* g = &vw->lay.geom[s];
    huff_t* h = (huff_t*)calloc(1, sizeof(huff_t));
    const uint32_t* cb = (const uint32_t*)(base + sh->off_codebook);
    for (uint32_t k = 0; k < sh->n_used; ++k) {
        const uint32_t e = cb[k];
        if ((e >> 16) == 0 || (e >> 16) > STZF_LMAX) { free(h); return 1; }
        h->len[e & 0xFFFF] = (uint8_t)(e >> 16);
    }
    huff_canonical(h);
    const uint64_t* cbit = (const uint64_t*)(base + sh->off_chunk_bit);
    const uint64_t* cun = (const uint64_t*)(base + sh->off_chunk_unpred);
    const uint32_t* csub = (const uint32_t*)(base + sh->off_chunk_sub);
    const uint16_t* sbits = (const uint16_t*)(base + sh->off_sub_bits);
    const uint16_t* sun = (const uint16_t*)(base + sh->off_sub_unpred);
    const uint8_t* un = base + sh->off_unpred;
    const uint32_t* words = (const uint32_t*)(base + sh->off_bits);
    const int64_t nc = list ? nlist : (int64_t)g->n_chunks;
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int64_t ci = 0; ci < nc; ++ci) {
        const uint32_t c = list ? (uint32_t)list[ci] : (uint32_t)ci;
        if (c >= g->n_chunks || csub[c] != g->chunk_sub[c]) { err = 1; continue; }
        uint64_t bit = cbit[c];
        uint64_t u = cun[c];
        int cerr = 0;
        for (uint32_t j = g->chunk_sub[c]; j < g->chunk_sub[c + 1] && !cerr; ++j) {
            const uint64_t p0 = g->chunk_start[c] + (uint64_t)(j - g->chunk_sub[c]) * STZF_SUB;
            uint64_t p1 = p0 + STZF_SUB;
            if (p1 > g->chunk_start[c + 1]) p1 = g->chunk_start[c + 1];
            const uint64_t bit0 = bit;
            uint32_t nz = 0;
            for (uint64_t p = p0; p < p1; ++p) {
                const uint32_t win = peek32(words, bit);
                int l, found = 0;
                for (l = 1; l <= STZF_LMAX; ++l) {
                    const uint32_t cd = win >> (32 - l);
                    if (cd - h->first_code[l] < h->count[l]) { found = 1; break; }
                }
                if (!found) { cerr = 1; break; }
                const uint16_t q = h->sorted[h->first_index[l] + (win >> (32 - l)) - h->first_code[l]];
                bit += (uint64_t)l;
                codes[p] = q;
                if (q == 0) {
                    memcpy(side + p * tsize, un + (u + nz) * tsize, tsize);
                    nz++;
                }
            }
            if (bit - bit0 != sbits[j] || nz != sun[j]) cerr = 1;
            u += nz;
        }
        err |= cerr;
    }
    free(h);
    return err;
}

/* ------------------------------------------------------------------------ */
/* ROI / slice planner (F5:1578-1830)                                        */
/* ------------------------------------------------------------------------ */
typedef struct { int x0, x1, y0, y1, z0, z1; } ora_box;
typedef struct { int valid; ora_box r; int64_t* chunks; int64_t n; } ora_bplan;
typedef struct {
    ora_box q;
    int is_slice;
    ora_bplan high[7], low[7];
    ora_box regions[8];
    int n_regions;
} ora_plan;

typedef struct { int b, e; } axr;

/* F5:1622-1635 normalize_axis */
static axr normalize_axis(int a, int b, int dim, int* is_slice)
{
    axr r;
    if (a == b) {
        *is_slice = 1;
        const int v = imax(0, imin(a, dim - 1));
        r.b = v; r.e = v;
        return r;
    }
    int lo = imax(0, imin(a, b));
    int hi = imin(dim - 1, imax(a, b));
    if (hi < lo) hi = lo;
    r.b = lo; r.e = hi;
    return r;
}

/* F5:1646-1660 range_for_parity */
static axr range_for_parity(axr f, int parity)
{
    axr r;
    int first = f.b;
    if ((first & 1) != parity) first++;
    if (first > f.e) { r.b = 1; r.e = 0; return r; }
    int last = f.e;
    if ((last & 1) != parity) last--;
    r.b = first / 2; r.e = last / 2;
    return r;
}

/* F5:1662-1687 chunks_for_box, with a partial last chunk per axis */
static int64_t chunks_for_box(int dx, int dy, int dz, int C, axr x, axr y, axr z, int64_t* out)
{
    if (x.b > x.e || y.b > y.e || z.b > z.e) return 0;
    const int ncx = (dx + C - 1) / C, ncy = (dy + C - 1) / C, ncz = (dz + C - 1) / C;
    const int cx0 = imax(0, x.b / C), cy0 = imax(0, y.b / C), cz0 = imax(0, z.b / C);
    const int cx1 = imin(ncx - 1, x.e / C), cy1 = imin(ncy - 1, y.e / C), cz1 = imin(ncz - 1, z.e / C);
    int64_t n = 0;
    for (int cz = cz0; cz <= cz1; ++cz)
        for (int cy = cy0; cy <= cy1; ++cy)
            for (int cx = cx0; cx <= cx1; ++cx) {
                if (out) out[n] = ((int64_t)cz * ncy + cy) * ncx + cx; /* F5:1577-1582 */
                n++;
            }
    return n;
}

/* F5:1726-1749 ref_range_for_used_axis */
static axr ref_range_for_used_axis(axr t, int dim)
{
    if (t.b > t.e) return t;
    int lo = t.b, hi = t.e;
    const int ib = imax(t.b, 1), ie = imin(t.e, dim - 3);
    if (ib <= ie) { lo = imin(lo, ib - 1); hi = imax(hi, ie + 2); }
    const int nb = imax(t.b, 0), ne = imin(t.e, dim - 2);
    if (nb <= ne) hi = imax(hi, ne + 1);
    axr r;
    r.b = imax(0, lo);
    r.e = imin(dim - 1, hi);
    return r;
}

static int i64_cmp(const void* a, const void* b)
{
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : x > y;
}

static void ora_free_plan(ora_plan* pl)
{
    for (int b = 0; b < 7; ++b) { free(pl->high[b].chunks); free(pl->low[b].chunks); }
}

static int ora_make_plan(ora_plan* pl, const ora_layout* L, int ax, int ay, int az, int bx, int by, int bz)
{
    memset(pl, 0, sizeof(*pl));
    const int mx = (int)(L->px / 2), my = (int)(L->py / 2), mz = (int)(L->pz / 2);
    const int lx = (int)(L->px / 4), ly = (int)(L->py / 4), lz = (int)(L->pz / 4);
    /* F5:1637-1644 make_query against the caller's (unpadded) dims */
    int sl = 0;
    axr qx = normalize_axis(ax, bx, (int)L->nx, &sl);
    axr qy = normalize_axis(ay, by, (int)L->ny, &sl);
    axr qz = normalize_axis(az, bz, (int)L->nz, &sl);
    pl->q.x0 = qx.b; pl->q.x1 = qx.e; pl->q.y0 = qy.b; pl->q.y1 = qy.e; pl->q.z0 = qz.b; pl->q.z1 = qz.e;
    pl->is_slice = sl;
    /* F5:1689-1711 build_residual_plan (high, mid-grid chunks) */
    for (int b = 0; b < 7; ++b) {
        const int k = b + 1;
        axr px = range_for_parity(qx, k & 1), py = range_for_parity(qy, (k >> 1) & 1),
            pz = range_for_parity(qz, (k >> 2) & 1);
        if (px.b > px.e || py.b > py.e || pz.b > pz.e) continue;
        ora_bplan* hp = &pl->high[b];
        hp->valid = 1;
        hp->r.x0 = px.b; hp->r.x1 = px.e; hp->r.y0 = py.b; hp->r.y1 = py.e; hp->r.z0 = pz.b; hp->r.z1 = pz.e;
        hp->n = chunks_for_box(mx, my, mz, L->chunk_high, px, py, pz, NULL);
        hp->chunks = (int64_t*)malloc((size_t)(hp->n + 1) * sizeof(int64_t));
        chunks_for_box(mx, my, mz, L->chunk_high, px, py, pz, hp->chunks);
        /* F5:1751-1770 build_high_ref_region */
        axr rx = (k & 1) ? ref_range_for_used_axis(px, mx) : px;
        axr ry = (k & 2) ? ref_range_for_used_axis(py, my) : py;
        axr rz = (k & 4) ? ref_range_for_used_axis(pz, mz) : pz;
        ora_box* R = &pl->regions[pl->n_regions++];
        R->x0 = rx.b; R->x1 = rx.e; R->y0 = ry.b; R->y1 = ry.e; R->z0 = rz.b; R->z1 = rz.e;
    }
    /* F5:1772-1782 build_query_sub0_region, appended as in F5:2464-2469 */
    {
        axr sx = range_for_parity(qx, 0), sy = range_for_parity(qy, 0), sz = range_for_parity(qz, 0);
        if (sx.b <= sx.e && sy.b <= sy.e && sz.b <= sz.e) {
            ora_box* R = &pl->regions[pl->n_regions++];
            R->x0 = sx.b; R->x1 = sx.e; R->y0 = sy.b; R->y1 = sy.e; R->z0 = sz.b; R->z1 = sz.e;
        }
    }
    /* F5:1784-1830 build_low_residual_plan: union of chunk sets, merged boxes */
    for (int b = 0; b < 7; ++b) {
        const int k = b + 1;
        int64_t cap = 0;
        for (int rg = 0; rg < pl->n_regions; ++rg) {
            const ora_box* R = &pl->regions[rg];
            axr rx = {R->x0, R->x1}, ry = {R->y0, R->y1}, rz = {R->z0, R->z1};
            if (rx.b > rx.e || ry.b > ry.e || rz.b > rz.e) continue;
            axr px = range_for_parity(rx, k & 1), py = range_for_parity(ry, (k >> 1) & 1),
                pz = range_for_parity(rz, (k >> 2) & 1);
            cap += chunks_for_box(lx, ly, lz, L->chunk_low, px, py, pz, NULL);
        }
        if (cap == 0) continue;
        ora_bplan* lp = &pl->low[b];
        lp->chunks = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
        int64_t n = 0;
        axr ux = {1, 0}, uy = {1, 0}, uz = {1, 0};
        for (int rg = 0; rg < pl->n_regions; ++rg) {
            const ora_box* R = &pl->regions[rg];
            axr rx = {R->x0, R->x1}, ry = {R->y0, R->y1}, rz = {R->z0, R->z1};
            if (rx.b > rx.e || ry.b > ry.e || rz.b > rz.e) continue;
            axr px = range_for_parity(rx, k & 1), py = range_for_parity(ry, (k >> 1) & 1),
                pz = range_for_parity(rz, (k >> 2) & 1);
            if (px.b > px.e || py.b > py.e || pz.b > pz.e) continue;
            n += chunks_for_box(lx, ly, lz, L->chunk_low, px, py, pz, lp->chunks + n);
            /* F5:1713-1724 merge_axis */
            if (ux.b > ux.e) ux = px; else { ux.b = imin(ux.b, px.b); ux.e = imax(ux.e, px.e); }
            if (uy.b > uy.e) uy = py; else { uy.b = imin(uy.b, py.b); uy.e = imax(uy.e, py.e); }
            if (uz.b > uz.e) uz = pz; else { uz.b = imin(uz.b, pz.b); uz.e = imax(uz.e, pz.e); }
        }
        if (n == 0) { free(lp->chunks); lp->chunks = NULL; continue; }
        qsort(lp->chunks, (size_t)n, sizeof(int64_t), i64_cmp);
        int64_t u = 0;
        for (int64_t i = 0; i < n; ++i)
            if (u == 0 || lp->chunks[u - 1] != lp->chunks[i]) lp->chunks[u++] = lp->chunks[i];
        lp->n = u;
        lp->valid = 1;
        lp->r.x0 = ux.b; lp->r.x1 = ux.e; lp->r.y0 = uy.b; lp->r.y1 = uy.e; lp->r.z0 = uz.b; lp->r.z1 = uz.e;
    }
    return 0;
}

int ora_plan_query(int fx, int fy, int fz, int chunk_high, int chunk_low, int ax, int ay, int az,
                   int bx, int by, int bz, int* query, int* high_plan, int64_t* high_chunks,
                   int* low_plan, int64_t* low_chunks, int64_t cap)
{
    ora_layout L;
    if (fx <= 0 || fy <= 0 || fz <= 0 || chunk_high <= 0 || chunk_low <= 0) return -1;
    ora_layout_dims(&L, (size_t)fx, (size_t)fy, (size_t)fz, pad4((size_t)fx), pad4((size_t)fy),
                    pad4((size_t)fz), chunk_high, chunk_low, ORA_BASE_CUBIC);
    ora_plan pl;
    ora_make_plan(&pl, &L, ax, ay, az, bx, by, bz);
    query[0] = pl.q.x0; query[1] = pl.q.x1; query[2] = pl.q.y0; query[3] = pl.q.y1;
    query[4] = pl.q.z0; query[5] = pl.q.z1; query[6] = pl.is_slice;
    int rc = 0;
    for (int b = 0; b < 7; ++b) {
        const ora_bplan* hp[2] = {&pl.high[b], &pl.low[b]};
        int* dst[2] = {high_plan + 8 * b, low_plan + 8 * b};
        int64_t* cd[2] = {high_chunks + cap * b, low_chunks + cap * b};
        for (int t = 0; t < 2; ++t) {
            dst[t][0] = hp[t]->valid;
            dst[t][1] = hp[t]->r.x0; dst[t][2] = hp[t]->r.x1;
            dst[t][3] = hp[t]->r.y0; dst[t][4] = hp[t]->r.y1;
            dst[t][5] = hp[t]->r.z0; dst[t][6] = hp[t]->r.z1;
            dst[t][7] = (int)hp[t]->n;
            if (hp[t]->n > cap) { rc = -2; continue; }
            for (int64_t i = 0; i < hp[t]->n; ++i) cd[t][i] = hp[t]->chunks[i];
        }
    }
    ora_free_plan(&pl);
    layout_free(&L);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* type-generic code                                                         */
/* ------------------------------------------------------------------------ */
#define T float
#define SFX _f32
#define DTYPE_ID ORA_F32
#include "stz_oracle_t.inc"
#undef T
#undef SFX
#undef DTYPE_ID

#define T double
#define SFX _f64
#define DTYPE_ID ORA_F64
#include "stz_oracle_t.inc"
#undef T
#undef SFX
#undef DTYPE_ID

/* ------------------------------------------------------------------------ */
/* exported entry points                                                     */
/* ------------------------------------------------------------------------ */
long long ora_compress(int dtype, const void* field, size_t nx, size_t ny, size_t nz, double eb,
                       int eb_mode, const ora_params* p, uint8_t** payload)
{
    ora_params d;
    if (!p) { ora_default_params(&d); p = &d; }
    if (!field || !payload || eb < 0 || p->alpha < 1.0 || p->beta < 1.0 || p->interp < 0 ||
        p->interp > ORA_BASE_LZREG || p->order < 0 || p->order > 5 || p->pred < 0 || p->pred > ORA_PRED_ALL ||
        ((p->pred & ORA_PRED_STAGED) && p->pred != ORA_PRED_STAGED))
        return -1;
    if (dtype == ORA_F32) return ora_compress_t_f32((const float*)field, nx, ny, nz, eb, eb_mode, p, payload);
    if (dtype == ORA_F64) return ora_compress_t_f64((const double*)field, nx, ny, nz, eb, eb_mode, p, payload);
    return -1;
}

int ora_decompress(const uint8_t* payload, size_t bytes, void* out)
{
    ora_view v;
    if (ora_open_view(&v, payload, bytes)) return -1;
    int rc = v.fh->dtype == ORA_F32 ? ora_decompress_t_f32(&v, (float*)out)
                                    : ora_decompress_t_f64(&v, (double*)out);
    layout_free(&v.lay);
    return rc;
}

int ora_decompress_roi(const uint8_t* payload, size_t bytes, int x0, int y0, int z0, int x1,
                       int y1, int z1, void* out, size_t* od)
{
    ora_view v;
    if (ora_open_view(&v, payload, bytes)) return -1;
    int rc = v.fh->dtype == ORA_F32
                 ? ora_decompress_roi_t_f32(&v, x0, y0, z0, x1, y1, z1, (float*)out, od)
                 : ora_decompress_roi_t_f64(&v, x0, y0, z0, x1, y1, z1, (double*)out, od);
    layout_free(&v.lay);
    return rc;
}

int ora_decompress_level(const uint8_t* payload, size_t bytes, int level, void* out, size_t* od)
{
    ora_view v;
    if (level < 0 || level > 3) return -1;
    if (ora_open_view(&v, payload, bytes)) return -1;
    int rc = v.fh->dtype == ORA_F32 ? ora_decompress_level_t_f32(&v, level, (float*)out, od)
                                    : ora_decompress_level_t_f64(&v, level, (double*)out, od);
    layout_free(&v.lay);
    return rc;
}

int ora_payload_info(const uint8_t* payload, size_t bytes, double* ebl, int64_t* sb, int64_t* nu)
{
    ora_view v;
    if (ora_open_view(&v, payload, bytes)) return -1;
    ebl[0] = v.fh->eb_base; ebl[1] = v.fh->eb_low; ebl[2] = v.fh->eb_high;
    for (int s = 0; s < STZF_NSTREAMS; ++s) {
        sb[s] = (int64_t)v.fh->stream_bytes[s];
        nu[s] = (int64_t)v.sh[s]->n_unpred;
    }
    layout_free(&v.lay);
    return 0;
}

void ora_free(void* p) { free(p); }

// ref_harness.cpp -- appended by oracle/build_ref.sh after the reference's own
// source lines (F5:1-17, F5:24-1253, F5:1577-1839) in one translation unit.
// TEST INFRASTRUCTURE: extern "C" entry points that call the reference's
// functions unchanged, so the tests can compare the oracle against them.
#include <cstdint>
#include <cstring>

namespace {
struct FtzScope {  // F5:3046-3047 flush-to-zero / denormals-are-zero
    unsigned saved;
    FtzScope() : saved(_mm_getcsr()) {
        _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
        _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
    }
    ~FtzScope() { _mm_setcsr(saved); }
};

template <typename T>
void h_quantize(const T* d, size_t n, double eb, int radius, int32_t* codes, T* unpred,
                size_t* nu)
{
    FtzScope f;
    // F5:2263-2266 dec_lut
    std::vector<T> dec_lut(static_cast<size_t>(radius) + 1);
    for (size_t h = 0; h < dec_lut.size(); ++h) dec_lut[h] = static_cast<T>(2 * static_cast<int>(h) * eb);
    std::vector<T> un;
    quantize_range_nodec(d, n, eb, radius, dec_lut.data(), codes, un);
    *nu = un.size();
    if (!un.empty()) std::memcpy(unpred, un.data(), un.size() * sizeof(T));
}

void put_plan(const std::vector<BlockAccessPlan>& plans, int* dst, int64_t* chunks, int64_t cap, int& rc)
{
    for (int b = 0; b < 7; ++b) dst[8 * b] = 0;
    for (const auto& p : plans) {
        int* o = dst + 8 * p.block;
        o[0] = 1;
        o[1] = p.x.begin; o[2] = p.x.end; o[3] = p.y.begin; o[4] = p.y.end;
        o[5] = p.z.begin; o[6] = p.z.end;
        o[7] = static_cast<int>(p.chunks.size());
        if (static_cast<int64_t>(p.chunks.size()) > cap) { rc = -2; continue; }
        for (size_t i = 0; i < p.chunks.size(); ++i) chunks[cap * p.block + i] = static_cast<int64_t>(p.chunks[i]);
    }
}
}  // namespace

extern "C" {

void ref_preprocess_f32(int b, const float* s, const float* r, float* d, int dx, int dy, int dz)
{ FtzScope f; preprocess_block_zslab<float>(b, s, r, d, dx, dy, dz, 0, dz); }
void ref_preprocess_f64(int b, const double* s, const double* r, double* d, int dx, int dy, int dz)
{ FtzScope f; preprocess_block_zslab<double>(b, s, r, d, dx, dy, dz, 0, dz); }

void ref_depreprocess_f32(int b, const float* s, const float* r, float* d, int dx, int dy, int dz)
{ FtzScope f; depreprocess_block_zslab<float>(b, s, r, d, dx, dy, dz, 0, dz); }
void ref_depreprocess_f64(int b, const double* s, const double* r, double* d, int dx, int dy, int dz)
{ FtzScope f; depreprocess_block_zslab<double>(b, s, r, d, dx, dy, dz, 0, dz); }

// ROI-restricted path (F5:978-999 -> depreprocess_block_fast)
void ref_depreprocess_roi_f32(int b, const float* s, const float* r, float* d, int dx, int dy, int dz,
                              int ex, int ey, int ez, int sx, int sy, int sz)
{ FtzScope f; depreprocess_block_split<float>(b, s, r, d, dx, dy, dz, ex, ey, ez, sx, sy, sz); }
void ref_depreprocess_roi_f64(int b, const double* s, const double* r, double* d, int dx, int dy, int dz,
                              int ex, int ey, int ez, int sx, int sy, int sz)
{ FtzScope f; depreprocess_block_split<double>(b, s, r, d, dx, dy, dz, ex, ey, ez, sx, sy, sz); }

// fused depreprocess + scatter into the full volume (F5:1096-1159)
void ref_depreprocess_to_full_f32(int b, const float* s, const float* r, float* full, int dx, int dy, int dz)
{ FtzScope f; depreprocess_block_zslab_to_full<float>(b, s, r, full, dx, dy, dz, 0, dz); }
void ref_depreprocess_to_full_f64(int b, const double* s, const double* r, double* full, int dx, int dy, int dz)
{ FtzScope f; depreprocess_block_zslab_to_full<double>(b, s, r, full, dx, dy, dz, 0, dz); }

void ref_quantize_f32(const float* d, size_t n, double eb, int radius, int32_t* codes, float* un, size_t* nu)
{ h_quantize<float>(d, n, eb, radius, codes, un, nu); }
void ref_quantize_f64(const double* d, size_t n, double eb, int radius, int32_t* codes, double* un, size_t* nu)
{ h_quantize<double>(d, n, eb, radius, codes, un, nu); }

void ref_split_f32(const float* full, float** sub, int nx, int ny, int nz)
{ slice_full_data<float>(full, sub, nx, ny, nz); }
void ref_split_parity_f32(const float* full, float** sub, int nx, int ny, int nz, int zp)
{ slice_full_data_parity<float>(full, sub, nx, ny, nz, zp); }
void ref_merge_f32(float** sub, float* full, int sx, int sy, int sz)
{ merge_sub_blocks_to_full<float>(sub, full, sx, sy, sz); }
void ref_split_f64(const double* full, double** sub, int nx, int ny, int nz)
{ slice_full_data<double>(full, sub, nx, ny, nz); }
void ref_merge_f64(double** sub, double* full, int sx, int sy, int sz)
{ merge_sub_blocks_to_full<double>(sub, full, sx, sy, sz); }

// F5:2460-2470 planning sequence; output layout matches ora_plan_query
int ref_plan_query(int fx, int fy, int fz, int chunk_high, int chunk_low, int ax, int ay, int az,
                   int bx, int by, int bz, int* query, int* high_plan, int64_t* high_chunks,
                   int* low_plan, int64_t* low_chunks, int64_t cap)
{
    set_full_dims(fx, fy, fz);
    roi_high_dim = chunk_high;
    roi_low_dim = chunk_low;
    const QueryBox q = make_query(ax, ay, az, bx, by, bz);
    const auto highPlans = build_residual_plan(q, dim_x, dim_y, dim_z, roi_high_dim);
    auto regions = build_high_ref_regions(highPlans);
    const QueryBox s0 = build_query_sub0_region(q);
    if (s0.x.begin <= s0.x.end && s0.y.begin <= s0.y.end && s0.z.begin <= s0.z.end) regions.push_back(s0);
    const auto lowPlans = build_low_residual_plan(regions);
    query[0] = q.x.begin; query[1] = q.x.end; query[2] = q.y.begin; query[3] = q.y.end;
    query[4] = q.z.begin; query[5] = q.z.end; query[6] = q.is_slice ? 1 : 0;
    int rc = 0;
    put_plan(highPlans, high_plan, high_chunks, cap, rc);
    put_plan(lowPlans, low_plan, low_chunks, cap, rc);
    return rc;
}

}  // extern "C"

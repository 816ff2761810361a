// ref3_harness.cpp -- appended by oracle/build_ref.sh after the SZ3-free
// lines of the reference's dependency.cpp (its standard #includes and
// globals, dependency.cpp:1-9 and 11-29, and the predictors and
// pre/de-preprocess functions, dependency.cpp:90-432) in one translation unit.
// TEST INFRASTRUCTURE: extern "C" entry points that call the reference's
// functions unchanged, so the tests can compare the oracle's staged
// inverse-variance predictors with them.
#include <xmmintrin.h>
#include <pmmintrin.h>

namespace {
struct Ftz {
    // flush-to-zero / denormals-are-zero, the mode the library's f32 kernels run in
    unsigned saved;
    Ftz() : saved(_mm_getcsr())
    {
        _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
        _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
    }
    ~Ftz() { _mm_setcsr(saved); }
};
}  // namespace

// dependency.cpp:333-380 (blocks 0, 1, 3 from the coarser grid)
extern "C" void ref3_preprocess_f32(int b, const float* s, const float* r, float* d, int dx, int dy, int dz)
{
    Ftz f;
    preprocess_block(b, s, r, d, dx, dy, dz);
}

extern "C" void ref3_depreprocess_f32(int b, const float* de, const float* r, float* out, int dx, int dy, int dz)
{
    Ftz f;
    depreprocess_block(b, de, r, out, dx, dy, dz);
}

// dependency.cpp:294-328 (blocks 2, 4, 5, 6 from the reconstructed blocks
// x, y, z, xy, xz, yz = 0, 1, 3, 2, 4, 5)
extern "C" void ref3_preprocess_staged_f32(int b, const float* s, const float* const* rb, float* d, int dx, int dy,
                                           int dz)
{
    Ftz f;
    preprocess_block_staged(b, s, rb[0], rb[1], rb[3], rb[2], rb[4], rb[5], d, dx, dy, dz);
}

extern "C" void ref3_depreprocess_staged_f32(int b, const float* de, const float* const* rb, float* out, int dx,
                                             int dy, int dz)
{
    Ftz f;
    depreprocess_block_staged(b, de, rb[0], rb[1], rb[3], rb[2], rb[4], rb[5], out, dx, dy, dz);
}

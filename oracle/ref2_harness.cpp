// ref2_harness.cpp -- appended by oracle/build_ref.sh after the SZ3-free
// lines of the reference's stz2.cpp (its standard #includes, stz2.cpp:1-8,
// and preprocess_block, stz2.cpp:82-306) in one translation unit.
// TEST INFRASTRUCTURE: an extern "C" entry point that calls the reference's
// function unchanged, so the tests can compare the oracle's stz2-style
// plane (12-point truncated bicubic) and cube (16-point) predictors with it.
#include <xmmintrin.h>
#include <pmmintrin.h>

extern "C" void ref2_preprocess_f32(int b, const float* s, const float* r, float* d, int dx, int dy, int dz)
{
    // flush-to-zero / denormals-are-zero, the mode the library's f32 kernels run in
    const unsigned saved = _mm_getcsr();
    _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
    _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
    preprocess_block(b, s, r, d, dx, dy, dz);
    _mm_setcsr(saved);
}

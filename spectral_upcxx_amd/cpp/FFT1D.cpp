// FFT1D.cpp -- plan handle over sh_fft1d_z2z (reference: src/FFT1D.cpp:14-101).
// No planning step is needed: the transform lengths are compiled kernels.
#include "FFT1D.H"

#include <cstdlib>
#include <iostream>

#include "spectral_hip.h"

static void fft1d_check(int rc) {
    if (rc != 0) {
        std::cout << "ERROR in FFT1D: " << sh_last_error() << std::endl;
        std::exit(1);
    }
}

FFT1D::FFT1D() : m_N(0), m_stride(1) {}

FFT1D::~FFT1D() {}

void FFT1D::init(int N, int stride, cdouble* input) {
    (void)input;
    m_N = N;
    m_stride = stride;
}

void FFT1D::forward(cdouble* input, int N, int stride) { transform(input, N, stride, -1); }

void FFT1D::inverse(cdouble* input, int N, int stride) {
    transform(input, N, stride, 1);
    // divide by N along the stride (one run when contiguous)
    if (stride == 1) {
        fft1d_check(sh_scale_z(input, 1.0 / N, input, N, nullptr));
    } else {
        for (int i = 0; i < N; i++) fft1d_check(sh_scale_z(input + (long)i * stride, 1.0 / N, input + (long)i * stride, 1, nullptr));
    }
}

void FFT1D::transform(cdouble* input, int N, int stride, int dir) {
    fft1d_check(sh_fft1d_z2z(input, N, stride, 1, 1, dir, nullptr));
}

void FFT1D::transform(cdouble* input, int DIR) {
    fft1d_check(sh_fft1d_z2z(input, m_N, m_stride, 1, 1, DIR == 1 ? 1 : -1, nullptr));
}

// FFT1D.cpp -- plan handle over sh_fft1d_z2z (reference: src/FFT1D.cpp:14-101).
// No planning step is needed: lengths are dispatched to compiled kernels.
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
    // input[i*stride] /= N for every i (src/FFT1D.cpp:62-65), one launch
    fft1d_check(sh_div_strided_z(input, (double)N, N, stride, nullptr));
}

void FFT1D::transform(cdouble* input, int N, int stride, int dir) {
    fft1d_check(sh_fft1d_z2z(input, N, stride, 1, 1, dir, nullptr));
}

void FFT1D::transform(cdouble* input, int DIR) {
    fft1d_check(sh_fft1d_z2z(input, m_N, m_stride, 1, 1, DIR == 1 ? 1 : -1, nullptr));
}

// FFT1D.H -- 1D FFT plan handle (the reference's FFTW wrapper, src/FFT1D.H:17-47),
// executing on device memory through sh_fft1d_z2z.
#ifndef SPECTRAL_FFT1D_H
#define SPECTRAL_FFT1D_H

#include <array>
#include <complex>
#include <cstdint>
#include <vector>

typedef std::complex<double> cdouble;
typedef uint64_t INT;
#ifndef DIM
#define DIM 3
#endif

/// in-place strided 1D C2C transform of one line, reusable for any line with
/// the same length and stride (FFT1D::init/transform, src/FFT1D.cpp:14-101)
class FFT1D {
public:
    FFT1D();
    /// set length and stride; `input` only documents the layout (device memory)
    void init(int N, int stride, cdouble* input);
    ~FFT1D();
    /// forward (sign -1) transform of device data of length N at `stride`
    void forward(cdouble* input, int N, int stride = 1);
    /// inverse (sign +1) transform, divided by N (src/FFT1D.cpp:60-67)
    void inverse(cdouble* input, int N, int stride = 1);
    /// unplanned transform with sign `dir`
    void transform(cdouble* input, int N, int stride, int dir);
    /// planned transform: DIR == 1 -> sign +1, else sign -1 (src/FFT1D.cpp:95-101)
    void transform(cdouble* input, int DIR);

private:
    int m_N;
    int m_stride;
};

#endif

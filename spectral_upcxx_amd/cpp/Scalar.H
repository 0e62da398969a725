// Scalar.H -- a complex<double> 3D field on one MI355X (src/Scalar.H:8-145).
//
// Same public interface and semantics as the reference's slab-distributed
// Scalar, with the whole domain on one device (p = 1: getPStart() = {0,0,0}).
// Data live in HBM (pooled device memory, layout n = i + Nx*j + Nx*Ny*k).
// Element access through operator[] / operator() / set / get goes through a
// host mirror that is synchronised lazily: a device operation uploads a
// host-modified mirror first, and host access downloads device results first.
// A reference obtained from a non-const operator[] marks the mirror as the
// current copy; it must not be held across a device operation.
#ifndef SPECTRAL_SCALAR_H
#define SPECTRAL_SCALAR_H

#include "FFT1D.H"

class Scalar {
public:
    /// default constructor (uninitialised; call init)
    Scalar();
    /// field of size N, physical (default) or Fourier space, zero-filled
    Scalar(std::array<INT, DIM> N, bool physical = true);
    ~Scalar();
    /// deep copy
    Scalar(const Scalar& scalar);
    /// deep copy; keeps the target's space flag if it is initialised (src/Scalar.cpp:59-66)
    void operator=(const Scalar& scalar);
    /// move from a temporary (the results of operations): takes its device
    /// buffer instead of copying it
    Scalar(Scalar&& scalar) noexcept;
    /// assignment from a temporary: the same result as the copy (the target
    /// keeps its space flag if it is initialised) without the device copy
    void operator=(Scalar&& scalar);
    /// tag: a field whose every element is written before it is read, so the
    /// zero fill of a new Scalar is skipped (operation results)
    struct NoFill {};
    Scalar(std::array<INT, DIM> N, bool physical, NoFill);
    void init(std::array<INT, DIM> N, bool physical = true);
    void set(std::array<INT, DIM> index, cdouble val);
    cdouble get(std::array<INT, DIM> index) const;
    cdouble& operator[](std::array<INT, DIM> index);
    cdouble operator[](std::array<INT, DIM> index) const;
    cdouble& operator()(INT a_i, INT a_j, INT a_k);
    cdouble operator()(INT a_i, INT a_j, INT a_k) const;
    cdouble& operator[](INT index);
    cdouble operator[](INT index) const;
    /// wavenumber of grid point ind (src/Scalar.cpp:198-206)
    std::array<double, DIM> freq(std::array<INT, DIM> ind) const;

    /// in place forward transform + zeroHighModes (src/Scalar.cpp:208-215)
    void forward();
    /// in place inverse transform, divided by Nx*Ny*Nz (src/Scalar.cpp:217-229)
    void inverse();
    /// out of place transforms (deep copy then in place)
    void forward(Scalar& a_scalar) const;
    void inverse(Scalar& a_scalar) const;

    /// i k f, returned in Fourier space
    Vector gradient() const;
    /// -|k|^2 f, returned in Fourier space
    Scalar laplacian() const;
    /// f /= (a - b |k|^2) where |k|^2 != 0
    void laplacian_inverse(double a = 0, double b = 1);
    /// zero the planes i = Nx/2, j = Ny/2, k = Nz/2
    void zeroHighModes();

    INT localSize() const;
    std::array<INT, DIM> getDims() const;
    std::array<INT, DIM> getPDims() const;
    std::array<INT, DIM> getPStart() const;
    std::array<INT, DIM> getPEnd() const;
    INT size() const;
    /// GB held by all live Scalars (printed on rank 0, as the reference)
    double memory() const;

    bool isPhysical() const;
    bool isFourier() const;
    void setFourier();
    void setPhysical();

    /// mean |f| (src/Scalar.cpp:15-22)
    double average() const;

    // ---- device access for code that calls the C ABI directly
    /// device pointer, current for reading
    const cdouble* deviceData() const;
    /// device pointer, current and marked as the authoritative copy
    cdouble* deviceDataMutable();

private:
    void init_impl(std::array<INT, DIM> N, bool physical, bool fill);
    void transform(int dir);
    void getIndex(std::array<INT, DIM> index, INT& rank, INT& loc) const;
    void host_current() const;
    void device_current() const;

    INT m_Nz, m_Ny, m_N, m_dSize;
    std::array<INT, DIM> m_dims, m_pdims;
    cdouble* m_dev;
    mutable std::vector<cdouble> m_host;
    // 0: device holds the data, 1: host mirror holds it, 2: both agree
    mutable int m_where;
    FFT1D m_fft_x, m_fft_y, m_fft_z;
    bool m_fourier;
    bool m_initialized;
    static int m_nscalars;
};

#endif

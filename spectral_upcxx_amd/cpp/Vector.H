// Vector.H -- three Scalars (src/Vector.H:4-70), same interface.
#ifndef SPECTRAL_VECTOR_H
#define SPECTRAL_VECTOR_H

class Vector {
public:
    Vector();
    Vector(std::array<INT, DIM> dims, bool physical = true);
    /// components whose every element is written before it is read (no zero fill)
    Vector(std::array<INT, DIM> dims, bool physical, Scalar::NoFill);
    void init(std::array<INT, DIM> dims, bool physical = true);
    Scalar& operator[](int a_i);
    const Scalar& operator[](int a_i) const;

    /// component-wise transforms
    void forward();
    void inverse();
    void forward(Vector& a_vector) const;
    void inverse(Vector& a_vector) const;

    /// (*this) x a_v computed in physical space, returned in Fourier space
    Vector cross(const Vector& a_v) const;
    /// i k x f
    Vector curl() const;
    /// sum_d i k_d f_d
    Scalar divergence() const;
    /// -|k|^2 f_d
    Vector laplacian() const;

    bool isFourier() const;
    /// always false, as the reference (src/Vector.cpp:191-197)
    bool isPhysical() const;

    INT localSize() const;
    std::array<INT, DIM> getDims() const;
    std::array<INT, DIM> getPStart() const;
    std::array<INT, DIM> getPEnd() const;
    std::array<INT, DIM> getPDims() const;

private:
    void setFourier();
    void setPhysical();

    std::array<INT, DIM> m_dims;
    INT m_N;
    std::array<Scalar, DIM> m_vector;
};

#endif

// Vector.cpp -- reference: src/Vector.cpp.  The Fourier-space operators are
// single fused kernels over the three components; cross follows the
// reference's sequence (deep copies, 6 inverse transforms, complex pointwise
// product, 3 forward transforms with Nyquist zeroing).
#include <cstdlib>
#include <iostream>

#include "DataObjects.H"
#include "spectral_hip.h"

#define ERROR(message)                                                                       \
    do {                                                                                     \
        std::cout << "ERROR in " << __func__ << " (" << __FILE__ << " " << __LINE__ << ", r" \
                  << upcxx::rank_me() << "): " << message << std::endl;                      \
        std::exit(1);                                                                        \
    } while (0)

#define CHECK(call)                                                                     \
    do {                                                                                \
        int _rc = (call);                                                               \
        if (_rc != 0) ERROR(#call << " failed (" << _rc << "): " << sh_last_error()); \
    } while (0)

Vector::Vector() : m_dims{0, 0, 0}, m_N(0) {}

Vector::Vector(std::array<INT, DIM> dims, bool physical) { init(dims, physical); }

Vector::Vector(std::array<INT, DIM> dims, bool physical, Scalar::NoFill) : m_dims(dims) {
    for (int i = 0; i < DIM; i++) m_vector[i] = Scalar(m_dims, physical, Scalar::NoFill());
    m_N = 1;
    for (int i = 0; i < DIM; i++) m_N *= m_dims[i];
}

void Vector::init(std::array<INT, DIM> dims, bool physical) {
    m_dims = dims;
    for (int i = 0; i < DIM; i++) m_vector[i].init(m_dims, physical);
    if (physical)
        setPhysical();
    else
        setFourier();
    m_N = 1;
    for (int i = 0; i < DIM; i++) m_N *= m_dims[i];
}

Scalar& Vector::operator[](int c) { return m_vector[c]; }
const Scalar& Vector::operator[](int c) const { return m_vector[c]; }

void Vector::forward() {
    for (int i = 0; i < DIM; i++) m_vector[i].forward();
    setFourier();
}

void Vector::inverse() {
    for (int i = 0; i < DIM; i++) m_vector[i].inverse();
    setPhysical();
}

void Vector::forward(Vector& a_vector) const {
    a_vector = (*this);
    a_vector.forward();
}

void Vector::inverse(Vector& a_vector) const {
    a_vector = (*this);
    a_vector.inverse();
}

Vector Vector::cross(const Vector& a_v) const {
    if (isPhysical() || a_v.isPhysical()) ERROR("must start in fourier space");
    Vector u;
    Vector v;
    inverse(u);
    a_v.inverse(v);
    Vector ret(m_dims, true, Scalar::NoFill());
    CHECK(sh_cross_pointwise_z(u[0].deviceData(), u[1].deviceData(), u[2].deviceData(),
                               v[0].deviceData(), v[1].deviceData(), v[2].deviceData(),
                               ret[0].deviceDataMutable(), ret[1].deviceDataMutable(),
                               ret[2].deviceDataMutable(), (long)u[0].localSize(), nullptr));
    ret.forward();
    return ret;
}

Vector Vector::curl() const {
    if (!isFourier()) ERROR("must start in fourier space");
    Vector c(m_dims, false, Scalar::NoFill());
    CHECK(sh_curl_z(m_vector[0].deviceData(), m_vector[1].deviceData(), m_vector[2].deviceData(),
                    c[0].deviceDataMutable(), c[1].deviceDataMutable(), c[2].deviceDataMutable(),
                    (int)m_dims[0], (int)m_dims[1], (int)m_dims[2], nullptr));
    return c;
}

Scalar Vector::divergence() const {
    if (!isFourier()) ERROR("must start in fourier space");
    Scalar div(m_dims, false, Scalar::NoFill());
    CHECK(sh_divergence_z(m_vector[0].deviceData(), m_vector[1].deviceData(), m_vector[2].deviceData(),
                          div.deviceDataMutable(), (int)m_dims[0], (int)m_dims[1], (int)m_dims[2],
                          nullptr));
    return div;
}

Vector Vector::laplacian() const {
    if (!isFourier()) ERROR("must start in fourier space");
    Vector lap(m_dims, false, Scalar::NoFill());
    CHECK(sh_vlaplacian_z(m_vector[0].deviceData(), m_vector[1].deviceData(), m_vector[2].deviceData(),
                          lap[0].deviceDataMutable(), lap[1].deviceDataMutable(),
                          lap[2].deviceDataMutable(), (int)m_dims[0], (int)m_dims[1], (int)m_dims[2],
                          nullptr));
    return lap;
}

bool Vector::isFourier() const {
    bool fourier = true;
    for (int i = 0; i < DIM; i++)
        if (!m_vector[i].isFourier()) fourier = false;
    return fourier;
}

bool Vector::isPhysical() const { return false; }

INT Vector::localSize() const { return m_vector[0].localSize(); }
std::array<INT, DIM> Vector::getDims() const { return m_vector[0].getDims(); }
std::array<INT, DIM> Vector::getPDims() const { return m_vector[0].getPDims(); }
std::array<INT, DIM> Vector::getPStart() const { return m_vector[0].getPStart(); }
std::array<INT, DIM> Vector::getPEnd() const { return m_vector[0].getPEnd(); }

void Vector::setPhysical() {
    for (int i = 0; i < DIM; i++) m_vector[i].setPhysical();
}

void Vector::setFourier() {
    for (int i = 0; i < DIM; i++) m_vector[i].setFourier();
}

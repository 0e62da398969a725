// Operators.cpp -- element-wise arithmetic as single device kernels
// (reference: src/Operators.cpp:9-99).  Results are new objects, as there;
// their memory comes from the device pool, so a temporary costs no hipMalloc.
#include <cstdlib>
#include <iostream>

#include "DataObjects.H"
#include "spectral_hip.h"

#define ERROR(message)                                                                          \
    do {                                                                                        \
        std::cout << "ERROR in " << __func__ << " (" << __FILE__ << " " << __LINE__ << "): " << message \
                  << std::endl;                                                                 \
        std::exit(1);                                                                           \
    } while (0)

#define CHECK(call)                                                                     \
    do {                                                                                \
        int _rc = (call);                                                               \
        if (_rc != 0) ERROR(#call << " failed (" << _rc << "): " << sh_last_error()); \
    } while (0)

Scalar operator*(const Scalar& a, const Scalar& b) {
    if (a.isPhysical() != b.isPhysical()) ERROR("space mismatch");
    Scalar ret(a.getDims(), a.isPhysical(), Scalar::NoFill());
    CHECK(sh_mul_z(ret.deviceDataMutable(), a.deviceData(), b.deviceData(), (long)a.localSize(), nullptr));
    return ret;
}

Scalar operator-(const Scalar& a, const Scalar& b) {
    if (a.isPhysical() != b.isPhysical()) ERROR("space mismatch");
    Scalar ret(a.getDims(), a.isPhysical(), Scalar::NoFill());
    CHECK(sh_sub_z(ret.deviceDataMutable(), a.deviceData(), b.deviceData(), (long)a.localSize(), nullptr));
    return ret;
}

Scalar operator+(const Scalar& a, const Scalar& b) {
    if (a.isPhysical() != b.isPhysical()) ERROR("space mismatch");
    Scalar ret(a.getDims(), a.isPhysical(), Scalar::NoFill());
    CHECK(sh_add_z(ret.deviceDataMutable(), a.deviceData(), b.deviceData(), (long)a.localSize(), nullptr));
    return ret;
}

Vector operator-(const Vector& a, const Vector& b) {
    if (a.isPhysical() != b.isPhysical()) ERROR("space mismatch");
    Vector ret(a.getDims(), a.isPhysical(), Scalar::NoFill());
    for (int d = 0; d < DIM; d++)
        CHECK(sh_sub_z(ret[d].deviceDataMutable(), a[d].deviceData(), b[d].deviceData(),
                       (long)a[0].localSize(), nullptr));
    return ret;
}

Vector operator+(const Vector& a, const Vector& b) {
    if (a.isPhysical() != b.isPhysical()) ERROR("space mismatch");
    Vector ret(a.getDims(), a.isPhysical(), Scalar::NoFill());
    for (int d = 0; d < DIM; d++)
        CHECK(sh_add_z(ret[d].deviceDataMutable(), a[d].deviceData(), b[d].deviceData(),
                       (long)a[0].localSize(), nullptr));
    return ret;
}

Scalar operator*(double alpha, const Scalar& s) {
    Scalar ret(s.getDims(), s.isPhysical(), Scalar::NoFill());
    CHECK(sh_scale_z(ret.deviceDataMutable(), alpha, s.deviceData(), (long)s.localSize(), nullptr));
    return ret;
}

Vector operator*(double alpha, const Vector& v) {
    Vector ret(v.getDims(), v.isPhysical(), Scalar::NoFill());
    for (int d = 0; d < DIM; d++)
        CHECK(sh_scale_z(ret[d].deviceDataMutable(), alpha, v[d].deviceData(), (long)v[0].localSize(),
                         nullptr));
    return ret;
}

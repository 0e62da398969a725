// DataObjects.H -- the drop-in API umbrella (src/DataObjects.H:9-14):
// Scalar, Vector and the Operators on MI355X device memory.
#ifndef SPECTRAL_DATA_OBJECTS_H
#define SPECTRAL_DATA_OBJECTS_H

#include <array>
#include <complex>
#include <vector>

#include <upcxx/upcxx.hpp>

class Vector;

#include "Scalar.H"
#include "Vector.H"
#include "Operators.H"

#endif

// test_freq -- prints Scalar::freq along x for N = 16 (behaviour of test/freq.cpp)
#include <iostream>

#include "DataObjects.H"

int main() {
    upcxx::init();
    INT N = 16;
    Scalar s({N, N, N}, true);
    for (INT i = 0; i < N; i++) std::cout << s.freq({i, 0, 0})[0] << std::endl;
    upcxx::finalize();
    return 0;
}

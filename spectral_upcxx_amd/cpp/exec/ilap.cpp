// ilap -- inverse-Laplacian timing (behaviour of exec/ilap.cpp):
//   ilap N [nruns]
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <limits>

#include "DataObjects.H"
#include "spectral_hip.h"

using namespace std;

int main(int argc, char* argv[]) {
    upcxx::init();
    INT N = argc > 1 ? atoi(argv[1]) : 32;
    int nruns = argc > 2 ? atoi(argv[2]) : 15;
    array<INT, DIM> dims = {N, N, N};
    Scalar s(dims, false);
    double best = numeric_limits<double>::max();
    for (int i = 0; i < nruns; i++) {
        sh_stream_synchronize(nullptr);
        auto t0 = chrono::steady_clock::now();
        s.laplacian_inverse();
        upcxx::barrier();
        sh_stream_synchronize(nullptr);
        double el = chrono::duration<double>(chrono::steady_clock::now() - t0).count();
        if (el < best) best = el;
    }
    cout << "n=" << upcxx::rank_n() << ", t=1, min time = " << best << " seconds" << endl;
    upcxx::finalize();
    return 0;
}

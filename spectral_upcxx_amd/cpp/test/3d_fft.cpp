// test_3d_fft -- 3D FFT round trip through the API on seeded data (behaviour
// of test/3d_fft.cpp), with a raw (unzeroed) transform so the round trip is
// exact to rounding.  Exit status 1 on failure.
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "DataObjects.H"
#include "spectral_hip.h"

using namespace std;

int main(int argc, char* argv[]) {
    upcxx::init();
    INT N = argc > 1 ? atoi(argv[1]) : 64;
    array<INT, DIM> dims = {N, N, N};
    Scalar d(dims), ans(dims);
    srand(upcxx::rank_me());
    for (INT i = 0; i < d.localSize(); i++) d[i] = cdouble((double)rand() / RAND_MAX, (double)rand() / RAND_MAX);
    ans = d;
    cdouble* p = d.deviceDataMutable();
    if (sh_fft3d_z2z(p, (int)N, (int)N, (int)N, -1, 0, nullptr) ||
        sh_fft3d_z2z(p, (int)N, (int)N, (int)N, 1, SH_NORMALIZE, nullptr)) {
        cout << "ERROR: " << sh_last_error() << endl;
        return 1;
    }
    bool wrong = false;
    for (INT i = 0; i < d.localSize(); i++)
        if (abs(d[i] - ans[i]) > 1e-13) wrong = true;
    // forward() then inverse() removes exactly the Nyquist planes
    Scalar z = ans;
    z.forward();
    z.inverse();
    Scalar zz = ans;
    cdouble* q = zz.deviceDataMutable();
    sh_fft3d_z2z(q, (int)N, (int)N, (int)N, -1, SH_ZERO_NYQUIST, nullptr);
    sh_fft3d_z2z(q, (int)N, (int)N, (int)N, 1, SH_NORMALIZE, nullptr);
    for (INT i = 0; i < z.localSize(); i++)
        if (abs(z[i] - zz[i]) > 1e-14) wrong = true;
    if (wrong) {
        cout << "WRONG!" << endl;
        return 1;
    }
    cout << "my man!" << endl;
    upcxx::finalize();
    return 0;
}

// test_cn_heat -- Crank-Nicolson heat equation through the API (behaviour of
// test/cn_heat.cpp), checked against the exact per-mode CN factor
// ((1 - K/2 k^2)/(1 + K/2 k^2))^Nt.  Exit status 1 on failure.
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "DataObjects.H"

using namespace std;

int main(int argc, char* argv[]) {
    upcxx::init();
    INT N = argc > 1 ? atoi(argv[1]) : 32;
    array<INT, DIM> dims = {N, N, N};
    Scalar u(dims, false);
    Scalar u0(dims, true);
    double h = 2 * M_PI / N;
    array<INT, DIM> ind;
    for (ind[2] = 0; ind[2] < N; ind[2]++)
        for (ind[1] = 0; ind[1] < N; ind[1]++)
            for (ind[0] = 0; ind[0] < N; ind[0]++) {
                double x = h * ind[0], y = h * ind[1], z = h * ind[2];
                u0[ind] = exp(-pow(x - M_PI, 2) / .5) * exp(-pow(y - M_PI, 2) / .5) * exp(-pow(z - M_PI, 2) / .5);
            }
    u0.forward();
    Scalar init = u0;
    int Nt = 200;
    double Tend = 0.5, K = Tend / Nt;
    for (int t = 1; t < Nt + 1; t++) {
        u = u0 + K / 2 * u0.laplacian();
        u.laplacian_inverse(1, -K / 2);
        u0 = u;
    }
    double err = 0, nrm = 0;
    for (ind[2] = 0; ind[2] < N; ind[2]++)
        for (ind[1] = 0; ind[1] < N; ind[1]++)
            for (ind[0] = 0; ind[0] < N; ind[0]++) {
                array<double, DIM> k = u.freq(ind);
                double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
                cdouble e = init.get(ind) * pow((1 - K / 2 * k2) / (1 + K / 2 * k2), Nt);
                err += norm(u.get(ind) - e);
                nrm += norm(e);
            }
    double rel = sqrt(err / nrm);
    cout << "cn heat relative error = " << rel << endl;
    if (rel > 1e-12) {
        cout << "cn heat failed" << endl;
        return 1;
    }
    cout << "cn heat passed" << endl;
    upcxx::finalize();
    return 0;
}

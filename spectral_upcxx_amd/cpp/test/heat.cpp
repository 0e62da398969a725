// test_heat -- exact heat solution through the API (behaviour of test/heat.cpp:
// u(T) = u0 exp(-|k|^2 T) set mode by mode with freq/set/operator[]), without
// the VTK writer.  Checked after the last time: the initial Gaussian is a
// product g(x) g(y) g(z), so u(T) = h(x) h(y) h(z) with h the 1D solution,
// computed here by a direct O(N^2) DFT on the host.  Exit status 1 on failure.
#include <cmath>
#include <complex>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "DataObjects.H"

using namespace std;

// 1D: h_j = (1/N) sum_k G_k exp(-k^2 T) e^{2 pi i k j / N}, G = DFT of g with
// the Nyquist entry zeroed (forward() zeroes the Nyquist planes)
static vector<double> heat1d(int N, double T) {
    const double h = 2. * M_PI / N;
    vector<complex<double>> G(N);
    for (int k = 0; k < N; k++) {
        complex<double> s = 0;
        for (int j = 0; j < N; j++) {
            double x = h * j;
            s += exp(-pow(x - M_PI, 2) / .01) * exp(complex<double>(0, -2. * M_PI * k * j / N));
        }
        G[k] = k == N / 2 ? 0. : s;
    }
    vector<double> out(N);
    for (int j = 0; j < N; j++) {
        complex<double> s = 0;
        for (int k = 0; k < N; k++) {
            double kk = k <= N / 2 ? k : k - N;
            s += G[k] * exp(-kk * kk * T) * exp(complex<double>(0, 2. * M_PI * k * j / N));
        }
        out[j] = s.real() / N;
    }
    return out;
}

int main(int argc, char* argv[]) {
    upcxx::init();
    INT N = 32;
    if (argc > 1) N = atoi(argv[1]);
    array<INT, DIM> dims = {N, N, N};
    Scalar u(dims, false);
    Scalar u0(dims);
    u.memory();
    double hx = 2. * M_PI / N;
    array<INT, DIM> ind = {0, 0, 0};
    for (ind[2] = 0; ind[2] < dims[2]; ind[2]++) {
        double z = hx * ind[2];
        for (ind[1] = 0; ind[1] < dims[1]; ind[1]++) {
            double y = hx * ind[1];
            for (ind[0] = 0; ind[0] < dims[0]; ind[0]++) {
                double x = hx * ind[0];
                u0.set(ind, exp(-pow(x - M_PI, 2) / .01) * exp(-pow(y - M_PI, 2) / .01) *
                                exp(-pow(z - M_PI, 2) / .01));
            }
        }
    }
    u0.forward();
    int Nt = 10;
    double Tend = 5, T = 0;
    for (int t = 1; t < Nt + 1; t++) {
        T = (double)t * Tend / Nt;
        for (ind[2] = 0; ind[2] < dims[2]; ind[2]++)
            for (ind[1] = 0; ind[1] < dims[1]; ind[1]++)
                for (ind[0] = 0; ind[0] < dims[0]; ind[0]++) {
                    array<double, DIM> k = u.freq(ind);
                    u.set(ind, u0[ind] * exp(-(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]) * T));
                }
        cout << t << endl;
    }
    // the last state against the separable solution
    u.inverse();
    vector<double> h = heat1d((int)N, T);
    double err = 0, nrm = 0;
    for (ind[2] = 0; ind[2] < dims[2]; ind[2]++)
        for (ind[1] = 0; ind[1] < dims[1]; ind[1]++)
            for (ind[0] = 0; ind[0] < dims[0]; ind[0]++) {
                double e = h[ind[0]] * h[ind[1]] * h[ind[2]];
                cdouble v = u.get(ind);
                err += norm(v - e);
                nrm += e * e;
            }
    double rel = sqrt(err / nrm);
    cout << "heat relative error = " << rel << endl;
    if (!(rel < 1e-12)) {
        cout << "heat failed" << endl;
        return 1;
    }
    cout << "heat passed" << endl;
    upcxx::finalize();
    return 0;
}

// test_particle -- marker particles through the API (src/Particle.H): 64
// Particle objects moved from 4 host threads at once (exec/ns2.cpp moves them
// from OpenMP threads) must equal a Particles set of the same tracers moved in
// one launch, bit for bit, and the energies must be 0.5 |v|^2.  Exit status 1
// on failure.
#include <cmath>
#include <iostream>
#include <thread>
#include <vector>

#include "DataObjects.H"
#include "Particle.H"

using namespace std;

int main() {
    upcxx::init();
    const INT N = 16;
    array<INT, DIM> dims = {N, N, N};
    // a smooth physical velocity field
    Vector u(dims, true);
    array<INT, DIM> ind;
    const double h = 2 * M_PI / N;
    for (ind[2] = 0; ind[2] < N; ind[2]++)
        for (ind[1] = 0; ind[1] < N; ind[1]++)
            for (ind[0] = 0; ind[0] < N; ind[0]++) {
                double x = h * ind[0], y = h * ind[1], z = h * ind[2];
                u[0][ind] = sin(x) * cos(y) + 0.3;
                u[1][ind] = -cos(x) * sin(y) + 0.2 * cos(z);
                u[2][ind] = 0.5 * sin(z);
            }
    // upload the host-written field once, before the threads read it (the
    // lazy host/device synchronisation of a Scalar is not thread-safe)
    for (int d = 0; d < DIM; d++) (void)u[d].deviceData();
    const int np = 64;
    vector<Particle> parts(np);
    Particles set(np, 1);
    const double K = 0.05;
    for (int step = 0; step < 3; step++) {
        vector<thread> th;
        for (int t = 0; t < 4; t++)
            th.emplace_back([&, t]() {
                for (int i = t; i < np; i += 4) parts[i].move(u, K);
            });
        for (auto& x : th) x.join();
        set.move(u, K);
    }
    bool ok = true;
    for (int i = 0; i < np && ok; i++) {
        array<double, DIM> a = parts[i].position(), b = set.position(i);
        ok = a == b && parts[i].energy() == set.energy(i);
    }
    cout << (ok ? "particle passed" : "particle failed") << endl;
    upcxx::finalize();
    return ok ? 0 : 1;
}

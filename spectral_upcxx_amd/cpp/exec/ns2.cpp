// ns2 -- the backward-Euler Navier-Stokes driver with marker particles on the
// drop-in API (behaviour of exec/ns2.cpp).
//
//   ns2 N [steps] [nu] [K] [mode] [np] [pdir] [out.bin]
//     steps : default T/K with T = 12 (exec/ns2.cpp:43-45)
//     mode  : api   -- the reference's operation sequence through Scalar/Vector/
//                      Operators (exec/ns2.cpp:132-204), one device kernel per op
//             fused -- the fused step of the C ABI (sh_ns_*, scheme 1)
//     np    : marker particles moved every step (exec/ns2.cpp:159, 191-204);
//             0 turns them off, as building without PARTICLES does
//     pdir  : if given, particle files pdir/p<n>.3D ("x y z E") every
//             steps/300 steps (exec/ns2.cpp:16-25, 200-202)
//     out   : writes the final velocity V1 (3 Fourier fields, complex128)
// Initial condition: the two Gaussian vortices of exec/ns2.cpp:97-130, with
// omega0 kept from the IC.  Prints the wall time of the step loop.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "DataObjects.H"
#include "Particle.H"
#include "spectral_hip.h"

using namespace std;

static void check(int rc, const char* what) {
    if (rc != 0) {
        cout << "ERROR: " << what << ": " << sh_last_error() << endl;
        exit(1);
    }
}

static double gaussian(double x, double y, double xc, double yc, double width) {
    return exp(-pow(x - xc, 2) / width) * exp(-pow(y - yc, 2) / width);
}

static void writeParticles(Particles& parts, const string& dir, int t) {
    ofstream out(dir + "/p" + to_string(t) + ".3D");
    out << "x y z E" << endl;
    for (INT i = 0; i < parts.size(); i++) {
        array<double, DIM> pos = parts.position(i);
        out << pos[0] << " " << pos[1] << " " << pos[2] << " " << parts.energy(i) << endl;
    }
}

int main(int argc, char* argv[]) {
    upcxx::init();
    INT N = argc > 1 ? atoi(argv[1]) : 16;
    const double T = 12;
    double nu = argc > 3 ? atof(argv[3]) : .001;
    double K = argc > 4 ? atof(argv[4]) : .001;
    int Nt = argc > 2 ? atoi(argv[2]) : (int)(T / K);
    string mode = argc > 5 ? argv[5] : "api";
    long Np = argc > 6 ? atol(argv[6]) : 10000;
    string pdir = argc > 7 ? argv[7] : "";
    string out = argc > 8 ? argv[8] : "";
    const int NSAVES = 300;
    int mod = Nt / NSAVES;
    if (mod < 1) mod = 1;
    array<INT, DIM> dims = {N, N, N};
    const long n = (long)N * N * N;

    // initial vorticity: two Gaussian vortices; psi; V0 = (-d_y psi, d_x psi, 0)
    Vector omega0(dims);
    {
        double width = M_PI / 8, mag = 50, dist = M_PI / 2;
        array<INT, DIM> ind;
        for (ind[2] = 0; ind[2] < dims[2]; ind[2]++)
            for (ind[1] = 0; ind[1] < dims[1]; ind[1]++)
                for (ind[0] = 0; ind[0] < dims[0]; ind[0]++) {
                    double x = 2. * M_PI / dims[0] * ind[0], y = 2. * M_PI / dims[1] * ind[1];
                    omega0[2][ind] = mag * gaussian(x, y, M_PI - dist, M_PI, width) +
                                     mag * gaussian(x, y, M_PI + dist, M_PI, width);
                }
    }
    omega0.forward();
    Vector V0(dims, false);
    {
        Scalar psi(dims, false);
        array<INT, DIM> ind;
        for (ind[2] = 0; ind[2] < dims[2]; ind[2]++)
            for (ind[1] = 0; ind[1] < dims[1]; ind[1]++)
                for (ind[0] = 0; ind[0] < dims[0]; ind[0]++) {
                    array<double, DIM> k = psi.freq(ind);
                    if (k[0] != 0 || k[1] != 0) psi[ind] = -1. / (k[0] * k[0] + k[1] * k[1]) * omega0[2].get(ind);
                }
        Vector grad = psi.gradient();
        V0[0] = -1. * grad[1];
        V0[1] = grad[0];
    }

    Particles parts(Np > 0 ? Np : 0);
    int nsave = 0;
    Vector V1;
    double wtime = 0, divavg = 0;
    if (mode == "api") {
        Vector omega1(dims, false), omega(dims, false), V(dims, false), Vhalf(dims, false);
        Scalar Pi0(dims, false), G(dims, false), div(dims, false);
        Pi0 = (V0.cross(omega0) + nu * V0.laplacian()).divergence();
        Pi0.laplacian_inverse();
        V1 = V0 + K * (V0.cross(omega0) + nu * V0.laplacian() - Pi0.gradient());
        omega1 = V1.curl();
        Vector F, cross1, lap1;
        Vector cross0 = V0.cross(omega0);
        check(sh_stream_synchronize(nullptr), "sync");
        auto t0 = chrono::steady_clock::now();
        for (int t = 1; t < Nt + 1; t++) {
            cross1 = V1.cross(omega1);
            F = K / 2 * (3. * cross1 - cross0);
            lap1 = nu * K * V1.laplacian();
            G = (V1 + lap1 + F).divergence();
            G.laplacian_inverse();
            Vhalf = V1 + F - G.gradient();
            V = Vhalf;
            for (int d = 0; d < DIM; d++) V[d].laplacian_inverse(1, -nu * K);
            if (Np > 0) {
                // a fresh vtmp each step, as exec/ns2.cpp:192: an initialised
                // target keeps its physical flag on assignment
                Vector vtmp;
                V.inverse(vtmp);
                parts.move(vtmp, K);
                if (!pdir.empty() && (t - 1) % mod == 0) writeParticles(parts, pdir, nsave++);
            }
            omega = V.curl();
            div = V.divergence();
            divavg = div.average();
            if (divavg > 1e-12) cout << "avg div = " << divavg << endl;
            V0 = V1;
            V1 = V;
            omega0 = omega1;
            omega1 = omega;
            cross0 = cross1;
        }
        check(sh_stream_synchronize(nullptr), "sync");
        wtime = chrono::duration<double>(chrono::steady_clock::now() - t0).count();
    } else if (mode == "fused") {
        void *v0 = nullptr, *w0 = nullptr;
        check(sh_malloc(&v0, 3 * n * sizeof(cdouble)), "sh_malloc");
        check(sh_malloc(&w0, 3 * n * sizeof(cdouble)), "sh_malloc");
        for (int d = 0; d < 3; d++) {
            check(sh_memcpy_d2d((cdouble*)v0 + d * n, V0[d].deviceData(), n * sizeof(cdouble), nullptr), "copy");
            check(sh_memcpy_d2d((cdouble*)w0 + d * n, omega0[d].deviceData(), n * sizeof(cdouble), nullptr), "copy");
        }
        sh_ns_t h = nullptr;
        check(sh_ns_create(&h, (int)N, (int)N, (int)N, nu, K, SH_F64, 1, nullptr), "sh_ns_create");
        check(sh_ns_set_scheme(h, 1), "sh_ns_set_scheme");
        check(sh_ns_set_initial(h, v0), "sh_ns_set_initial");
        check(sh_ns_set_vorticity(h, w0), "sh_ns_set_vorticity");
        check(sh_ns_startup(h), "sh_ns_startup");
        Vector vtmp(dims, true);
        check(sh_stream_synchronize(nullptr), "sync");
        auto t0 = chrono::steady_clock::now();
        if (Np > 0) {
            // the particles need the physical velocity after every step
            for (int t = 1; t < Nt + 1; t++) {
                check(sh_ns_steps(h, 1), "sh_ns_steps");
                check(sh_ns_get_velocity(h, v0), "velocity");
                check(sh_fft3d_many_z2z(v0, 3, n, (int)N, (int)N, (int)N, +1, SH_NORMALIZE, nullptr), "inverse");
                for (int d = 0; d < 3; d++)
                    check(sh_memcpy_d2d(vtmp[d].deviceDataMutable(), (cdouble*)v0 + d * n, n * sizeof(cdouble),
                                        nullptr),
                          "copy");
                parts.move(vtmp, K);
                if (!pdir.empty() && (t - 1) % mod == 0) writeParticles(parts, pdir, nsave++);
            }
        } else {
            check(sh_ns_steps(h, Nt), "sh_ns_steps");
        }
        check(sh_stream_synchronize(nullptr), "sync");
        wtime = chrono::duration<double>(chrono::steady_clock::now() - t0).count();
        check(sh_ns_div_average(h, &divavg), "div");
        check(sh_ns_get_velocity(h, v0), "velocity");
        V1.init(dims, false);
        for (int d = 0; d < 3; d++)
            check(sh_memcpy_d2d(V1[d].deviceDataMutable(), (cdouble*)v0 + d * n, n * sizeof(cdouble), nullptr),
                  "copy");
        sh_ns_destroy(h);
        sh_free(v0);
        sh_free(w0);
    } else {
        cout << "unknown mode " << mode << endl;
        return 2;
    }
    cout << "n=" << upcxx::rank_n() << ", t=1, Wtime = " << wtime << endl;
    cout << "steps/s = " << (wtime > 0 ? Nt / wtime : 0) << ", div avg = " << divavg << endl;
    if (!out.empty()) {
        FILE* f = fopen(out.c_str(), "wb");
        if (!f) {
            cout << "cannot open " << out << endl;
            return 1;
        }
        std::vector<cdouble> hbuf(n);
        for (int d = 0; d < 3; d++) {
            check(sh_memcpy_d2h(hbuf.data(), V1[d].deviceData(), n * sizeof(cdouble)), "d2h");
            fwrite(hbuf.data(), sizeof(cdouble), n, f);
        }
        fclose(f);
    }
    upcxx::finalize();
    return 0;
}

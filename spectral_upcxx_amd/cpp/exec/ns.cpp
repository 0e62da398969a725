// ns -- Navier-Stokes driver on the drop-in API (behaviour of exec/ns.cpp).
//
//   ns N [steps] [ic] [nu] [K] [mode] [out.bin]
//     ic   : tg (Taylor-Green) | sol:<seed>:<k0> (random solenoidal, u_rms 1)
//            | vortex (three Gaussian vortices, exec/ns.cpp:97-131; omega0 is
//              the vortices' vorticity, as there, not curl(V0))
//     mode : api   -- the reference's operation sequence through Scalar/Vector/
//                     Operators (exec/ns.cpp:132-217), one device kernel per op
//            fused -- the five-kernel step of the C ABI (sh_ns_*)
//     out  : writes the final velocity V1 (3 Fourier fields, complex128)
// Prints the wall time of the step loop (exec/ns.cpp:162-163, 224-228).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "DataObjects.H"
#include "spectral_hip.h"

using namespace std;

static double gaussian(double x, double y, double xc, double yc, double width) {
    return exp(-pow(x - xc, 2) / width) * exp(-pow(y - yc, 2) / width);
}

static void check(int rc, const char* what) {
    if (rc != 0) {
        cout << "ERROR: " << what << ": " << sh_last_error() << endl;
        exit(1);
    }
}

// initial velocity V0 in Fourier space; omega0 is set (and true returned) when
// the IC defines the initial vorticity itself
static bool initial_velocity(Vector& V0, Vector& omega0_out, const string& ic, array<INT, DIM> dims) {
    const int nx = (int)dims[0], ny = (int)dims[1], nz = (int)dims[2];
    if (ic == "tg") {
        Vector u(dims, true);
        array<INT, DIM> ind;
        for (ind[2] = 0; ind[2] < dims[2]; ind[2]++)
            for (ind[1] = 0; ind[1] < dims[1]; ind[1]++)
                for (ind[0] = 0; ind[0] < dims[0]; ind[0]++) {
                    double x = 2. * M_PI / nx * ind[0], y = 2. * M_PI / ny * ind[1], z = 2. * M_PI / nz * ind[2];
                    u[0][ind] = sin(x) * cos(y) * cos(z);
                    u[1][ind] = -cos(x) * sin(y) * cos(z);
                }
        u.forward();
        V0 = u;
        return false;
    } else if (ic.rfind("sol", 0) == 0) {
        unsigned long long seed = 2024;
        double k0 = 4.0;
        sscanf(ic.c_str(), "sol:%llu:%lf", &seed, &k0);
        void* buf = nullptr;
        const long n = (long)nx * ny * nz;
        check(sh_malloc(&buf, 3 * n * sizeof(cdouble)), "sh_malloc");
        check(sh_solenoidal_ic_z(buf, nx, ny, nz, seed, k0, 1.0, nullptr), "sh_solenoidal_ic_z");
        for (int d = 0; d < 3; d++)
            check(sh_memcpy_d2d(V0[d].deviceDataMutable(), (cdouble*)buf + d * n, n * sizeof(cdouble), nullptr),
                  "copy");
        sh_free(buf);
        return false;
    } else if (ic == "vortex") {
        // vorticity of three Gaussian vortices, streamfunction, V0 = (-d_y psi, d_x psi, 0)
        Vector omega0(dims);
        double width = M_PI / 8, mag = 50, dist = M_PI / 2;
        array<INT, DIM> ind;
        for (ind[2] = 0; ind[2] < dims[2]; ind[2]++)
            for (ind[1] = 0; ind[1] < dims[1]; ind[1]++)
                for (ind[0] = 0; ind[0] < dims[0]; ind[0]++) {
                    double x = 2. * M_PI / nx * ind[0], y = 2. * M_PI / ny * ind[1];
                    omega0[2][ind] = mag * gaussian(x, y, M_PI - dist, M_PI - dist, width) +
                                     mag * gaussian(x, y, M_PI + dist, M_PI - dist, width) +
                                     mag * gaussian(x, y, M_PI, M_PI + dist, width);
                }
        omega0.forward();
        Scalar psi(dims, false);
        for (ind[2] = 0; ind[2] < dims[2]; ind[2]++)
            for (ind[1] = 0; ind[1] < dims[1]; ind[1]++)
                for (ind[0] = 0; ind[0] < dims[0]; ind[0]++) {
                    array<double, DIM> k = psi.freq(ind);
                    if (k[0] != 0 || k[1] != 0) psi[ind] = -1. / (k[0] * k[0] + k[1] * k[1]) * omega0[2].get(ind);
                }
        Vector grad = psi.gradient();
        V0[0] = -1. * grad[1];
        V0[1] = grad[0];
        omega0_out = omega0;
        return true;
    }
    cout << "unknown ic " << ic << endl;
    exit(2);
}

int main(int argc, char* argv[]) {
    upcxx::init();
    INT N = argc > 1 ? atoi(argv[1]) : 16;
    int Nt = argc > 2 ? atoi(argv[2]) : 100;
    string ic = argc > 3 ? argv[3] : "tg";
    double nu = argc > 4 ? atof(argv[4]) : 0.01;
    double K = argc > 5 ? atof(argv[5]) : 1e-3;
    string mode = argc > 6 ? argv[6] : "api";
    string out = argc > 7 ? argv[7] : "";
    array<INT, DIM> dims = {N, N, N};
    const long n = (long)N * N * N;

    Vector V0(dims, false), omega_ic;
    const bool has_omega = initial_velocity(V0, omega_ic, ic, dims);

    Vector V1;
    double wtime = 0, divavg = 0;
    if (mode == "api") {
        Vector omega0 = has_omega ? omega_ic : V0.curl();
        Vector omega1(dims, false), omega(dims, false), V(dims, false), Vhalf(dims, false);
        Scalar Pi0(dims, false), G(dims, false), div(dims, false);
        // start-up: Pi0, Forward Euler, omega1, cross0 (exec/ns.cpp:132-153)
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
            lap1 = nu * K / 2 * V1.laplacian();
            G = (V1 + lap1 + F).divergence();
            G.laplacian_inverse();
            Vhalf = V1 + F - G.gradient();
            V = Vhalf + lap1;
            for (int d = 0; d < DIM; d++) V[d].laplacian_inverse(1, -nu * K / 2);
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
        void* v0 = nullptr;
        check(sh_malloc(&v0, 3 * n * sizeof(cdouble)), "sh_malloc");
        for (int d = 0; d < 3; d++)
            check(sh_memcpy_d2d((cdouble*)v0 + d * n, V0[d].deviceData(), n * sizeof(cdouble), nullptr), "copy");
        sh_ns_t h = nullptr;
        check(sh_ns_create(&h, (int)N, (int)N, (int)N, nu, K, SH_F64, 1, nullptr), "sh_ns_create");
        check(sh_ns_set_initial(h, v0), "sh_ns_set_initial");
        if (has_omega) {
            void* w0 = nullptr;
            check(sh_malloc(&w0, 3 * n * sizeof(cdouble)), "sh_malloc");
            for (int d = 0; d < 3; d++)
                check(sh_memcpy_d2d((cdouble*)w0 + d * n, omega_ic[d].deviceData(), n * sizeof(cdouble), nullptr),
                      "copy");
            check(sh_ns_set_vorticity(h, w0), "sh_ns_set_vorticity");
            check(sh_stream_synchronize(nullptr), "sync");
            sh_free(w0);
        }
        check(sh_ns_startup(h), "sh_ns_startup");
        check(sh_stream_synchronize(nullptr), "sync");
        auto t0 = chrono::steady_clock::now();
        check(sh_ns_steps(h, Nt), "sh_ns_steps");
        check(sh_stream_synchronize(nullptr), "sync");
        wtime = chrono::duration<double>(chrono::steady_clock::now() - t0).count();
        check(sh_ns_div_average(h, &divavg), "div");
        check(sh_ns_get_velocity(h, v0), "velocity");
        V1.init(dims, false);
        for (int d = 0; d < 3; d++)
            check(sh_memcpy_d2d(V1[d].deviceDataMutable(), (cdouble*)v0 + d * n, n * sizeof(cdouble), nullptr), "copy");
        sh_ns_destroy(h);
        sh_free(v0);
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
        std::vector<cdouble> h(n);
        for (int d = 0; d < 3; d++) {
            check(sh_memcpy_d2h(h.data(), V1[d].deviceData(), n * sizeof(cdouble)), "d2h");
            fwrite(h.data(), sizeof(cdouble), n, f);
        }
        fclose(f);
    }
    upcxx::finalize();
    return 0;
}

// test_operators -- the checks of test/operators.cpp on the drop-in API, with
// relative tolerance 1e-12 (the reference uses abs 1e-3) and a non-zero exit
// status when any check fails.
//   test_operators [N]
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "DataObjects.H"

using namespace std;

static int failures = 0;

static void report(const string& name, bool wrong) {
    cout << name << (wrong ? " failed" : " passed") << endl;
    if (wrong) failures++;
}

int main(int argc, char* argv[]) {
    upcxx::init();
    INT N = argc > 1 ? atoi(argv[1]) : 32;
    array<INT, DIM> dims = {N, N, N};
    const double tol = 1e-12;
    Scalar s(dims);
    array<INT, DIM> start = s.getPStart(), end = s.getPEnd(), ind = {0, 0, 0};
#define FOR_ALL                                                   \
    for (ind[2] = start[2]; ind[2] < end[2]; ind[2]++)            \
        for (ind[1] = start[1]; ind[1] < end[1]; ind[1]++)        \
            for (ind[0] = start[0]; ind[0] < end[0]; ind[0]++)
    FOR_ALL s[ind] = sin(2 * M_PI / N * ind[0]);
    Scalar s_f(dims);
    s.forward(s_f);

    bool wrong = false;
    Scalar sum = s + s;
    FOR_ALL if (abs(sum[ind].real() - 2 * sin(2 * M_PI / N * ind[0])) > tol) wrong = true;
    report("addition", wrong);

    wrong = false;
    Scalar sub = s - s;
    FOR_ALL if (abs(sub[ind]) != 0) wrong = true;
    report("subtraction", wrong);

    wrong = false;
    Scalar prod = s * s;
    FOR_ALL if (abs(prod[ind].real() - pow(sin(2 * M_PI / N * ind[0]), 2)) > tol) wrong = true;
    report("multiplication", wrong);

    wrong = false;
    Vector grad = s_f.gradient();
    grad.inverse();
    FOR_ALL {
        if (abs(grad[0][ind].real() - cos(2 * M_PI / N * ind[0])) > tol) wrong = true;
        if (abs(grad[1][ind]) > tol || abs(grad[2][ind]) > tol) wrong = true;
    }
    report("gradient", wrong);

    wrong = false;
    Scalar lap = s_f.laplacian();
    lap.inverse();
    FOR_ALL if (abs(lap[ind].real() + sin(2 * M_PI / N * (double)ind[0])) > tol) wrong = true;
    report("laplacian", wrong);

    wrong = false;
    Scalar lap_inv;
    lap.forward(lap_inv);
    lap_inv.laplacian_inverse();
    lap_inv.inverse();
    FOR_ALL if (abs(lap_inv[ind].real() - s[ind].real()) > tol) wrong = true;
    report("laplacian inverse", wrong);

    wrong = false;
    Vector v(dims);
    v[0] = s;
    v.forward();
    Scalar div = v.divergence();
    div.inverse();
    FOR_ALL if (abs(div[ind].real() - cos(2 * M_PI / N * (double)ind[0])) > tol) wrong = true;
    report("divergence", wrong);

    wrong = false;
    Vector vlap = v.laplacian();
    vlap.inverse();
    FOR_ALL if (abs(vlap[0][ind].real() + sin(2 * M_PI / N * (double)ind[0])) > tol) wrong = true;
    report("vector laplacian", wrong);

    wrong = false;
    Vector v1(dims), v2(dims);
    for (INT i = 0; i < v1[0].localSize(); i++) {
        v1[2][i] = 1.;
        v2[1][i] = 1.;
    }
    v1.forward();
    v2.forward();
    Vector cross = v1.cross(v2);
    cross.inverse();
    FOR_ALL {
        if (abs(cross[0][ind].real() + 1) > tol) wrong = true;
        if (abs(cross[1][ind]) > tol) wrong = true;
        if (abs(cross[2][ind]) > tol) wrong = true;
    }
    report("cross", wrong);

    wrong = false;
    grad.forward();
    Vector curl = grad.curl();
    curl.inverse();
    for (int d = 0; d < DIM; d++)
        for (INT i = 0; i < curl[0].localSize(); i++)
            if (abs(curl[d][i]) > tol) wrong = true;
    report("curl", wrong);

    wrong = false;
    Vector g2 = s_f.gradient();
    Vector mult = 2. * g2;
    FOR_ALL for (int d = 0; d < DIM; d++) if (mult[d][ind] != 2. * g2[d][ind]) wrong = true;
    report("double * vector", wrong);

    wrong = false;
    Vector diff = mult - g2;
    FOR_ALL for (int d = 0; d < DIM; d++) if (abs(diff[d][ind] - g2[d][ind]) > tol * abs(g2[d][ind]) + 1e-300) wrong = true;
    report("vector subtraction", wrong);

    // the reference quirks kept by the API
    report("vector isPhysical is always false", v1.isPhysical());
    Scalar phys(dims, true), four(dims, false);
    four = phys;  // operator= keeps the initialised target's flag
    report("operator= keeps the target flag", !four.isFourier());

    cout << (failures ? "SOME CHECKS FAILED" : "all operator checks passed") << endl;
    upcxx::finalize();
    return failures ? 1 : 0;
}

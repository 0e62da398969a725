// test_constructor -- Vector construction and copy (behaviour of
// test/constructor.cpp), with the checks the reference leaves implicit: the
// copy is deep, zero-filled and keeps the source's space flags.  Exit status 1
// on failure.
#include <iostream>

#include "DataObjects.H"

using namespace std;

int main() {
    upcxx::init();
    INT N = 16;
    array<INT, DIM> dims = {N, N, N};
    Vector x(dims);
    Vector y = x;
    bool ok = !y.isFourier() && y.getDims() == dims && y.localSize() == x.localSize();
    for (int d = 0; d < DIM && ok; d++)
        for (INT i = 0; i < y[d].localSize() && ok; i++) ok = y[d][i] == cdouble(0, 0);
    // deep: changing the copy leaves the source alone
    y[0][5] = cdouble(1, 2);
    ok = ok && x[0].get({5, 0, 0}) == cdouble(0, 0) && y[0].get({5, 0, 0}) == cdouble(1, 2);
    Vector f(dims, false);
    Vector g = f;
    ok = ok && g.isFourier();
    cout << (ok ? "constructor passed" : "constructor failed") << endl;
    upcxx::finalize();
    return ok ? 0 : 1;
}

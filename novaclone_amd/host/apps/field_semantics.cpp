// field_semantics.cpp -- value semantics of the C++ drop-in's Field algebra
// and host mirror, checked on a small periodic grid:
//   * assigning a named FieldPtr copies it (the FieldPtr keeps its value when
//     the target changes afterwards), also when two FieldPtrs share it;
//   * assigning an unshared expression result takes over its buffer and
//     gives the expression's value, also when the expression reads the
//     target itself (U = U + U);
//   * host writes through operator()(i,j,k) reach the next device operation,
//     and a device result is read back through the (page-locked) mirror.
// The reference's Field (src/Field/Field.h:160-367) has these value
// semantics: every operator returns a new temporary and assignment copies.
// Values are compared to 1e-13 (the device and the host round sums of
// products alike but not always identically).  Prints "field_semantics: ok"
// and exits 0, or names the first failed check and exits 1.
//
// usage: field_semantics [N]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "nova/nova.hpp"

namespace {

int fails = 0;
// the checks compare against host arithmetic on the initial field; a copy
// that shared a buffer, or a takeover that lost one, differs by O(1)
constexpr double tol = 1e-13;

void check(bool ok, const std::string &what)
{
    if (!ok && fails++ == 0) std::fprintf(stderr, "field_semantics: FAILED %s\n", what.c_str());
}

// max |a - s b| over the interior nodes (host mirror)
template <class F1, class F2>
double maxdiff(const F1 &a, const F2 &b, double s)
{
    const int h = settings::m() / 2;
    double m = 0.0;
    for (int i = h; i < a.ni() - h; i++)
        for (int j = h; j < a.nj() - h; j++)
            for (int k = h; k < a.nk() - h; k++) {
                const vector &x = a(i, j, k);
                const vector &y = b(i, j, k);
                for (int c = 0; c < 3; c++) m = std::fmax(m, std::fabs(x[c] - s * y[c]));
            }
    return m;
}

}  // namespace

int main(int argc, char *argv[])
{
    try {
        const int N = argc > 1 ? std::atoi(argv[1]) : 16;
        Time time(0.01, 0.005, 1);
        const scalar L = 2.0 * tools::pi;
        Mesh mesh({N, N, N}, {L, L, L}, time);
        const int h = settings::m() / 2;
        Field<vector> U(mesh, vector(0, 0, 0));
        for (int i = h; i < mesh.ni() - h; i++)
            for (int j = h; j < mesh.nj() - h; j++)
                for (int k = h; k < mesh.nk() - h; k++)
                    U(i, j, k) = vector(std::sin(0.3 * i + 0.1 * k), std::cos(0.2 * j), 0.01 * (i + 2 * j - k));
        U.correctBoundaryConditions();
        Field<vector> U0(U);  // a copy of the initial field, made on the device from the host writes

        // host writes reached the device: 2U computed on the device equals 2 x the host values
        {
            FieldPtr<vector> t = 2.0 * U;
            check(maxdiff(*t, U0, 2.0) <= tol, "host writes through operator() reach the device");
        }
        // a named FieldPtr is copied: changing the target leaves it as it was
        {
            Field<vector> V(mesh, vector(0, 0, 0));
            FieldPtr<vector> t = 3.0 * U;
            V = t;
            V += U;
            check(maxdiff(*t, U0, 3.0) <= tol, "a named FieldPtr keeps its value after V = t; V += U");
            check(maxdiff(V, U0, 4.0) <= tol, "V = t copies t's value");
            FieldPtr<vector> t2 = t;  // shared by two names
            V = std::move(t);
            V += U;
            check(maxdiff(*t2, U0, 3.0) <= tol, "a shared FieldPtr moved into V keeps its value");
        }
        // an unshared expression result: V takes its buffer and its value
        {
            Field<vector> V(mesh, vector(1, 1, 1));
            V = U + U;
            check(maxdiff(V, U0, 2.0) <= tol, "V = U + U");
            Field<vector> W(U);
            W = W + W;  // reads the target
            check(maxdiff(W, U0, 2.0) <= tol, "W = W + W");
            W = W - U;
            check(maxdiff(W, U0, 1.0) <= tol, "W = W - U after the buffer exchange");
            check(maxdiff(U, U0, 1.0) <= tol, "U unchanged by the assignments to V and W");
        }
        // a device result read back through the mirror after host reads
        {
            double s = 0.0;
            for (int i = h; i < mesh.ni() - h; i++) s += U(i, h, h)[0];
            Field<vector> V(mesh, vector(0, 0, 0));
            V = 0.5 * U;
            check(maxdiff(V, U0, 0.5) <= tol, "V = 0.5 U after host reads of U");
            check(std::isfinite(s), "host reads");
        }
        if (fails) return 1;
        std::printf("field_semantics: ok\n");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "field_semantics: %s\n", e.what());
        return 1;
    }
    return 0;
}

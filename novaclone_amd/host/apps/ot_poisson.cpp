// ot_poisson.cpp -- the incompressible Orszag-Tang case of the reference's
// OT_Poisson app (app/OT_Poisson/OT.cpp:30-190 with UEqn.H and BEqn.H)
// written against the C++ drop-in API with the app's own expressions: the
// frozen-stage RK4 of U with the Lorentz terms div(B, B) - grad(0.5*(B&B))
// and its projection, the RK4 of B with U frozen and B reassigned every
// stage, the cleaning of B, J = ex::curl(B), tools::CFL, and the per-step
// diagnostics Ek, Em, Jav, Jmax, omega, epsilon, Umax of OT.cpp:134-176
// written as "t Ek Em Jav Jmax omega epsilon -dEk/dt dEm/dt Umax".
//
// usage: ot_poisson N steps dt mu eta [out.dat]
#include <cstdio>
#include <iomanip>
#include <iostream>

#include "nova/nova.hpp"

int main(int argc, char *argv[])
{
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s N steps dt mu eta [out.dat]\n", argv[0]);
        return 2;
    }
    try {
        const int N = std::atoi(argv[1]);
        const int steps = std::atoi(argv[2]);
        const scalar dt = std::atof(argv[3]);
        const scalar mu = std::atof(argv[4]);
        const scalar eta = std::atof(argv[5]);
        Time time(dt, steps * dt - 0.5 * dt, 1);
        const scalar L = 2.0 * tools::pi;
        Mesh mesh({N, N, N}, {L, L, L}, time);
        Field<vector> U(mesh, vector(0, 0, 0), "U");
        Field<vector> Ustar(mesh, vector(0, 0, 0), "U");
        Field<vector> B(mesh, vector(0, 0, 0), "B");
        Field<vector> Bstar(mesh, vector(0, 0, 0), "B");
        Field<vector> J(mesh, vector(0, 0, 0), "B");
        Field<vector> omegav(mesh, vector(0, 0, 0), "U");
        auto p_ptr = std::make_shared<Field<scalar>>(mesh, 0.0, "p");
        auto &p = *p_ptr;
        auto pB_ptr = std::make_shared<Field<scalar>>(mesh, 0.0, "pB");
        auto &pB = *pB_ptr;
        poisson pEqn(p_ptr);
        poisson pBEqn(pB_ptr);
        const int h = settings::m() / 2;
        // OT.cpp:73-88
        for (int i = h; i < mesh.ni() - h; i++)
            for (int j = h; j < mesh.nj() - h; j++)
                for (int k = h; k < mesh.nk() - h; k++) {
                    const scalar x = (i - h) * mesh.dx() + mesh.origin().x();
                    const scalar y = (j - h) * mesh.dy() + mesh.origin().y();
                    const scalar z = (k - h) * mesh.dz() + mesh.origin().z();
                    U(i, j, k) = vector((-2.0) * std::sin(y), 2.0 * std::sin(x), 0.0);
                    const vector b(((-2.0) * std::sin(2.0 * y)) + std::sin(z), ((2.0 * std::sin(x)) + std::sin(z)),
                                   (std::sin(x) + std::sin(y)));
                    B(i, j, k) = vector(0.8 * b.x(), 0.8 * b.y(), 0.8 * b.z());
                }
        U.correctBoundaryConditions();
        p.correctBoundaryConditions();
        B.correctBoundaryConditions();
        pB.correctBoundaryConditions();
        std::ofstream data(argc > 6 ? argv[6] : "dat.dat");
        scalar EkOld = 0.0, EmOld = 0.0;
        while (time.run()) {
            time++;
            {  // UEqn.H
                std::array<FieldPtr<vector>, 4> dU;
                Field<vector> UOld(U);
                for (int rk = 0; rk < 4; rk++) {
                    dU[rk] = time.dt() * (-ex::div(U, U) + ex::laplacian(mu, U) + ex::div(B, B) -
                                          ex::grad(0.5 * (B & B)));
                    Ustar = UOld;
                    for (int jj = 0; jj <= rk; jj++) Ustar = Ustar + settings::RK4a()[rk][jj] * dU[jj];
                }
                Ustar.correctBoundaryConditions();
                pEqn.rhs(ex::div(Ustar) / time.dt());
                pEqn.solve();
                p.correctBoundaryConditions();
                U = Ustar - time.dt() * ex::grad(p);
                U.correctBoundaryConditions();
            }
            {  // BEqn.H
                std::array<FieldPtr<vector>, 4> dB;
                Field<vector> Bold(B);
                for (int rk = 0; rk < 4; rk++) {
                    dB[rk] = time.dt() * (-ex::div(U, B) + ex::div(B, U) + ex::laplacian(eta, B));
                    Bstar = Bold;
                    for (int jj = 0; jj <= rk; jj++) Bstar = Bstar + settings::RK4a()[rk][jj] * dB[jj];
                    B = Bstar;
                    B.correctBoundaryConditions();
                }
                pBEqn.rhs(ex::div(B));
                pBEqn.solve();
                pB.correctBoundaryConditions();
                B -= ex::grad(pB);
                B.correctBoundaryConditions();
                J = ex::curl(B);
                J.correctBoundaryConditions();
            }
            if (time.writelog()) tools::CFL(U, mesh);
            // OT.cpp:134-176
            scalar Ek = 0.0, Em = 0.0, Jmax = 0.0, Umax = 0.0, Jav = 0.0, omega = 0.0;
            int n = 0;
            omegav = ex::curl(U);
            for (int i = h; i < mesh.ni() - h - 1; i++)
                for (int j = h; j < mesh.nj() - h - 1; j++)
                    for (int k = h; k < mesh.nk() - h - 1; k++) {
                        const vector u = U(i, j, k), b = B(i, j, k), jj = J(i, j, k), w = omegav(i, j, k);
                        Ek += 0.5 * (u.x() * u.x() + u.y() * u.y() + u.z() * u.z());
                        Em += 0.5 * (b.x() * b.x() + b.y() * b.y() + b.z() * b.z());
                        Jmax = std::max(Jmax, std::sqrt(jj.x() * jj.x() + jj.y() * jj.y() + jj.z() * jj.z()));
                        Umax = std::max(Umax, std::sqrt(u.x() * u.x() + u.y() * u.y() + u.z() * u.z()));
                        omega += std::sqrt(w.x() * w.x() + w.y() * w.y() + w.z() * w.z());
                        Jav += std::sqrt(jj.x() * jj.x() + jj.y() * jj.y() + jj.z() * jj.z());
                        n++;
                    }
            Ek /= n;
            Em /= n;
            Jav /= n;
            omega /= n;
            const scalar epsilon = -mu * (omega * omega) - eta * (Jav * Jav);
            data << time.curTime() << " " << std::setprecision(15) << Ek << " " << Em << " " << Jav << " " << Jmax
                 << " " << omega << " " << epsilon << " " << -(Ek - EkOld) / time.dt() << " "
                 << (Em - EmOld) / time.dt() << " " << Umax << "\n";
            EkOld = Ek;
            EmOld = Em;
        }
        std::cout << "done " << time.timeStep() << " steps" << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "ot_poisson: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

// tg_wc.cpp -- the weakly-compressible Taylor-Green case of the reference's
// TG app (app/TG/TG.cpp:25-120 with rhoEqn.H, pEqn.H and UEqn.H) written
// against the C++ drop-in API with the app's own expressions: a classic RK4
// of rho with U frozen (-ex::div(rho*U)), p = 50.0*(rho*rho), and a classic
// RK4 of rho U with rho and p frozen (-ex::div(rho*U, U) - ex::grad(p) +
// ex::laplacian(mu, U), U = rhoUOld/rho + c*(drhoU/rho)), then the kinetic
// energy history "t Ek -dEk/dt" of TG.cpp:91-113.
//
// usage: tg_wc N steps dt mu [out.dat]
#include <cstdio>
#include <iomanip>
#include <iostream>

#include "nova/nova.hpp"

int main(int argc, char *argv[])
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s N steps dt mu [out.dat]\n", argv[0]);
        return 2;
    }
    try {
        const int N = std::atoi(argv[1]);
        const int steps = std::atoi(argv[2]);
        const scalar dt = std::atof(argv[3]);
        const scalar mu = std::atof(argv[4]);
        Time time(dt, steps * dt - 0.5 * dt, 1);
        const scalar pi = 3.1415926536;  // TG.cpp:25
        const scalar L = 2.0 * tools::pi;
        Mesh mesh({N, N, N}, {L, L, L}, time);
        Field<vector> U(mesh, vector(1, 0, 0), "U");
        Field<scalar> rho(mesh, 1.0, "rho");
        Field<scalar> p(mesh, 0, "p");
        const int h = settings::m() / 2;
        for (int i = h; i < mesh.ni() - h; i++)  // TG.cpp:38-52
            for (int j = h; j < mesh.nj() - h; j++)
                for (int k = h; k < mesh.nk() - h; k++) {
                    const scalar x = -pi + (i - h) * mesh.dx() + mesh.origin().x();
                    const scalar y = -pi + (j - h) * mesh.dy() + mesh.origin().y();
                    const scalar z = -pi + (k - h) * mesh.dz() + mesh.origin().z();
                    U(i, j, k) = vector(std::sin(x) * std::cos(y) * std::cos(z),
                                        -std::cos(x) * std::sin(y) * std::cos(z), 0);
                    p(i, j, k) = 50 + 1.0 / 16.0 * (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2);
                    rho(i, j, k) = std::pow(p(i, j, k) / 50.0, 0.5);
                }
        U.correctBoundaryConditions();
        p.correctBoundaryConditions();
        rho.correctBoundaryConditions();
        const scalar RK4[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
        std::ofstream data(argc > 5 ? argv[5] : "dat.dat");
        scalar EkOld = 0.0;
        while (time.run()) {
            time++;
            {  // rhoEqn.H
                auto drho = nova::tmp<scalar>(mesh);
                auto rhoc = nova::tmp<scalar>(mesh);
                auto rhoOld = nova::tmp<scalar>(mesh);
                *rhoOld = rho;
                *rhoc = rho;
                for (int rk = 0; rk < 4; rk++) {
                    *drho = time.dt() * (-ex::div(rho * U));
                    *rhoc = rhoc + RK4[rk] * drho;
                    if (rk < 3) rho = rhoOld + (int(rk / 2) + 1) / 2.0 * drho;
                    else rho = rhoc;
                    rho.correctBoundaryConditions();
                }
            }
            {  // pEqn.H
                p = 50.0 * (rho * rho);
            }
            {  // UEqn.H
                auto drhoU = nova::tmp<vector>(mesh);
                auto rhoUOld = nova::tmp<vector>(mesh);
                auto rhoUc = nova::tmp<vector>(mesh);
                *rhoUc = rho * U;
                *rhoUOld = rho * U;
                for (int rk = 0; rk < 4; rk++) {
                    *drhoU = time.dt() * (-ex::div(rho * U, U) - ex::grad(p) + ex::laplacian(mu, U));
                    *rhoUc = rhoUc + RK4[rk] * drhoU;
                    if (rk < 3) U = rhoUOld / rho + (int(rk / 2) + 1) / 2.0 * (drhoU / rho);
                    else U = rhoUc / rho;
                    U.correctBoundaryConditions();
                }
            }
            if (time.writelog()) tools::CFL(U, mesh);
            scalar Ek = 0.0;
            int n = 0;
            for (int i = h; i < mesh.ni() - h - 1; i++)
                for (int j = h; j < mesh.nj() - h - 1; j++)
                    for (int k = h; k < mesh.nk() - h - 1; k++) {
                        const vector u = U(i, j, k);
                        Ek += 0.5 * (u.x() * u.x() + u.y() * u.y() + u.z() * u.z());
                        n++;
                    }
            reduce(Ek, plusOp<scalar>());
            reduce(n, plusOp<int>());
            Ek /= n;
            data << time.curTime() << " " << std::setprecision(15) << Ek << " " << std::setprecision(15)
                 << -(Ek - EkOld) / time.dt() << "\n";
            EkOld = Ek;
        }
        std::cout << "done " << time.timeStep() << " steps" << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "tg_wc: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

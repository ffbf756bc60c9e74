// mhdjet.cpp -- the reference's mhdjet app (app/mhdjet/mhdjet.cpp:134-470
// with UEqn.H and BEqn.H) written against the C++ drop-in API: a Bickley
// jet u = sech^2(y) with random perturbations drawn as the app draws them
// (std::default_random_engine seeded with the rank, uniform on [-1, 1], the
// same calls in the same order) under a uniform field B = (0.015, 0, 0);
// per step the OT_Poisson U and B blocks, tools::CFL, and at the steps the
// clock writes, the HDF5 dumps ./h5data/{Ux, Uy_, Uz_, p_, Bx_, By_, Bz_,
// Jx_, Jy_, Jz_}<t>.h5 (t = curTime - dt, std::to_string).
//
// usage: mhdjet Nx Ny Nz steps dt writeEvery [ic_dir]
// (ic_dir: the initial U and B are also dumped there, ic_U{x,y,z}.h5 and
// ic_B{x,y,z}.h5, so that a checker can start from the same state.)
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <random>
#include <sys/stat.h>

#include "nova/nova.hpp"

int main(int argc, char *argv[])
{
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s Nx Ny Nz steps dt writeEvery [ic_dir]\n", argv[0]);
        return 2;
    }
    try {
        const int N[3] = {std::atoi(argv[1]), std::atoi(argv[2]), std::atoi(argv[3])};
        const int steps = std::atoi(argv[4]);
        const scalar dt = std::atof(argv[5]);
        const int writeEvery = std::atoi(argv[6]);
        Time time(dt, steps * dt - 0.5 * dt, writeEvery);
        Mesh mesh({N[0], N[1], N[2]}, {13.96, 20.0, 13.96}, time);  // examples/mhdjet/mesh
        Field<vector> U(mesh, vector(0, 0, 0), "U");
        Field<vector> Ustar(mesh, vector(0, 0, 0), "U");
        Field<vector> B(mesh, vector(0, 0, 0), "B");
        Field<vector> Bstar(mesh, vector(0, 0, 0), "B");
        Field<vector> J(mesh, vector(0, 0, 0), "B");
        auto p_ptr = std::make_shared<Field<scalar>>(mesh, 0.0, "p");
        auto &p = *p_ptr;
        auto pB_ptr = std::make_shared<Field<scalar>>(mesh, 0.0, "pB");
        auto &pB = *pB_ptr;
        const scalar mu = 0.000285714, eta = 0.000285714;
        poisson pEqn(p_ptr);
        poisson pBEqn(pB_ptr);
        std::default_random_engine eng(0);  // parallelCom::myProcNo() on one rank
        std::uniform_real_distribution<double> dist(-1.0, 1.0);
        const int h = settings::m() / 2;
        for (int i = h; i < mesh.ni() - h; i++)  // mhdjet.cpp:181-218
            for (int j = h; j < mesh.nj() - h; j++)
                for (int k = h; k < mesh.nk() - h; k++) {
                    scalar x = (i - h) * mesh.dx() + mesh.origin().x();
                    scalar y = -10 + (j - h) * mesh.dy() + mesh.origin().y();
                    scalar z = (k - h) * mesh.dz() + mesh.origin().z();
                    // the app draws these sums (and discards them): the draws
                    // advance the engine
                    int perU = 0.0;
                    for (int px = 0; px < 2; px++) perU += dist(eng) * sin((px * x) - (dist(eng) * 13.96));
                    for (int px = 3; px < mesh.ni() / 3; px++) perU += dist(eng) * sin((px * x) - (dist(eng) * 13.96));
                    int perZ = 0.0;
                    for (int pz = 0; pz < 2; pz++) perZ += dist(eng) * sin((pz * z) - (dist(eng) * 13.96));
                    for (int pz = 3; pz < mesh.nk() / 3; pz++) perZ += dist(eng) * sin((pz * z) - (dist(eng) * 13.96));
                    (void)perU;
                    (void)perZ;
                    U(i, j, k) = vector(1.0 / (std::cosh(y) * std::cosh(y)), 0.0, 0.0);
                    U(i, j, k).y() += (0.0001 * dist(eng)) * std::exp(-(y * y));
                    U(i, j, k).x() += (0.001 * std::sin(x * 0.9) + 0.0001 * dist(eng)) * std::exp(-(y * y));
                    U(i, j, k).z() += (0.0001 * dist(eng)) * std::exp(-(y * y));
                    B(i, j, k) = vector(0.015, 0.0, 0.0);
                }
        U.correctBoundaryConditions();
        p.correctBoundaryConditions();
        B.correctBoundaryConditions();
        pB.correctBoundaryConditions();
        if (argc > 7) {
            const std::string d = argv[7];
            const char *xyz[3] = {"x", "y", "z"};
            for (int c = 0; c < 3; c++) {
                nova::writeHDF5(d + "/ic_U" + xyz[c] + ".h5", U, c);
                nova::writeHDF5(d + "/ic_B" + xyz[c] + ".h5", B, c);
            }
        }
        mkdir("h5data", 0755);
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
            if (time.write()) {  // mhdjet.cpp:270-404
                const std::string t = std::to_string(time.curTime() - time.dt()) + ".h5";
                nova::writeHDF5("./h5data/Ux" + t, U, 0);
                nova::writeHDF5("./h5data/Uy_" + t, U, 1);
                nova::writeHDF5("./h5data/Uz_" + t, U, 2);
                nova::writeHDF5("./h5data/p_" + t, p, 0);
                nova::writeHDF5("./h5data/Bx_" + t, B, 0);
                nova::writeHDF5("./h5data/By_" + t, B, 1);
                nova::writeHDF5("./h5data/Bz_" + t, B, 2);
                nova::writeHDF5("./h5data/Jx_" + t, J, 0);
                nova::writeHDF5("./h5data/Jy_" + t, J, 1);
                nova::writeHDF5("./h5data/Jz_" + t, J, 2);
                std::cout << "Vector field written to vector_field.h5" << std::endl;
            }
        }
        std::cout << "done " << time.timeStep() << " steps" << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "mhdjet: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

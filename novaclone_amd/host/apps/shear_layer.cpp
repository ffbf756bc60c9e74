// shear_layer.cpp -- the reference's shearLayer app (app/shearLayer/
// shearLayer.cpp:19-200 with UEqn.H; TEqn and BEqn switched off, as there)
// written against the C++ drop-in API: the mesh and the BC files of a zone
// (examples/shearLayer: y walled, u and v neumann 0 and w dirichlet 0 on
// the walls, p neumann 0), the double shear layer of shearLayer.cpp:86-136,
// and per step tools::CFL and UEqn.H's four stages, each projected with
// RK4c and followed by the fills of U*, p and U.  B (periodic in every
// direction by its BBC file) does not evolve: its force div(B, B) -
// grad(0.5*(B&B)) is formed once on the periodic mesh that file describes
// and added to every stage, where UEqn.H forms it anew each time.
//
// usage: shear_layer zone steps [state.bin]
//   writes U (3 components) and p on the stored nodes, k fastest, as
//   native doubles after the last step.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <random>

#include "nova/nova.hpp"

int main(int argc, char *argv[])
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s zone steps [state.bin]\n", argv[0]);
        return 2;
    }
    try {
        const std::string zone = argv[1];
        const int steps = std::atoi(argv[2]);
        const scalar dt = 0.001;  // Time time(0.001, 420, 1), shearLayer.cpp:24
        Time time(dt, steps * dt - 0.5 * dt, 1);
        Mesh mesh(zone + "/mesh", time);
        Field<vector> U(mesh, vector(0, 0, 0), "U");
        Field<vector> Ustar(mesh, vector(0, 0, 0), "U");
        // B on the periodic grid of its BC file (the same cells and spacings)
        Mesh meshB({mesh.ni() - settings::m() - 1, mesh.nj() - settings::m() - 1, mesh.nk() - settings::m() - 1},
                   {mesh.lx(), mesh.ly(), mesh.lz()}, time, {0, 0, 0}, false, 0, zone);
        Field<vector> B(meshB, vector(0, 0, 0), "B");
        const scalar mu = 1.0 / 28000.0, u0 = 1.0, delta = 1.0 / 28.0, cn = 0.001;  // shearLayer.cpp:65-69
        auto p_ptr = std::make_shared<Field<scalar>>(mesh, (0.87), "p");
        auto &p = *p_ptr;
        poisson pEqn(p_ptr);
        const int h = settings::m() / 2;
        const scalar pi = tools::pi;
        // shearLayer.cpp:86-123 (U on the walled mesh, B on the periodic one)
        for (int i = h; i < mesh.ni() - h; i++)
            for (int k = h; k < mesh.nk() - h; k++)
                for (int j = h; j < mesh.nj() - h; j++) {
                    const scalar x = (i - h) * mesh.dx() + mesh.origin().x();
                    const scalar y = (j - h) * mesh.dy() + mesh.origin().y();
                    const scalar psi = u0 * std::exp(-(((y - 0.5) * (y - 0.5)) / (delta * delta))) *
                                       (std::cos(8.0 * pi * x) + std::cos(20.0 * pi * x));
                    const scalar psi2 = u0 * std::exp(-(((y - 1.5) * (y - 1.5)) / (delta * delta))) *
                                        (std::cos(8.0 * pi * x) + std::cos(20.0 * pi * x));
                    vector u, b;
                    if (y < 1.0) {
                        u = vector((u0 * std::tanh(((4.0 * y) - 2.0) / delta)), 0, 0);
                        u.x() += (cn * mesh.dy() * psi);
                        u.y() += -cn * mesh.dx() * psi;
                        b = vector((u0 * std::tanh(((4.0 * y) - 2.0) / delta)), 0, 0);
                    } else {
                        u = vector(-u0 * std::tanh(((4.0 * y) - 6.0) / delta), 0, 0);
                        u.x() += (cn * mesh.dy() * psi2);
                        u.y() += -cn * mesh.dx() * psi2;
                        b = vector(-u0 * std::tanh(((4.0 * y) - 6.0) / delta), 0, 0);
                    }
                    U(i, j, k) = u;
                    if (j < meshB.nj() - h) B(i, j, k) = b;
                }
        U.correctBoundaryConditions();
        p.correctBoundaryConditions();
        B.correctBoundaryConditions();
        // the force of the frozen B, carried to the walled mesh (its node N_y
        // is the periodic image of node 0)
        auto FB = ex::div(B, B) - ex::grad(0.5 * (B & B));
        Field<vector> F(mesh, vector(0, 0, 0));
        for (int i = h; i < mesh.ni() - h; i++)
            for (int j = h; j < mesh.nj() - h; j++)
                for (int k = h; k < mesh.nk() - h; k++) {
                    const int jb = j < meshB.nj() - h - 1 ? j : h;
                    F(i, j, k) = (*FB)(i, jb, k);
                }
        const scalar RK4[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};  // shearLayer.cpp:145
        while (time.run()) {
            time++;
            std::cout << "Step: " << time.timeStep() << ".              Time: " << time.curTime() << std::endl;
            tools::CFL(U, mesh);
            {  // UEqn.H:1-54
                Field<vector> UOld(U);
                for (int rk = 0; rk < 4; rk++) {
                    auto dU = time.dt() * (-ex::div(U, U) + ex::laplacian(mu, U) + F);
                    Ustar = UOld;
                    Ustar = Ustar + RK4[rk] * dU;
                    Ustar.correctBoundaryConditions();
                    pEqn.rhs(ex::div(Ustar) / time.dt() / settings::RK4c()[rk]);
                    pEqn.solve();
                    p.correctBoundaryConditions();
                    U = Ustar - settings::RK4c()[rk] * time.dt() * ex::grad(p);
                    U.correctBoundaryConditions();
                }
                auto divU = ex::div(U);
                std::cout << "Velocity divergence: " << tools::maximum(divU) << std::endl;
            }
        }
        if (argc > 3) {
            std::ofstream out(argv[3], std::ios::binary);
            int ns[3];  // stored nodes: N + 1 on a walled direction, N on a periodic one
            const int np[3] = {mesh.ni(), mesh.nj(), mesh.nk()};
            for (int a = 0; a < 3; a++) ns[a] = np[a] - 2 * h - (mesh.walled(a) ? 0 : 1);
            for (int c = 0; c < 4; c++)
                for (int i = 0; i < ns[0]; i++)
                    for (int j = 0; j < ns[1]; j++)
                        for (int k = 0; k < ns[2]; k++) {
                            const double v = c < 3 ? U(i + h, j + h, k + h)[c] : p(i + h, j + h, k + h);
                            out.write(reinterpret_cast<const char *>(&v), sizeof v);
                        }
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

// tg_poisson.cpp -- the incompressible Taylor-Green case driven through the
// C++ drop-in API, with the time loop of the reference's TG_Poisson app
// (app/TG_Piosson/TG.cpp:76-182 and RKLoop.H:1-54): four RK4 stages whose
// right-hand sides are all evaluated at the same U, one pressure projection
// per step, and the kinetic-energy history written as "t Ek -dEk/dt".
//
// With an output directory, the mesh, U and p are written after every step
// in EnSight Gold, as app/TG_Piosson/TG.cpp:78-80 does.  At the end it
// prints the wall time per step of the device work (the RK stages, the
// projection and the fills, synchronised) and of the host kinetic-energy
// loop (which reads U through the host mirror, as the reference app does;
// the mirror's download, part of it, is also reported on its own),
// averaged over the steps after the first (which loads the kernels and
// allocates the operators' temporaries).
//
// usage: tg_poisson N steps dt nu [out.dat [output_dir]]
#include <chrono>
#include <cstdio>
#include <iomanip>
#include <iostream>

#include "nova/nova.hpp"

int main(int argc, char *argv[])
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s N steps dt nu [out.dat [output_dir]]\n", argv[0]);
        return 2;
    }
    try {
        const int N = std::atoi(argv[1]);
        const int steps = std::atoi(argv[2]);
        const scalar dt = std::atof(argv[3]);
        const scalar mu = std::atof(argv[4]);
        Time time(dt, steps * dt - 0.5 * dt, 1);
        const scalar L = 2.0 * tools::pi;
        Mesh mesh({N, N, N}, {L, L, L}, time);
        Field<vector> U(mesh, vector(0, 0, 0), "U");
        Field<vector> Ustar(mesh, vector(0, 0, 0), "U");
        auto p_ptr = std::make_shared<Field<scalar>>(mesh, 0.0, "p");
        poisson pEqn(p_ptr);
        const int h = settings::m() / 2;
        // initial condition on the interior nodes, then the BC fill
        for (int i = h; i < mesh.ni() - h; i++)
            for (int j = h; j < mesh.nj() - h; j++)
                for (int k = h; k < mesh.nk() - h; k++) {
                    const scalar x = (i - h) * mesh.dx(), y = (j - h) * mesh.dy(), z = (k - h) * mesh.dz();
                    U(i, j, k) = vector(std::sin(x) * std::cos(y) * std::cos(z),
                                        -std::cos(x) * std::sin(y) * std::cos(z), 0.0);
                }
        U.correctBoundaryConditions();
        std::ofstream out(argc > 5 ? argv[5] : "tg_poisson.dat");
        scalar EkOld = 0.0;
        using clk = std::chrono::steady_clock;
        double t_dev = 0.0, t_ke = 0.0, t_dl = 0.0;
        nova::check(nova_ctx_sync(mesh.ctx()), "nova_ctx_sync");
        while (time.run()) {
            const auto t0 = clk::now();
            time++;
            Field<vector> UOld(U);
            std::array<FieldPtr<vector>, 4> dU;
            for (int rk = 0; rk < 4; rk++) {
                dU[rk] = time.dt() * (-ex::div(U, U) + ex::laplacian(mu, U));
                Ustar = UOld;
                for (int jj = 0; jj <= rk; jj++) Ustar = Ustar + settings::RK4a()[rk][jj] * dU[jj];
            }
            Ustar.correctBoundaryConditions();
            pEqn.rhs(ex::div(Ustar) / time.dt());
            pEqn.solve();
            p_ptr->correctBoundaryConditions();
            U = Ustar - time.dt() * ex::grad(*p_ptr);
            U.correctBoundaryConditions();
            if (argc > 6) {
                mesh.write(argv[6]);
                U.write(argv[6], "U");
                p_ptr->write(argv[6], "p");
            }
            if (time.writelog()) std::cout << "Velocity divergence: " << tools::maximum(ex::div(U)) << std::endl;
            nova::check(nova_ctx_sync(mesh.ctx()), "nova_ctx_sync");
            const auto t1 = clk::now();
            // kinetic energy over the unique nodes, on the host mirror
            // (its first access downloads U)
            U(h, h, h);
            const auto t1b = clk::now();
            scalar Ek = 0.0;
            int n = 0;
            for (int i = h; i < mesh.ni() - h - 1; i++)
                for (int j = h; j < mesh.nj() - h - 1; j++)
                    for (int k = h; k < mesh.nk() - h - 1; k++) {
                        const vector &u = U(i, j, k);
                        Ek += 0.5 * (u.x() * u.x() + u.y() * u.y() + u.z() * u.z());
                        n++;
                    }
            reduce(Ek, plusOp<scalar>());
            Ek /= n;
            out << time.curTime() << " " << std::setprecision(15) << Ek << " " << std::setprecision(15)
                << -(Ek - EkOld) / time.dt() << "\n";
            EkOld = Ek;
            const auto t2 = clk::now();
            if (time.timeStep() > 1) {
                t_dev += std::chrono::duration<double>(t1 - t0).count();
                t_ke += std::chrono::duration<double>(t2 - t1).count();
                t_dl += std::chrono::duration<double>(t1b - t1).count();
            }
        }
        mesh.flushOutput();
        const int ns = time.timeStep() > 1 ? time.timeStep() - 1 : 1;
        std::cout << "done " << time.timeStep() << " steps" << std::endl;
        std::cout << "timing: device_ms_per_step " << 1e3 * t_dev / ns << " host_ke_ms_per_step " << 1e3 * t_ke / ns
                  << " mirror_download_ms_per_step " << 1e3 * t_dl / ns << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "tg_poisson: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}

// poiseuille_driver.cpp -- plane Poiseuille flow written the way the reference's own test is
// (src/Tests/2DTests/PoiseuilleChanel.h:36-120 RunSingleExperiment, tests.h:1617 RunPoiseuille2D),
// against the HipExplicitRungeKuttaFVM host class: a channel periodic along x between two walls,
// the y-axis compressed towards both walls by two geometric blocks, viscous fluxes on and the
// flow driven by the potential force Sigma = (sigma, 0, 0).  The steady centre velocity is
// Uc = sigma Ly^2 / (8 mu) (PoiseuilleChanel.h:44-47).
//
// usage: poiseuille_driver nx ny mu sigma pin t_end q_com init rk out.bin
//   init: 0 = the exact profile, 1 = 1.1 x exact, 2 = 0.9 x exact (InitExact / InitHigher /
//   InitLower, PoiseuilleChanel.h:135-199), set as point values at the cell centres.
// out.bin: int32 nXAll, nYAll, iteration, pad; double time, dt, Uc;
//          CoordinateX[nXAll], hx[nXAll], CoordinateY[nYAll], hy[nYAll], values[nYAll*nXAll*5].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "ts_kernel.hpp"

using namespace tsh;

int main(int argc, char** argv) {
    if (argc != 11) {
        std::fprintf(stderr, "usage: %s nx ny mu sigma pin t_end q_com init rk out.bin\n", argv[0]);
        return 2;
    }
    const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]);
    const double mu = std::atof(argv[3]), sigma = std::atof(argv[4]), pin = std::atof(argv[5]);
    const double t_end = std::atof(argv[6]), q = std::atof(argv[7]);
    const int init = std::atoi(argv[8]), rk = std::atoi(argv[9]);
    const char* out = argv[10];
    const double gamma = 1.4, lx = 1.0, ly = 0.5, ro = 1.0;  // DefaultSettings (PoiseuilleChanel.h:30-40)
    try {
        KernelConfiguration conf;
        conf.nDims = 2;
        conf.nX = nx;
        conf.nY = ny;
        conf.LX = lx;
        conf.LY = ly;
        conf.isPeriodicY = false;
        conf.Gamma = gamma;
        conf.IsViscousFlow = true;
        conf.Viscosity = mu;
        conf.Sigma = Vector(sigma, 0, 0);
        conf.IsExternalForceRequared = true;

        BlockNode y_bot, y_cen;
        y_bot.N_cells = ny / 2;
        y_bot.q_com = q;
        conf.CompressionY[0] = y_bot;
        y_cen.pos = 0.5 * ly;
        y_cen.q_com = 1.0 / y_bot.q_com;
        y_cen.N_cells = ny - y_bot.N_cells;
        conf.CompressionY.push_back(y_cen);

        conf.MyConditions[1] = BoundaryConditionConfiguration(BoundaryConditionType::Wall);
        conf.yLeftBoundary.SetMarker(1);
        conf.yRightBoundary.SetMarker(1);

        conf.methodConfiguration.CFL = 0.45;
        conf.methodConfiguration.RungeKuttaOrder = rk;
        conf.methodConfiguration.Eps = 0.05;
        conf.methodConfiguration.RiemannProblemSolver = RPSolver::RoePikeSolver;
        conf.methodConfiguration.ReconstructionType = Reconstruction::Linear2psLim;
        conf.DummyLayerSize = 1;
        conf.MaxTime = t_end;
        conf.MaxIteration = 10000000;

        const double Uc = 0.125 * sigma * ly * ly / mu;
        const double scale = init == 1 ? 1.1 : (init == 2 ? 0.9 : 1.0);
        auto kernel = CreateKernel(conf);
        kernel->Init(conf);
        kernel->SetInitialConditions([&](const Vector& r) {
            const double y = 1.0 - std::abs(2 * r.y / ly - 1.0);
            const double u = scale * Uc * y * (2.0 - y);
            return std::vector<double>{ro, ro * u, 0.0, 0.0, pin / (gamma - 1.0) + 0.5 * ro * u * u};
        });
        kernel->Run();
        kernel->SyncValues();

        FILE* f = std::fopen(out, "wb");
        if (!f) throw std::runtime_error("cannot open output file");
        int hdr[4] = {kernel->nXAll, kernel->nYAll, kernel->stepInfo.Iteration, 0};
        double st[3] = {kernel->stepInfo.Time, kernel->stepInfo.TimeStep, Uc};
        std::fwrite(hdr, sizeof(int), 4, f);
        std::fwrite(st, sizeof(double), 3, f);
        std::fwrite(kernel->CoordinateX.data(), sizeof(double), kernel->nXAll, f);
        std::fwrite(kernel->hx.data(), sizeof(double), kernel->nXAll, f);
        std::fwrite(kernel->CoordinateY.data(), sizeof(double), kernel->nYAll, f);
        std::fwrite(kernel->hy.data(), sizeof(double), kernel->nYAll, f);
        std::fwrite(&kernel->values[0], sizeof(double), kernel->values.size(), f);
        std::fclose(f);
        // centre velocity: the two middle rows of the middle column
        const int i = kernel->nXAll / 2, j = kernel->nYAll / 2;
        auto vel = [&](int jj) {
            const double* U = &kernel->values[((size_t)jj * kernel->nXAll + i) * 5];
            return U[1] / U[0];
        };
        std::printf("steps %d time %.17g u_centre/Uc %.6f\n", kernel->stepInfo.Iteration, kernel->stepInfo.Time,
                    0.5 * (vel(j - 1) + vel(j)) / Uc);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

// shock_tube_driver.cpp -- 1-D shock-tube runs written the way the reference's own drivers are
// (src/Tests/tests.h:13-80 RunSODTestRoe1D, src/Tests/1DTests/ToroTests.h), against the
// HipExplicitRungeKuttaFVM host class: KernelConfiguration, CreateKernel, Init,
// SetInitialConditions, Run.
//
// usage: shock_tube_driver --reference-sod out_dir out.bin
//   RunSODTestRoe1D exactly as the reference configures it (tests.h:13-80): 100 cells on the
//   two-block compressed grid, Natural boundaries, gamma 1.4, CFL 0.5, forward Euler, Eps 0.05,
//   Linear2psLim<limVenkatar>, Roe-Pike, MaxTime 0.25, SaveSolutionTime 0.25,
//   ResidualOutputIters 100; init.dat, sol.sol (saved and loaded back), init2.dat and the
//   snapshot "solution, t = 0.25.dat" are written into out_dir.
//
// usage: shock_tube_driver rhoL uL pL rhoR uR pR x0 t_end nx recon riemann rk compress
//                          save_time residual_norms out.bin [save.sol load.sol [probe_dir]]
//   a parametrised shock tube with CFL 0.4, Eps 0, the Barth-Jespersen limiter for recon 3 and
//   no snapshot files (the parity tests' cases)
//   recon / riemann: the reference's enum values; rk: 1 or 3; compress: 1 = the two-block
//   geometric grid of RunSODTestRoe1D (q = 1/1.05 then 1.05, node at 0.5 L), 0 = uniform;
//   save_time > 0 clamps dt onto its multiples (SaveSolutionTime); save.sol / load.sol ("-" for
//   none): Kernel::SaveSolution after the run / Kernel::LoadSolution before it (the restart path);
//   probe_dir: history.dat every 5 iterations and, every 7, a density sensor at the middle cell,
//   a (rho, u, p) sensor at the first cell and the position of the pressure maximum (first and
//   second order), written there.
// out.bin: int32 nXAll, int32 snapshots, int32 iteration, pad; double time, dt, residual[5];
//          CoordinateX[nXAll], hx[nXAll], values[nXAll*5].
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <vector>

#include "ts_kernel.hpp"

using namespace tsh;

static void write_out(const HipExplicitRungeKuttaFVM& k, int snapshots, const char* out) {
    FILE* f = std::fopen(out, "wb");
    if (!f) throw std::runtime_error("cannot open output file");
    int hdr[4] = {k.nXAll, snapshots, k.stepInfo.Iteration, 0};
    double st[7] = {k.stepInfo.Time, k.stepInfo.TimeStep};
    for (int v = 0; v < 5; v++) st[2 + v] = k.stepInfo.Residual[v];
    std::fwrite(hdr, sizeof(int), 4, f);
    std::fwrite(st, sizeof(double), 7, f);
    std::fwrite(k.CoordinateX.data(), sizeof(double), k.nXAll, f);
    std::fwrite(k.hx.data(), sizeof(double), k.nXAll, f);
    std::fwrite(&k.values[0], sizeof(double), k.values.size(), f);
    std::fclose(f);
}

// RunSODTestRoe1D (tests.h:13-80) as the reference configures and runs it
static int reference_sod(const std::string& dir, const char* out) {
    KernelConfiguration conf;
    conf.nDims = 1;
    conf.LX = 1.0;
    conf.nX = 100;
    conf.isPeriodicX = false;
    conf.DummyLayerSize = 1;
    BlockNode nleft, ncenter;  // grid compression (tests.h:25-32)
    nleft.N_cells = conf.nX / 2;
    nleft.q_com = 1.0 / 1.05;
    ncenter.pos = 0.5 * conf.LX;
    ncenter.N_cells = conf.nX - nleft.N_cells;
    ncenter.q_com = 1.0 / nleft.q_com;
    conf.CompressionX[0] = nleft;
    conf.CompressionX.push_back(ncenter);
    conf.MyConditions[1] = BoundaryConditionConfiguration(BoundaryConditionType::Natural);
    conf.xLeftBoundary.SetMarker(1);
    conf.xRightBoundary.SetMarker(1);
    conf.Gamma = 1.4;
    conf.IsViscousFlow = false;
    conf.methodConfiguration.CFL = 0.5;
    conf.methodConfiguration.RungeKuttaOrder = 1;
    conf.methodConfiguration.Eps = 0.05;
    conf.methodConfiguration.ReconstructionType = Reconstruction::Linear2psLim;
    conf.methodConfiguration.GeneralLimitter = LimiterType::Venkatakrishnan;  // Linear2psLim<limVenkatar>
    conf.methodConfiguration.RiemannProblemSolver = RPSolver::RoePikeSolver;
    conf.MaxTime = 0.25;
    conf.MaxIteration = 1000000;
    conf.SaveSolutionTime = 0.25;
    conf.SaveSolutionIters = 0;
    conf.ResidualOutputIters = 100;
    auto kernel = CreateKernel(conf);
    kernel->Init(conf);
    kernel->OutputDirectory = dir;
    const double roL = 1.0, PL = 1.0, uL = 0, roR = 0.125, PR = 0.1, uR = 0, x0 = 0.5;
    kernel->SetInitialConditions([&](const Vector& r) {
        const bool left = r.x < x0;
        const double ro = left ? roL : roR, u = left ? uL : uR, p = left ? PL : PR;
        return std::vector<double>{ro, ro * u, 0.0, 0.0, p / (conf.Gamma - 1.0) + 0.5 * ro * u * u};
    });
    kernel->SaveSolutionToTecplot(dir + "/init.dat");
    kernel->SaveSolution(dir + "/sol.sol");
    kernel->LoadSolution(dir + "/sol.sol");
    kernel->SaveSolutionToTecplot(dir + "/init2.dat");
    int snapshots = 0;
    kernel->OnSolutionSnapshot = [&](HipExplicitRungeKuttaFVM&) { snapshots++; };
    kernel->Run();
    kernel->SyncValues();
    write_out(*kernel, snapshots, out);
    std::printf("steps %d time %.17g snapshots %d\n", kernel->stepInfo.Iteration, kernel->stepInfo.Time, snapshots);
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && std::string(argv[1]) == "--reference-sod") {
        try {
            return reference_sod(argv[2], argv[3]);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "error: %s\n", e.what());
            return 1;
        }
    }
    if (argc != 17 && argc != 19 && argc != 20) {
        std::fprintf(stderr,
                     "usage: %s rhoL uL pL rhoR uR pR x0 t_end nx recon riemann rk compress save_time norms out.bin"
                     " [save.sol load.sol [probe_dir]]\n",
                     argv[0]);
        return 2;
    }
    const std::string save_sol = argc >= 19 ? argv[17] : "-", load_sol = argc >= 19 ? argv[18] : "-";
    const std::string probe_dir = argc == 20 ? argv[19] : "-";
    const double roL = std::atof(argv[1]), uL = std::atof(argv[2]), PL = std::atof(argv[3]);
    const double roR = std::atof(argv[4]), uR = std::atof(argv[5]), PR = std::atof(argv[6]);
    const double x0 = std::atof(argv[7]), t_end = std::atof(argv[8]);
    const int nx = std::atoi(argv[9]), recon = std::atoi(argv[10]), riemann = std::atoi(argv[11]);
    const int rk = std::atoi(argv[12]), compress = std::atoi(argv[13]);
    const double save_time = std::atof(argv[14]);
    const int norms = std::atoi(argv[15]);
    const char* out = argv[16];
    try {
        KernelConfiguration conf;
        conf.nDims = 1;
        conf.LX = 1.0;
        conf.nX = nx;
        conf.isPeriodicX = false;
        conf.DummyLayerSize = 1;
        if (compress) {
            BlockNode nleft, ncenter;
            nleft.N_cells = conf.nX / 2;
            nleft.q_com = 1.0 / 1.05;
            ncenter.pos = 0.5 * conf.LX;
            ncenter.N_cells = conf.nX - nleft.N_cells;
            ncenter.q_com = 1.0 / nleft.q_com;
            conf.CompressionX[0] = nleft;
            conf.CompressionX.push_back(ncenter);
        }
        conf.MyConditions[1] = BoundaryConditionConfiguration(BoundaryConditionType::Natural);
        conf.xLeftBoundary.SetMarker(1);
        conf.xRightBoundary.SetMarker(1);
        conf.Gamma = 1.4;
        conf.methodConfiguration.CFL = 0.4;
        conf.methodConfiguration.RungeKuttaOrder = rk;
        conf.methodConfiguration.Eps = 0.0;
        conf.methodConfiguration.ReconstructionType = static_cast<Reconstruction>(recon);
        conf.methodConfiguration.GeneralLimitter = LimiterType::BarsJespersen;
        conf.methodConfiguration.RiemannProblemSolver = static_cast<RPSolver>(riemann);
        conf.MaxTime = t_end;
        conf.MaxIteration = 1000000;
        conf.SaveSolutionTime = save_time;

        auto kernel = CreateKernel(conf);
        kernel->Init(conf);
        kernel->ComputeResidualNorms = norms != 0;
        kernel->WriteSnapshotFiles = false;
        const double g1 = conf.Gamma - 1.0;
        kernel->SetInitialConditions([&](const Vector& r) {
            const bool left = r.x < x0;  // tests.h:85
            const double ro = left ? roL : roR, u = left ? uL : uR, p = left ? PL : PR;
            return std::vector<double>{ro, ro * u, 0.0, 0.0, p / g1 + 0.5 * ro * u * u};
        });
        if (load_sol != "-") kernel->LoadSolution(load_sol);
        if (probe_dir != "-") {
            kernel->SaveHistoryFileIters = 5;
            kernel->OutputDirectory = probe_dir;
            kernel->isSensorEnable = true;
            kernel->SaveSensorRecordIters = 7;
            auto rho = [](const std::valarray<double>& U) { return U[0]; };
            auto vel = [](const std::valarray<double>& U) { return U[1] / U[0]; };
            auto prs = [g1](const std::valarray<double>& U) { return g1 * (U[4] - 0.5 * U[1] * U[1] / U[0]); };
            auto c1 = std::make_unique<CellSensor>(probe_dir + "/cell.dat", rho);
            c1->SetSensor(*kernel, nx / 2 + 1, 0, 0);
            auto c2 = std::make_unique<CellSensor>(probe_dir + "/multi.dat", std::vector<sen_func>{rho, vel, prs});
            c2->SetSensor(*kernel, 1, 0, 0);
            auto m1 = std::make_unique<MValuePosXSensor>(probe_dir + "/pmax1.dat", prs, false);
            m1->SetSensor(*kernel, 0, 0);
            auto m2 = std::make_unique<MValuePosXSensor>(probe_dir + "/pmax2.dat", prs, true);
            m2->SetSensor(*kernel, 0, 0);
            kernel->Sensors.push_back(std::move(c1));
            kernel->Sensors.push_back(std::move(c2));
            kernel->Sensors.push_back(std::move(m1));
            kernel->Sensors.push_back(std::move(m2));
        }
        int snapshots = 0;
        kernel->OnSolutionSnapshot = [&](HipExplicitRungeKuttaFVM&) { snapshots++; };
        kernel->Run();
        kernel->SyncValues();
        if (save_sol != "-") kernel->SaveSolution(save_sol);

        write_out(*kernel, snapshots, out);
        std::printf("steps %d time %.17g snapshots %d\n", kernel->stepInfo.Iteration, kernel->stepInfo.Time, snapshots);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

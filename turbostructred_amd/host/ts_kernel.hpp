// ts_kernel.hpp -- C++ host class over libts_fvm.so with the shape of TurboStructured's
// Kernel / ExplicitRungeKuttaFVM<R> (src/kernel/kernel.h:42-1850, src/Methods/ExplicitRungeKuttaFVM.h).
//
// A host program written against the reference builds a KernelConfiguration, calls
// CreateKernel, Init, SetInitialConditions and Run, and reads `values` / `stepInfo`.  The same
// program runs on an MI355X by switching its includes to this header: configuration fields,
// enum members, the step information and the termination rule of Run keep the reference's
// names and meaning.  The step itself (ghost fill, reconstruction, Riemann fluxes, time step,
// update) runs in libts_fvm.so; the state lives in HBM and `values` is a host mirror refreshed
// on demand (SyncValues) and before every snapshot hook.
//
// Errors: the reference reports failures with exceptions and asserts; here every failing
// library call throws std::runtime_error carrying the library's message.
#pragma once

#include <cmath>
#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <functional>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <valarray>
#include <vector>

#include "../../include/ts_fvm.h"

namespace tsh {

// utility/Vector.h
struct Vector {
    double x = 0, y = 0, z = 0;
    Vector() = default;
    Vector(double a, double b, double c) : x(a), y(b), z(c) {}
};

// src/Methods/MethodConfiguration.h:5-23; HLLCSolver is an extension (no reference member)
enum class Reconstruction { PiecewiseConstant, Linear2PointsStencil, ENO2PointsStencil, Linear2psLim };
enum class RPSolver { GodunovSolver, RoePikeSolver, NoSolver, HLLCSolver };
enum class LimiterType { BarsJespersen, Venkatakrishnan };

// src/Methods/MethodConfiguration.h:26-37 (+ arithmetic of the residual, an extension)
struct MethodConfiguration {
    RPSolver RiemannProblemSolver{RPSolver::NoSolver};
    Reconstruction ReconstructionType{Reconstruction::PiecewiseConstant};
    LimiterType GeneralLimitter{LimiterType::BarsJespersen};
    int RungeKuttaOrder{0};  // 0 or 1 = forward Euler (the reference's step), 3 = SSP-RK3
    double CFL{0.25};
    double Eps{0.0};
    int Precision{TS_FP64};  // TS_FP64 or TS_FP32_RESIDUAL
};

// src/BoundaryConditions/BoundaryConditionConfiguration.h:12-69
enum class BoundaryConditionType { Wall, Symmetry, MovingWall, SubsonicInlet, SubsonicOutlet, Natural };

struct BoundaryConditionConfiguration {
    BoundaryConditionType BCType{BoundaryConditionType::Natural};
    Vector Velocity;
    double Pstatic{0};
    double Density{0};
    BoundaryConditionConfiguration() = default;
    explicit BoundaryConditionConfiguration(BoundaryConditionType t) : BCType(t) {}
};

class DomainBCinfo {
   public:
    bool isComplex{false};
    std::function<int(const Vector r)> getMarker;
    void SetMarker(int m) { bcMarker = m; }
    int GetMarker() const { return bcMarker; }

   private:
    int bcMarker{0};
};

// src/kernel/KernelConfiguration.h:10-18
struct BlockNode {
    double pos{0.0};
    double q_com{1.0};
    int N_cells{-1};
    BlockNode() = default;
    explicit BlockNode(double p) : pos(p) {}
};

// src/kernel/KernelConfiguration.h:21-93: the fields of the explicit path
struct KernelConfiguration {
    int nDims{1};
    int nX{1}, nY{1}, nZ{1};
    bool isPeriodicX{true}, isPeriodicY{true}, isPeriodicZ{true};
    double LX{1.0}, LY{1.0}, LZ{1.0};
    std::vector<BlockNode> CompressionX{std::vector<BlockNode>(1)};
    std::vector<BlockNode> CompressionY{std::vector<BlockNode>(1)};
    std::vector<BlockNode> CompressionZ{std::vector<BlockNode>(1)};
    double Gamma{1.4};
    double Viscosity{0.0};
    bool IsViscousFlow{false};
    bool IsExternalForceRequared{false};
    bool IsUnifromAccelerationRequared{false};
    Vector Sigma{Vector(0, 0, 0)};                // potential force (pressure gradient)
    Vector UniformAcceleration{Vector(0, 0, 0)};  // uniform gravity field
    int DummyLayerSize{1};
    MethodConfiguration methodConfiguration;
    int MaxIteration{1000};
    double MaxTime{1.0};
    double SaveSolutionTime{0};
    double SaveSliceTime{0};
    int SaveSolutionIters{0};
    int SaveSliceIters{0};
    int SaveBinarySolIters{0};
    double MolarMass{0.0};                 // temperature column of the Tecplot writers
    double UniversalGasConstant{8.3144598};
    int ResidualOutputIters{0};
    int ChangeHistoryFileIters{5000000};
    // history.dat every SaveHistoryFileIters iterations; the reference's default is 10, here it
    // is 0 (no file) so that a library run writes nothing unless asked
    int SaveHistoryFileIters{0};
    std::map<int, BoundaryConditionConfiguration> MyConditions;
    DomainBCinfo xLeftBoundary, xRightBoundary, yLeftBoundary, yRightBoundary, zLeftBoundary, zRightBoundary;
    int Device{0};  // HIP device ordinal (extension)
};

// src/kernel/kernel.h:31-39
struct StepInfo {
    double Time{0};
    double TimeStep{0};
    int Iteration{0};
    std::vector<double> Residual{std::vector<double>(5, 0.0)};
    double NextSolutionSnapshotTime{0};
    double NextSliceSnapshotTime{0};
};

// Cell-centre coordinates and sizes along one axis, ghost layers included, from the block
// compression description: the construction of Grid::InitLocal (src/kernel/grid.h:164-218),
// positions advanced as running sums from the left ghost layer.
inline void BuildAxis(std::vector<BlockNode> blocks, int n, double L, int ghost, std::vector<double>& X,
                      std::vector<double>& H) {
    X.assign(n + 2 * ghost, 0.0);
    H.assign(n + 2 * ghost, 0.0);
    blocks.push_back(BlockNode(L));
    double stp;
    if (blocks[0].N_cells < 1) {
        stp = L / n;
        blocks[0].N_cells = n;
    } else {
        const double len = blocks[1].pos - blocks[0].pos;
        const int nc = blocks[0].N_cells;
        stp = blocks[0].q_com == 1.0 ? len / nc : len * (1.0 - blocks[0].q_com) / (1.0 - std::pow(blocks[0].q_com, nc));
    }
    double xc = -(ghost * stp) + 0.5 * stp;
    int c = 0;
    for (; c < ghost; c++) {
        X[c] = xc;
        H[c] = stp;
        xc += stp;
    }
    for (size_t b = 1; b < blocks.size(); b++) {
        const BlockNode& nd = blocks[b - 1];
        for (int i = 0; i < nd.N_cells; i++) {
            X[c + i] = xc;
            H[c + i] = stp;
            xc += 0.5 * stp * (1.0 + nd.q_com);
            stp *= nd.q_com;
        }
        c += nd.N_cells;
        if (blocks[b].N_cells > 0) {
            const double len = blocks[b + 1].pos - blocks[b].pos;
            const int nc = blocks[b].N_cells;
            const double q = blocks[b].q_com;
            stp = q == 1.0 ? len / nc : len * (1.0 - q) / (1.0 - std::pow(q, nc));
            xc = blocks[b].pos + 0.5 * stp;
        }
    }
    if (c != n + ghost) throw std::runtime_error("grid blocks do not add up to the cell count");
    for (int i = 0; i < ghost; i++) {
        X[c + i] = xc;
        H[c + i] = stp;
        xc += stp;
    }
}

// ---- equation of state (src/EquationsOfState/generalEOS.h:6-23, idealGasEOS.h:7-46) ------------
// The host side of the EquationOfState interface.  The device solvers take the same ideal-gas
// law as a functor (tsd::IdealGasEOS in the library); both are the perfect gas p = (g - 1) rho e
// of the reference's Roe and Godunov solvers.
class GeneralEOS {
   public:
    virtual ~GeneralEOS() = default;
    virtual double GetPressure(double ro, double e) = 0;
    virtual double GetSoundSpeed(double ro, double e) = 0;
    virtual double GetInternalEnergy(double ro, double pressure) = 0;
    virtual double GetPressureEnergyDerivative(double ro, double e) = 0;
    virtual double GetPressureDensityDerivative(double ro, double e) = 0;
};

class IdealGasEOS : public GeneralEOS {
   public:
    double gamma;
    explicit IdealGasEOS(double g) : gamma(g) {}
    void SetGamma(double g) { gamma = g; }
    double GetPressure(double ro, double e) override { return (gamma - 1.0) * ro * e; }
    double GetSoundSpeed(double ro, double e) override { return std::sqrt(gamma * ((gamma - 1.0) * ro * e) / ro); }
    double GetInternalEnergy(double ro, double pressure) override { return pressure / ((gamma - 1.0) * ro); }
    double GetPressureEnergyDerivative(double ro, double) override { return (gamma - 1.0) * ro; }
    double GetPressureDensityDerivative(double, double e) override { return (gamma - 1.0) * e; }
};

// ---- cell averages by Gauss quadrature (src/utility/NumericQuadrature.h:18-226) --------------------
// Gauss-Legendre nodes and weights are computed here (Newton's method on P_n in long double, then
// rounded), not tabulated; the orders the reference tabulates are accepted (1-12 and 32).  Two
// behaviours of the reference are kept because they decide the initial state it produces:
//  * order 10 takes the 12-point half table through the 10-point mirror fill (the missing
//    `break` of NumericQuadrature.h:183-191): its weights sum to 2 - 2 w_1 of the 12-point rule;
//  * IntegrateGroap samples point (x_i, y_i, z_i) for the weight wx_i wy_j wz_k (the y and z
//    coordinates take the x index, NumericQuadrature.h:97-99), so in 2-D and 3-D it integrates
//    along the cell's diagonal.  Below 3-D the reference reads y/z nodes past their one-element
//    arrays there; here those coordinates stay at the cell centre.
class NumericQuadrature {
   public:
    NumericQuadrature(int n, int dims) : n_(n), dims_(dims) {
        const bool ok = (n >= 1 && n <= 12) || n == 32;
        if (!ok) throw std::invalid_argument("Gauss quadrature order must be 1-12 or 32: " + std::to_string(n));
        Fill();
    }
    int Order() const { return n_; }
    const std::vector<double>& Nodes() const { return x_; }
    const std::vector<double>& Weights() const { return wx_; }

    // integral of f (nvars values) over the cell, IntegrateGroap (NumericQuadrature.h:86-118)
    template <typename F>
    std::valarray<double> IntegrateGroap(double cx, double cy, double cz, double hx, double hy, double hz, F&& f,
                                         size_t nvars) const {
        std::valarray<double> res(0.0, nvars);
        const int ny = dims_ > 1 ? n_ : 1, nz = dims_ > 2 ? n_ : 1;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < n_; i++) {
                    const double rx = cx + x_[i] * 0.5 * hx;
                    const double ry = cy + (dims_ > 1 ? x_[i] : 0.0) * 0.5 * hy;
                    const double rz = cz + (dims_ > 2 ? x_[i] : 0.0) * 0.5 * hz;
                    const std::vector<double> vars = f(Vector(rx, ry, rz));
                    const double w = wx_[i] * (dims_ > 1 ? wx_[j] : 1.0) * (dims_ > 2 ? wx_[k] : 1.0);
                    for (size_t v = 0; v < nvars; v++) res[v] += w * vars[v];
                }
        double J = 0.5 * hx;  // ComputeJacobian (NumericQuadrature.h:31-37)
        if (dims_ > 1) J *= 0.5 * hy;
        if (dims_ > 2) J *= 0.5 * hz;
        res *= J;
        return res;
    }

   private:
    int n_, dims_;
    std::vector<double> x_, wx_;

    // nonnegative roots of P_m in ascending order and their weights
    static void HalfTable(int m, std::vector<double>& r, std::vector<double>& w) {
        std::vector<std::pair<long double, long double>> rw;
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int i = 1; i <= m; i++) {
            long double x = std::cos(pi * (i - 0.25L) / (m + 0.5L)), dp = 0;
            for (int it = 0; it < 100; it++) {
                long double p0 = 1, p1 = x;
                for (int k = 2; k <= m; k++) {
                    const long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                if (m == 1) p0 = 1;
                dp = m * (x * p1 - p0) / (x * x - 1);
                const long double dx = p1 / dp;
                x -= dx;
                if (std::fabs(dx) < 1e-19L) break;
            }
            if (x >= -1e-30L) rw.push_back({std::fabs(x) < 1e-30L ? 0.0L : x, 2 / ((1 - x * x) * dp * dp)});
        }
        std::sort(rw.begin(), rw.end());
        r.clear();
        w.clear();
        for (auto& q : rw) {
            r.push_back((double)q.first);
            w.push_back((double)q.second);
        }
    }
    // FillCoefficients (NumericQuadrature.h:121-226): the mirror fill from the half table
    void Fill() {
        std::vector<double> r, w;
        HalfTable(n_ == 10 ? 12 : n_, r, w);
        x_.assign(n_, 0.0);
        wx_.assign(n_, 0.0);
        const int ws = (int)w.size();
        for (int i = 0; i < ws; i++) {
            wx_[i] = w[ws - 1 - i];
            wx_[n_ - 1 - i] = wx_[i];
            x_[n_ - 1 - i] = r[ws - 1 - i];
            x_[i] = -x_[n_ - 1 - i];
        }
    }
};

// ---- Tecplot and slice writers (kernel.h:1081-1632) ---------------------------------------------
// The grid (ghost layer included) and state as the writers read them: the reference's arrays.
struct TecplotView {
    int nDims{1}, nX{1}, nY{1}, nZ{1}, nXAll{1}, nYAll{1};
    const double *X{nullptr}, *Y{nullptr}, *Z{nullptr}, *hx{nullptr}, *hy{nullptr}, *hz{nullptr};
    const double* values{nullptr};  // AoS, 5 per cell, serial index (k nYAll + j) nXAll + i
    double gamma{1.4}, molarMass{0.0}, universalGasConstant{8.3144598}, time{0.0};
    const double* U(int i, int j, int k) const { return values + (((size_t)k * nYAll + j) * nXAll + i) * 5; }
    double cx(int i) const { return X ? X[i] : 0.0; }
    double cy(int j) const { return nDims > 1 ? Y[j] : 0.0; }
    double cz(int k) const { return nDims > 2 ? Z[k] : 0.0; }
};

// utility/Slices.h: a line or plane of cells; -1 marks a free index
struct Slice {
    int nDims{0}, i{-1}, j{-1}, k{-1};
    Slice() = default;
    Slice(int _i, int _j, int _k) : i(_i), j(_j), k(_k) { nDims = (i == -1) + (j == -1) + (k == -1); }
};

// Kernel::SaveSolutionToTecplot (kernel.h:1081-1416) as one rank writes it
inline void WriteTecplotSolution(const std::string& fname, const TecplotView& v) {
    std::ofstream ofs(fname);
    if (!ofs) throw std::runtime_error("cannot open " + fname);
    ofs << std::scientific;
    const int gx = 1, gy = v.nDims > 1 ? 1 : 0, gz = v.nDims > 2 ? 1 : 0;
    const double g1 = v.gamma - 1.0;
    if (v.nDims == 1) {
        ofs << "VARIABLES = \"X\" \"ro\" \"u\" \"v\" \"w\" \"P\" \"e\" " << std::endl;
        for (int i = gx; i < gx + v.nX; i++) {
            const double* U = v.U(i, 0, 0);
            const double ro = U[0], u = U[1] / ro, vv = U[2] / ro, w = U[3] / ro;
            const double e = U[4] / ro - 0.5 * (u * u + vv * vv + w * w);
            const double P = g1 * ro * e;
            ofs << v.cx(i) << " " << ro << " " << u << " " << vv << " " << w << " " << P << " " << e << " " << std::endl;
        }
        return;
    }
    const int nd = v.nDims;
    const long long nodes = (long long)(v.nX + 1) * (v.nY + 1) * (nd > 2 ? v.nZ + 1 : 1);
    const long long cells = (long long)v.nX * v.nY * (nd > 2 ? v.nZ : 1);
    ofs << (nd == 2 ? "VARIABLES = \"X\" \"Y\" " : "VARIABLES = \"X\" \"Y\" \"Z\" ");
    ofs << "\"Rho\" \"u\" \"v\" \"w\" \"P\" \"e\"\"T\"" << std::endl;
    ofs << "ZONE T=\"" << v.time << "\"" << std::endl;
    ofs << "N=" << nodes << ", E=" << cells << ", F=FEBLOCK, ET=" << (nd == 2 ? "QUADRILATERAL" : "BRICK") << std::endl;
    ofs << (nd == 2 ? "VARLOCATION = (NODAL, NODAL" : "VARLOCATION = (NODAL, NODAL, NODAL");
    for (int q = 0; q < 7; q++) ofs << (q == 6 ? ", CELLCENTERED)" : ", CELLCENTERED");
    ofs << std::endl;
    const int k0 = gz, k1 = gz + (nd > 2 ? v.nZ : 0);
    for (int axis = 0; axis < nd; axis++)
        for (int k = k0; k <= k1; k++)
            for (int j = gy; j <= gy + v.nY; j++)
                for (int i = gx; i <= gx + v.nX; i++) {
                    if (axis == 0) ofs << v.X[i] - 0.5 * v.hx[i];
                    else if (axis == 1) ofs << v.Y[j] - 0.5 * v.hy[j];
                    else ofs << v.Z[k] - 0.5 * v.hz[k];
                    ofs << std::endl;
                }
    // cell-centred blocks: rho, u, v, w, P, e, T, each over k, j, i (i fastest)
    const size_t nc = (size_t)cells;
    std::vector<double> q[7];
    for (auto& a : q) a.resize(nc);
    const double cond = g1 * v.molarMass / v.universalGasConstant;
    size_t c = 0;
    for (int k = gz; k < gz + (nd > 2 ? v.nZ : 1); k++)
        for (int j = gy; j < gy + v.nY; j++)
            for (int i = gx; i < gx + v.nX; i++, c++) {
                const double* U = v.U(i, j, k);
                const double rho = U[0];
                double u, vv, w, e, P;
                if (nd == 2) {  // kernel.h:1205-1225
                    u = U[1] / rho, vv = U[2] / rho, w = U[3] / rho;
                    e = U[4] / rho;
                    e = e - 0.5 * (u * u + vv * vv + w * w);
                    P = g1 * (rho * e);
                } else {  // kernel.h:1339-1360
                    const double ke = 0.5 * (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / rho;
                    u = U[1] / rho, vv = U[2] / rho, w = U[3] / rho;
                    e = U[4] - ke;
                    P = g1 * e;
                    e /= rho;
                }
                q[0][c] = rho, q[1][c] = u, q[2][c] = vv, q[3][c] = w, q[4][c] = P, q[5][c] = e, q[6][c] = cond * e;
            }
    for (auto& a : q)
        for (double x : a) ofs << x << '\n';
    // connectivity, Tecplot node numbers from 1
    for (int k = gz; k < gz + (nd > 2 ? v.nZ : 1); k++)
        for (int j = gy; j < gy + v.nY; j++) {
            if (nd == 2) {
                int bl = (v.nX + 1) * (j - gy) + 1;
                for (int i = 0; i < v.nX; i++, bl++) {
                    const int br = bl + 1, ul = bl + (v.nX + 1), ur = ul + 1;
                    ofs << bl << ' ' << ul << ' ' << ur << ' ' << br << std::endl;
                }
            } else {
                int fbl = (v.nX + 1) * (v.nY + 1) * (k - gz) + (v.nX + 1) * (j - gy) + 1;
                for (int i = 0; i < v.nX; i++, fbl++) {
                    const int fbr = fbl + 1, ful = fbl + (v.nX + 1), fur = ful + 1;
                    const int bbl = fbl + (v.nX + 1) * (v.nY + 1), bbr = bbl + 1, bul = bbl + (v.nX + 1), bur = bul + 1;
                    ofs << bbr << ' ' << bur << ' ' << bul << ' ' << bbl << ' ' << fbr << ' ' << fur << ' ' << ful << ' '
                        << fbl << std::endl;
                }
            }
        }
}

// Kernel::Save2DSliceToTecplot (kernel.h:1416-1511, with_temperature = false) and
// Kernel::SaveSliceToTecplot (kernel.h:1513-1632, with_temperature = true): the cells with i == I
// (j == J, k == K) unless that index is -1, visited i, then j, then k (k fastest)
inline void WriteTecplotSlice(const std::string& fname, const TecplotView& v, int I, int J, int K,
                              bool with_temperature) {
    std::ofstream ofs(fname);
    if (!ofs) throw std::runtime_error("cannot open " + fname);
    ofs << std::scientific;
    ofs << "VARIABLES = ";
    if (I == -1) ofs << "\"X\" ";
    if (J == -1) ofs << "\"Y\" ";
    if (K == -1) ofs << "\"Z\" ";
    ofs << "\"ro\" \"u\" \"v\" \"w\" \"P\" \"e\" ";
    if (with_temperature) {
        ofs << "\"T\" ";
        if (I != -1) ofs << "\"x_cnst\" ";
        if (J != -1) ofs << "\"y_cnst\" ";
        if (K != -1) ofs << "\"z_cnst\" ";
    }
    ofs << std::endl;
    ofs << "ZONE T=\"1\"" << std::endl;
    ofs << "I=" << (I == -1 ? v.nX : 1) << "J=" << (J == -1 ? v.nY : 1) << "K=" << (K == -1 ? v.nZ : 1);
    ofs << "F=POINT\n";
    ofs << (with_temperature ? "DT=(SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE)\n"
                             : "DT=(SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE SINGLE)\n");
    const int gy = v.nDims > 1 ? 1 : 0, gz = v.nDims > 2 ? 1 : 0;
    const double g1 = v.gamma - 1.0;
    for (int i = 1; i <= v.nX; i++) {
        if (I != -1 && i != I) continue;
        for (int j = gy; j < gy + (v.nDims > 1 ? v.nY : 1); j++) {
            if (J != -1 && j != J) continue;
            for (int k = gz; k < gz + (v.nDims > 2 ? v.nZ : 1); k++) {
                if (K != -1 && k != K) continue;
                const double x = v.cx(i), y = v.cy(j), z = v.cz(k);
                const double* U = v.U(i, j, k);
                const double ro = U[0], u = U[1] / ro, vv = U[2] / ro, w = U[3] / ro;
                const double e = U[4] / ro - 0.5 * (u * u + vv * vv + w * w);
                const double P = g1 * ro * e;
                if (I == -1) ofs << x << " ";
                if (J == -1) ofs << y << " ";
                if (K == -1) ofs << z << " ";
                ofs << ro << " " << u << " " << vv << " " << w << " " << P << " " << e << " ";
                if (with_temperature) {
                    const double T = g1 * v.molarMass * e / v.universalGasConstant;
                    ofs << T << " ";
                    if (I != -1) ofs << x << " ";
                    if (J != -1) ofs << y << " ";
                    if (K != -1) ofs << z << " ";
                }
                ofs << std::endl;
            }
        }
    }
}

class HipExplicitRungeKuttaFVM;

// src/Sensors/Sensor.h: a probe processed every SaveSensorRecordIters iterations.  Here a sensor
// reads only the cells it needs from the device (ts_get_cells) instead of the whole state.
class Sensor {
   public:
    int iteration{0};
    double timer{0};
    std::string filename;
    explicit Sensor(std::string fname) : filename(std::move(fname)) {}
    virtual ~Sensor() = default;
    virtual void Process(HipExplicitRungeKuttaFVM& kernel) = 0;
    void UpdateIteration(int it) { iteration = it; }
    void UpdateTimer(double t) { timer = t; }
};
using sen_func = std::function<double(const std::valarray<double>&)>;

// The explicit finite-volume method on one GPU: Kernel + ExplicitRungeKuttaFVM<R> with the
// reconstruction chosen at run time (the reference picks it as a template argument in
// CreateKernel, ExplicitRungeKuttaFVM.h:1495-1515).
class HipExplicitRungeKuttaFVM {
   public:
    int nDims{1};
    int nVariables{5};
    int nX{1}, nY{1}, nZ{1};
    int nXAll{1}, nYAll{1}, nZAll{1};
    std::vector<double> CoordinateX, CoordinateY, CoordinateZ, hx, hy, hz;
    StepInfo stepInfo;
    std::valarray<double> values;  // host mirror of the state (reference AoS layout)
    int MaxIteration{1000};
    double MaxTime{1.0};
    double SaveSolutionTime{0};
    double SaveSliceTime{0};
    int SaveSolutionIters{0};
    int SaveSliceIters{0};
    int SaveBinarySolIters{0};
    double Gamma{1.4}, MolarMass{0.0}, UniversalGasConstant{8.3144598};
    std::vector<Slice> slices;  // written by Run at the slice snapshot times / iterations
    // Run writes the reference's Tecplot snapshot files ("solution, t = ....dat", "slice<i>,T=....dat",
    // kernel.h:1723-1797) into OutputDirectory; the snapshot hooks below are called as well
    bool WriteSnapshotFiles{true};
    bool ComputeResidualNorms{false};
    int ResidualOutputIters{0};
    int SaveHistoryFileIters{0};
    int ChangeHistoryFileIters{5000000};
    int NextHistoryFileIters{5000000};
    std::string OutputDirectory{"."};  // where history files go
    bool isSensorEnable{false};
    int SaveSensorRecordIters{1};
    std::vector<std::unique_ptr<Sensor>> Sensors;
    // called at every solution / slice snapshot time with `values` synchronised
    std::function<void(HipExplicitRungeKuttaFVM&)> OnSolutionSnapshot, OnSliceSnapshot;

    HipExplicitRungeKuttaFVM() = default;
    HipExplicitRungeKuttaFVM(const HipExplicitRungeKuttaFVM&) = delete;
    HipExplicitRungeKuttaFVM& operator=(const HipExplicitRungeKuttaFVM&) = delete;
    ~HipExplicitRungeKuttaFVM() {
        if (ctx_) ts_destroy(ctx_);
    }

    // Kernel::Init (kernel.h:178-281) + InitializeMethod: grid, boundary conditions, method,
    // run parameters; creates the device context.
    void Init(KernelConfiguration& config) {
        nDims = config.nDims;
        nX = config.nX;
        nY = nDims > 1 ? config.nY : 1;
        nZ = nDims > 2 ? config.nZ : 1;
        const int g = config.DummyLayerSize;
        nXAll = nX + 2 * g;
        nYAll = nDims > 1 ? nY + 2 * g : 1;
        nZAll = nDims > 2 ? nZ + 2 * g : 1;
        BuildAxis(config.CompressionX, nX, config.LX, g, CoordinateX, hx);
        if (nDims > 1) BuildAxis(config.CompressionY, nY, config.LY, g, CoordinateY, hy);
        if (nDims > 2) BuildAxis(config.CompressionZ, nZ, config.LZ, g, CoordinateZ, hz);

        ts_config c{};
        c.ndims = nDims;
        c.nx = nX;
        c.ny = nY;
        c.nz = nZ;
        c.ghost = g;
        c.lx = config.LX;
        c.ly = config.LY;
        c.lz = config.LZ;
        c.periodic[0] = config.isPeriodicX;
        c.periodic[1] = nDims > 1 && config.isPeriodicY;
        c.periodic[2] = nDims > 2 && config.isPeriodicZ;
        c.coord[0] = CoordinateX.data();
        c.h[0] = hx.data();
        if (nDims > 1) c.coord[1] = CoordinateY.data(), c.h[1] = hy.data();
        if (nDims > 2) c.coord[2] = CoordinateZ.data(), c.h[2] = hz.data();
        const MethodConfiguration& m = config.methodConfiguration;
        c.gamma = config.Gamma;
        c.cfl = m.CFL;
        c.eps = m.Eps;
        c.riemann = static_cast<int>(m.RiemannProblemSolver);
        c.recon = static_cast<int>(m.ReconstructionType);
        c.limiter = static_cast<int>(m.GeneralLimitter);
        c.rk_stages = m.RungeKuttaOrder == 3 ? 3 : 1;
        c.precision = m.Precision;
        c.device = config.Device;
        c.viscous = config.IsViscousFlow;
        c.viscosity = config.Viscosity;
        c.external_force = config.IsExternalForceRequared;
        c.sigma[0] = config.Sigma.x;
        c.sigma[1] = config.Sigma.y;
        c.sigma[2] = config.Sigma.z;
        c.external_accel = config.IsUnifromAccelerationRequared;
        c.accel[0] = config.UniformAcceleration.x;
        c.accel[1] = config.UniformAcceleration.y;
        c.accel[2] = config.UniformAcceleration.z;

        // boundary conditions: the reference's markers (any int) become table slots 0..7
        std::map<int, int> slot;
        for (auto& kv : config.MyConditions) {
            if (c.n_bc == TS_MAX_BC) throw std::runtime_error("more boundary conditions than the library's table holds");
            ts_bc_config& b = c.bc[c.n_bc];
            b.type = static_cast<int>(kv.second.BCType);
            b.velocity[0] = kv.second.Velocity.x;
            b.velocity[1] = kv.second.Velocity.y;
            b.velocity[2] = kv.second.Velocity.z;
            b.pstatic = kv.second.Pstatic;
            b.density = kv.second.Density;
            slot[kv.first] = c.n_bc++;
        }
        auto lookup = [&](int marker) {
            auto it = slot.find(marker);
            if (it == slot.end()) throw std::runtime_error("boundary marker without a condition: " + std::to_string(marker));
            return it->second;
        };
        const DomainBCinfo* sides[6] = {&config.xLeftBoundary, &config.xRightBoundary, &config.yLeftBoundary,
                                        &config.yRightBoundary, &config.zLeftBoundary, &config.zRightBoundary};
        const bool periodic[3] = {c.periodic[0] != 0, c.periodic[1] != 0, c.periodic[2] != 0};
        for (int s = 0; s < 6; s++) {
            const int axis = s / 2;
            if (axis >= nDims || periodic[axis]) continue;
            const DomainBCinfo& d = *sides[s];
            if (!d.isComplex) {
                c.side_marker[s] = lookup(d.GetMarker());
                continue;
            }
            // composite side: the marker function at every boundary face centre (kernel.h:742,759)
            std::vector<uint8_t>& fm = face_markers_[s];
            const int na = axis == 0 ? (nDims > 1 ? nY : 1) : nX;
            const int nb = axis == 2 ? (nDims > 1 ? nY : 1) : (nDims > 2 ? nZ : 1);
            fm.assign((size_t)na * nb, 0);
            const double xb = s == 0 ? CoordinateX[g] - 0.5 * hx[g] : CoordinateX[g + nX - 1] + 0.5 * hx[g + nX - 1];
            for (int b2 = 0; b2 < nb; b2++)
                for (int a2 = 0; a2 < na; a2++) {
                    Vector r;
                    if (axis == 0) {
                        r = Vector(xb, nDims > 1 ? CoordinateY[g + a2] : 0.0, nDims > 2 ? CoordinateZ[g + b2] : 0.0);
                    } else if (axis == 1) {
                        const double yb = s == 2 ? CoordinateY[g] - 0.5 * hy[g] : CoordinateY[g + nY - 1] + 0.5 * hy[g + nY - 1];
                        r = Vector(CoordinateX[g + a2], yb, nDims > 2 ? CoordinateZ[g + b2] : 0.0);
                    } else {
                        const double zb = s == 4 ? CoordinateZ[g] - 0.5 * hz[g] : CoordinateZ[g + nZ - 1] + 0.5 * hz[g + nZ - 1];
                        r = Vector(CoordinateX[g + a2], CoordinateY[g + b2], zb);
                    }
                    fm[(size_t)b2 * na + a2] = (uint8_t)lookup(d.getMarker(r));
                }
            c.face_marker[s] = fm.data();
        }

        MaxIteration = config.MaxIteration;
        MaxTime = config.MaxTime;
        ResidualOutputIters = config.ResidualOutputIters;
        SaveHistoryFileIters = config.SaveHistoryFileIters;
        ChangeHistoryFileIters = config.ChangeHistoryFileIters;
        NextHistoryFileIters = ChangeHistoryFileIters;  // kernel.h:252-254
        SaveSolutionTime = config.SaveSolutionTime;
        SaveSliceTime = config.SaveSliceTime;
        SaveSolutionIters = config.SaveSolutionIters;
        SaveSliceIters = config.SaveSliceIters;
        SaveBinarySolIters = config.SaveBinarySolIters;
        Gamma = config.Gamma;
        MolarMass = config.MolarMass;  // kernel.h:214-215
        UniversalGasConstant = config.UniversalGasConstant;
        stepInfo = StepInfo();
        stepInfo.NextSolutionSnapshotTime = SaveSolutionTime;  // kernel.h:261-262
        stepInfo.NextSliceSnapshotTime = SaveSliceTime;

        if (ctx_) ts_destroy(ctx_), ctx_ = nullptr;
        int rc = ts_create(&c, &ctx_);
        if (rc != TS_OK) {
            std::string msg = ts_last_error(ctx_);  // with no context: this thread's create error
            if (ctx_) ts_destroy(ctx_), ctx_ = nullptr;
            throw std::runtime_error("ts_create failed (" + std::to_string(rc) + "): " + msg);
        }
        values.resize((size_t)nXAll * nYAll * nZAll * nVariables, 0.0);
        check(ts_set_time(ctx_, 0.0, 0));
    }

    // Kernel::SetInitialConditions with cell-centre point values (kernel.h:125-146)
    void SetInitialConditions(const std::function<std::vector<double>(const Vector& r)>& initF) {
        const int g = 1;
        for (int k = 0; k < nZ; k++)
            for (int j = 0; j < nY; j++)
                for (int i = 0; i < nX; i++) {
                    const int I = i + g, J = nDims > 1 ? j + g : 0, K = nDims > 2 ? k + g : 0;
                    Vector r(CoordinateX[I], nDims > 1 ? CoordinateY[J] : 0.0, nDims > 2 ? CoordinateZ[K] : 0.0);
                    std::vector<double> U = initF(r);
                    if ((int)U.size() < nVariables) throw std::runtime_error("initial condition returned too few values");
                    const size_t s = (((size_t)K * nYAll + J) * nXAll + I) * nVariables;
                    for (int v = 0; v < nVariables; v++) values[s + v] = U[v];
                }
        UploadValues();
    }

    // Kernel::SetInitialConditions with cell averages (kernel.h:149-175): the quadrature's
    // integral of initF over each interior cell divided by the cell volume hx hy hz
    void SetInitialConditions(const std::function<std::vector<double>(const Vector& r)>& initF,
                              NumericQuadrature& Int) {
        const int g = 1;
        for (int k = 0; k < nZ; k++)
            for (int j = 0; j < nY; j++)
                for (int i = 0; i < nX; i++) {
                    const int I = i + g, J = nDims > 1 ? j + g : 0, K = nDims > 2 ? k + g : 0;
                    const double x = CoordinateX[I], y = nDims > 1 ? CoordinateY[J] : 0.0;
                    const double z = nDims > 2 ? CoordinateZ[K] : 0.0;
                    const double hxc = hx[I], hyc = nDims > 1 ? hy[J] : 1.0, hzc = nDims > 2 ? hz[K] : 1.0;
                    const double cellV = hxc * hyc * hzc;
                    const std::valarray<double> avg =
                        Int.IntegrateGroap(x, y, z, hxc, hyc, hzc, initF, (size_t)nVariables) / cellV;
                    const size_t s = (((size_t)K * nYAll + J) * nXAll + I) * nVariables;
                    for (int v = 0; v < nVariables; v++) values[s + v] = avg[v];
                }
        UploadValues();
    }

    // the grid and the host mirror as the Tecplot writers read them
    TecplotView View() const {
        TecplotView v;
        v.nDims = nDims, v.nX = nX, v.nY = nY, v.nZ = nZ, v.nXAll = nXAll, v.nYAll = nYAll;
        v.X = CoordinateX.data(), v.hx = hx.data();
        if (nDims > 1) v.Y = CoordinateY.data(), v.hy = hy.data();
        if (nDims > 2) v.Z = CoordinateZ.data(), v.hz = hz.data();
        v.values = &values[0];
        v.gamma = Gamma, v.molarMass = MolarMass, v.universalGasConstant = UniversalGasConstant;
        v.time = stepInfo.Time;
        return v;
    }
    // Kernel::SaveSolutionToTecplot (kernel.h:1081-1416): the state comes from the device first
    void SaveSolutionToTecplot(const std::string& fname) {
        SyncValues();
        WriteTecplotSolution(fname, View());
    }
    // Kernel::Save2DSliceToTecplot (kernel.h:1416-1511); I/J/K with the ghost offset, -1 = free
    void Save2DSliceToTecplot(const std::string& fname, int I, int J, int K) {
        SyncValues();
        WriteTecplotSlice(fname, View(), I, J, K, false);
    }
    // Kernel::SaveSliceToTecplot (kernel.h:1513-1632)
    void SaveSliceToTecplot(const std::string& fname, const Slice& s) {
        SyncValues();
        WriteTecplotSlice(fname, View(), s.i, s.j, s.k, true);
    }

    // push the host mirror to the device (after editing `values` by hand)
    void UploadValues() { check(ts_set_state(ctx_, &values[0], values.size())); }
    // refresh the host mirror from the device
    void SyncValues() { check(ts_get_state(ctx_, &values[0], values.size())); }

    // ExplicitRungeKuttaFVM::IterationStep (ERK:1440-1491): one step
    void IterationStep() { Advance(1); }

    // Kernel::Run (kernel.h:1701-1850): steps until Iteration == MaxIteration or
    // Time >= MaxTime, with dt clamped onto the solution / slice snapshot times (ERK:1433-1434)
    // and the snapshot hooks called there.  Steps between snapshots run on the device without
    // a host round trip.
    //
    // Per iteration the reference (kernel.h:1701-1850) prints residuals every
    // ResidualOutputIters, processes the sensors every SaveSensorRecordIters, handles the
    // solution/slice snapshots, and appends a history.dat line every SaveHistoryFileIters
    // (starting history<n>.dat past each ChangeHistoryFileIters).  Here the device runs
    // uninterrupted up to the next such event; the residual norms of the step that ends a
    // device run are computed on the device when an event needs them.
    void Run() {
        const bool hist = SaveHistoryFileIters > 0;
        std::ofstream hofs;
        if (hist) OpenHistory(hofs, OutputDirectory + "/history.dat");
        const bool norms = ComputeResidualNorms || hist || ResidualOutputIters > 0;
        SetNorms(norms);
        const bool sensing = isSensorEnable && !Sensors.empty() && SaveSensorRecordIters > 0;
        while (true) {
            // the reference steps first and tests afterwards, so a run always takes a step, and
            // one started past MaxIteration only ends on time
            const int it = stepInfo.Iteration;
            long long left = it < MaxIteration ? MaxIteration - it : 1 << 30;
            auto upto = [&](int every) {
                if (every > 0) left = std::min<long long>(left, every - ((it % every) + every) % every);
            };
            if (hist) upto(SaveHistoryFileIters);
            if (sensing) upto(SaveSensorRecordIters);
            if (ResidualOutputIters > 0) upto(ResidualOutputIters);
            if (WriteSnapshotFiles) {
                upto(SaveSolutionIters);
                if (!slices.empty()) upto(SaveSliceIters), upto(SaveBinarySolIters);
            }
            if (hist && NextHistoryFileIters >= it) left = std::min<long long>(left, NextHistoryFileIters + 1 - it);
            if (left < 1) left = 1;
            ts_step_info info{};
            check(ts_step(ctx_, (int)left, MaxTime, NextSnapshot(), &info));
            Absorb(info);
            if (info.steps_done == 0) break;
            const int now = stepInfo.Iteration;
            if (ResidualOutputIters > 0 && now % ResidualOutputIters == 0)
                std::cout << "Iteration = " << now << "; Total time = " << stepInfo.Time
                          << "; Time step = " << stepInfo.TimeStep << "; RMSrou = " << stepInfo.Residual[1] << std::endl;
            if (sensing && now % SaveSensorRecordIters == 0)
                for (auto& sn : Sensors) {
                    sn->UpdateIteration(now);
                    sn->UpdateTimer(stepInfo.Time);
                    sn->Process(*this);
                }
            if (WriteSnapshotFiles) IterationSnapshots(now);
            Snapshots();
            if (hist) {
                if (now > NextHistoryFileIters) {
                    const int n_his = now / ChangeHistoryFileIters;
                    NextHistoryFileIters += ChangeHistoryFileIters;
                    hofs.close();
                    OpenHistory(hofs, OutputDirectory + "/history" + std::to_string(n_his) + ".dat");
                }
                if (now % SaveHistoryFileIters == 0) {
                    hofs << stepInfo.Time << " " << stepInfo.Residual[0] << " " << stepInfo.Residual[1] << " ";
                    if (nDims > 1) hofs << stepInfo.Residual[2] << " ";
                    if (nDims > 2) hofs << stepInfo.Residual[3] << " ";
                    hofs << stepInfo.Residual[4] << " " << now << std::endl;
                }
            }
            if (stepInfo.Iteration == MaxIteration || stepInfo.Time >= MaxTime) break;
        }
    }

    // values of interior cells (reference indices i, j, k with the ghost offset) from the device
    std::vector<std::valarray<double>> GetCells(const std::vector<int>& ijk) {
        const size_t n = ijk.size() / 3;
        std::vector<double> raw(n * 5);
        check(ts_get_cells(ctx_, ijk.data(), n, raw.data()));
        std::vector<std::valarray<double>> out(n, std::valarray<double>(5));
        for (size_t c = 0; c < n; c++)
            for (int v = 0; v < 5; v++) out[c][v] = raw[c * 5 + v];
        return out;
    }

    // Kernel::SaveSolution (kernel.h:992-1029), as one rank writes it: the time (double), the
    // rank (int 0), then every interior cell's 5 doubles in k, j, i order (i fastest).  The
    // state comes from the device first.
    void SaveSolution(const std::string& fname) {
        SyncValues();
        std::FILE* f = std::fopen(fname.c_str(), "wb");
        if (!f) throw std::runtime_error("cannot open " + fname + " for writing");
        const int rank = 0;
        bool ok = std::fwrite(&stepInfo.Time, sizeof(double), 1, f) == 1 && std::fwrite(&rank, sizeof(int), 1, f) == 1;
        ForInterior([&](size_t s) { ok = ok && std::fwrite(&values[s], sizeof(double), nVariables, f) == (size_t)nVariables; });
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) throw std::runtime_error("writing " + fname + " failed");
    }

    // Kernel::LoadSolution (kernel.h:1031-1079) for a file written by one rank: restores the
    // time and the interior state (ghosts are regenerated on the device), and moves the next
    // snapshot times past the restored time as kernel.h:1043-1044 do.  The iteration counter
    // is not stored in the format and keeps its value.
    void LoadSolution(const std::string& fname) {
        std::FILE* f = std::fopen(fname.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open " + fname);
        double t = 0;
        int rank = -1;
        bool ok = std::fread(&t, sizeof(double), 1, f) == 1 && std::fread(&rank, sizeof(int), 1, f) == 1;
        ok = ok && rank == 0;
        ForInterior([&](size_t s) { ok = ok && std::fread(&values[s], sizeof(double), nVariables, f) == (size_t)nVariables; });
        char extra;
        ok = ok && std::fread(&extra, 1, 1, f) == 0;  // exactly one rank's block: nothing left over
        std::fclose(f);
        if (!ok) throw std::runtime_error(fname + " is not a one-rank solution file of this grid");
        stepInfo.Time = t;
        if (SaveSolutionTime > 0) stepInfo.NextSolutionSnapshotTime = (int(t / SaveSolutionTime) + 1) * SaveSolutionTime;
        if (SaveSliceTime > 0) stepInfo.NextSliceSnapshotTime = (int(t / SaveSliceTime) + 1) * SaveSliceTime;
        UploadValues();
        check(ts_set_time(ctx_, stepInfo.Time, stepInfo.Iteration));
    }

    ts_ctx* Context() const { return ctx_; }

   private:
    ts_ctx* ctx_ = nullptr;
    // visit the interior cells in the reference's k, j, i order; s = offset in `values`
    template <typename F> void ForInterior(F&& f) {
        const int gy = nDims > 1 ? 1 : 0, gz = nDims > 2 ? 1 : 0;
        for (int k = 0; k < nZ; k++)
            for (int j = 0; j < nY; j++)
                for (int i = 0; i < nX; i++)
                    f((((size_t)(k + gz) * nYAll + (j + gy)) * nXAll + (i + 1)) * nVariables);
    }
    bool norms_on_ = false;
    std::vector<uint8_t> face_markers_[6];

    void check(int rc) {
        if (rc != TS_OK) throw std::runtime_error("libts_fvm: " + std::to_string(rc) + ": " + ts_last_error(ctx_));
    }
    void EnableNorms() { SetNorms(ComputeResidualNorms); }
    void SetNorms(bool on) {
        if (on != norms_on_) {
            check(ts_enable_residual_norms(ctx_, on ? 1 : 0));
            norms_on_ = on;
        }
    }
    // history file header (kernel.h:1646-1657, 1808-1819)
    void OpenHistory(std::ofstream& ofs, const std::string& path) {
        ofs.open(path, std::ios_base::out);
        if (!ofs) throw std::runtime_error("cannot open " + path);
        ofs << "VARIABLES = \"time\" \"roR\" \"rouR\" ";
        if (nDims > 1) ofs << "\"rovR\" ";
        if (nDims > 2) ofs << "\"rowR\" ";
        ofs << "\"roeR\" \"iter\" " << std::endl;
    }
    double NextSnapshot() const {
        double t = 0.0;
        if (SaveSolutionTime > 0) t = stepInfo.NextSolutionSnapshotTime;
        if (SaveSliceTime > 0 && (t == 0.0 || stepInfo.NextSliceSnapshotTime < t)) t = stepInfo.NextSliceSnapshotTime;
        return t;
    }
    void Absorb(const ts_step_info& info) {
        stepInfo.Time = info.time;
        stepInfo.TimeStep = info.dt;
        stepInfo.Iteration = (int)info.iteration;
        for (int v = 0; v < nVariables; v++) stepInfo.Residual[v] = info.residual_norm[v];
    }
    void Advance(int n) {
        EnableNorms();
        ts_step_info info{};
        check(ts_step(ctx_, n, 1e300, NextSnapshot(), &info));
        Absorb(info);
        Snapshots();
    }
    void Snapshots() {
        const bool sol = SaveSolutionTime > 0 && stepInfo.NextSolutionSnapshotTime == stepInfo.Time;
        const bool slc = SaveSliceTime > 0 && stepInfo.NextSliceSnapshotTime == stepInfo.Time;
        if (!sol && !slc) return;
        SyncValues();
        if (sol) {
            if (WriteSnapshotFiles) {  // kernel.h:1762-1777
                std::stringstream fn;
                fn << std::fixed;
                fn.precision(TimePrecision(SaveSolutionTime));
                fn << "solution, t = " << stepInfo.Time << ".dat";
                WriteTecplotSolution(OutputDirectory + "/" + fn.str(), View());
            }
            if (OnSolutionSnapshot) OnSolutionSnapshot(*this);
            stepInfo.NextSolutionSnapshotTime += SaveSolutionTime;
        }
        if (slc) {
            if (WriteSnapshotFiles)  // kernel.h:1778-1795
                for (size_t i = 0; i < slices.size(); i++) {
                    std::stringstream fn;
                    fn << "slice" << i;
                    fn << std::fixed;
                    fn.precision(TimePrecision(SaveSliceTime));
                    fn << ",T=" << stepInfo.Time << ".dat";
                    const Slice& s = slices[i];
                    WriteTecplotSlice(OutputDirectory + "/" + fn.str(), View(), s.i, s.j, s.k, true);
                }
            if (OnSliceSnapshot) OnSliceSnapshot(*this);
            stepInfo.NextSliceSnapshotTime += SaveSliceTime;
        }
    }
    // digits of the snapshot times in file names (kernel.h:1636-1644)
    static int TimePrecision(double every) { return static_cast<int>(1 - std::floor(std::log10(every))); }
    // the snapshots every few iterations (kernel.h:1723-1760)
    void IterationSnapshots(int it) {
        const bool sol = SaveSolutionIters != 0 && it % SaveSolutionIters == 0;
        const bool slc = SaveSliceIters != 0 && it % SaveSliceIters == 0 && !slices.empty();
        const bool bin = SaveBinarySolIters != 0 && it % SaveBinarySolIters == 0 && !slices.empty();
        if (!sol && !slc && !bin) return;
        SyncValues();
        if (sol) {
            std::stringstream fn;
            fn << "solution, It = " << it << ".dat";
            WriteTecplotSolution(OutputDirectory + "/" + fn.str(), View());
        }
        if (slc)
            for (size_t i = 0; i < slices.size(); i++) {
                std::stringstream fn;
                fn << "slice" << i << ", It =" << it << ".dat";
                const Slice& s = slices[i];
                WriteTecplotSlice(OutputDirectory + "/" + fn.str(), View(), s.i, s.j, s.k, true);
            }
        // the reference saves solution.sol once per slice (kernel.h:1752-1759)
        if (bin)
            for (size_t i = 0; i < slices.size(); i++) SaveSolution(OutputDirectory + "/solution.sol");
    }
};

// src/Sensors/CellSensor.h and CellMultiSensor.h: function(s) of the state of one cell, one line
// "iteration value(s) time" per record; the file starts with the cell centre (Sensor.h:55-63)
class CellSensor : public Sensor {
   public:
    CellSensor(std::string fname, sen_func f) : Sensor(std::move(fname)), fs_{std::move(f)} {}
    CellSensor(std::string fname, std::vector<sen_func> fs) : Sensor(std::move(fname)), fs_(std::move(fs)) {
        multi_ = true;
    }
    void SetSensor(const HipExplicitRungeKuttaFVM& k, int i, int j, int kk);
    void Process(HipExplicitRungeKuttaFVM& k) override {
        if (!active_) return;
        std::valarray<double> U = k.GetCells(ijk_)[0];
        std::ofstream ofs(filename, std::ios_base::app);
        ofs << iteration << ' ';
        if (multi_) {
            for (auto& f : fs_) ofs << f(U) << ' ';
        } else {
            ofs << fs_[0](U) << ' ';
        }
        ofs << timer << std::endl;
    }

   private:
    std::vector<sen_func> fs_;
    std::vector<int> ijk_;
    bool active_{false}, multi_{false};
};

// src/Sensors/MValuePosXSensor.h / MValuePosXSensor2.h: position of the maximum of a function
// along the x-line (j, k); `second_order` places it at the vertex of the parabola through the
// maximum cell and its neighbours on the (non-uniform) grid, the end cells at their outer faces
class MValuePosXSensor : public Sensor {
   public:
    MValuePosXSensor(std::string fname, sen_func f, bool second_order = false)
        : Sensor(std::move(fname)), f_(std::move(f)), second_(second_order) {}
    void SetSensor(const HipExplicitRungeKuttaFVM& k, int j_slice, int k_slice);
    void Process(HipExplicitRungeKuttaFVM& k) override;

   private:
    sen_func f_;
    bool second_;
    std::vector<int> ijk_;
};

inline void CellSensor::SetSensor(const HipExplicitRungeKuttaFVM& k, int i, int j, int kk) {
    ijk_ = {i, j, kk};
    active_ = true;
    std::ofstream ofs(filename);
    ofs << "TITLE = \"X=" << k.CoordinateX[i] << ", Y=" << (k.nDims > 1 ? k.CoordinateY[j] : 0.0)
        << ", Z=" << (k.nDims > 2 ? k.CoordinateZ[kk] : 0.0) << '"' << std::endl;
}

inline void MValuePosXSensor::SetSensor(const HipExplicitRungeKuttaFVM& k, int j_slice, int k_slice) {
    // j_slice, k_slice: interior indices without the ghost offset (MValuePosXSensor.h:22-30)
    ijk_.clear();
    const int J = k.nDims > 1 ? j_slice + 1 : 0, K = k.nDims > 2 ? k_slice + 1 : 0;
    for (int i = 1; i <= k.nX; i++) ijk_.insert(ijk_.end(), {i, J, K});
}

inline void MValuePosXSensor::Process(HipExplicitRungeKuttaFVM& k) {
    if (ijk_.empty()) return;
    const std::vector<std::valarray<double>> row = k.GetCells(ijk_);
    double uMax = -std::numeric_limits<double>::max();
    int iMax = 0;
    std::vector<double> u(row.size());
    for (size_t i = 0; i < row.size(); i++) {
        u[i] = f_(row[i]);
        if (u[i] > uMax) uMax = u[i], iMax = (int)i;
    }
    const int I = 1 + iMax;  // index with the ghost offset
    double xmax = k.CoordinateX[I];
    if (second_) {
        if (iMax == 0) {
            xmax = k.CoordinateX[I] - 0.5 * k.hx[I];
        } else if (iMax == k.nX - 1) {
            xmax = k.CoordinateX[I] + 0.5 * k.hx[I];
        } else {
            const double xi = k.CoordinateX[I], dxp = k.CoordinateX[I + 1] - xi, dxm = xi - k.CoordinateX[I - 1];
            const double v2 = uMax - u[iMax - 1], v3 = u[iMax + 1] - u[iMax - 1];
            const double adxm = v3 * dxm / ((dxm + dxp) * dxp) - v2 / dxp;
            const double bda = dxm + v2 / adxm;
            xmax = xi - 0.5 * bda;
        }
    }
    std::ofstream ofs(filename, std::ios_base::app);
    ofs << iteration << ' ' << xmax << ' ' << timer << std::endl;
}

// CreateKernel (ExplicitRungeKuttaFVM.h:1495-1515)
inline std::unique_ptr<HipExplicitRungeKuttaFVM> CreateKernel(KernelConfiguration&) {
    return std::unique_ptr<HipExplicitRungeKuttaFVM>(new HipExplicitRungeKuttaFVM());
}

}  // namespace tsh

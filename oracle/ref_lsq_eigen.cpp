/*
 * ref_lsq_eigen.cpp -- TEST INFRASTRUCTURE ONLY.  Builds oracle/_ref/libref_lsq_eigen.so against
 * the Eigen 3.2.5 headers the reference vendors (libs/eigen), read where they lie: the
 * least-squares gradient of GradientComputer::ExecuteByLeastSquares (GradientComputer.h:19-53)
 * solved by Eigen's JacobiSVD, as the reference solves it.  Used only to check the oracle's
 * normal-equations restatement of that solve.  (GradientComputer.h itself does not compile with
 * g++: it includes <Eigen\Dense> with a backslash and builds std::exception from a string, both
 * MSVC-only, so the call is written here against Eigen directly.)
 */
#include <Eigen/Dense>

extern "C" int ref_lsq_eigen(int nd, const double* point, double value, const double* pts, const double* vals,
                             int m, double* grad) {
    Eigen::MatrixXd A(m, nd);
    Eigen::VectorXd rhs(m);
    for (int i = 0; i < m; i++) {
        for (int e = 0; e < nd; e++) A(i, e) = point[e] - pts[3 * i + e];
        rhs(i) = value - vals[i];
    }
    Eigen::JacobiSVD<Eigen::MatrixXd> svd(A, Eigen::ComputeThinU | Eigen::ComputeThinV);
    Eigen::VectorXd x = svd.solve(rhs);
    for (int e = 0; e < 3; e++) grad[e] = e < nd ? x(e) : 0.0;
    return svd.rank() < nd ? 1 : 0;
}

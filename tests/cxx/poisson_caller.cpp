// A C++ caller written against the mrcpp names (as CoulombPotential.cpp:69-90 and
// qmfunction_utils.cpp:88-263 use them), compiled against the facade over libmwgreen.so.
// Projects three normalised Gaussian charges, applies the Poisson operator and checks the
// potential against the closed form sum_i q_i erf(sqrt(a_i) r_i) / r_i; then exercises
// multiply, add, copy_func, build_grid, dot, saveTree / loadTree and the refusals.
// Prints one line per check; exit status 0 when every check holds.  Needs a GPU.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "mwg_mrcpp.hpp"

namespace {
int fails = 0;
void expect(bool ok, const char *what, double a = 0.0, double b = 0.0) {
    std::printf("%s %s (%.6e, %.6e)\n", ok ? "ok  " : "FAIL", what, a, b);
    if (!ok) fails++;
}
}  // namespace

int main(int argc, char **argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const double prec = 1e-4;
    int corner[3] = {-1, -1, -1};
    int boxes[3] = {2, 2, 2};
    mrcpp::BoundingBox<3> world(-4, corner, boxes);  // [-16, 16]^3
    mrcpp::LegendreBasis basis(5);
    mrcpp::MultiResolutionAnalysis<3> MRA(world, basis, 16);

    // rho = sum_i q_i (a_i / pi)^{3/2} exp(-a_i |r - c_i|^2)
    const double q[3] = {1.0, 0.7, 1.5}, a[3] = {4.0, 12.0, 30.0};
    const double c[3][3] = {{0.3, -0.5, 0.2}, {-1.0, 0.8, -0.4}, {0.9, 1.1, -1.3}};
    mrcpp::GaussExp<3> f;
    for (int i = 0; i < 3; i++)
        f.append(mrcpp::GaussFunc<3>(a[i], q[i] * std::pow(a[i] / M_PI, 1.5), {c[i][0], c[i][1], c[i][2]}));
    mrcpp::FunctionTree<3> rho(MRA);
    mrcpp::project(prec, rho, f);

    mrcpp::PoissonOperator P(MRA, prec);
    mrcpp::FunctionTree<3> V(MRA);
    mrcpp::apply(prec, V, P, rho);

    // closed form at pseudo-random points (the L2 error estimate of the Python tests)
    uint64_t s = 12345;
    auto uni = [&]() {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((s >> 11) * (1.0 / 9007199254740992.0)) * 32.0 - 16.0;
    };
    const int npts = 50000;
    std::vector<double> xyz(3 * (size_t)npts), val((size_t)npts);
    for (double &v : xyz) v = uni();
    mrcpp::detail::check(mwg_tree_eval(V.handle(), npts, xyz.data(), val.data()));  // one batched evaluation
    double se = 0.0;
    for (int p = 0; p < npts; p++) {
        const double *r = &xyz[3 * (size_t)p];
        double exact = 0.0;
        for (int i = 0; i < 3; i++) {
            const double dx = r[0] - c[i][0], dy = r[1] - c[i][1], dz = r[2] - c[i][2];
            const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
            exact += d > 0.0 ? q[i] * std::erf(std::sqrt(a[i]) * d) / d : q[i] * 2.0 * std::sqrt(a[i] / M_PI);
        }
        const double e = val[(size_t)p] - exact;
        se += e * e;
    }
    const double l2 = std::sqrt(se / npts * 32.0 * 32.0 * 32.0), vnorm = std::sqrt(V.getSquareNorm());
    expect(l2 < prec * vnorm, "poisson apply vs erf closed form (L2 error, prec ||V||)", l2, prec * vnorm);
    // <rho, V> = sum_ij q_i q_j erf(sqrt(a_i a_j / (a_i + a_j)) R_ij) / R_ij (Coulomb energy x 2)
    double e2 = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double pij = std::sqrt(a[i] * a[j] / (a[i] + a[j]));
            const double dx = c[i][0] - c[j][0], dy = c[i][1] - c[j][1], dz = c[i][2] - c[j][2];
            const double R = std::sqrt(dx * dx + dy * dy + dz * dz);
            e2 += q[i] * q[j] * (R > 0.0 ? std::erf(pij * R) / R : 2.0 * pij / std::sqrt(M_PI));
        }
    const double rv = mrcpp::dot(rho, V);
    expect(std::fabs(rv - e2) < 10.0 * prec * e2, "dot(rho, V) vs the Gaussian Coulomb integral", rv, e2);

    // copy_func, add (V - copy = 0), saveTree / loadTree
    mrcpp::FunctionTree<3> W(MRA), D(MRA), L(MRA);
    mrcpp::copy_func(W, V);
    const double vv = V.getSquareNorm();
    expect(W.getNNodes() == V.getNNodes() && std::fabs(mrcpp::dot(W, V) - vv) <= 1e-12 * vv, "copy_func",
           mrcpp::dot(W, V), vv);
    mrcpp::add(prec, D, 1.0, V, -1.0, W);
    expect(D.getSquareNorm() == 0.0, "add: V - copy(V) = 0", D.getSquareNorm());
    V.saveTree(dir + "/V");
    L.loadTree(dir + "/V");
    expect(L.getNNodes() == V.getNNodes() && std::fabs(mrcpp::dot(L, V) - vv) <= 1e-12 * vv, "saveTree / loadTree",
           mrcpp::dot(L, V), vv);
    // FunctionTreeVector: 2 V - V - V = 0
    mrcpp::FunctionTreeVector<3> vec{{2.0, &V}, {-1.0, &W}, {-1.0, &L}};
    mrcpp::FunctionTree<3> Z(MRA);
    mrcpp::add(prec, Z, vec);
    expect(Z.getSquareNorm() < 1e-24 * V.getSquareNorm(), "add(FunctionTreeVector)", Z.getSquareNorm());
    // multiply: <rho^2, 1> through dot(rho, rho)
    mrcpp::FunctionTree<3> rho2(MRA);
    mrcpp::multiply(prec, rho2, 1.0, rho, rho);
    expect(rho2.getNNodes() >= rho.getNNodes(), "multiply grid covers the factor", rho2.getNNodes(), rho.getNNodes());
    // build_grid: rho on the union grid with V keeps rho
    mrcpp::FunctionTree<3> R(MRA);
    mrcpp::copy_func(R, rho);
    mrcpp::build_grid(R, V);
    expect(R.getNNodes() >= V.getNNodes() && std::fabs(mrcpp::dot(R, rho) - rho.getSquareNorm()) <=
                                                 1e-12 * rho.getSquareNorm(),
           "build_grid keeps the function", mrcpp::dot(R, rho), rho.getSquareNorm());
    // refusals: absolute precision, the mixed path at k = 5
    bool thrown = false;
    try {
        mrcpp::FunctionTree<3> X(MRA);
        mrcpp::apply(prec, X, P, rho, -1, true);
    } catch (const mrcpp::Error &) {
        thrown = true;
    }
    expect(thrown, "absPrec = true is refused");
    thrown = false;
    try {
        mrcpp::FunctionTree<3> X(MRA);
        mrcpp::apply(prec, X, P, rho, mrcpp::ApplyMode::Mixed);
    } catch (const mrcpp::Error &) {
        thrown = true;
    }
    expect(thrown, "mixed mode at k = 5 is refused");
    std::printf("%s\n", fails ? "poisson caller: FAILED" : "poisson caller: ok");
    return fails ? 1 : 0;
}

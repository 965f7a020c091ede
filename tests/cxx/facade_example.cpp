// A caller written against the mrcpp names (as in src/scf_solver/HelmholtzVector.cpp:168-184
// and src/qmoperators/two_electron/CoulombPotential.cpp:69-90), compiled against the
// facade over libmwgreen.so.  Run with a GPU: prints nodes and norms of a Poisson and a
// Helmholtz apply of an empty-grid function (exercises the error path) .
#include <cmath>
#include <cstdio>

#include "mwg_mrcpp.hpp"

int main() {
    int corner[3] = {-1, -1, -1};
    int boxes[3] = {2, 2, 2};
    mrcpp::BoundingBox<3> world(-4, corner, boxes);
    mrcpp::LegendreBasis basis(5);
    mrcpp::MultiResolutionAnalysis<3> MRA(world, basis, 16);
    mrcpp::PoissonOperator P(MRA, 1e-4);
    mrcpp::HelmholtzOperator H(MRA, 1.0, 1e-4);
    mrcpp::FunctionTree<3> rho(MRA), V(MRA);
    try {
        mrcpp::apply(1e-4, V, P, rho);  // rho has no nodes: rejected like MRCPP's MSG_ERROR
        std::printf("unexpected success\n");
        return 1;
    } catch (const mrcpp::Error &e) {
        std::printf("error path ok: %s\n", e.what());
    }
    std::printf("terms: poisson %d helmholtz %d\n", P.size(), H.size());
    return 0;
}

// Host-code exercise for an AddressSanitizer + UndefinedBehaviorSanitizer build: the FP64
// oracle (oracle/mwref.cpp, through its C API) and the product's host numerics
// (mrchem_amd/csrc/mwg_core.cpp) are linked into one executable and driven through
// projection, Poisson and Helmholtz applies, products, dot, import/export, every basis and
// the tree-layout check.  Exits non-zero on a wrong result; the sanitizers abort on the
// first memory or undefined-behaviour error.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../mrchem_amd/csrc/mwg_core.h"

extern "C" {
struct mwref_stats {
    int64_t tuples, gen_nodes, nodes_out, iters;
    double t_total;
    int64_t tuples_iter[64];
    int64_t nodes_iter[64];
};
void *mwref_mra_new(int n0, const int *corner, const int *boxes, double sfac, int k, int basis_type, int max_depth);
void mwref_mra_free(void *m);
void *mwref_tree_new(void *m);
void mwref_tree_free(void *t);
int mwref_tree_nnodes(void *t);
double mwref_tree_sqnorm(void *t);
void mwref_project_gauss(void *t, double prec, int n, const double *coef, const double *alpha, const double *center,
                         const double *poly);
double mwref_dot(void *a, void *b);
void mwref_tree_export(void *t, int *depth, int64_t *lr, double *coefs);
int mwref_tree_import(void *t, int n, const int *depth, const int64_t *lr, const double *coefs);
void *mwref_oper_new(void *m, double mu, double prec);
void mwref_oper_free(void *o);
int mwref_apply(double prec, void *out, void *oper, void *inp, int max_iter, mwref_stats *st);
void mwref_multiply(void *out, double prec, double coef, void *a, void *b);
void mwref_set_threads(int n);
}

static int fails = 0;
static void check(bool ok, const char *what) {
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s\n", what);
        fails++;
    }
}

int main() {
    mwref_set_threads(2);
    // host numerics of the product: every order and basis, a kernel fit, the layout check
    for (int k = 3; k <= 9; k++)
        for (int type = 0; type < 2; type++) {
            mwg::Basis B;
            mwg::build_basis(B, k, type);
            check((int)B.W.size() == 4 * B.K * B.K, "two-scale matrix size");
        }
    const mwg::KernelFit fit = mwg::fit_kernel(0.0, 1e-4, 1e-4, 60.0);
    check(!fit.a.empty() && fit.err <= 1e-4 * 1.01, "Poisson kernel fit within eps");
    const int boxes[3] = {2, 2, 2};
    std::vector<int> depth(8, 0);
    std::vector<int64_t> lr;
    for (int z = 0; z < 2; z++)
        for (int y = 0; y < 2; y++)
            for (int x = 0; x < 2; x++) lr.insert(lr.end(), {x, y, z});
    check(mwg::check_tree_layout(boxes, 10, 8, depth.data(), lr.data()).empty(), "valid root layout");
    lr[0] = 5;
    check(!mwg::check_tree_layout(boxes, 10, 8, depth.data(), lr.data()).empty(), "out-of-box root refused");
    depth.push_back(1);
    lr.insert(lr.end(), {0, 0, 0});
    check(!mwg::check_tree_layout(boxes, 10, 9, depth.data(), lr.data()).empty(), "incomplete group refused");

    // the oracle: a charge density, Poisson and Helmholtz applies, products, dot, round trip
    const int corner[3] = {-1, -1, -1}, rb[3] = {2, 2, 2};
    void *mra = mwref_mra_new(-3, corner, rb, 1.0, 5, 0, 12);
    const double coef[2] = {1.0, 0.5}, alpha[2] = {4.0, 9.0}, center[6] = {0.1, -0.2, 0.3, -0.4, 0.2, 0.1};
    std::vector<double> poly(30, 0.0);
    poly[0] = poly[5] = poly[10] = poly[15] = poly[20] = poly[25] = 1.0;
    const double prec = 1e-3;
    void *f = mwref_tree_new(mra);
    mwref_project_gauss(f, prec, 2, coef, alpha, center, poly.data());
    check(mwref_tree_nnodes(f) > 8 && mwref_tree_sqnorm(f) > 0.0, "projection");
    for (double mu : {0.0, 1.0}) {
        void *op = mwref_oper_new(mra, mu, prec);
        void *g = mwref_tree_new(mra);
        mwref_stats st{};
        mwref_apply(prec, g, op, f, -1, &st);
        check(st.tuples > 0 && st.nodes_out == mwref_tree_nnodes(g), "apply statistics");
        check(mwref_dot(g, f) > 0.0, "positive-definite kernel: <G f, f> > 0");
        mwref_tree_free(g);
        mwref_oper_free(op);
    }
    void *p = mwref_tree_new(mra);
    mwref_multiply(p, prec, 1.0, f, f);
    check(mwref_tree_nnodes(p) > 0, "product");
    const int n = mwref_tree_nnodes(f);
    std::vector<int> d(n);
    std::vector<int64_t> l(3 * (size_t)n);
    std::vector<double> c((size_t)n * 8 * 216);
    mwref_tree_export(f, d.data(), l.data(), c.data());
    void *f2 = mwref_tree_new(mra);
    check(mwref_tree_import(f2, n, d.data(), l.data(), c.data()) == 0, "import of an export");
    check(std::fabs(mwref_dot(f2, f) - mwref_tree_sqnorm(f)) <= 1e-12 * mwref_tree_sqnorm(f), "round trip");
    mwref_tree_free(f2);
    mwref_tree_free(p);
    mwref_tree_free(f);
    mwref_mra_free(mra);
    if (fails) return 1;
    std::printf("sanitized host run: ok\n");
    return 0;
}

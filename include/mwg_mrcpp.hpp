// mwg_mrcpp.hpp — header-only C++ facade with the mrcpp names MRChem's hot path uses,
// implemented over the C ABI of libmwgreen.so (include/mwg.h).
//
// Drop-in for the subset of MRCPP called by
//   src/scf_solver/HelmholtzVector.cpp:171-181
//       mrcpp::HelmholtzOperator H(*MRA, mu_i.real(), this->prec);
//       mrcpp::apply(this->prec, out.real(), H, phi.real());
//   src/qmoperators/two_electron/CoulombPotential.cpp:85
//       mrcpp::apply(abs_prec, V.real(), P, rho.real());
//   src/qmoperators/two_electron/ExchangePotential.cpp:259,318 (pair Poissons)
//   src/SCFDriver.cpp:279     new mrcpp::PoissonOperator(*MRA, rel_prec)
//   src/mrenv.cpp:46-72       BoundingBox<3>, LegendreBasis / InterpolatingBasis,
//                             MultiResolutionAnalysis<3>(world, basis, max_depth)
// and the tree utilities around every apply (src/qmfunctions/qmfunction_utils.cpp:88-263,
// QMFunction.cpp:157-200, Orbital.cpp:118-190): project (Gaussian expansions), multiply,
// add (two terms or a FunctionTreeVector), build_grid, copy_grid, copy_func, dot,
// FunctionTree::saveTree / loadTree / setZero.
// Errors: MRCPP aborts (MSG_ERROR / MSG_FATAL); this facade throws mrcpp::Error
// carrying mwg_last_error().
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <tuple>
#include <string>
#include <utility>
#include <vector>

#include "mwg.h"

namespace mrcpp {

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

namespace detail {
inline mwg_ctx context() {
    static thread_local std::unique_ptr<mwg_ctx_s, int (*)(mwg_ctx)> ctx(nullptr, mwg_ctx_destroy);
    if (!ctx) {
        mwg_ctx c = nullptr;
        if (mwg_ctx_create(0, &c) != MWG_OK) throw Error("mwg_ctx_create failed");
        ctx.reset(c);
    }
    return ctx.get();
}
inline void check(int rc) {
    if (rc != MWG_OK) throw Error(std::string("mwg: ") + mwg_last_error(context()));
}
}  // namespace detail

template <int D>
class BoundingBox;
template <>
class BoundingBox<3> {
   public:
    BoundingBox(int min_scale, const int *corner, const int *boxes, double scaling_factor = 1.0)
        : scale_(min_scale), sfac_(scaling_factor) {
        for (int d = 0; d < 3; d++) {
            corner_[d] = corner[d];
            boxes_[d] = boxes[d];
        }
    }
    int getScale() const { return scale_; }
    const int *getCorner() const { return corner_; }
    const int *getBoxes() const { return boxes_; }
    double getScalingFactor() const { return sfac_; }

   private:
    int scale_;
    int corner_[3];
    int boxes_[3];
    double sfac_;
};

class ScalingBasis {
   public:
    ScalingBasis(int order, int type) : order_(order), type_(type) {}
    int getScalingOrder() const { return order_; }
    int getScalingType() const { return type_; }

   private:
    int order_, type_;
};
class LegendreBasis : public ScalingBasis {
   public:
    explicit LegendreBasis(int order) : ScalingBasis(order, MWG_BASIS_LEGENDRE) {}
};
class InterpolatingBasis : public ScalingBasis {
   public:
    explicit InterpolatingBasis(int order) : ScalingBasis(order, MWG_BASIS_INTERPOLATING) {}
};

template <int D>
class MultiResolutionAnalysis;
template <>
class MultiResolutionAnalysis<3> {
   public:
    MultiResolutionAnalysis(const BoundingBox<3> &box, const ScalingBasis &basis, int max_depth) {
        detail::check(mwg_mra_create(detail::context(), box.getScale(), box.getCorner(), box.getBoxes(),
                                     box.getScalingFactor(), basis.getScalingOrder(), basis.getScalingType(),
                                     max_depth, &h_));
    }
    ~MultiResolutionAnalysis() { mwg_mra_destroy(h_); }
    MultiResolutionAnalysis(const MultiResolutionAnalysis &) = delete;
    MultiResolutionAnalysis &operator=(const MultiResolutionAnalysis &) = delete;
    mwg_mra handle() const { return h_; }

   private:
    mwg_mra h_ = nullptr;
};

template <int D>
class FunctionTree;
template <>
class FunctionTree<3> {
   public:
    explicit FunctionTree(const MultiResolutionAnalysis<3> &mra) { detail::check(mwg_tree_create(mra.handle(), &h_)); }
    explicit FunctionTree(mwg_mra mra) { detail::check(mwg_tree_create(mra, &h_)); }
    ~FunctionTree() { mwg_tree_destroy(h_); }
    FunctionTree(const FunctionTree &) = delete;
    FunctionTree &operator=(const FunctionTree &) = delete;
    int getNNodes() const {
        int64_t n = 0;
        detail::check(mwg_tree_nnodes(h_, &n));
        return (int)n;
    }
    double getSquareNorm() const {
        double v = 0.0;
        detail::check(mwg_tree_sqnorm(h_, &v));
        return v;
    }
    void rescale(double c) { detail::check(mwg_tree_rescale(h_, c)); }
    void setZero() { detail::check(mwg_tree_rescale(h_, 0.0)); }
    // FunctionTree::saveTree / loadTree: "<file>.tree" (Orbital.cpp:138, 180)
    void saveTree(const std::string &file) { detail::check(mwg_tree_save(h_, file.c_str(), nullptr)); }
    void loadTree(const std::string &file) { detail::check(mwg_tree_load(h_, file.c_str())); }
    double evalf(const double *r) const {
        double v = 0.0;
        detail::check(mwg_tree_eval(h_, 1, r, &v));
        return v;
    }
    mwg_tree handle() const { return h_; }

   private:
    mwg_tree h_ = nullptr;
};

template <int D>
class ConvolutionOperator;
template <>
class ConvolutionOperator<3> {
   public:
    virtual ~ConvolutionOperator() { mwg_oper_destroy(h_); }
    ConvolutionOperator(const ConvolutionOperator &) = delete;
    ConvolutionOperator &operator=(const ConvolutionOperator &) = delete;
    int size() const {
        int n = 0;
        detail::check(mwg_oper_nterms(h_, &n));
        return n;
    }
    mwg_oper handle() const { return h_; }

   protected:
    ConvolutionOperator() = default;
    mwg_oper h_ = nullptr;
};

// kernel 1/r (CoulombPotential.cpp:85 uses P[rho] directly as J)
class PoissonOperator : public ConvolutionOperator<3> {
   public:
    PoissonOperator(const MultiResolutionAnalysis<3> &mra, double prec) {
        detail::check(mwg_oper_poisson(mra.handle(), prec, &h_));
    }
};

// kernel exp(-mu r)/r (the caller rescales by -1/(2 pi), HelmholtzVector.cpp:81,135)
class HelmholtzOperator : public ConvolutionOperator<3> {
   public:
    HelmholtzOperator(const MultiResolutionAnalysis<3> &mra, double mu, double prec) {
        detail::check(mwg_oper_helmholtz(mra.handle(), mu, prec, &h_));
    }
};

// out = oper * inp, built adaptively to relative precision prec; out must be empty.  The
// MRCPP signature: absPrec = true (absolute precision) is not supported and throws.
template <int D>
inline void apply(double prec, FunctionTree<D> &out, ConvolutionOperator<D> &oper, FunctionTree<D> &inp,
                  int maxIter = -1, bool absPrec = false) {
    if (absPrec) throw Error("mrcpp::apply: absolute precision is not supported");
    detail::check(mwg_apply(prec, out.handle(), oper.handle(), inp.handle(), maxIter, MWG_APPLY_FP64, nullptr));
}

// the same with a contraction mode of this library (no MRCPP counterpart)
enum class ApplyMode : unsigned {
    FP64 = MWG_APPLY_FP64,
    Mixed = MWG_APPLY_MIXED,
    MixedPredict = MWG_APPLY_MIXED | MWG_APPLY_PREDICT,
};
template <int D>
inline void apply(double prec, FunctionTree<D> &out, ConvolutionOperator<D> &oper, FunctionTree<D> &inp,
                  ApplyMode mode, int maxIter = -1) {
    detail::check(mwg_apply(prec, out.handle(), oper.handle(), inp.handle(), maxIter, (unsigned)mode, nullptr));
}

inline double dot(FunctionTree<3> &a, FunctionTree<3> &b) {
    double v = 0.0;
    detail::check(mwg_tree_dot(a.handle(), b.handle(), &v));
    return v;
}

// ---- Gaussian expansions and projection (qmfunction_utils.cpp:88-120 projects analytic
// functions; MRCPP's GaussFunc<3>(alpha, coef, pos, power) = coef * prod_d (x_d - pos_d)^power_d
// * exp(-alpha |x - pos|^2), powers up to 4)
template <int D>
class GaussFunc;
template <>
class GaussFunc<3> {
   public:
    GaussFunc(double alpha, double coef, const std::array<double, 3> &pos = {0.0, 0.0, 0.0},
              const std::array<int, 3> &power = {0, 0, 0})
        : alpha_(alpha), coef_(coef), pos_(pos), power_(power) {
        for (int p : power)
            if (p < 0 || p > 4) throw Error("GaussFunc: powers 0..4 only");
    }
    double alpha() const { return alpha_; }
    double coef() const { return coef_; }
    const std::array<double, 3> &pos() const { return pos_; }
    const std::array<int, 3> &power() const { return power_; }

   private:
    double alpha_, coef_;
    std::array<double, 3> pos_;
    std::array<int, 3> power_;
};
template <int D>
class GaussExp;
template <>
class GaussExp<3> {
   public:
    void append(const GaussFunc<3> &g) { terms_.push_back(g); }
    int size() const { return (int)terms_.size(); }
    const GaussFunc<3> &getFunc(int i) const { return terms_[(size_t)i]; }

   private:
    std::vector<GaussFunc<3>> terms_;
};

// out = the function projected adaptively to relative precision prec
inline void project(double prec, FunctionTree<3> &out, const GaussExp<3> &f) {
    const int n = f.size();
    std::vector<double> coef(n), alpha(n), center(3 * (size_t)n), poly(15 * (size_t)n, 0.0);
    for (int i = 0; i < n; i++) {
        const GaussFunc<3> &g = f.getFunc(i);
        coef[i] = g.coef();
        alpha[i] = g.alpha();
        for (int d = 0; d < 3; d++) {
            center[3 * i + d] = g.pos()[d];
            poly[15 * i + 5 * d + g.power()[d]] = 1.0;
        }
    }
    detail::check(mwg_project_gauss(prec, out.handle(), n, coef.data(), alpha.data(), center.data(), poly.data()));
}
inline void project(double prec, FunctionTree<3> &out, const GaussFunc<3> &g) {
    GaussExp<3> f;
    f.append(g);
    project(prec, out, f);
}

// out = coef * a * b (the union grid of a and b, refined to prec)
inline void multiply(double prec, FunctionTree<3> &out, double coef, FunctionTree<3> &a, FunctionTree<3> &b,
                     int maxIter = -1) {
    if (maxIter >= 0) throw Error("mrcpp::multiply: maxIter is not supported");
    detail::check(mwg_multiply(prec, out.handle(), coef, a.handle(), b.handle()));
}

// out = a * inpA + b * inpB on the union grid of the inputs (exact: the sum of two trees
// is represented without truncation, so prec does not change the result)
inline void add(double prec, FunctionTree<3> &out, double a, FunctionTree<3> &inpA, double b, FunctionTree<3> &inpB,
                int maxIter = -1) {
    (void)prec;
    (void)maxIter;
    detail::check(mwg_add(out.handle(), a, inpA.handle(), b, inpB.handle()));
}

// MRCPP's FunctionTreeVector<3>: (coefficient, tree) terms
template <int D>
using FunctionTreeVector = std::vector<std::tuple<double, FunctionTree<D> *>>;

// out = sum_i c_i f_i (QMFunction::add / qmfunction::linear_combination), by pairwise
// exact union-grid sums
inline void add(double prec, FunctionTree<3> &out, FunctionTreeVector<3> &inp, int maxIter = -1) {
    (void)maxIter;
    if (inp.empty()) throw Error("mrcpp::add: empty FunctionTreeVector");
    if (inp.size() == 1) {
        // c * f: f + 0 f on f's grid, then the scale
        detail::check(mwg_add(out.handle(), std::get<0>(inp[0]), std::get<1>(inp[0])->handle(), 0.0,
                              std::get<1>(inp[0])->handle()));
        return;
    }
    std::vector<std::unique_ptr<mwg_tree_s, int (*)(mwg_tree)>> tmp;
    mwg_tree acc = std::get<1>(inp[0])->handle();
    double cacc = std::get<0>(inp[0]);
    for (size_t i = 1; i < inp.size(); i++) {
        mwg_tree t = (i + 1 == inp.size()) ? out.handle() : nullptr;
        if (!t) {
            mwg_mra m = nullptr;
            detail::check(mwg_tree_mra(acc, &m));
            detail::check(mwg_tree_create(m, &t));
            tmp.emplace_back(t, mwg_tree_destroy);
        }
        detail::check(mwg_add(t, cacc, acc, std::get<0>(inp[i]), std::get<1>(inp[i])->handle()));
        acc = t;
        cacc = 1.0;
    }
    (void)prec;
}

// copy_func: out = inp (grid and coefficients); copy_grid: inp's grid, zero coefficients;
// build_grid: out's grid extended by inp's (out keeps its function)
inline void copy_func(FunctionTree<3> &out, FunctionTree<3> &inp) {
    int64_t n = 0;
    detail::check(mwg_tree_nnodes(inp.handle(), &n));
    mwg_mra m = nullptr;
    detail::check(mwg_tree_mra(inp.handle(), &m));
    int k = 0;
    detail::check(mwg_mra_order(m, &k));
    const size_t K3 = (size_t)(k + 1) * (k + 1) * (k + 1);
    std::vector<int> depth((size_t)n);
    std::vector<int64_t> lr(3 * (size_t)n);
    std::vector<double> coefs((size_t)n * 8 * K3);
    detail::check(mwg_tree_export(inp.handle(), depth.data(), lr.data(), coefs.data()));
    detail::check(mwg_tree_import(out.handle(), n, depth.data(), lr.data(), coefs.data()));
}
inline void copy_grid(FunctionTree<3> &out, FunctionTree<3> &inp) {
    copy_func(out, inp);
    out.setZero();
}
inline void build_grid(FunctionTree<3> &out, FunctionTree<3> &inp) {
    int64_t n = 0;
    detail::check(mwg_tree_nnodes(out.handle(), &n));
    if (n == 0) {
        copy_grid(out, inp);
        return;
    }
    mwg_mra m = nullptr;
    detail::check(mwg_tree_mra(out.handle(), &m));
    FunctionTree<3> tmp(m);
    detail::check(mwg_add(tmp.handle(), 1.0, out.handle(), 0.0, inp.handle()));
    copy_func(out, tmp);
}

}  // namespace mrcpp

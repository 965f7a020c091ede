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
// Errors: MRCPP aborts (MSG_ERROR / MSG_FATAL); this facade throws mrcpp::Error
// carrying mwg_last_error().
#pragma once

#include <memory>
#include <stdexcept>
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

// out = oper * inp, built adaptively to relative precision prec; out must be empty
template <int D>
inline void apply(double prec, FunctionTree<D> &out, ConvolutionOperator<D> &oper, FunctionTree<D> &inp,
                  int maxIter = -1, unsigned flags = MWG_APPLY_FP64) {
    detail::check(mwg_apply(prec, out.handle(), oper.handle(), inp.handle(), maxIter, flags, nullptr));
}

inline double dot(FunctionTree<3> &a, FunctionTree<3> &b) {
    double v = 0.0;
    detail::check(mwg_tree_dot(a.handle(), b.handle(), &v));
    return v;
}

}  // namespace mrcpp

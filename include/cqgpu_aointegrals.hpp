// cqgpu_aointegrals.hpp -- header-only C++ layer with ChronusQ's operator signature.
//
// The body of AOIntegrals::newTwoEContractDirect<Op,T>
// (include/aointegrals/aointegrals_contract.h:24-194) becomes a call to
// cqgpu::newTwoEContractDirect(ctx, X, AX, contractionList, scalingFactors): same argument
// meaning, same accumulate semantics, errors raised as std::runtime_error like the
// reference's CErr (src/cerr/cerr.cpp).  Op is any Eigen-like column-major matrix or map
// with Scalar, data(), rows(), cols() and outerStride() (RealMatrix / RealMap,
// include/global.h:131-140).  Complex Op (ComplexMatrix / ComplexMap,
// src/aointegrals/aointegrals_impl.cpp:31-45) is rejected at compile time.
#ifndef CQGPU_AOINTEGRALS_HPP
#define CQGPU_AOINTEGRALS_HPP

#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "cqgpu.h"

namespace cqgpu {

inline void check(int rc, const cq_ctx_t* ctx) {
  if (rc != CQ_OK) throw std::runtime_error(std::string("cqgpu: ") + cq_last_error(ctx));
}

// RAII owner of a context built from the reference's shell list (libint2::Shell after
// BasisSet::renormShells): l, pure, exponents, coefficients, origin (bohr).
class Context {
 public:
  Context(const std::vector<int>& l, const std::vector<int>& pure, const std::vector<int>& nprim,
          const std::vector<double>& alpha, const std::vector<double>& coef, const std::vector<double>& origin) {
    check(cq_ctx_create(&ctx_, (int)l.size(), l.data(), pure.data(), nprim.data(), alpha.data(), coef.data(),
                        origin.data()),
          nullptr);
  }
  ~Context() { if (ctx_) cq_ctx_destroy(ctx_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  cq_ctx_t* get() const { return ctx_; }
  void setThresholdSchwartz(double tau) { check(cq_set_schwarz_threshold(ctx_, tau), ctx_); }

 private:
  cq_ctx_t* ctx_ = nullptr;
};

// ERI_CONTRACTION_TYPE values (include/aointegrals.h:468-472) map 1:1 on CQ_COULOMB /
// CQ_EXCHANGE; PAIR (2) is not handled by the direct path and raises.
template <typename Op, typename T, typename ContractionType>
void newTwoEContractDirect(cq_ctx_t* ctx, const std::vector<std::reference_wrapper<Op>>& X,
                           std::vector<std::reference_wrapper<Op>>& AX,
                           std::vector<ContractionType>& contractionList, std::vector<T>& scalingFactors) {
  static_assert(std::is_same<typename std::remove_cv<typename Op::Scalar>::type, double>::value,
                "the GPU direct contraction accepts real double matrices only");
  const size_t nMat = X.size();
  if (AX.size() != nMat || contractionList.size() != nMat || scalingFactors.size() != nMat)
    throw std::runtime_error("cqgpu: X, AX, contractionList and scalingFactors differ in length");
  // an empty list contracts nothing (the reference's loops do not run)
  if (nMat == 0) return;
  std::vector<const double*> xs(nMat);
  std::vector<double*> axs(nMat);
  std::vector<int> types(nMat);
  std::vector<double> scales(nMat);
  const int n = (int)X[0].get().rows();
  const int ld = (int)X[0].get().outerStride();
  for (size_t i = 0; i < nMat; ++i) {
    if (X[i].get().rows() != n || X[i].get().cols() != n || AX[i].get().rows() != n || AX[i].get().cols() != n ||
        X[i].get().outerStride() != ld || AX[i].get().outerStride() != ld)
      throw std::runtime_error("cqgpu: all matrices must be n x n with one leading dimension");
    xs[i] = X[i].get().data();
    axs[i] = AX[i].get().data();
    types[i] = static_cast<int>(contractionList[i]);
    scales[i] = static_cast<double>(scalingFactors[i]);
  }
  check(cq_contract_direct(ctx, (int)nMat, xs.data(), axs.data(), types.data(), scales.data(), n, ld), ctx);
}

}  // namespace cqgpu

#endif  // CQGPU_AOINTEGRALS_HPP

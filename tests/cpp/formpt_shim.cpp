// ChronusQ's formPT call (include/singleslater/singleslater_fock.h:30-76) written against the
// header-only layer include/cqgpu_aointegrals.hpp, the way the reference's SingleSlater would
// reach it once AOIntegrals::newTwoEContractDirect forwards there.  The matrices are stand-ins
// for Eigen::Map<RealMatrix> (column-major over caller memory, outer stride ld >= n), the
// lists are std::vector<std::reference_wrapper<Op>> and a local ERI_CONTRACTION_TYPE enum, as
// in include/aointegrals.h:468-480.  Built as a small shared library that the tests call.
#include <string>

#include "cqgpu_aointegrals.hpp"

namespace {

struct MapStandIn {
  typedef double Scalar;
  double* p;
  int n, ld;
  double* data() { return p; }
  const double* data() const { return p; }
  int rows() const { return n; }
  int cols() const { return n; }
  int outerStride() const { return ld; }
};

enum ERI_CONTRACTION_TYPE { COULOMB, EXCHANGE, PAIR };

thread_local std::string g_err;

void zero(MapStandIn& m) {
  for (int j = 0; j < m.n; ++j)
    for (int i = 0; i < m.n; ++i) m.p[(long)j * m.ld + i] = 0.0;
}

}  // namespace

extern "C" {

const char* shim_last_error() { return g_err.c_str(); }

// formPT with the reference's lists: X = {P, P (, Mz)}, AX = {PT, PT (, PTMz)},
// {COULOMB, EXCHANGE (, EXCHANGE)}, {1, xHF (, xHF)}; PT (and PTMz) are zeroed first, as
// formPT does.  Mz == nullptr: closed shell.  Returns 0, or -1 with shim_last_error() set.
int shim_form_pt(cq_ctx_t* ctx, int n, int ld, double* P, double* Mz, double* PT, double* PTMz, double xHF) {
  try {
    MapStandIn mP{P, n, ld}, mPT{PT, n, ld};
    MapStandIn mMz{Mz, n, ld}, mPTMz{PTMz, n, ld};
    std::vector<std::reference_wrapper<MapStandIn>> mats, ax;
    std::vector<ERI_CONTRACTION_TYPE> contList;
    std::vector<double> scalingFactors;
    mats.emplace_back(mP);
    mats.emplace_back(mP);
    ax.emplace_back(mPT);
    ax.emplace_back(mPT);
    contList.push_back(COULOMB);
    contList.push_back(EXCHANGE);
    scalingFactors.push_back(1.0);
    scalingFactors.push_back(xHF);
    zero(mPT);
    if (Mz) {
      mats.emplace_back(mMz);
      ax.emplace_back(mPTMz);
      contList.push_back(EXCHANGE);
      scalingFactors.push_back(xHF);
      zero(mPTMz);
    }
    cqgpu::newTwoEContractDirect(ctx, mats, ax, contList, scalingFactors);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// One contraction of a caller-chosen type code (PAIR = 2 must raise) or an empty list (nmat 0
// contracts nothing).  Returns 0, or -1 with shim_last_error() set.
int shim_contract_type(cq_ctx_t* ctx, int n, int ld, double* X, double* AX, int type, int nmat) {
  try {
    MapStandIn mX{X, n, ld}, mAX{AX, n, ld};
    std::vector<std::reference_wrapper<MapStandIn>> mats, ax;
    std::vector<ERI_CONTRACTION_TYPE> contList;
    std::vector<double> scalingFactors;
    if (nmat > 0) {
      mats.emplace_back(mX);
      ax.emplace_back(mAX);
      contList.push_back(static_cast<ERI_CONTRACTION_TYPE>(type));
      scalingFactors.push_back(1.0);
    }
    cqgpu::newTwoEContractDirect(ctx, mats, ax, contList, scalingFactors);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

}  // extern "C"

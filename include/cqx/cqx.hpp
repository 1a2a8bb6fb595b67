// C++ adapter over the cqx C ABI, shaped like the ChronusQ classes it
// replaces on the DIRECT path:
//
//   TwoBodyContraction<T>             include/particleintegrals/twopints.hpp:59-75
//   TPIContractions<MatsT,IntsT>      include/particleintegrals/twopints.hpp:172-260
//   GTODirectTPIContraction           include/particleintegrals/twopints/gtodirecttpi.hpp:143-228
//   PauliSpinorMatrices, SpinScatter  include/matrix/paulispinormatrices.hpp:30-392,
//                                     src/cqlinalg/blasutil.cxx:71-189
//   FockBuilder::formGD / formRawGDInBatches / formRawGDSCRSizePerBatch /
//   formFock                          include/fockbuilder.hpp:66-77, fockbuilder/impl.hpp:75-319
//   SingleSlater::computeEnergy       include/singleslater/quantum.hpp:147-212
//
// Header-only; link against libcqx.so.  A non-zero status from the C ABI is
// raised as std::runtime_error("FATAL: ..."), the exception ChronusQ's CErr
// throws (include/cerr.hpp:33-51).  Without CQ_ENABLE_MPI, MPI_Comm is the
// reference's own single-rank stand-in (include/util/mpi.hpp:33-67).
#pragma once

#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#ifdef CQ_ENABLE_MPI
#include <mpi.h>
#endif

#include "../cqx.h"

namespace cqx_cpp {

using dcomplex = std::complex<double>;

#ifndef CQ_ENABLE_MPI
// the reference's MPI_Comm of a build without MPI: one rank
struct MPI_Comm {
  int internal = 0;
  MPI_Comm(int c) : internal(c) {}
  MPI_Comm() : MPI_Comm(0) {}
  static int size() { return 1; }
  static int rank() { return 0; }
  operator int() const { return internal; }
};
static const MPI_Comm MPI_COMM_WORLD{0};
#else
using ::MPI_Comm;
#endif

inline void CErr(const std::string &msg) { throw std::runtime_error("FATAL: " + msg); }

inline void check(int status, const cqx_ctx *ctx = nullptr) {
  if (status != CQX_OK) {
    const char *m = cqx_last_error(ctx);
    CErr(m ? m : "cqx error");
  }
}

enum TWOBODY_CONTRACTION_TYPE { COULOMB = CQX_COULOMB, EXCHANGE = CQX_EXCHANGE };

// {X, AX, HER, contType}: X and AX are NB x NB column-major buffers
template <typename T>
struct TwoBodyContraction {
  T *X;
  T *AX;
  bool HER;
  TWOBODY_CONTRACTION_TYPE contType;
};

// EMPerturbation (include/fields.hpp:124-195), dipole fields only
enum EMFieldTyp { Electric, Magnetic };
struct EMPerturbation {
  std::vector<std::pair<EMFieldTyp, std::array<double, 3>>> fields;
  void addField(EMFieldTyp t, const std::array<double, 3> &amp) { fields.emplace_back(t, amp); }
  bool hasField(EMFieldTyp t) const {
    for (auto &f : fields)
      if (f.first == t) return true;
    return false;
  }
  std::array<double, 3> getDipoleAmp(EMFieldTyp t) const {
    std::array<double, 3> a{0.0, 0.0, 0.0};
    for (auto &f : fields)
      if (f.first == t)
        for (int i = 0; i < 3; ++i) a[i] += f.second[i];
    return a;
  }
};

// Shell set built by the engine's G94 reader (or from arrays that a caller's
// own BasisSet already holds, normalised the libint2 way).
class BasisSet {
 public:
  BasisSet(const std::string &g94, const std::vector<int> &Z, const std::vector<double> &xyz,
           bool forceCart = false) {
    check(cqx_basis_from_g94(g94.c_str(), (int)Z.size(), Z.data(), xyz.data(), forceCart, &b_));
    check(cqx_basis_dims(b_, &nShell, &nBasis, &maxPrim, &maxL));
  }
  ~BasisSet() { cqx_basis_destroy(b_); }
  BasisSet(const BasisSet &) = delete;
  BasisSet &operator=(const BasisSet &) = delete;
  const cqx_basis *handle() const { return b_; }
  int nShell = 0, nBasis = 0, maxPrim = 0, maxL = 0;

 private:
  cqx_basis *b_ = nullptr;
};

// DirectTPI<double>: basis + Schwarz threshold, device-resident pair data
class DirectTPI {
 public:
  DirectTPI(const BasisSet &bs, double threshSchwarz = 1e-12, int device = -1)
      : nBasis_(bs.nBasis), thresh_(threshSchwarz) {
    check(cqx_ctx_create(bs.handle(), threshSchwarz, device, &ctx_));
  }
  ~DirectTPI() { cqx_ctx_destroy(ctx_); }
  DirectTPI(const DirectTPI &) = delete;
  DirectTPI &operator=(const DirectTPI &) = delete;
  cqx_ctx *handle() const { return ctx_; }
  int nBasis() const { return nBasis_; }
  double threshSchwarz() const { return thresh_; }
  // engine options (cqx_set_option: "graph", "fp32_cut", "list_cap", ...)
  void setOption(const std::string &name, double value) { check(cqx_set_option(ctx_, name.c_str(), value), ctx_); }
  // statistics of the last contraction (quartets, timings, graph replay)
  cqx_stats lastStats() const {
    cqx_stats st{};
    check(cqx_last_stats(ctx_, &st), ctx_);
    return st;
  }

 private:
  cqx_ctx *ctx_ = nullptr;
  int nBasis_;
  double thresh_;
};

// The virtual interface the reference's callers see (twopints.hpp:172-260)
template <typename MatsT, typename IntsT>
class TPIContractions {
 public:
  virtual ~TPIContractions() = default;
  virtual void twoBodyContract(MPI_Comm, const bool, std::vector<TwoBodyContraction<MatsT>> &,
                               EMPerturbation &) const = 0;
  void twoBodyContract(MPI_Comm comm, const bool screen, std::vector<TwoBodyContraction<MatsT>> &contList) const {
    EMPerturbation pert;
    twoBodyContract(comm, screen, contList, pert);
  }
  void twoBodyContract(MPI_Comm comm, std::vector<TwoBodyContraction<MatsT>> &contList, EMPerturbation &pert) const {
    twoBodyContract(comm, true, contList, pert);
  }
  void twoBodyContract(MPI_Comm comm, std::vector<TwoBodyContraction<MatsT>> &contList) const {
    twoBodyContract(comm, true, contList);
  }
  // device workspace (in MatsT elements) of one contraction of each kind
  // listed; the reference's directScaffoldNewSCRSize for the DIRECT class
  virtual size_t scratchSize(const std::vector<TWOBODY_CONTRACTION_TYPE> &types, bool HerDen) const = 0;
  bool contractSecond = false;
  bool printContractionTiming = false;
};

// GPU replacement of GTODirectTPIContraction<MatsT, IntsT> for host matrices
// (MatsT = double or std::complex<double>, IntsT = double)
template <typename MatsT, typename IntsT = double>
class GpuDirectTPIContraction : public TPIContractions<MatsT, IntsT> {
  static_assert(std::is_same<MatsT, double>::value || std::is_same<MatsT, dcomplex>::value,
                "MatsT must be double or std::complex<double>");
  static_assert(std::is_same<IntsT, double>::value, "the engine evaluates real integrals (IntsT = double)");

 public:
  explicit GpuDirectTPIContraction(std::shared_ptr<DirectTPI> tpi) : tpi_(std::move(tpi)) {}
  using TPIContractions<MatsT, IntsT>::twoBodyContract;
  // the perturbation is not read, as the reference's DIRECT contraction
  // ignores it (gtodirecttpi.hpp:199-206); results are complete on the
  // calling rank
  void twoBodyContract(MPI_Comm, const bool screen, std::vector<TwoBodyContraction<MatsT>> &list,
                       EMPerturbation &) const override {
    if (list.empty()) CErr("Nothing specified to compute");
    std::vector<cqx_contraction> c(list.size());
    for (size_t i = 0; i < list.size(); ++i)
      c[i] = {list[i].X, list[i].AX, list[i].HER ? 1 : 0, (int)list[i].contType};
    check(cqx_two_body_contract(tpi_->handle(), screen, (int)c.size(), c.data(), isComplex(), CQX_HOST),
          tpi_->handle());
  }
  size_t scratchSize(const std::vector<TWOBODY_CONTRACTION_TYPE> &types, bool HerDen) const override {
    std::vector<cqx_contraction> c(types.size());
    for (size_t i = 0; i < types.size(); ++i) c[i] = {nullptr, nullptr, HerDen ? 1 : 0, (int)types[i]};
    size_t bytes = 0;
    check(cqx_scratch_size(tpi_->handle(), (int)c.size(), c.data(), isComplex(), &bytes), tpi_->handle());
    return (bytes + sizeof(MatsT) - 1) / sizeof(MatsT);
  }
  // directScaffoldNewSCRSize (direct.hpp:2125-2180): one Hermitian Coulomb
  // contraction, in MatsT elements
  size_t directScaffoldNewSCRSize() const { return scratchSize({COULOMB}, true); }
  const std::shared_ptr<DirectTPI> &ints() const { return tpi_; }

 private:
  static int isComplex() { return std::is_same<MatsT, dcomplex>::value ? 1 : 0; }
  std::shared_ptr<DirectTPI> tpi_;
};

// GradContractions<double, double> for the DIRECT algorithm
// (particleintegrals/gradints/direct.hpp:31-76): cList[g][m], g = 3 atom +
// xyz, one density per matrix m shared by every g (FockBuilder::getGDGrad).
class GpuDirectGradContraction {
 public:
  explicit GpuDirectGradContraction(std::shared_ptr<DirectTPI> tpi) : tpi_(std::move(tpi)) {}
  void gradTwoBodyContract(MPI_Comm, std::vector<std::vector<TwoBodyContraction<double>>> &cList) const {
    if (cList.empty() || cList[0].empty()) CErr("Nothing specified to compute");
    const size_t nmat = cList[0].size();
    std::vector<cqx_contraction> c;
    c.reserve(cList.size() * nmat);
    for (auto &row : cList) {
      if (row.size() != nmat) CErr("ragged gradient contraction list");
      for (auto &x : row) c.push_back({x.X, x.AX, x.HER ? 1 : 0, (int)x.contType});
    }
    check(cqx_grad_two_body_contract(tpi_->handle(), (int)cList.size(), (int)nmat, c.data(), CQX_HOST),
          tpi_->handle());
  }

 private:
  std::shared_ptr<DirectTPI> tpi_;
};

// cqmatrix::PauliSpinorMatrices<MatsT> (paulispinormatrices.hpp:30-392):
// the S component and, where present, Z (hasZ) and Y, X (hasXY), NB x NB
// column-major each.
template <typename MatsT>
class PauliSpinorMatrices {
 public:
  PauliSpinorMatrices(size_t N, bool hasXY = true, bool hasZ = true)
      : N_(N), hasZ_(hasZ || hasXY), hasXY_(hasXY), comp_(1 + (hasZ_ ? 1 : 0) + (hasXY_ ? 2 : 0),
                                                          std::vector<MatsT>(N * N, MatsT(0.))) {}
  size_t dimension() const { return N_; }
  bool hasZ() const { return hasZ_; }
  bool hasXY() const { return hasXY_; }
  size_t nComponent() const { return comp_.size(); }
  MatsT *S() { return comp_[0].data(); }
  MatsT *Z() { return hasZ_ ? comp_[1].data() : missing("Z"); }
  MatsT *Y() { return hasXY_ ? comp_[2].data() : missing("Y"); }
  MatsT *X() { return hasXY_ ? comp_[3].data() : missing("X"); }
  const MatsT *component(size_t c) const { return comp_.at(c).data(); }
  MatsT *component(size_t c) { return comp_.at(c).data(); }
  void clear() {
    for (auto &v : comp_) std::fill(v.begin(), v.end(), MatsT(0.));
  }
  // this += s * other, componentwise over the components both hold
  void axpy(MatsT s, const PauliSpinorMatrices &o) {
    for (size_t c = 0; c < std::min(comp_.size(), o.comp_.size()); ++c)
      for (size_t k = 0; k < comp_[c].size(); ++k) comp_[c][k] += s * o.comp_[c][k];
  }
  void scale(MatsT s) {
    for (auto &v : comp_)
      for (auto &x : v) x *= s;
  }
  // spinScatter: a 2N x 2N spinor matrix to its components (SpinScatter on
  // the context's GPU)
  static PauliSpinorMatrices spinScatter(const DirectTPI &tpi, const MatsT *A, size_t N, bool hasXY = true,
                                         bool hasZ = true, bool zeroABBA = false, bool BBeqAA = false) {
    PauliSpinorMatrices P(N, hasXY, hasZ);
    const int flags = (zeroABBA ? CQX_SPIN_ZERO_ABBA : 0) | (BBeqAA ? CQX_SPIN_BB_EQ_AA : 0);
    check(cqx_spin_scatter(tpi.handle(), (int)N, A, (int)(2 * N), P.S(), P.hasZ_ ? P.Z() : nullptr,
                           P.hasXY_ ? P.Y() : nullptr, P.hasXY_ ? P.X() : nullptr, isComplex(), flags, CQX_HOST),
          tpi.handle());
    return P;
  }
  // spinGather: the 2N x 2N spinor matrix of these components
  std::vector<MatsT> spinGather(const DirectTPI &tpi) const {
    std::vector<MatsT> A(4 * N_ * N_);
    const int flags = (hasZ_ ? 0 : CQX_SPIN_ZERO_Z) | (hasXY_ ? 0 : CQX_SPIN_ZERO_XY);
    check(cqx_spin_gather(tpi.handle(), (int)N_, A.data(), (int)(2 * N_), component(0),
                          hasZ_ ? component(1) : nullptr, hasXY_ ? component(2) : nullptr,
                          hasXY_ ? component(3) : nullptr, isComplex(), flags, CQX_HOST),
          tpi.handle());
    return A;
  }
  // hasNaN (paulispinormatrices.hpp:383-388), on the context's GPU
  bool hasNaN(const DirectTPI &tpi) const {
    for (auto &v : comp_) {
      size_t n = 0;
      check(cqx_count_nonfinite(tpi.handle(), reinterpret_cast<const double *>(v.data()),
                                v.size() * sizeof(MatsT) / sizeof(double), CQX_HOST, &n),
            tpi.handle());
      if (n) return true;
    }
    return false;
  }

 private:
  static int isComplex() { return std::is_same<MatsT, dcomplex>::value ? 1 : 0; }
  MatsT *missing(const char *c) {
    CErr(std::string("PauliSpinorMatrices has no ") + c + " component");
    return nullptr;
  }
  size_t N_;
  bool hasZ_, hasXY_;
  std::vector<std::vector<MatsT>> comp_;
};

// The parts of SingleSlater<MatsT,IntsT> that FockBuilder reads and writes
// (onePDM, deltaOnePDM, coulombMatrix, exchangeMatrix, twoeH, coreH,
// fockMatrix, GDDur, nC, TPI, comm) and computeEnergy.
template <typename MatsT, typename IntsT = double>
struct SingleSlater {
  using PSM = PauliSpinorMatrices<MatsT>;
  SingleSlater(std::shared_ptr<TPIContractions<MatsT, IntsT>> tpi, size_t NB, int nC_ = 1)
      : nC(nC_), TPI(std::move(tpi)) {
    const bool two = nC > 1;
    onePDM = std::make_shared<PSM>(NB, two, two);
    deltaOnePDM = std::make_shared<PSM>(NB, two, two);
    coulombMatrix = std::make_shared<std::vector<MatsT>>(NB * NB, MatsT(0.));
    exchangeMatrix = std::make_shared<PSM>(NB, two, two);
    twoeH = std::make_shared<PSM>(NB, two, two);
    coreH = std::make_shared<PSM>(NB, two, two);
    fockMatrix = std::make_shared<PSM>(NB, two, two);
  }
  size_t nBasis() const { return onePDM->dimension(); }
  // SingleSlater::computeEnergy (quantum.hpp:147-212) on the GPU of tpi:
  // OBEnergy = 1/2 sum_c Re<D_c, H_c>, MBEnergy = 1/4 sum_c Re<D_c, G_c>
  void computeEnergy(const DirectTPI &tpi, double nucRepEnergy = 0.0) {
    const int nc = (int)onePDM->nComponent();
    std::vector<const void *> D(nc), H(nc), G(nc);
    for (int c = 0; c < nc; ++c) {
      D[c] = onePDM->component(c);
      H[c] = c < (int)coreH->nComponent() ? coreH->component(c) : nullptr;
      G[c] = c < (int)twoeH->nComponent() ? twoeH->component(c) : nullptr;
    }
    double out[3];
    check(cqx_compute_energy(tpi.handle(), (int)nBasis(), nc, D.data(), H.data(), G.data(),
                             std::is_same<MatsT, dcomplex>::value ? 1 : 0, CQX_HOST, nucRepEnergy, out),
          tpi.handle());
    OBEnergy = out[0];
    MBEnergy = out[1];
    totalEnergy = out[2];
  }

  int nC = 1;
  int particleCharge = -1;  // electrons
  MPI_Comm comm{0};
  std::shared_ptr<TPIContractions<MatsT, IntsT>> TPI;
  std::shared_ptr<PSM> onePDM, deltaOnePDM, exchangeMatrix, twoeH, coreH, fockMatrix;
  std::shared_ptr<std::vector<MatsT>> coulombMatrix;
  // length-gauge electric dipole integrals (x, y, z), for formFock under an
  // electric field
  std::vector<std::vector<MatsT>> lenElectric;
  double GDDur = 0.0, OBEnergy = 0.0, MBEnergy = 0.0, totalEnergy = 0.0;
};

template <typename MatsT, typename IntsT = double>
class FockBuilder {
 public:
  using PSM = PauliSpinorMatrices<MatsT>;
  using PSMList = std::vector<std::shared_ptr<PSM>>;
  virtual ~FockBuilder() = default;

  // impl.hpp:75-102: G[D] into ss.twoeH / ss.exchangeMatrix /
  // ss.coulombMatrix; with increment, G and K accumulate the build of
  // ss.deltaOnePDM and coulombMatrix receives J[deltaOnePDM]
  virtual void formGD(SingleSlater<MatsT, IntsT> &ss, EMPerturbation &pert, bool increment = false,
                      double xHFX = 1., bool HerDen = true) {
    PSMList onePDMs{increment ? ss.deltaOnePDM : ss.onePDM};
    PSMList exchangeMatrices{ss.exchangeMatrix}, twoeHs{ss.twoeH};
    PSMList coulombMatrices{std::make_shared<PSM>(ss.nBasis(), false, false)};
    formRawGDInBatches(ss, pert, increment, xHFX, HerDen, onePDMs, coulombMatrices, exchangeMatrices, twoeHs);
    std::copy(coulombMatrices[0]->S(), coulombMatrices[0]->S() + ss.nBasis() * ss.nBasis(),
              ss.coulombMatrix->begin());
  }

  // impl.hpp:109-244: every batch's contractions in one twoBodyContract call
  virtual void formRawGDInBatches(SingleSlater<MatsT, IntsT> &ss, EMPerturbation &pert, bool increment, double xHFX,
                                  bool HerDen, PSMList &onePDMs, PSMList &coulombMatrices, PSMList &exchangeMatrices,
                                  PSMList &twoeHs) {
    const size_t nBatch = onePDMs.size();
    const bool computeCoulomb = coulombMatrices.size() > 0;
    const bool computeExchange = std::abs(xHFX) > 1e-12 && exchangeMatrices.size() > 0;
    const bool computeTwoeHs = twoeHs.size() > 0;
    if (!computeCoulomb && !computeExchange && !computeTwoeHs)
      CErr("Nothing specified to compute in FockBuilder::formRawGDInBatches");
    if (ss.nC == 4) CErr("4C formGD is implemented in class FourCompFock.");
    if (!increment)
      for (size_t i = 0; i < nBatch; ++i) {
        if (computeCoulomb) coulombMatrices[i]->clear();
        if (computeExchange) exchangeMatrices[i]->clear();
        if (computeTwoeHs) twoeHs[i]->clear();
      }
    std::vector<TwoBodyContraction<MatsT>> contract;
    if (computeCoulomb || computeTwoeHs) {
      auto &cc = computeCoulomb ? coulombMatrices : twoeHs;
      for (size_t i = 0; i < nBatch; ++i) contract.push_back({onePDMs[i]->S(), cc[i]->S(), HerDen, COULOMB});
    }
    if (computeExchange)
      for (size_t i = 0; i < nBatch; ++i) {
        contract.push_back({onePDMs[i]->S(), exchangeMatrices[i]->S(), HerDen, EXCHANGE});
        if (exchangeMatrices[i]->hasZ() && ss.particleCharge < 0)
          contract.push_back({onePDMs[i]->Z(), exchangeMatrices[i]->Z(), HerDen, EXCHANGE});
        if (exchangeMatrices[i]->hasXY()) {
          contract.push_back({onePDMs[i]->Y(), exchangeMatrices[i]->Y(), HerDen, EXCHANGE});
          contract.push_back({onePDMs[i]->X(), exchangeMatrices[i]->X(), HerDen, EXCHANGE});
        }
      }
    ss.TPI->twoBodyContract(ss.comm, contract, pert);
    if (computeTwoeHs)
      for (size_t i = 0; i < nBatch; ++i) {
        if (computeCoulomb)
          twoeHs[i]->axpy(MatsT(2.0), *coulombMatrices[i]);
        else
          for (size_t k = 0; k < ss.nBasis() * ss.nBasis(); ++k) twoeHs[i]->S()[k] *= MatsT(2.0);
        if (computeExchange) twoeHs[i]->axpy(MatsT(-xHFX), *exchangeMatrices[i]);
      }
  }

  // impl.hpp:252-269: device workspace of one batch, in MatsT elements
  // (J, and K on every component the density holds)
  virtual size_t formRawGDSCRSizePerBatch(SingleSlater<MatsT, IntsT> &ss, bool computeExchange, bool HerDen) const {
    std::vector<TWOBODY_CONTRACTION_TYPE> types{COULOMB};
    if (computeExchange)
      for (size_t c = 0; c < ss.onePDM->nComponent(); ++c) types.push_back(EXCHANGE);
    return ss.TPI->scratchSize(types, HerDen);
  }

  // impl.hpp:280-319: formGD (timed into GDDur), F = H + G, minus
  // 2 E_i mu_i for an electric dipole field
  virtual void formFock(SingleSlater<MatsT, IntsT> &ss, EMPerturbation &pert, bool increment = false,
                        double xHFX = 1.) {
    const auto t0 = std::chrono::steady_clock::now();
    formGD(ss, pert, increment, xHFX);
    ss.GDDur = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const size_t n2 = ss.nBasis() * ss.nBasis();
    for (size_t c = 0; c < ss.fockMatrix->nComponent(); ++c)
      for (size_t k = 0; k < n2; ++k)
        ss.fockMatrix->component(c)[k] = ss.coreH->component(c)[k] + ss.twoeH->component(c)[k];
    if (pert.hasField(Electric)) {
      if (ss.lenElectric.size() != 3) CErr("formFock: an electric field needs the length-gauge dipole integrals");
      const auto amp = pert.getDipoleAmp(Electric);
      for (int i = 0; i < 3; ++i)
        for (size_t k = 0; k < n2; ++k) ss.fockMatrix->S()[k] -= MatsT(2.0 * amp[i]) * ss.lenElectric[i][k];
    }
  }
};

}  // namespace cqx_cpp

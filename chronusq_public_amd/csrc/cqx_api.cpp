// C-ABI of the MI355X direct J/K engine: context set-up (shell pairs,
// primitive pairs, Schwarz bounds on the GPU) and the two-body contraction
// driver that replaces GTODirectTPIContraction::directScaffoldNew
// (include/particleintegrals/contract/direct.hpp:1412-2123).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "cqx.h"
#include "basis.hpp"
#include "device_types.hpp"
#include "kernels.hpp"
#include "onebody.hpp"
#include "ctx_access.hpp"
#include "limits.hpp"

namespace {

const double kRysTable[] = {
#include "rys_table_data.inc"
};

thread_local std::string g_error;  // error of the last failing call of this thread

}  // namespace

// The workspace of cqx_two_body_contract_noalloc: one fixed region per GPU,
// loaded with the code object and held until the process ends.  In host
// code `scratch` names a host-side stand-in, never device memory: the
// region's device address is looked up once per device (scratch_region)
// and only that address is handed to kernels and copies.
static __device__ __align__(256) char scratch[CQX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current
// device; resolved on the first call per device (from cqx_ctx_create).
void *scratch_region(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

}  // namespace

struct cqx_basis {
  cqx::Basis b;
};

using cqx::PairRec;
using cqx::PrimPair;
using cqx::Quartet;

namespace {

constexpr int kNumClass = (cqx::kMaxLDev + 1) * (cqx::kMaxLDev + 2) / 2;  // (la, lb), la >= lb
inline int class_of(int la, int lb) { return la * (la + 1) / 2 + lb; }
inline void class_ls(int c, int &la, int &lb) {
  la = 0;
  while ((la + 1) * (la + 2) / 2 <= c) ++la;
  lb = c - la * (la + 1) / 2;
}

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Sizes the device kernels cannot take (index packing, launch limits, a
// scratch region too small): reported as CQX_ERR_UNSUPPORTED.
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};


#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));               \
  } while (0)

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void ensure(size_t m) {
    if (m <= n) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (m == 0) return;
    HIPCHK(hipMalloc((void **)&p, sizeof(T) * m));
    n = m;
  }
  void upload(const std::vector<T> &v, hipStream_t s) {
    ensure(v.size());
    if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

struct cqx_ctx {
  cqx::Basis basis;
  double thresh = 1e-12;
  double prim_thresh = 1e-18;
  int fast_max_nc = 27;   // classes up to this Cartesian size use the lane-group kernels
  int use_wave = 1;       // templated wave-per-quartet kernels for the other s/p/d classes
  int fast_lanes = 0;     // lanes per quartet of the lane-group kernels (0: per launch)
  int lds_max = 65536;    // LDS bytes a block may use on the device
  bool autotune_pending = false;  // time the lane counts of each class on the next call
  std::map<int, int> tuned_lanes; // (class pair, precision) -> lanes per quartet, or quartets per wave
  DevBuf<double> d_acc_tune;      // scratch accumulators of the timing launches
  DevBuf<cqx::MolPlanes> d_molplanes_tune;
  double fp32_cut = 0.0;  // quartets with Q_ab Q_cd below this are evaluated in FP32
  long long list_cap = 1ll << 28;  // quartet-list entries per batch (8 B each, x2 when mixed)
  int device = 0;
  hipStream_t stream = nullptr;
  void *scratch = nullptr;  // this device's allocation-free workspace region
  std::string err;

  int ns = 0, nb = 0, nbc = 0;
  int nmol = 1;                                // molecules of a batched context
  std::vector<std::pair<int, int>> ref_pairs;  // (s1 >= s2) reference order
  std::vector<std::vector<PairRec>> cls;       // per class, sorted by (molecule, Q desc)
  std::vector<std::vector<int>> seg;           // per class: molecule m's pairs [seg[m], seg[m + 1])
  std::vector<long long> mol_boff;             // molecule m's NS_m^2 block-norm offset
  std::vector<PrimPair> prims;
  std::vector<double> prim_alpha;  // exponent of the first shell of each primitive pair (gradients)
  std::vector<cqx::PrimC32> primc32;  // the same pairs as the FP32 lists read them
  std::vector<double> schwarz;                 // NS x NS

  DevBuf<PairRec> d_pairs[kNumClass];
  DevBuf<cqx::PairScr> d_scr[kNumClass];
  DevBuf<double> d_pn[kNumClass];  // per build: each pair's own block norms
  DevBuf<PrimPair> d_prims;
  DevBuf<double> d_prim_alpha;
  DevBuf<cqx::PrimC32> d_primc32;
  // nuclear-gradient contraction: shell atoms, Cartesian densities,
  // accumulators and outputs
  DevBuf<int> d_shell_atom;
  DevBuf<double> d_gX, d_gacc, d_gAX;
  DevBuf<double> d_rys, d_sph;
  DevBuf<float> d_rys32;
  DevBuf<int> d_cart_shell, d_cart_comp, d_sh_bf, d_sh_bfc, d_sh_l, d_sh_pure, d_sh_nf,
      d_bf_shell, d_bf_comp;
  cqx::CartMap cmap{};

  // per-call workspace
  DevBuf<double> d_X, d_AX, d_Bm, d_Bg, d_planes, d_acc;
  DevBuf<int> d_her;
  DevBuf<long long> d_cand, d_mol_boff;
  DevBuf<int> d_mol_ns, d_ket_lo, d_seg;
  DevBuf<char> d_scan_tmp;
  DevBuf<cqx::ListLaunch> d_lists;
  DevBuf<long long> d_lstart;
  DevBuf<double> d_molmax;
  DevBuf<cqx::MolPlanes> d_molplanes;
  DevBuf<Quartet> d_quart;
  DevBuf<unsigned long long> d_count;
  DevBuf<double> d_psum;
  DevBuf<double> d_store;
  DevBuf<long long> d_store_off;
  DevBuf<long long> d_tidx, d_tot;  // candidate totals of an eager build, gathered for one copy
  // The workspace of the contraction group in flight: views of the buffers
  // above (allocated on demand) or of a fixed scratch region
  struct Ws {
    double *X, *AX, *Bm, *Bg, *molmax, *planes, *acc;
    double *pn[kNumClass];
    int *her, *ket_lo;
    long long *cand, *tidx, *tot, *lstart;
    void *scan;
    size_t scan_bytes;
    cqx::MolPlanes *molplanes;
    Quartet *quart;
    unsigned long long *count;
    double *psum;
    cqx::ListLaunch *lists;
  } w{};
  size_t ws_scratch_peak = 0;  // bytes of the scratch region the last call used
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
  // Concurrent class-pair launches (option "streams"): a build, eager or
  // recorded, deals its ERI launches round-robin over the context stream and
  // nstreams - 1 side streams, forked from and joined back into the context
  // stream, so that one launch's last waves overlap the next launch's first
  // (the launches only add into the accumulators, with atomics)
  static constexpr int kMaxSide = 3;
  int nstreams = 4;
  hipStream_t side[kMaxSide] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kMaxSide] = {};

  // Captured build (option "graph"): the launches of one build, from the
  // shell-block norms to the accumulation into AX, recorded once into a HIP
  // graph and replayed by the next builds with the same contraction
  // signature.  Every list and ERI launch reads its count from the device;
  // the capacities come from the build that recorded the graph.
  struct GraphSlot { int la, lb, lc, ld, f32; };
  int use_graph = 0;
  double graph_margin = 1.25;      // candidate capacity / candidates of the recording build
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  std::string graph_key;           // contraction signature of the recording build
  int g_nmat = 0, g_nJ = 0, g_nK = 0;
  bool g_pure_d = false;
  std::vector<size_t> g_xoff, g_mnb2;  // per molecule: offset and size of its matrices
  std::vector<long long> g_cap;        // per class pair: candidate capacity
  std::vector<long long> g_lcap;       // per ERI slot: list capacity
  std::vector<GraphSlot> g_slots;      // per ERI slot (class pair, precision); la < 0: not launched
  DevBuf<cqx::ListLaunch> d_glists;
  DevBuf<long long> d_glstart, d_gidx, d_gchk;

  // Tiles (option "tiles", default 0): the lane-group class pairs of a
  // build whose planes are all Hermitian go through tile lists instead of
  // quartet lists (1); 2 does so only when the build has at most one
  // Hermitian J and one Hermitian K plane (a real RHF J / K pair).  Per
  // class, its member array (pair indices grouped by molecule and first
  // shell); per tile class pair, its sub-blocks, a range of d_sub.
  int use_tiles = 0;
  // lane-group classes that take tiles (bit k: k-th of ss|ss, ps|ss, ps|ps,
  // pp|ss, pp|ps, ds|ss, ds|ps, dp|ss); (ss|ss) and (dp|ss) are faster
  // through quartet lists
  int tile_classes = 0x7E;
  int rys_lds = 1;  // lane-group kernels stage their Rys table section in LDS
  DevBuf<int> d_members[kNumClass];
  struct TileCP { int cb, ck; long long off, n; };
  std::vector<TileCP> tile_cps;
  DevBuf<cqx::SubBlock> d_sub;
  long long nsub_total = 0;
  DevBuf<cqx::Tile> d_tiles;                    // per build: the tile lists
  DevBuf<unsigned long long> d_tcount;          // per tile class pair: tiles and quartets, FP64 / FP32
  DevBuf<double> d_tpsum;                       // per tile class pair: sums of Ka Kb Kc Kd, FP64 / FP32
  DevBuf<cqx::TileListLaunch> d_tll;
  DevBuf<long long> d_tstart;

  cqx_stats stats{};
};

namespace {

int fail(cqx_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  g_error = msg;
  return code;
}

// A captured build is dropped whenever the workspace it refers to may be
// reallocated (any build launched eagerly, options, gradient contractions).
void drop_graph(cqx_ctx *c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_key.clear();
}

// contraction signature: what the recorded launches depend on besides the
// densities themselves
std::string contraction_key(int screen, int is_complex, int ncon, const cqx_contraction *list) {
  std::string k = std::to_string(screen) + ":" + std::to_string(is_complex) + ":";
  for (int i = 0; i < ncon; ++i) k += (char)('0' + 2 * list[i].type + (list[i].her ? 1 : 0));
  return k;
}

// Cartesian -> pure transform of a contracted Cartesian quartet (host side,
// used for the Schwarz diagonals and the ERI inspection entry point).
std::vector<double> cart_to_pure(const cqx::Basis &B, const int sh[4], const double *cart) {
  int l[4], nc[4], nf[4];
  for (int k = 0; k < 4; ++k) {
    l[k] = B.shells[sh[k]].l;
    nc[k] = cqx::ncart(l[k]);
    nf[k] = B.shells[sh[k]].nfunc();
  }
  std::vector<double> cur(cart, cart + (size_t)nc[0] * nc[1] * nc[2] * nc[3]);
  int dims[4] = {nc[0], nc[1], nc[2], nc[3]};
  for (int k = 0; k < 4; ++k) {
    if (!B.shells[sh[k]].pure) continue;
    auto T = cqx::solid_harmonics(l[k]);
    int pre = 1, post = 1;
    for (int q = 0; q < k; ++q) pre *= dims[q];
    for (int q = k + 1; q < 4; ++q) post *= dims[q];
    std::vector<double> nxt((size_t)pre * nf[k] * post, 0.0);
    for (int a = 0; a < pre; ++a)
      for (int m = 0; m < nf[k]; ++m)
        for (int c = 0; c < nc[k]; ++c) {
          const double t = T[(size_t)m * nc[k] + c];
          if (t == 0.0) continue;
          for (int p = 0; p < post; ++p)
            nxt[((size_t)a * nf[k] + m) * post + p] += t * cur[((size_t)a * nc[k] + c) * post + p];
        }
    cur.swap(nxt);
    dims[k] = nf[k];
  }
  return cur;
}

// Algorithmic flop model of one class pair (Rys quadrature form):
//   per primitive quartet  n_r (30 + 12 (la+lb+1)(lc+ld+1)
//                                + 6 (la+1)(lb+1)(lc+1)(ld+1) + 3 N_c) + 2 N_c
//   per contracted quartet 8 N_c N_s^(1/4) + N_s w (4 n_J + 8 n_K)
// with N_c / N_s the Cartesian / pure quartet sizes, w = 2 for complex.
double class_pair_flops(int la, int lb, int lc, int ld, bool pure_d, double prim_sum, double nq,
                        int w, int nJ, int nK) {
  const int L = la + lb + lc + ld, nr = L / 2 + 1;
  auto nc = [](int l) { return (double)cqx::ncart(l); };
  auto nf = [&](int l) { return (l > 1 && pure_d) ? 2.0 * l + 1.0 : nc(l); };
  const double Nc = nc(la) * nc(lb) * nc(lc) * nc(ld);
  const double Ns = nf(la) * nf(lb) * nf(lc) * nf(ld);
  const double per_prim = nr * (30.0 + 12.0 * (la + lb + 1) * (lc + ld + 1) +
                                6.0 * (la + 1) * (lb + 1) * (lc + 1) * (ld + 1) + 3.0 * Nc) +
                          2.0 * Nc;
  const double per_q = 8.0 * Nc * std::pow(Ns, 0.25) + Ns * w * (4.0 * nJ + 8.0 * nK);
  return prim_sum * per_prim + nq * per_q;
}

// primitive pairs of shells (a, b) in kernel orientation
void make_prims(const cqx::Shell &A, const cqx::Shell &B, std::vector<PrimPair> &out,
                std::vector<double> *bound, std::vector<double> *alpha_a = nullptr,
                std::vector<cqx::PrimC32> *compact = nullptr) {
  double AB2 = 0;
  for (int d = 0; d < 3; ++d) AB2 += (A.origin[d] - B.origin[d]) * (A.origin[d] - B.origin[d]);
  for (int i = 0; i < A.nprim(); ++i)
    for (int j = 0; j < B.nprim(); ++j) {
      const double a = A.alpha[i], b = B.alpha[j], p = a + b;
      PrimPair pp;
      pp.p = p;
      for (int d = 0; d < 3; ++d) {
        pp.P[d] = (a * A.origin[d] + b * B.origin[d]) / p;
        pp.PA[d] = pp.P[d] - A.origin[d];
      }
      pp.K = A.coef[i] * B.coef[j] * std::exp(-a * b / p * AB2);
      out.push_back(pp);
      if (alpha_a) alpha_a->push_back(a);
      if (compact) compact->push_back({(float)p, (float)(b / p), pp.K});
      // |(pp|qq)| <= U_p U_q with U = sqrt(2) pi^(5/4) |K| p^(-5/4)
      if (bound) bound->push_back(std::sqrt(2.0) * std::pow(M_PI, 1.25) * std::fabs(pp.K) * std::pow(p, -1.25));
    }
}

// Run the generic kernel in STORE mode over (bra[q] | ket[q]) and return
// the contracted Cartesian quartets back to back.
std::vector<double> eri_store(cqx_ctx *c, const std::vector<PairRec> &bra,
                              const std::vector<PairRec> &ket, int la, int lb, int lc, int ld,
                              const std::vector<PrimPair> &prims) {
  const long long nq = (long long)bra.size();
  const int ncq = cqx::ncart(la) * cqx::ncart(lb) * cqx::ncart(lc) * cqx::ncart(ld);
  DevBuf<PairRec> db, dk;
  DevBuf<PrimPair> dp;
  DevBuf<Quartet> dq;
  DevBuf<long long> doff;
  DevBuf<double> dout;
  db.upload(bra, c->stream);
  dk.upload(ket, c->stream);
  dp.upload(prims, c->stream);
  std::vector<Quartet> q(nq);
  std::vector<long long> off(nq);
  for (long long i = 0; i < nq; ++i) {
    q[i].bra = (uint32_t)i;
    q[i].ket_mask = (uint32_t)i;
    off[i] = i * ncq;
  }
  dq.upload(q, c->stream);
  doff.upload(off, c->stream);
  dout.ensure((size_t)nq * ncq);
  cqx::EriLaunch L{};
  L.la = la; L.lb = lb; L.lc = lc; L.ld = ld;
  L.quartets = dq.p;
  L.nq = nq;
  L.bra = db.p;
  L.ket = dk.p;
  L.prims = dp.p;
  L.rys = c->d_rys.p;
  L.store = dout.p;
  L.store_off = doff.p;
  L.store_mode = 1;
  HIPCHK(cqx::launch_eri_team(L, c->stream));
  std::vector<double> h((size_t)nq * ncq);
  HIPCHK(hipMemcpyAsync(h.data(), dout.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return h;
}

PairRec make_pair_rec(const cqx::Basis &B, int a, int b, int s1, int s2, int pp_off, int npp) {
  PairRec r{};
  const int m = B.shells[a].mol;
  r.mol = m;
  r.a = a; r.b = b;
  r.s1 = s1 - B.mol_sh0[m]; r.s2 = s2 - B.mol_sh0[m];  // numbered within the molecule
  r.pp_off = pp_off; r.npp = npp;
  r.bfa = B.bf_cart[a] - B.mol_bfc0[m]; r.bfb = B.bf_cart[b] - B.mol_bfc0[m];
  r.Q = 0.0;
  for (int d = 0; d < 3; ++d) r.AB[d] = B.shells[a].origin[d] - B.shells[b].origin[d];
  r.kk = B.shells[a].nprim() * B.shells[b].nprim();
  for (int d = 0; d < 3; ++d) r.A[d] = B.shells[a].origin[d];
  return r;
}

void setup_device_maps(cqx_ctx *c) {
  const cqx::Basis &B = c->basis;
  std::vector<int> cart_shell, cart_comp, bf_shell, bf_comp, sh_bf, sh_bfc, sh_l, sh_pure, sh_nf;
  for (int s = 0; s < c->ns; ++s) {
    const auto &sh = B.shells[s];
    for (int k = 0; k < cqx::ncart(sh.l); ++k) { cart_shell.push_back(s); cart_comp.push_back(k); }
    for (int k = 0; k < sh.nfunc(); ++k) { bf_shell.push_back(s); bf_comp.push_back(k); }
    sh_bf.push_back(B.bf[s]);
    sh_bfc.push_back(B.bf_cart[s]);
    sh_l.push_back(sh.l);
    sh_pure.push_back(sh.pure ? 1 : 0);
    sh_nf.push_back(sh.nfunc());
  }
  std::vector<double> sph;
  for (int l = 0; l <= cqx::kMaxLDev; ++l) {
    auto T = cqx::solid_harmonics(l);
    sph.insert(sph.end(), T.begin(), T.end());
  }
  c->d_cart_shell.upload(cart_shell, c->stream);
  c->d_cart_comp.upload(cart_comp, c->stream);
  c->d_bf_shell.upload(bf_shell, c->stream);
  c->d_bf_comp.upload(bf_comp, c->stream);
  c->d_sh_bf.upload(sh_bf, c->stream);
  c->d_sh_bfc.upload(sh_bfc, c->stream);
  c->d_sh_l.upload(sh_l, c->stream);
  c->d_sh_pure.upload(sh_pure, c->stream);
  c->d_sh_nf.upload(sh_nf, c->stream);
  c->d_sph.upload(sph, c->stream);
  c->cmap.cart_shell = c->d_cart_shell.p;
  c->cmap.cart_comp = c->d_cart_comp.p;
  c->cmap.sh_bf = c->d_sh_bf.p;
  c->cmap.sh_bfc = c->d_sh_bfc.p;
  c->cmap.sh_l = c->d_sh_l.p;
  c->cmap.sh_pure = c->d_sh_pure.p;
  c->cmap.sph = c->d_sph.p;
  c->cmap.bf_shell = c->d_bf_shell.p;
  c->cmap.bf_comp = c->d_bf_comp.p;
  c->cmap.nb = c->nb;
  c->cmap.nbc = c->nbc;
  c->cmap.ns = c->ns;
  c->cmap.bf0 = 0;
  c->cmap.bfc0 = 0;
  // per-molecule block-norm layout
  std::vector<int> mns(c->nmol);
  c->mol_boff.assign(c->nmol + 1, 0);
  for (int m = 0; m < c->nmol; ++m) {
    mns[m] = B.mol_sh0[m + 1] - B.mol_sh0[m];
    c->mol_boff[m + 1] = c->mol_boff[m] + (long long)mns[m] * mns[m];
  }
  c->d_mol_ns.upload(mns, c->stream);
  c->d_mol_boff.upload(c->mol_boff, c->stream);
}

void upload_pairs(cqx_ctx *c);
void build_tiles(cqx_ctx *c);

// Significant pairs, primitive pairs and GPU Schwarz bounds.
void build_pairs(cqx_ctx *c) {
  const cqx::Basis &B = c->basis;
  c->ref_pairs = B.significant_pairs();
  // primitive-pair bound scale
  std::vector<PrimPair> tmp;
  std::vector<double> bnd, alp;
  double umax = 0.0;
  for (auto &pr : c->ref_pairs) {
    tmp.clear();
    bnd.clear();
    make_prims(B.shells[pr.first], B.shells[pr.second], tmp, &bnd);
    for (double u : bnd) umax = std::max(umax, u);
  }
  c->cls.assign(kNumClass, {});
  c->prims.clear();
  c->prim_alpha.clear();
  c->primc32.clear();
  std::vector<cqx::PrimC32> cmp;
  for (auto &pr : c->ref_pairs) {
    int s1 = pr.first, s2 = pr.second;
    int a = s1, b = s2;
    if (B.shells[s1].l < B.shells[s2].l) std::swap(a, b);
    tmp.clear();
    bnd.clear();
    alp.clear();
    cmp.clear();
    make_prims(B.shells[a], B.shells[b], tmp, &bnd, &alp, &cmp);
    int off = (int)c->prims.size(), n = 0;
    for (size_t k = 0; k < tmp.size(); ++k)
      if (bnd[k] * umax >= c->prim_thresh) {
        c->prims.push_back(tmp[k]);
        c->prim_alpha.push_back(alp[k]);
        c->primc32.push_back(cmp[k]);
        ++n;
      }
    c->cls[class_of(B.shells[a].l, B.shells[b].l)].push_back(make_pair_rec(B, a, b, s1, s2, off, n));
  }
  // a quartet carries its ket's index in 24 bits (Quartet::ket_mask)
  for (int k = 0; k < kNumClass; ++k)
    if (!cqx::ket_index_fits(c->cls[k].size()))
      throw Unsupported("a shell-pair class holds " + std::to_string(c->cls[k].size()) +
                        " pairs, beyond the 2^24 a quartet can index");
  // Schwarz: (ab|ab) for every significant pair, with all primitives
  // (the reference evaluates these at precision 0, aointegrals_utilities.cxx:70)
  if (c->nmol == 1) c->schwarz.assign((size_t)c->ns * c->ns, 0.0);
  for (int k = 0; k < kNumClass; ++k) {
    auto &v = c->cls[k];
    if (v.empty()) continue;
    int la, lb;
    class_ls(k, la, lb);
    std::vector<PairRec> full;
    std::vector<PrimPair> fp;
    for (auto &r : v) {
      int off = (int)fp.size();
      make_prims(B.shells[r.a], B.shells[r.b], fp, nullptr);
      full.push_back(make_pair_rec(B, r.a, r.b, r.s1, r.s2, off, (int)fp.size() - off));
    }
    auto ints = eri_store(c, full, full, la, lb, la, lb, fp);
    const int ncq = cqx::ncart(la) * cqx::ncart(lb) * cqx::ncart(la) * cqx::ncart(lb);
    for (size_t q = 0; q < v.size(); ++q) {
      auto &r = v[q];
      int sh[4] = {r.a, r.b, r.a, r.b};
      auto pure = cart_to_pure(B, sh, ints.data() + q * ncq);
      const int na = B.shells[r.a].nfunc(), nbf = B.shells[r.b].nfunc();
      // rows of the infinity norm run over the functions of s1
      const bool a_is_s1 = (r.a == r.s1 + B.mol_sh0[r.mol]);  // s1 is numbered within the molecule
      const int n1 = a_is_s1 ? na : nbf, n2 = a_is_s1 ? nbf : na;
      double nrm = 0.0;
      for (int i = 0; i < n1; ++i) {
        double row = 0.0;
        for (int j = 0; j < n2; ++j) {
          const int ia = a_is_s1 ? i : j, ib = a_is_s1 ? j : i;
          const size_t ab = (size_t)ia * nbf + ib;
          row += std::fabs(pure[ab * na * nbf + ab]);
        }
        nrm = std::max(nrm, row);
      }
      r.Q = std::sqrt(nrm);
      if (c->nmol == 1) {
        c->schwarz[r.s1 + (size_t)r.s2 * c->ns] = r.Q;
        c->schwarz[r.s2 + (size_t)r.s1 * c->ns] = r.Q;
      }
    }
    std::stable_sort(v.begin(), v.end(), [](const PairRec &x, const PairRec &y) {
      return x.mol != y.mol ? x.mol < y.mol : x.Q > y.Q;
    });
  }
  upload_pairs(c);
}

// Device copies of the pair data (class arrays, screening records,
// primitive pairs) and the molecule segments of every class; the class
// arrays must be sorted by (molecule, Q descending).
void upload_pairs(cqx_ctx *c) {
  // molecule segments of every class
  c->seg.assign(kNumClass, std::vector<int>(c->nmol + 1, 0));
  for (int k = 0; k < kNumClass; ++k) {
    auto &sg = c->seg[k];
    const auto &v = c->cls[k];
    size_t i = 0;
    for (int m = 0; m <= c->nmol; ++m) {
      while (i < v.size() && v[i].mol < m) ++i;
      sg[m] = (int)i;
    }
  }
  std::vector<int> segs;
  for (int k = 0; k < kNumClass; ++k) segs.insert(segs.end(), c->seg[k].begin(), c->seg[k].end());
  c->d_seg.upload(segs, c->stream);
  c->d_prims.upload(c->prims, c->stream);
  c->d_prim_alpha.upload(c->prim_alpha, c->stream);
  c->d_primc32.upload(c->primc32, c->stream);
  for (int k = 0; k < kNumClass; ++k) {
    c->d_pairs[k].upload(c->cls[k], c->stream);
    std::vector<cqx::PairScr> scr;
    for (const auto &r : c->cls[k]) {
      if (r.s1 > 0xFFFF || r.s2 > 0xFFFF || r.kk > 0xFFFF || r.mol > 0xFFFF)
        throw std::runtime_error("shell, contraction or molecule index beyond 16 bits");
      cqx::PairScr q;
      q.s12 = (uint32_t)r.s1 | ((uint32_t)r.s2 << 16);
      q.kk = (uint16_t)r.kk;
      q.mol = (uint16_t)r.mol;
      q.Q = r.Q;
      scr.push_back(q);
    }
    c->d_scr[k].upload(scr, c->stream);
  }
  build_tiles(c);
  HIPCHK(hipStreamSynchronize(c->stream));
}

// Member arrays and sub-block tables of the tile path.  A class's pairs are
// grouped by (molecule, first shell in kernel orientation), each group in
// class order (Schwarz Q descending); a sub-block takes up to kTileDim
// consecutive members of a bra group against up to kTileDim of a ket group of
// the same molecule (bra class >= ket class; in one class, ket group <= bra
// group, and in one group ket chunk <= bra chunk, the diagonal chunk keeping
// i >= j).  Every pair of pairs of a tile class pair is thus in exactly one
// sub-block, as in the quartet lists.
void build_tiles(cqx_ctx *c) {
  struct Group { int m0, n, mol; };
  std::vector<std::vector<Group>> groups(kNumClass);
  std::vector<std::vector<int>> mem_host(kNumClass);
  for (int k = 0; k < kNumClass; ++k) {
    const auto &v = c->cls[k];
    auto &mem = mem_host[k];
    mem.resize(v.size());
    std::iota(mem.begin(), mem.end(), 0);
    // by molecule, then first shell; class order (Q descending) within a
    // group, so a chunk's first member has its largest Q
    std::stable_sort(mem.begin(), mem.end(), [&](int x, int y) {
      return v[x].mol != v[y].mol ? v[x].mol < v[y].mol : v[x].a < v[y].a;
    });
    for (size_t i = 0; i < mem.size();) {
      size_t j = i + 1;
      while (j < mem.size() && v[mem[j]].mol == v[mem[i]].mol && v[mem[j]].a == v[mem[i]].a) ++j;
      groups[k].push_back({(int)i, (int)(j - i), v[mem[i]].mol});
      i = j;
    }
    c->d_members[k].upload(mem, c->stream);
  }
  std::vector<cqx::SubBlock> sub;
  c->tile_cps.clear();
  for (int cb = 0; cb < kNumClass; ++cb)
    for (int ck = 0; ck <= cb; ++ck) {
      if (c->cls[cb].empty() || c->cls[ck].empty()) continue;
      int la, lb, lc, ld;
      class_ls(cb, la, lb);
      class_ls(ck, lc, ld);
      if (!cqx::eri_fast_available(la, lb, lc, ld)) continue;
      cqx_ctx::TileCP tc{cb, ck, (long long)sub.size(), 0};
      const auto &GB = groups[cb], &GK = groups[ck];
      const auto &VB = c->cls[cb], &VK = c->cls[ck];
      size_t y0 = 0;  // first ket group of the bra group's molecule
      for (size_t x = 0; x < GB.size(); ++x) {
        const int mol = GB[x].mol;
        while (y0 < GK.size() && GK[y0].mol < mol) ++y0;
        for (size_t y = y0; y < GK.size() && GK[y].mol == mol; ++y) {
          if (cb == ck && y > x) break;
          const Group &gx = GB[x], &gy = GK[y];
          for (int u = 0; u * cqx::kTileDim < gx.n; ++u)
            for (int w = 0; w * cqx::kTileDim < gy.n; ++w) {
              const bool diag = cb == ck && x == y && u == w;
              if (cb == ck && x == y && w > u) break;
              cqx::SubBlock sb{};
              sb.bra_m0 = gx.m0 + u * cqx::kTileDim;
              sb.ket_m0 = gy.m0 + w * cqx::kTileDim;
              sb.nb = (unsigned short)std::min(cqx::kTileDim, gx.n - u * cqx::kTileDim);
              sb.nk = (unsigned short)std::min(cqx::kTileDim, gy.n - w * cqx::kTileDim);
              sb.diag = diag ? 1 : 0;
              sb.qq = VB[mem_host[cb][sb.bra_m0]].Q * VK[mem_host[ck][sb.ket_m0]].Q;
              sub.push_back(sb);
            }
        }
      }
      tc.n = (long long)sub.size() - tc.off;
      if (!tc.n) continue;
      // padded to whole list-kernel chunks with empty sub-blocks
      while (((long long)sub.size() - tc.off) % cqx::kTileChunk) {
        cqx::SubBlock e{};
        sub.push_back(e);
      }
      tc.n = (long long)sub.size() - tc.off;
      c->tile_cps.push_back(tc);
    }
  c->nsub_total = (long long)sub.size();
  c->d_sub.upload(sub, c->stream);
}

// Replay of the captured build for the densities of `list`: stage X and AX,
// launch the graph, read back the candidate totals and list counts in one
// synchronisation.  Returns false, with the caller's AX untouched, when a
// class pair had more candidates, or a list more quartets, than its
// capacity (the lists or the evaluation would be incomplete); the caller
// then builds eagerly.
bool replay_graph(cqx_ctx *c, int ncon, const cqx_contraction *list, int is_complex, int mem) {
  hipStream_t s = c->stream;
  const int nmol = c->nmol, nper = ncon / nmol, nmat = c->g_nmat;
  const size_t ew = is_complex ? 2 : 1;
  const hipMemcpyKind hk = mem == CQX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  for (int m = 0; m < nmol; ++m)
    for (int j = 0; j < nmat; ++j) {
      const cqx_contraction &cc = list[m * nper + j];
      const size_t o = c->g_xoff[m] + (size_t)j * c->g_mnb2[m] * ew, n = sizeof(double) * c->g_mnb2[m] * ew;
      HIPCHK(hipMemcpyAsync(c->d_X.p + o, cc.X, n, hk, s));
      HIPCHK(hipMemcpyAsync(c->d_AX.p + o, cc.AX, n, hk, s));
    }
  HIPCHK(hipEventRecord(c->ev0, s));
  HIPCHK(hipGraphLaunch(c->graph_exec, s));
  HIPCHK(hipEventRecord(c->ev1, s));
  const size_t ncp = c->g_cap.size();
  std::vector<long long> tot(ncp);
  std::vector<unsigned long long> counts(2 * ncp);
  std::vector<double> psum(2 * ncp);
  HIPCHK(hipMemcpyAsync(tot.data(), c->d_gchk.p, sizeof(long long) * ncp, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(counts.data(), c->d_count.p, sizeof(unsigned long long) * 2 * ncp, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(psum.data(), c->d_psum.p, sizeof(double) * 2 * ncp, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (size_t k = 0; k < ncp; ++k)
    if (tot[k] > c->g_cap[k]) return false;
  for (size_t kk = 0; kk < 2 * ncp; ++kk)
    if ((long long)counts[kk] > c->g_lcap[kk]) return false;
  const hipMemcpyKind dk = mem == CQX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  for (int m = 0; m < nmol; ++m)
    for (int j = 0; j < nmat; ++j)
      HIPCHK(hipMemcpyAsync(list[m * nper + j].AX, c->d_AX.p + c->g_xoff[m] + (size_t)j * c->g_mnb2[m] * ew,
                            sizeof(double) * c->g_mnb2[m] * ew, dk, s));
  HIPCHK(hipStreamSynchronize(s));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->stats.ms_eri = ms;  // the whole recorded build
  for (size_t k = 0; k < ncp; ++k) c->stats.n_candidates += tot[k];
  for (size_t kk = 0; kk < 2 * ncp; ++kk) {
    const auto &G = c->g_slots[kk];
    if (G.la < 0) continue;
    c->stats.n_launch_eri++;
    if (!counts[kk]) continue;
    const double fl = class_pair_flops(G.la, G.lb, G.lc, G.ld, c->g_pure_d, psum[kk], (double)counts[kk],
                                       is_complex ? 2 : 1, c->g_nJ, c->g_nK);
    c->stats.n_quartets += (long long)counts[kk];
    c->stats.n_prim_quartets_max += (long long)psum[kk];
    c->stats.flops_model += fl;
    if (G.f32) {
      c->stats.n_quartets_f32 += (long long)counts[kk];
      c->stats.flops_model_f32 += fl;
    }
  }
  c->stats.graph = 2;
  return true;
}

// Everything of a new context but its pair data: basis, device, streams,
// events, scratch region, Rys tables and the Cartesian maps.
void init_ctx(cqx_ctx *c, const cqx::Basis &basis, double thresh, int device) {
  c->basis = basis;
  c->thresh = thresh;
  if (device >= 0) HIPCHK(hipSetDevice(device));
  HIPCHK(hipGetDevice(&c->device));
  c->scratch = scratch_region(c->device);
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&c->ev0));
  HIPCHK(hipEventCreate(&c->ev1));
  HIPCHK(hipEventCreate(&c->ev_t0));
  HIPCHK(hipEventCreate(&c->ev_t1));
  HIPCHK(hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
  c->ns = (int)c->basis.shells.size();
  c->nb = c->basis.nbasis;
  c->nbc = c->basis.ncart_total;
  c->nmol = c->basis.nmol;
  std::vector<double> rys(kRysTable, kRysTable + sizeof(kRysTable) / sizeof(double));
  c->d_rys.upload(rys, c->stream);
  std::vector<float> rys32(rys.begin(), rys.end());
  c->d_rys32.upload(rys32, c->stream);
  HIPCHK(hipStreamSynchronize(c->stream));
  setup_device_maps(c);
}

}  // namespace

namespace cqx {

// A batched context over the molecules of one-molecule contexts, built from
// their shell pairs, primitive pairs and Schwarz bounds (no ERI is
// evaluated): molecule m is parts[m].  The parts keep their own data.
int ctx_merge(cqx_ctx *const *parts, int n, cqx_ctx **out) {
  if (!parts || n <= 0 || !out) return fail(nullptr, CQX_ERR_ARG, "ctx_merge: bad arguments");
  std::vector<const Basis *> bases;
  for (int m = 0; m < n; ++m) {
    if (!parts[m] || parts[m]->nmol != 1) return fail(nullptr, CQX_ERR_ARG, "ctx_merge: parts must be one-molecule contexts");
    if (parts[m]->thresh != parts[0]->thresh || parts[m]->device != parts[0]->device ||
        parts[m]->prim_thresh != parts[0]->prim_thresh)
      return fail(nullptr, CQX_ERR_ARG, "ctx_merge: parts differ in device or thresholds");
    bases.push_back(&parts[m]->basis);
  }
  auto c = std::make_unique<cqx_ctx>();
  try {
    init_ctx(c.get(), Basis::batch(bases), parts[0]->thresh, parts[0]->device);
    c->prim_thresh = parts[0]->prim_thresh;
    c->cls.assign(kNumClass, {});
    int prim0 = 0;
    for (int m = 0; m < n; ++m) {
      const cqx_ctx *p = parts[m];
      const int sh0 = c->basis.mol_sh0[m];
      for (const auto &pr : p->ref_pairs) c->ref_pairs.emplace_back(pr.first + sh0, pr.second + sh0);
      for (int k = 0; k < kNumClass; ++k)
        for (PairRec r : p->cls[k]) {
          r.a += sh0;
          r.b += sh0;
          r.pp_off += prim0;
          r.mol = m;
          c->cls[k].push_back(r);
        }
      c->prims.insert(c->prims.end(), p->prims.begin(), p->prims.end());
      c->prim_alpha.insert(c->prim_alpha.end(), p->prim_alpha.begin(), p->prim_alpha.end());
      c->primc32.insert(c->primc32.end(), p->primc32.begin(), p->primc32.end());
      prim0 = (int)c->prims.size();
    }
    for (int k = 0; k < kNumClass; ++k)
      if (!ket_index_fits(c->cls[k].size()))
        throw Unsupported("a shell-pair class holds " + std::to_string(c->cls[k].size()) +
                          " pairs, beyond the 2^24 a quartet can index");
    upload_pairs(c.get());
  } catch (const Unsupported &e) {
    return fail(nullptr, CQX_ERR_UNSUPPORTED, e.what());
  } catch (const std::exception &e) {
    return fail(nullptr, CQX_ERR_HIP, e.what());
  }
  *out = c.release();
  return CQX_OK;
}

}  // namespace cqx

extern "C" {

const char *cqx_last_error(const cqx_ctx *ctx) { return ctx ? ctx->err.c_str() : g_error.c_str(); }


int cqx_basis_create(int nshell, const int *l, const int *pure, const int *nprim, const double *alpha,
                     const double *coef, const double *center, const int *atom, cqx_basis **out) {
  if (!out || nshell <= 0 || !l || !pure || !nprim || !alpha || !coef || !center || !atom)
    return fail(nullptr, CQX_ERR_ARG, "cqx_basis_create: bad arguments");
  auto b = std::make_unique<cqx_basis>();
  size_t off = 0;
  for (int s = 0; s < nshell; ++s) {
    cqx::Shell sh;
    sh.l = l[s];
    sh.pure = pure[s] != 0;
    sh.atom = atom[s];
    if (sh.l < 0 || sh.l > cqx::kMaxL || nprim[s] <= 0)
      return fail(nullptr, CQX_ERR_ARG, "cqx_basis_create: bad shell " + std::to_string(s));
    sh.alpha.assign(alpha + off, alpha + off + nprim[s]);
    sh.coef.assign(coef + off, coef + off + nprim[s]);
    off += nprim[s];
    for (int d = 0; d < 3; ++d) sh.origin[d] = center[3 * s + d];
    b->b.shells.push_back(sh);
  }
  b->b.finalise();
  *out = b.release();
  return CQX_OK;
}

int cqx_basis_from_g94(const char *text, int natom, const int *Z, const double *xyz, int force_cart,
                       cqx_basis **out) {
  if (!text || !out || natom <= 0 || !Z || !xyz) return fail(nullptr, CQX_ERR_ARG, "cqx_basis_from_g94: bad arguments");
  try {
    auto b = std::make_unique<cqx_basis>();
    b->b = cqx::Basis::from_g94(text, std::vector<int>(Z, Z + natom),
                                std::vector<double>(xyz, xyz + 3 * natom), force_cart != 0);
    *out = b.release();
    return CQX_OK;
  } catch (const std::exception &e) {
    return fail(nullptr, CQX_ERR_BASIS, e.what());
  }
}

int cqx_basis_batch(int nmol, const cqx_basis *const *parts, cqx_basis **out) {
  if (nmol <= 0 || !parts || !out) return fail(nullptr, CQX_ERR_ARG, "cqx_basis_batch: bad arguments");
  try {
    std::vector<const cqx::Basis *> v;
    for (int m = 0; m < nmol; ++m) {
      if (!parts[m]) return fail(nullptr, CQX_ERR_ARG, "cqx_basis_batch: null basis");
      if (parts[m]->b.nmol != 1) return fail(nullptr, CQX_ERR_ARG, "cqx_basis_batch: parts must be single molecules");
      v.push_back(&parts[m]->b);
    }
    auto b = std::make_unique<cqx_basis>();
    b->b = cqx::Basis::batch(v);
    *out = b.release();
    return CQX_OK;
  } catch (const std::exception &e) {
    return fail(nullptr, CQX_ERR_BASIS, e.what());
  }
}

int cqx_basis_destroy(cqx_basis *b) {
  delete b;
  return CQX_OK;
}

int cqx_basis_dims(const cqx_basis *b, int *nshell, int *nbasis, int *max_prim, int *max_l) {
  if (!b) return fail(nullptr, CQX_ERR_ARG, "null basis");
  if (nshell) *nshell = (int)b->b.shells.size();
  if (nbasis) *nbasis = b->b.nbasis;
  if (max_prim) *max_prim = b->b.max_prim;
  if (max_l) *max_l = b->b.max_l;
  return CQX_OK;
}

int cqx_one_body(const cqx_basis *b, int kind, int ncharge, const double *Z, const double *xyz,
                 double *out) {
  if (!b || !out || kind < CQX_OVERLAP || kind > CQX_POTENTIAL ||
      (kind == CQX_POTENTIAL && (ncharge < 0 || (ncharge && (!Z || !xyz)))))
    return fail(nullptr, CQX_ERR_ARG, "cqx_one_body: bad arguments");
  try {
    std::vector<double> Zc, C;
    if (kind == CQX_POTENTIAL) {
      Zc.assign(Z, Z + ncharge);
      C.assign(xyz, xyz + 3 * (size_t)ncharge);
    }
    const auto M = cqx::one_body_matrix(b->b, kind, Zc, C);
    std::copy(M.begin(), M.end(), out);
  } catch (const std::exception &e) {
    return fail(nullptr, CQX_ERR_ARG, e.what());
  }
  return CQX_OK;
}

int cqx_basis_shell(const cqx_basis *b, int s, int *l, int *pure, int *nprim, int *bf, int *atom) {
  if (!b || s < 0 || s >= (int)b->b.shells.size()) return fail(nullptr, CQX_ERR_ARG, "bad shell index");
  const auto &sh = b->b.shells[s];
  if (l) *l = sh.l;
  if (pure) *pure = sh.pure;
  if (nprim) *nprim = sh.nprim();
  if (bf) *bf = b->b.bf[s];
  if (atom) *atom = sh.atom;
  return CQX_OK;
}

int cqx_basis_shell_prims(const cqx_basis *b, int s, double *alpha, double *coef, double *origin) {
  if (!b || s < 0 || s >= (int)b->b.shells.size()) return fail(nullptr, CQX_ERR_ARG, "bad shell index");
  const auto &sh = b->b.shells[s];
  if (alpha) std::copy(sh.alpha.begin(), sh.alpha.end(), alpha);
  if (coef) std::copy(sh.coef.begin(), sh.coef.end(), coef);
  if (origin) std::copy(sh.origin, sh.origin + 3, origin);
  return CQX_OK;
}

int cqx_ctx_create(const cqx_basis *b, double thresh, int device, cqx_ctx **out) {
  if (!b || !out) return fail(nullptr, CQX_ERR_ARG, "cqx_ctx_create: bad arguments");
  if (b->b.max_l > cqx::kMaxLDev)
    return fail(nullptr, CQX_ERR_UNSUPPORTED, "device kernels cover shells up to l = 4 (g)");
  auto c = std::make_unique<cqx_ctx>();
  try {
    init_ctx(c.get(), b->b, thresh, device);
    build_pairs(c.get());
  } catch (const Unsupported &e) {
    return fail(nullptr, CQX_ERR_UNSUPPORTED, e.what());
  } catch (const std::exception &e) {
    return fail(nullptr, CQX_ERR_HIP, e.what());
  }
  *out = c.release();
  return CQX_OK;
}

int cqx_ctx_destroy(cqx_ctx *c) {
  if (!c) return CQX_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->ev_t0) (void)hipEventDestroy(c->ev_t0);
  if (c->ev_t1) (void)hipEventDestroy(c->ev_t1);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  for (int i = 0; i < cqx_ctx::kMaxSide; ++i) {
    if (c->side[i]) (void)hipStreamSynchronize(c->side[i]);
    if (c->side[i]) (void)hipStreamDestroy(c->side[i]);
    if (c->ev_join[i]) (void)hipEventDestroy(c->ev_join[i]);
  }
  hipStream_t s = c->stream;
  delete c;
  if (s) (void)hipStreamDestroy(s);
  return CQX_OK;
}

int cqx_ctx_schwarz(cqx_ctx *c, double *Q) {
  if (!c || !Q) return fail(c, CQX_ERR_ARG, "cqx_ctx_schwarz: bad arguments");
  if (c->nmol != 1) return fail(c, CQX_ERR_UNSUPPORTED, "cqx_ctx_schwarz: not kept for a batched context");
  std::copy(c->schwarz.begin(), c->schwarz.end(), Q);
  return CQX_OK;
}

int cqx_ctx_shell_pairs(cqx_ctx *c, int *npairs, int *s1, int *s2) {
  if (!c || !npairs) return fail(c, CQX_ERR_ARG, "cqx_ctx_shell_pairs: bad arguments");
  *npairs = (int)c->ref_pairs.size();
  if (s1 && s2)
    for (size_t k = 0; k < c->ref_pairs.size(); ++k) {
      s1[k] = c->ref_pairs[k].first;
      s2[k] = c->ref_pairs[k].second;
    }
  return CQX_OK;
}

int cqx_set_option(cqx_ctx *c, const char *name, double value) {
  if (!c || !name) return fail(c, CQX_ERR_ARG, "cqx_set_option: bad arguments");
  drop_graph(c);  // every option changes what a recorded build would launch
  if (!std::strcmp(name, "graph")) {
    c->use_graph = value != 0.0;
    return CQX_OK;
  }
  if (!std::strcmp(name, "graph_margin")) {
    if (!(value >= 1.0)) return fail(c, CQX_ERR_ARG, "graph_margin must be >= 1");
    c->graph_margin = value;
    return CQX_OK;
  }
  if (!std::strcmp(name, "prim_screen")) {
    c->prim_thresh = value;
    try {
      build_pairs(c);
    } catch (const std::exception &e) {
      return fail(c, CQX_ERR_HIP, e.what());
    }
    return CQX_OK;
  }
  if (!std::strcmp(name, "fast_max_nc")) {
    c->fast_max_nc = (int)value;
    return CQX_OK;
  }
  if (!std::strcmp(name, "use_wave")) {
    c->use_wave = value != 0.0;
    return CQX_OK;
  }
  if (!std::strcmp(name, "fast_lanes")) {
    const int g = (int)value;
    if (g != 0 && g != 1 && g != 2 && g != 4 && g != 8 && g != 16)
      return fail(c, CQX_ERR_ARG, "fast_lanes must be 0, 1, 2, 4, 8 or 16");
    c->fast_lanes = g;
    return CQX_OK;
  }
  if (!std::strcmp(name, "autotune")) {
    // 1: the next contraction times every lane count of every lane-group
    // launch, and every quartets-per-wave count of every wave launch, on
    // scratch accumulators and keeps the fastest per class pair and
    // precision; 0: forget the choices
    c->autotune_pending = value != 0.0;
    if (value == 0.0) c->tuned_lanes.clear();
    return CQX_OK;
  }
  if (!std::strcmp(name, "tiles")) {
    // 1: lane-group class pairs of Hermitian builds through tile lists
    // (shared J / K targets summed per tile); 2: only builds with at most
    // one Hermitian J and one Hermitian K plane; 0 (default): quartet lists
    if (!(value == 0.0 || value == 1.0 || value == 2.0)) return fail(c, CQX_ERR_ARG, "tiles must be 0, 1 or 2");
    c->use_tiles = (int)value;
    return CQX_OK;
  }
  if (!std::strcmp(name, "tile_classes")) {
    // bit k set: the k-th lane-group class (ss|ss), (ps|ss), (ps|ps), (pp|ss),
    // (pp|ps), (ds|ss), (ds|ps), (dp|ss) takes tiles when "tiles" is on
    const int m = (int)value;
    if (!(value == (double)m && m >= 0 && m <= 0xFF)) return fail(c, CQX_ERR_ARG, "tile_classes must be 0-255");
    c->tile_classes = m;
    return CQX_OK;
  }
  if (!std::strcmp(name, "rys_lds")) {
    // 1 (default): the lane-group kernels read their Rys table section from
    // an LDS copy per block; 0: from global memory
    c->rys_lds = value != 0.0;
    return CQX_OK;
  }
  if (!std::strcmp(name, "streams")) {
    const int n = (int)value;
    if (!(value == (double)n && n >= 1 && n <= cqx_ctx::kMaxSide + 1))
      return fail(c, CQX_ERR_ARG, "streams must be 1, 2, 3 or 4");
    c->nstreams = n;
    return CQX_OK;
  }
  if (!std::strcmp(name, "list_cap")) {
    if (!(value >= 1.0)) return fail(c, CQX_ERR_ARG, "list_cap must be >= 1");
    c->list_cap = (long long)value;
    return CQX_OK;
  }
  if (!std::strcmp(name, "fp32_cut")) {
    if (!(value >= 0.0)) return fail(c, CQX_ERR_ARG, "fp32_cut must be >= 0");
    c->fp32_cut = value;
    return CQX_OK;
  }
  return fail(c, CQX_ERR_ARG, std::string("unknown option ") + name);
}

int cqx_last_stats(const cqx_ctx *c, cqx_stats *st) {
  if (!c || !st) return fail(nullptr, CQX_ERR_ARG, "cqx_last_stats: bad arguments");
  *st = c->stats;
  return CQX_OK;
}

int cqx_eri_quartets(cqx_ctx *c, int nq, const int *s1, const int *s2, const int *s3, const int *s4,
                     double *out, size_t *offsets) {
  if (!c || nq < 0 || (nq && (!s1 || !s2 || !s3 || !s4 || !out || !offsets)))
    return fail(c, CQX_ERR_ARG, "cqx_eri_quartets: bad arguments");
  try {
    HIPCHK(hipSetDevice(c->device));
    const cqx::Basis &B = c->basis;
    size_t pos = 0;
    for (int q = 0; q < nq; ++q) {
      int sh[4] = {s1[q], s2[q], s3[q], s4[q]};
      for (int k = 0; k < 4; ++k)
        if (sh[k] < 0 || sh[k] >= c->ns) return fail(c, CQX_ERR_ARG, "shell index out of range");
      std::vector<PrimPair> fp;
      make_prims(B.shells[sh[0]], B.shells[sh[1]], fp, nullptr);
      const int n12 = (int)fp.size();
      make_prims(B.shells[sh[2]], B.shells[sh[3]], fp, nullptr);
      std::vector<PairRec> bra{make_pair_rec(B, sh[0], sh[1], sh[0], sh[1], 0, n12)};
      std::vector<PairRec> ket{make_pair_rec(B, sh[2], sh[3], sh[2], sh[3], n12, (int)fp.size() - n12)};
      auto cart = eri_store(c, bra, ket, B.shells[sh[0]].l, B.shells[sh[1]].l, B.shells[sh[2]].l,
                            B.shells[sh[3]].l, fp);
      auto pure = cart_to_pure(B, sh, cart.data());
      offsets[q] = pos;
      std::copy(pure.begin(), pure.end(), out + pos);
      pos += pure.size();
    }
    offsets[nq] = pos;
  } catch (const std::exception &e) {
    return fail(c, CQX_ERR_HIP, e.what());
  }
  return CQX_OK;
}

}  // extern "C"

namespace {

// A bump allocator over a fixed device region, 256-byte aligned pieces.
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  template <class T>
  T *take(size_t n, const char *what) {
    const size_t off = (used + 255) & ~(size_t)255;
    const size_t b = sizeof(T) * std::max<size_t>(n, 1);
    if (off > cap || b > cap - off)
      throw Unsupported(std::string("cqx_two_body_contract_noalloc: the scratch region (") + std::to_string(cap) +
                        " bytes) cannot hold " + what);
    used = off + b;
    return reinterpret_cast<T *>(base + off);
  }
  size_t left() const {
    const size_t off = (used + 255) & ~(size_t)255;
    return off < cap ? cap - off : 0;
  }
};

// Contractions processed together (sharing one ERI evaluation): at most
// kMaxMat per molecule and kMaxPlanes real planes.
int group_size(const cqx_contraction *L, int left, int is_complex, int *nplane_out) {
  int nmat = 0, nplane = 0;
  while (nmat < left && nmat < cqx::kMaxMat) {
    const auto &cc = L[nmat];
    const int np = (!is_complex || (cc.her && cc.type == CQX_COULOMB)) ? 1 : 2;
    if (nplane + np > cqx::kMaxPlanes) break;
    nplane += np;
    ++nmat;
  }
  if (nplane_out) *nplane_out = nplane;
  return nmat;
}

// The density-independent workspace of one contraction group (elements of
// each buffer), computed the same way for a call and for the scratch query.
struct GroupPlan {
  int nmat = 0;
  size_t x = 0;       // doubles of the staged densities (and of the staged outputs)
  size_t bm = 0, bg = 0;
  size_t pn[kNumClass] = {};
  size_t planes = 0;  // doubles of the Cartesian planes (and of the accumulators)
  size_t cand = 0;    // candidate offsets (long long) and first kets (int)
  size_t ncp = 0;     // class pairs with pairs on both sides
  size_t scan_bytes = 0;
  long long max_ket = 0;   // largest ket class of a class pair
  long long cand_all = 0;  // candidates of every class pair without the Schwarz pre-cut
  int kcnt[4] = {0, 0, 0, 0};
  int kps[4] = {1, 1, 1, 1};
};

GroupPlan plan_group(const cqx_ctx *c, const cqx_contraction *L, int nmat, int is_complex, hipStream_t s) {
  GroupPlan g;
  g.nmat = nmat;
  const cqx::Basis &BS = c->basis;
  const size_t ew = is_complex ? 2 : 1;
  size_t nb2 = 0, nbc2 = 0;
  for (int m = 0; m < c->nmol; ++m) {
    const size_t nb = BS.mol_bf0[m + 1] - BS.mol_bf0[m], nbc = BS.mol_bfc0[m + 1] - BS.mol_bfc0[m];
    nb2 += nb * nb;
    nbc2 += nbc * nbc;
  }
  g.x = (size_t)nmat * nb2 * ew;
  g.bm = (size_t)nmat * c->mol_boff[c->nmol];
  g.bg = c->mol_boff[c->nmol];
  for (int k = 0; k < kNumClass; ++k) g.pn[k] = c->cls[k].size() * (size_t)(nmat + 1);
  for (int j = 0; j < nmat; ++j) {
    const bool two = is_complex && !(L[j].her && L[j].type == CQX_COULOMB);
    const int kind = L[j].type == CQX_COULOMB ? (L[j].her ? cqx::kJHer : cqx::kJNonHer)
                                              : (L[j].her ? cqx::kKHer : cqx::kKNonHer);
    g.kcnt[kind] += two ? 2 : 1;
  }
  int kpsum = 0;
  for (int k = 0; k < 4; ++k) {
    while (g.kps[k] < g.kcnt[k]) g.kps[k] *= 2;
    if (g.kcnt[k]) kpsum += g.kps[k];
  }
  g.planes = (size_t)kpsum * nbc2;
  for (int cb = 0; cb < kNumClass; ++cb)
    for (int ck = 0; ck <= cb; ++ck) {
      if (c->cls[cb].empty() || c->cls[ck].empty()) continue;
      const long long nbra = (long long)c->cls[cb].size(), nket = (long long)c->cls[ck].size();
      g.cand += (size_t)nbra + 1;
      ++g.ncp;
      g.max_ket = std::max(g.max_ket, nket);
      g.cand_all += nbra * nket;
      cqx::CandLaunch CL{};
      CL.nbra = (int)nbra;
      size_t b = 0;
      HIPCHK(cqx::launch_candidates(CL, nullptr, b, s));  // size query only
      g.scan_bytes = std::max(g.scan_bytes, b);
    }
  return g;
}

// Quartet-list entries of one batch: list_cap, or more when one bra pair
// alone has more candidates (a piece holds at least one bra pair), and never
// more than every candidate.
long long list_entries_bound(const GroupPlan &g, long long list_cap) {
  return std::max<long long>(1, std::min(std::max(list_cap, g.max_ket), g.cand_all));
}

// Device bytes of the per-call workspace a context holds (every buffer a
// contraction may grow).
size_t workspace_bytes(const cqx_ctx *c) {
  size_t b = 0;
  auto add = [&](const auto &buf) { b += buf.n * sizeof(*buf.p); };
  add(c->d_X); add(c->d_AX); add(c->d_Bm); add(c->d_Bg); add(c->d_planes); add(c->d_acc);
  add(c->d_her); add(c->d_cand); add(c->d_ket_lo); add(c->d_scan_tmp); add(c->d_lists);
  add(c->d_lstart); add(c->d_molmax); add(c->d_molplanes); add(c->d_quart); add(c->d_count);
  add(c->d_psum); add(c->d_tidx); add(c->d_tot); add(c->d_acc_tune); add(c->d_molplanes_tune);
  add(c->d_glists); add(c->d_glstart); add(c->d_gidx); add(c->d_gchk);
  add(c->d_tiles); add(c->d_tcount); add(c->d_tpsum); add(c->d_tll); add(c->d_tstart);
  for (int k = 0; k < kNumClass; ++k) add(c->d_pn[k]);
  return b;
}

// Upload a host array into a workspace view (the context's buffer, grown
// on demand, or a piece of the scratch region).
template <class T>
T *ws_upload(DevBuf<T> &buf, Arena *scr, const std::vector<T> &v, hipStream_t s, const char *what) {
  T *p;
  if (scr) {
    p = scr->take<T>(v.size(), what);
  } else {
    buf.ensure(std::max<size_t>(v.size(), 1));
    p = buf.p;
  }
  if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
  return p;
}
template <class T>
T *ws_take(DevBuf<T> &buf, Arena *scr, size_t n, const char *what) {
  if (scr) return scr->take<T>(n, what);
  buf.ensure(std::max<size_t>(n, 1));
  return buf.p;
}

// The direct J/K build of a contraction list (directScaffoldNew,
// direct.hpp:1412-2123).  scr == nullptr: the workspace lives in the
// context's own buffers, grown on demand (and a build may be recorded into
// a graph and replayed); otherwise every per-call buffer is carved from
// the scratch region, X / AX are device pointers, AX is accumulated in
// place, and nothing is allocated, recorded or created.
int contract(cqx_ctx *c, int screen, int ncon, const cqx_contraction *list, int is_complex, int mem, Arena *scr) {
  if (!c || ncon < 0 || (ncon && !list)) return fail(c, CQX_ERR_ARG, "cqx_two_body_contract: bad arguments");
  if (ncon == 0) return fail(c, CQX_ERR_ARG, "Nothing specified to compute");
  for (int i = 0; i < ncon; ++i)
    if (!list[i].X || !list[i].AX || (list[i].type != CQX_COULOMB && list[i].type != CQX_EXCHANGE))
      return fail(c, CQX_ERR_ARG, "invalid contraction " + std::to_string(i));
  const int nmol = c->nmol;
  if (ncon % nmol)
    return fail(c, CQX_ERR_ARG, "a batched context takes the same number of contractions for every molecule");
  const int nper = ncon / nmol;
  for (int m = 1; m < nmol; ++m)
    for (int i = 0; i < nper; ++i)
      if (list[m * nper + i].type != list[i].type || (list[m * nper + i].her != 0) != (list[i].her != 0))
        return fail(c, CQX_ERR_ARG, "batched molecules must share the contraction types and HER flags");
  if (scr && mem != CQX_DEVICE) return fail(c, CQX_ERR_ARG, "the allocation-free contraction takes device pointers");
  const auto t_start = std::chrono::steady_clock::now();
  c->stats = cqx_stats{};
  const std::string key = contraction_key(screen, is_complex, ncon, list);
  try {
    HIPCHK(hipSetDevice(c->device));
    if (!scr && c->use_graph && c->graph_exec && c->graph_key == key) {
      if (replay_graph(c, ncon, list, is_complex, mem)) {
        c->stats.ms_total =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        c->stats.workspace_bytes = (long long)workspace_bytes(c);
        return CQX_OK;
      }
      c->stats = cqx_stats{};  // capacities exceeded: build eagerly and record again
    }
    // an eager build may reallocate the workspace a recorded graph refers to
    if (!scr) drop_graph(c);
    hipStream_t s = c->stream;
    const cqx::Basis &BS = c->basis;
    const size_t ew = is_complex ? 2 : 1;  // doubles per element
    // per-molecule sizes
    std::vector<int> mnb(nmol), mnbc(nmol), mns(nmol);
    std::vector<size_t> mnb2(nmol), mnbc2(nmol);
    for (int m = 0; m < nmol; ++m) {
      mnb[m] = BS.mol_bf0[m + 1] - BS.mol_bf0[m];
      mnbc[m] = BS.mol_bfc0[m + 1] - BS.mol_bfc0[m];
      mns[m] = BS.mol_sh0[m + 1] - BS.mol_sh0[m];
      mnb2[m] = (size_t)mnb[m] * mnb[m];
      mnbc2[m] = (size_t)mnbc[m] * mnbc[m];
    }
    float eri_ms_total = 0.0f;
    bool pure_d = false;
    for (auto &sh : BS.shells)
      if (sh.l > 1 && sh.pure) pure_d = true;
    cqx_ctx::Ws &w = c->w;
    size_t scratch_peak = 0;
    // groups of at most kMaxMat contractions / kMaxPlanes planes (per molecule)
    int start = 0;
    while (start < nper) {
      int nplane = 0;
      const int nmat = group_size(list + start, nper - start, is_complex, &nplane);
      const cqx_contraction *L = list + start;  // molecule 0's contractions of the group
      const GroupPlan gp = plan_group(c, L, nmat, is_complex, s);
      if (scr) scr->used = 0;  // each group reuses the whole region
      // densities (and, when allocating, the outputs) on the device:
      // molecule m's nmat matrices at xoff[m]
      std::vector<size_t> xoff(nmol + 1, 0);
      for (int m = 0; m < nmol; ++m) xoff[m + 1] = xoff[m] + (size_t)nmat * mnb2[m] * ew;
      w.X = ws_take(c->d_X, scr, gp.x, "the staged densities");
      w.AX = scr ? nullptr : ws_take(c->d_AX, scr, gp.x, "the staged outputs");
      const hipMemcpyKind hk = mem == CQX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
      for (int m = 0; m < nmol; ++m)
        for (int j = 0; j < nmat; ++j) {
          const cqx_contraction &cc = list[m * nper + start + j];
          const size_t o = xoff[m] + (size_t)j * mnb2[m] * ew, n = sizeof(double) * mnb2[m] * ew;
          HIPCHK(hipMemcpyAsync(w.X + o, cc.X, n, hk, s));
          if (!scr) HIPCHK(hipMemcpyAsync(w.AX + o, cc.AX, n, hk, s));
        }
      // --- plan: every upload and allocation of the build comes first, so
      // that the launches below can also be recorded into a graph ---
      std::vector<int> her(nmat);
      for (int j = 0; j < nmat; ++j) her[j] = L[j].her ? 1 : 0;
      w.her = ws_upload(c->d_her, scr, her, s, "the HER flags");
      const long long nbl = c->mol_boff[nmol];
      w.Bm = ws_take(c->d_Bm, scr, gp.bm, "the shell-block norms");
      w.Bg = ws_take(c->d_Bg, scr, gp.bg, "the shell-block norms");
      w.molmax = ws_take(c->d_molmax, scr, (size_t)nmol, "the molecule norms");
      for (int k = 0; k < kNumClass; ++k)
        w.pn[k] = c->cls[k].empty() ? nullptr : ws_take(c->d_pn[k], scr, gp.pn[k], "the pair norms");
      cqx::PlaneSet ps{};
      // planes (one per real part), ordered by kind; slot = position, so the
      // planes that one digestion term visits sit in consecutive slots
      struct PlaneDesc { int m, part, kind; double sign; };
      std::vector<PlaneDesc> pd;
      for (int j = 0; j < nmat; ++j) {
        const bool two = is_complex && !(L[j].her && L[j].type == CQX_COULOMB);
        int kind = L[j].type == CQX_COULOMB ? (L[j].her ? cqx::kJHer : cqx::kJNonHer)
                                            : (L[j].her ? cqx::kKHer : cqx::kKNonHer);
        for (int k = 0; k < (two ? 2 : 1); ++k)
          pd.push_back({j, k, kind, (k == 1 && L[j].her) ? -1.0 : 1.0});  // conj of Im X
      }
      // the kind with the most planes first, so that its slots start on an
      // interleave-group boundary and the lanes digesting one element
      // touch as few 64-byte segments as possible
      int krank[4];
      const int *kcnt = gp.kcnt;
      for (int k = 0; k < 4; ++k) {
        krank[k] = 0;
        for (int j = 0; j < 4; ++j)
          if (kcnt[j] > kcnt[k] || (kcnt[j] == kcnt[k] && j < k)) ++krank[k];
      }
      std::stable_sort(pd.begin(), pd.end(),
                       [&](const PlaneDesc &x, const PlaneDesc &y) { return krank[x.kind] < krank[y.kind]; });
      // Cartesian density planes and accumulators: per molecule, one
      // interleaved array per kind with stride kps[k] (its plane count
      // rounded up to a power of two)
      int kpsum = 0;
      for (int k = 0; k < 4; ++k) {
        ps.kps[k] = gp.kps[k];
        if (kcnt[k]) kpsum += ps.kps[k];
      }
      std::vector<size_t> poff(nmol + 1, 0);
      for (int m = 0; m < nmol; ++m) poff[m + 1] = poff[m] + (size_t)kpsum * mnbc2[m];
      w.planes = ws_take(c->d_planes, scr, gp.planes, "the Cartesian density planes");
      w.acc = ws_take(c->d_acc, scr, gp.planes, "the Cartesian accumulators");
      std::vector<cqx::MolPlanes> mp(nmol);
      for (int m = 0; m < nmol; ++m) {
        size_t o = poff[m];
        for (int k = 0; k < 4; ++k) {
          mp[m].kx[k] = w.planes + o;
          mp[m].ka[k] = w.acc + o;
          if (kcnt[k]) o += (size_t)ps.kps[k] * mnbc2[m];
        }
        mp[m].nbc = mnbc[m];
      }
      w.molplanes = ws_upload(c->d_molplanes, scr, mp, s, "the molecule plane table");
      ps.mols = w.molplanes;
      ps.m0 = mp[0];
      ps.nmol = nmol;
      std::vector<int> slot_re(nmat, -1), slot_im(nmat, -1), pidx(nplane);
      for (int k = 0; k < 4; ++k) { ps.kstart[k] = 0; ps.kcount[k] = 0; }
      for (int p = 0; p < (int)pd.size(); ++p) {
        cqx::Plane &P = ps.p[p];
        const int k = pd[p].kind;
        pidx[p] = ps.kcount[k];  // position within its kind
        P.kind = k;
        P.sign = pd[p].sign;
        P.mat = pd[p].m;
        (pd[p].part ? slot_im : slot_re)[pd[p].m] = p;
        if (ps.kcount[k]++ == 0) ps.kstart[k] = p;
        ps.mats |= (unsigned long long)pd[p].m << (4 * p);
        if (P.sign < 0.0) ps.negs |= 1u << p;
      }
      ps.n = nplane;
      // Cartesian planes of every molecule (real and imaginary planes of a
      // contraction share a kind, hence a stride)
      auto mol_map = [&](int m) {
        cqx::CartMap M = c->cmap;
        M.cart_shell += BS.mol_bfc0[m];
        M.cart_comp += BS.mol_bfc0[m];
        M.bf_shell += BS.mol_bf0[m];
        M.bf_comp += BS.mol_bf0[m];
        M.nb = mnb[m];
        M.nbc = mnbc[m];
        M.ns = mns[m];
        M.bf0 = BS.mol_bf0[m];
        M.bfc0 = BS.mol_bfc0[m];
        return M;
      };
      int nJ = 0, nK = 0;
      for (int j = 0; j < nmat; ++j) (L[j].type == CQX_COULOMB ? nJ : nK)++;
      // candidate lists per class pair (bra class >= ket class): bra i of
      // molecule m meets a prefix of molecule m's kets (Q descending); the
      // counts and offsets are computed on the device
      // tile class pairs: the lane-group class pairs with a sub-block table,
      // when every plane of the group is Hermitian and the build is neither
      // recorded nor allocation-free; their quartets come from tile lists
      const bool tiles_ok = c->use_tiles && !scr && !c->use_graph && !c->tile_cps.empty() &&
                            gp.kcnt[cqx::kJNonHer] == 0 && gp.kcnt[cqx::kKNonHer] == 0 &&
                            (c->use_tiles == 1 || (gp.kcnt[cqx::kJHer] <= 1 && gp.kcnt[cqx::kKHer] <= 1));
      std::vector<const cqx_ctx::TileCP *> tcps;
      auto lane_group = [&](int cb, int ck) {
        int la_, lb_, lc_, ld_;
        class_ls(cb, la_, lb_);
        class_ls(ck, lc_, ld_);
        const int ncq = cqx::ncart(la_) * cqx::ncart(lb_) * cqx::ncart(lc_) * cqx::ncart(ld_);
        const int idx = cqx::fast_class_index(la_, lb_, lc_, ld_);
        return ncq <= c->fast_max_nc && idx >= 0 && ((c->tile_classes >> idx) & 1);
      };
      if (tiles_ok)
        for (const auto &tc : c->tile_cps)
          if (lane_group(tc.cb, tc.ck)) tcps.push_back(&tc);
      auto is_tile_cp = [&](int cb, int ck) {
        for (const auto *tc : tcps)
          if (tc->cb == cb && tc->ck == ck) return true;
        return false;
      };
      struct CP { int cb, ck, nbra; long long ncand; size_t coff; std::vector<long long> off; };
      std::vector<CP> cps;
      size_t cand_total = 0;
      for (int cb = 0; cb < kNumClass; ++cb)
        for (int ck = 0; ck <= cb; ++ck) {
          if (c->cls[cb].empty() || c->cls[ck].empty() || is_tile_cp(cb, ck)) continue;
          CP cp{cb, ck, (int)c->cls[cb].size(), 0, cand_total, {}};
          cand_total += (size_t)cp.nbra + 1;
          cps.push_back(std::move(cp));
        }
      w.cand = ws_take(c->d_cand, scr, cand_total, "the candidate offsets");
      w.ket_lo = ws_take(c->d_ket_lo, scr, cand_total, "the candidate offsets");
      w.scan = ws_take(c->d_scan_tmp, scr, gp.scan_bytes, "the scan workspace");
      w.scan_bytes = std::max<size_t>(gp.scan_bytes, 1);
      std::vector<cqx::CandLaunch> cls_launch(cps.size());
      for (size_t k = 0; k < cps.size(); ++k) {
        const auto &cp = cps[k];
        cqx::CandLaunch &CL = cls_launch[k];
        CL = cqx::CandLaunch{};
        CL.bra = c->d_pairs[cp.cb].p;
        CL.ket = c->d_pairs[cp.ck].p;
        CL.nbra = cp.nbra;
        CL.bra_base = 0;
        CL.seg = c->d_seg.p + (size_t)cp.ck * (nmol + 1);
        CL.molmax = w.molmax;
        CL.thresh = c->thresh;
        CL.screen = screen;
        CL.same_class = cp.cb == cp.ck;
        CL.off = w.cand + cp.coff;
        CL.klo = w.ket_lo + cp.coff;
      }
      std::vector<long long> tidx(cps.size());
      for (size_t k = 0; k < cps.size(); ++k) tidx[k] = (long long)(cps[k].coff + cps[k].nbra);
      if (!cps.empty()) {
        w.tidx = ws_upload(c->d_tidx, scr, tidx, s, "the candidate totals");
        w.tot = ws_take(c->d_tot, scr, cps.size(), "the candidate totals");
      }
      // AX of the outputs: the staged copies, or the caller's own device
      // matrices for the allocation-free entry point
      auto ax_of = [&](int m, int j) -> double * {
        if (scr) return static_cast<double *>(list[m * nper + start + j].AX);
        return w.AX + xoff[m] + (size_t)j * mnb2[m] * ew;
      };
      // --- launches up to the candidate counts: shell-block norms, per
      // molecule, and each molecule's largest one; the pairs' own norms;
      // cleared accumulators; Cartesian density planes; candidate offsets ---
      auto front = [&]() {
        for (int m = 0; m < nmol; ++m)
          HIPCHK(cqx::launch_block_norms(w.X + xoff[m], nmat, is_complex, w.her, mnb[m], mns[m],
                                         c->d_sh_bf.p + BS.mol_sh0[m], BS.mol_bf0[m], c->d_sh_nf.p + BS.mol_sh0[m],
                                         w.Bm + (size_t)c->mol_boff[m] * nmat, w.Bg + c->mol_boff[m], s));
        HIPCHK(cqx::launch_mol_max(w.Bg, c->d_mol_boff.p, c->d_mol_ns.p, nmol, w.molmax, s));
        for (int k = 0; k < kNumClass; ++k)
          if (!c->cls[k].empty())
            HIPCHK(cqx::launch_pair_norms(c->d_scr[k].p, (int)c->cls[k].size(), w.Bg, w.Bm, c->d_mol_boff.p,
                                          c->d_mol_ns.p, nmat, w.pn[k], s));
        HIPCHK(cqx::launch_zero(w.acc, poff[nmol], s));
        for (int m = 0; m < nmol; ++m) {
          const cqx::CartMap M = mol_map(m);
          for (int j = 0; j < nmat; ++j) {
            const int pr = slot_re[j], pi = slot_im[j];
            double *xre = const_cast<double *>(mp[m].kx[pd[pr].kind]) + pidx[pr];
            double *xim = pi >= 0 ? const_cast<double *>(mp[m].kx[pd[pi].kind]) + pidx[pi] : nullptr;
            HIPCHK(cqx::launch_to_cart(M, w.X + xoff[m] + (size_t)j * mnb2[m] * ew, is_complex, xre, xim,
                                       ps.kps[pd[pr].kind], s));
          }
        }
        for (auto &CL : cls_launch) {
          size_t b = w.scan_bytes;
          HIPCHK(cqx::launch_candidates(CL, w.scan, b, s));
        }
      };
      front();
      const bool mixed = c->fp32_cut > 0.0;
      const int qmul = mixed ? 2 : 1;  // quartet lists per class pair
      // tile lists of the tile class pairs: one launch over their sub-blocks;
      // each sub-block gives at most one tile per precision
      const int nt = (int)tcps.size();
      std::vector<unsigned long long> tcount(3 * (size_t)nt);  // quartets FP64, FP32, candidates
      std::vector<double> tpsum(2 * (size_t)nt);
      std::vector<long long> tcap_off(nt + 1, 0);
      std::vector<cqx::TileListLaunch> tll(nt);  // alive until the synchronisation below
      cqx::Tile *tiles = nullptr;
      if (nt) {
        for (int k = 0; k < nt; ++k) tcap_off[k + 1] = tcap_off[k] + tcps[k]->n;
        tiles = ws_take(c->d_tiles, nullptr, (size_t)tcap_off[nt] * qmul, "the tile lists");
        unsigned long long *tcnt = ws_take(c->d_tcount, nullptr, 3 * (size_t)nt, "the tile counters");
        double *tps = ws_take(c->d_tpsum, nullptr, 2 * (size_t)nt, "the tile counters");
        for (int k = 0; k < nt; ++k) {
          const auto &tc = *tcps[k];
          cqx::TileListLaunch &T = tll[k];
          T = cqx::TileListLaunch{};
          T.sub = c->d_sub.p + tc.off;
          T.bra = c->d_scr[tc.cb].p;
          T.ket = c->d_scr[tc.ck].p;
          T.bmem = c->d_members[tc.cb].p;
          T.kmem = c->d_members[tc.ck].p;
          T.bra_pn = w.pn[tc.cb];
          T.ket_pn = w.pn[tc.ck];
          T.Bg = w.Bg;
          T.Bm = w.Bm;
          T.mol_boff = c->d_mol_boff.p;
          T.mol_ns = c->d_mol_ns.p;
          T.molmax = w.molmax;
          T.nmat = nmat;
          T.thresh = c->thresh;
          T.screen = screen;
          T.out = tiles + tcap_off[k];
          T.nquart = tcnt + 3 * k;
          T.ncand = tcnt + 3 * k + 2;
          T.prim_sum = tps + 2 * k;
          T.cut32 = mixed ? c->fp32_cut : 0.0;
          T.out32 = mixed ? tiles + tcap_off[nt] + tcap_off[k] : nullptr;
          T.nquart32 = tcnt + 3 * k + 1;
          T.prim_sum32 = tps + 2 * k + 1;
        }
        c->d_tll.upload(tll, s);
        c->d_tstart.upload(tcap_off, s);
        HIPCHK(cqx::launch_zero(reinterpret_cast<double *>(tcnt), 3 * (size_t)nt, s));
        HIPCHK(cqx::launch_zero(tps, 2 * (size_t)nt, s));
        HIPCHK(cqx::launch_tile_lists(c->d_tll.p, c->d_tstart.p, nt, tcap_off[nt], s));
        HIPCHK(hipMemcpyAsync(tcount.data(), tcnt, sizeof(unsigned long long) * tcount.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(tpsum.data(), tps, sizeof(double) * tpsum.size(), hipMemcpyDeviceToHost, s));
      }
      // every class pair's candidate total (the last entry of its offsets),
      // gathered on the device and read back in one copy
      std::vector<long long> totals(cps.size());
      if (!cps.empty()) {
        HIPCHK(cqx::launch_gather(w.cand, w.tidx, (int)cps.size(), w.tot, s));
        HIPCHK(hipMemcpyAsync(totals.data(), w.tot, sizeof(long long) * cps.size(), hipMemcpyDeviceToHost, s));
      }
      HIPCHK(hipStreamSynchronize(s));
      // entries of one list batch: the option, or what the scratch region
      // leaves after this group's fixed workspace (less a reserve for the
      // per-batch tables)
      long long list_cap = c->list_cap;
      if (scr) {
        const size_t reserve = 4096 + (cps.size() + 64) * (2 * sizeof(cqx::ListLaunch) + 64);
        const size_t left = scr->left() > reserve ? scr->left() - reserve : 0;
        list_cap = std::min<long long>(list_cap, (long long)(left / (sizeof(Quartet) * qmul)));
        if (list_cap < 1) throw Unsupported("cqx_two_body_contract_noalloc: the scratch region leaves no room for quartet lists");
      }
      long long total = 0;
      for (size_t k = 0; k < cps.size(); ++k) {
        cps[k].ncand = totals[k];
        total += totals[k];
        if (totals[k] > list_cap) {  // split into bra ranges on the host
          cps[k].off.resize((size_t)cps[k].nbra + 1);
          HIPCHK(hipMemcpyAsync(cps[k].off.data(), w.cand + cps[k].coff, sizeof(long long) * cps[k].off.size(),
                                hipMemcpyDeviceToHost, s));
          HIPCHK(hipStreamSynchronize(s));
        }
      }
      c->stats.n_candidates += total;
      for (int k = 0; k < nt; ++k) c->stats.n_candidates += (long long)tcount[3 * k + 2];
      // autotuned shapes: key of a class and precision, and the value tag of
      // a wave shape for the classes built in both families
      constexpr int kB = cqx::kMaxLDev + 1;
      constexpr int kWaveTag = 64;
      auto class_key = [&](int la, int lb, int lc, int ld, int f32) {
        return ((((la * kB + lb) * kB + lc) * kB + ld) << 1) | (f32 ? 1 : 0);
      };
      // classes with both a lane-group and a wave kernel, whose family the
      // autotuner picks (without a tuned choice: the lane-group kernels up to
      // fast_max_nc components)
      auto dual = [&](int la, int lb, int lc, int ld) {
        return c->use_wave && cqx::eri_lane_group_dual(la, lb, lc, ld) && cqx::eri_wave_available(la, lb, lc, ld);
      };
      // kernel family of a class: 0 lane-group, 1 wave-per-quartet, 2 generic;
      // f32 < 0 asks for the family without a tuned choice
      auto route = [&](int la, int lb, int lc, int ld, int f32 = -1) {
        const int ncq = cqx::ncart(la) * cqx::ncart(lb) * cqx::ncart(lc) * cqx::ncart(ld);
        if (f32 >= 0 && dual(la, lb, lc, ld)) {
          auto it = c->tuned_lanes.find(class_key(la, lb, lc, ld, f32));
          if (it != c->tuned_lanes.end()) return it->second >= kWaveTag ? 1 : 0;
        }
        if (ncq <= c->fast_max_nc && cqx::eri_lane_group_built(la, lb, lc, ld)) return 0;
        if (c->use_wave && cqx::eri_wave_available(la, lb, lc, ld)) return 1;
        return 2;
      };
      // classes with an FP32 kernel: the lane-group and wave kernels of the
      // s/p/d classes (quartets with an f or g shell stay in FP64)
      auto fp32_class = [&](int la, int lb, int lc, int ld) {
        return route(la, lb, lc, ld) < 2 && std::max(std::max(la, lb), std::max(lc, ld)) <= 2;
      };
      // the ERI launch of class pair k's bras from i0 on, in one precision
      // (the count, list and lane choice are set by the caller)
      auto eri_base = [&](const CP &cp, int i0, bool f32) {
        cqx::EriLaunch E{};
        class_ls(cp.cb, E.la, E.lb);
        class_ls(cp.ck, E.lc, E.ld);
        E.bra = c->d_pairs[cp.cb].p + i0;
        E.ket = c->d_pairs[cp.ck].p;
        E.prims = c->d_prims.p;
        E.primc32 = c->d_primc32.p;
        E.rys = c->d_rys.p;
        E.rys32 = c->d_rys32.p;
        E.fp32 = f32 ? 1 : 0;
        E.force_g = c->fast_lanes;
        E.rys_lds = c->rys_lds;
        E.planes = ps;
        E.store_mode = 0;
        return E;
      };
      // autotuned shape of a launch, per class pair and precision: the lanes
      // per quartet of a lane-group kernel, the quartets per wave of a wave
      // kernel
      // (a class built in both families keeps its choice under the plain key,
      // tagged kWaveTag + QT for a wave shape)
      auto tune_key = [&](const cqx::EriLaunch &E) {
        const int k = class_key(E.la, E.lb, E.lc, E.ld, E.fp32);
        if (dual(E.la, E.lb, E.lc, E.ld)) return k;
        return route(E.la, E.lb, E.lc, E.ld) == 1 ? k + 2 * kB * kB * kB * kB : k;
      };
      auto apply_tuned = [&](cqx::EriLaunch &E) {
        const int r = route(E.la, E.lb, E.lc, E.ld, E.fp32);
        auto it = c->tuned_lanes.find(tune_key(E));
        if (it == c->tuned_lanes.end()) return;
        const int v = it->second >= kWaveTag ? it->second - kWaveTag : it->second;
        if (r == 0 && !c->fast_lanes) E.force_g = v;
        if (r == 1) E.force_qt = v;
      };
      // one list in launches of at most kLaunchPiece quartets, so that no
      // grid reaches 2^32 work-items (limits.hpp), on the kernels of family r
      auto launch_family = [&](const cqx::EriLaunch &E, int r, hipStream_t st) {
        if (r == 2 && E.fp32) throw std::runtime_error("FP32 quartet list for a class without an FP32 kernel");
        // the shape a launch picks for itself is at worst the smallest tile
        const int shape = r == 0 ? (E.force_g > 0 ? E.force_g : 16) : r == 1 ? (E.force_qt > 0 ? E.force_qt : 1) : 1;
        if (!cqx::launch_fits(E.nq, r, shape)) throw Unsupported("an ERI launch beyond 2^32 work-items");
        for (long long b = 0; b < E.nq; b += cqx::kLaunchPiece) {
          cqx::EriLaunch P = E;
          P.quartets = E.quartets + b;
          P.nq = std::min(cqx::kLaunchPiece, E.nq - b);
          P.nq_base = b;
          HIPCHK(r == 0 ? cqx::launch_eri_fast(P, st) : r == 1 ? cqx::launch_eri_wave(P, st) : cqx::launch_eri_team(P, st));
        }
      };
      auto launch_eri = [&](const cqx::EriLaunch &E, hipStream_t st) {
        launch_family(E, route(E.la, E.lb, E.lc, E.ld, E.fp32), st);
      };
      // side streams for concurrent class-pair launches: created outside any
      // capture, forked from the context stream before the launches and
      // joined back into it after them (also inside a recorded build); the
      // allocation-free path launches on the context stream alone
      auto ensure_side = [&](int nst) {
        for (int i = 0; i < nst - 1; ++i)
          if (!c->side[i]) {
            HIPCHK(hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&c->ev_join[i], hipEventDisableTiming));
          }
        if (nst > 1 && !c->ev_fork) HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
      };
      auto fork_side = [&](int nst) {
        if (nst <= 1) return;
        HIPCHK(hipEventRecord(c->ev_fork, s));
        for (int i = 0; i < nst - 1; ++i) HIPCHK(hipStreamWaitEvent(c->side[i], c->ev_fork, 0));
      };
      auto join_side = [&](int nst) {
        for (int i = 0; i < nst - 1; ++i) {
          HIPCHK(hipEventRecord(c->ev_join[i], c->side[i]));
          HIPCHK(hipStreamWaitEvent(s, c->ev_join[i], 0));
        }
      };
      auto stream_of = [&](int launch, int nst) { return launch % nst == 0 ? s : c->side[launch % nst - 1]; };
      // --- back to pure functions, symmetrise, accumulate into AX: the
      // contractions' transforms are independent and are dealt over nst
      // streams like the ERI launches (side streams already created) ---
      auto back = [&](int nst) {
        fork_side(nst);
        int q = 0;
        for (int m = 0; m < nmol; ++m) {
          const cqx::CartMap M = mol_map(m);
          for (int j = 0; j < nmat; ++j) {
            const bool two = is_complex && !(L[j].her && L[j].type == CQX_COULOMB);
            const int pr = slot_re[j], pi = slot_im[j];
            const double *are = mp[m].ka[pd[pr].kind] + pidx[pr];
            const double *aim = two ? mp[m].ka[pd[pi].kind] + pidx[pi] : nullptr;
            HIPCHK(cqx::launch_from_cart(M, are, aim, ps.kps[pd[pr].kind], L[j].her, is_complex, ax_of(m, j),
                                         stream_of(q++, nst)));
          }
        }
        join_side(nst);
      };
      // bra-range pieces of at most list_cap candidates, grouped into batches
      // of at most list_cap candidates: the quartet lists of one batch are
      // built, counted (one host synchronisation) and evaluated before the
      // next, so the list workspace stays bounded for any molecule
      struct Piece { int cp; int i0, i1; long long base, ncand, seg; };
      std::vector<std::vector<Piece>> batches(1);
      long long in_batch = 0, cap_need = 1;
      auto add_piece = [&](int k, int i0, int i1, long long base, long long n) {
        if (n <= 0) return;
        if (in_batch > 0 && in_batch + n > list_cap) {
          batches.emplace_back();
          in_batch = 0;
        }
        batches.back().push_back({k, i0, i1, base, n, in_batch});
        in_batch += n;
        cap_need = std::max(cap_need, in_batch);
      };
      for (int k = 0; k < (int)cps.size(); ++k) {
        const auto &cp = cps[k];
        if (cp.off.empty()) {  // the whole class pair in one piece
          add_piece(k, 0, cp.nbra, 0, cp.ncand);
          continue;
        }
        int i0 = 0;
        while (i0 < cp.nbra) {
          int i1 = i0 + 1;  // at least one bra pair per piece
          while (i1 < cp.nbra && cp.off[i1 + 1] - cp.off[i0] <= list_cap) ++i1;
          add_piece(k, i0, i1, cp.off[i0], cp.off[i1] - cp.off[i0]);
          i0 = i1;
        }
      }
      size_t max_np = 1;
      for (const auto &b : batches) max_np = std::max(max_np, b.size());
      // list workspace: the per-batch tables first, then the lists
      w.count = ws_take(c->d_count, scr, 2 * max_np, "the list counters");
      w.psum = ws_take(c->d_psum, scr, 2 * max_np, "the list counters");
      w.lists = ws_take(c->d_lists, scr, max_np, "the list launch table");
      w.lstart = ws_take(c->d_lstart, scr, max_np + 1, "the list launch table");
      w.quart = ws_take(c->d_quart, scr, (size_t)cap_need * qmul, "the quartet lists");
      // per class pair and precision: quartets and primitive quartets of this
      // build (they size the launches of a recorded graph)
      std::vector<double> cp_count(2 * cps.size(), 0.0), cp_psum(2 * cps.size(), 0.0);
      bool tune_ready = false;
      const bool tuning = c->autotune_pending && !scr;  // the allocation-free path never times
      cqx::PlaneSet ps_tune = ps;
      // scratch accumulators for the timing launches of the autotuner
      auto prepare_tune = [&]() {
        if (tune_ready) return;
        c->d_acc_tune.ensure(std::max<size_t>(poff[nmol], 1));
        std::vector<cqx::MolPlanes> mt = mp;
        for (int m = 0; m < nmol; ++m)
          for (int k = 0; k < 4; ++k) mt[m].ka[k] = c->d_acc_tune.p + (mp[m].ka[k] - w.acc);
        c->d_molplanes_tune.upload(mt, s);
        HIPCHK(hipStreamSynchronize(s));  // mt is uploaded from this frame
        ps_tune = ps;
        ps_tune.mols = c->d_molplanes_tune.p;
        ps_tune.m0 = mt[0];
        tune_ready = true;
      };
      // time a launch's shapes on scratch accumulators (results discarded)
      // and return the fastest
      // Each shape is timed with the launch replicated on the build's
      // streams at once: a build overlaps the launches of several class
      // pairs, and the shape that is fastest alone on the device is not
      // always the fastest beside others.
      auto time_shapes = [&](const std::vector<int> &shapes, const std::function<void(int, hipStream_t)> &launch) {
        float best = 1e30f;
        int best_g = shapes.empty() ? 0 : shapes[0];
        const int nrep = c->nstreams;
        ensure_side(nrep);
        for (int g : shapes) {
          float ms = 0.0f;
          for (int rep = 0; rep < 2; ++rep) {  // the second run is timed
            HIPCHK(hipEventRecord(c->ev_t0, s));
            fork_side(nrep);
            for (int i = 0; i < nrep; ++i) launch(g, stream_of(i, nrep));
            join_side(nrep);
            HIPCHK(hipEventRecord(c->ev_t1, s));
            HIPCHK(hipEventSynchronize(c->ev_t1));
            HIPCHK(hipEventElapsedTime(&ms, c->ev_t0, c->ev_t1));
          }
          if (ms < best) {
            best = ms;
            best_g = g;
          }
        }
        return best_g;
      };
      // ERI + digestion of the tile lists: launched with the first batch's
      // ERI launches (or alone when no quartet list remains), each over its
      // list's capacity (empty tiles exit at once)
      const int nst_all = (tuning || scr) ? 1 : c->nstreams;
      bool tiles_done = nt == 0;
      auto launch_tiles = [&](int nst, int &nl) {
        tiles_done = true;
        for (int f = 0; f < qmul; ++f)
          for (int k = 0; k < nt; ++k) {
            const unsigned long long nq = tcount[3 * k + f];
            if (!nq) continue;
            const auto &tc = *tcps[k];
            cqx::TileEriLaunch T{};
            class_ls(tc.cb, T.la, T.lb);
            class_ls(tc.ck, T.lc, T.ld);
            T.tiles = tiles + (f ? tcap_off[nt] : 0) + tcap_off[k];
            T.ntile = tc.n;
            T.bmem = c->d_members[tc.cb].p;
            T.kmem = c->d_members[tc.ck].p;
            T.bra = c->d_pairs[tc.cb].p;
            T.ket = c->d_pairs[tc.ck].p;
            T.prims = c->d_prims.p;
            T.primc32 = c->d_primc32.p;
            T.rys = c->d_rys.p;
            T.rys32 = c->d_rys32.p;
            T.fp32 = f;
            T.mean_prims = tpsum[2 * k + f] / (double)nq;
            T.force_g = c->fast_lanes;
            T.rys_lds = c->rys_lds;
            T.planes = ps;
            if ((T.ntile + 3) / 4 >= (1ll << 31)) throw Unsupported("a tile launch beyond 2^31 blocks");
            const int key = (((((T.la * 4 + T.lb) * 4 + T.lc) * 4 + T.ld) << 1) | f) + 8192;
            if (!c->fast_lanes) {
              if (tuning && !c->tuned_lanes.count(key)) {
                prepare_tune();
                const int nc = cqx::ncart(T.la) * cqx::ncart(T.lb) * cqx::ncart(T.lc) * cqx::ncart(T.ld);
                std::vector<int> shapes;
                for (int g : {1, 2, 4, 8, 16})
                  if (!(g == 1 && nc > 9)) shapes.push_back(g);
                cqx::TileEriLaunch TT = T;
                TT.planes = ps_tune;
                c->tuned_lanes[key] = time_shapes(shapes, [&](int g, hipStream_t st) {
                  TT.force_g = g;
                  HIPCHK(cqx::launch_eri_tile(TT, st));
                });
              }
              auto it = c->tuned_lanes.find(key);
              if (it != c->tuned_lanes.end()) T.force_g = it->second;
            }
            HIPCHK(cqx::launch_eri_tile(T, stream_of(nl++, nst)));
            const double fl = class_pair_flops(T.la, T.lb, T.lc, T.ld, pure_d, tpsum[2 * k + f], (double)nq,
                                               is_complex ? 2 : 1, nJ, nK);
            c->stats.n_quartets += (long long)nq;
            c->stats.n_prim_quartets_max += (long long)tpsum[2 * k + f];
            c->stats.flops_model += fl;
            if (f) {
              c->stats.n_quartets_f32 += (long long)nq;
              c->stats.flops_model_f32 += fl;
            }
            c->stats.n_launch_eri++;
          }
      };
      for (const auto &batch : batches) {
        const size_t np = batch.size();
        if (!np) continue;
        const size_t nslot = 2 * np;
        // counters cleared by a fill kernel (no memset nodes if the caller captures a graph)
        HIPCHK(cqx::launch_zero(reinterpret_cast<double *>(w.count), nslot, s));
        HIPCHK(cqx::launch_zero(w.psum, nslot, s));
        std::vector<cqx::ListLaunch> lls(np);
        std::vector<long long> lstart(np + 1, 0);
        for (size_t k = 0; k < np; ++k) {
          const Piece &pc = batch[k];
          const auto &cp = cps[pc.cp];
          cqx::ListLaunch &LL = lls[k];
          LL = cqx::ListLaunch{};
          LL.bra = c->d_scr[cp.cb].p + pc.i0;
          LL.ket = c->d_scr[cp.ck].p;
          LL.bra_pn = w.pn[cp.cb] + (size_t)pc.i0 * (nmat + 1);
          LL.ket_pn = w.pn[cp.ck];
          LL.same_class = cp.cb == cp.ck;
          LL.cand_off = w.cand + cp.coff + pc.i0;
          LL.cand_base = pc.base;
          LL.ket_lo = w.ket_lo + cp.coff + pc.i0;
          LL.nbra = pc.i1 - pc.i0;
          LL.ncand = pc.ncand;
          LL.Bg = w.Bg;
          LL.Bm = w.Bm;
          LL.mol_boff = c->d_mol_boff.p;
          LL.mol_ns = c->d_mol_ns.p;
          LL.nmat = nmat;
          LL.thresh = c->thresh;
          LL.screen = screen;
          LL.out = w.quart + pc.seg;
          LL.count = w.count + k;
          LL.prim_sum = w.psum + k;
          LL.cap = pc.ncand;
          LL.cap32 = pc.ncand;
          int la_, lb_, lc_, ld_;
          class_ls(cp.cb, la_, lb_);
          class_ls(cp.ck, lc_, ld_);
          if (mixed && fp32_class(la_, lb_, lc_, ld_)) {
            LL.cut32 = c->fp32_cut;
            LL.out32 = w.quart + cap_need + pc.seg;
            LL.count32 = w.count + np + k;
            LL.prim_sum32 = w.psum + np + k;
          }
          lstart[k + 1] = lstart[k] + (pc.ncand + cqx::kListChunk - 1) / cqx::kListChunk;
        }
        HIPCHK(hipMemcpyAsync(w.lists, lls.data(), sizeof(cqx::ListLaunch) * np, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(w.lstart, lstart.data(), sizeof(long long) * (np + 1), hipMemcpyHostToDevice, s));
        HIPCHK(cqx::launch_quartet_lists(w.lists, w.lstart, (int)np, lstart[np], s));
        std::vector<unsigned long long> counts(nslot);
        std::vector<double> psum(nslot);
        HIPCHK(hipMemcpyAsync(counts.data(), w.count, sizeof(unsigned long long) * nslot, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(psum.data(), w.psum, sizeof(double) * nslot, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        // ERI + digestion: one launch per class pair and precision for every
        // molecule of the batch (and, with the first batch, the tile lists')
        HIPCHK(hipEventRecord(c->ev0, s));
        // autotuning times launches one by one on the context stream
        const int nst = nst_all;
        ensure_side(nst);
        fork_side(nst);
        int nlaunched = 0;
        if (!tiles_done) launch_tiles(nst, nlaunched);
        for (size_t kk = 0; kk < nslot; ++kk) {
          if (!counts[kk]) continue;
          const size_t k = kk % np;
          const bool f32 = kk >= np;
          const Piece &pc = batch[k];
          const auto &cp = cps[pc.cp];
          cp_count[pc.cp + (f32 ? cps.size() : 0)] += (double)counts[kk];
          cp_psum[pc.cp + (f32 ? cps.size() : 0)] += psum[kk];
          cqx::EriLaunch E = eri_base(cp, pc.i0, f32);
          E.quartets = w.quart + (f32 ? cap_need : 0) + pc.seg;
          E.nq = (long long)counts[kk];
          E.mean_prims = psum[kk] / (double)counts[kk];
          const int rt = route(E.la, E.lb, E.lc, E.ld);
          if (rt <= 1) {
            const int key = tune_key(E);
            const bool both = dual(E.la, E.lb, E.lc, E.ld);  // the family is tuned too
            if (tuning && (both || !(rt == 0 && c->fast_lanes)) && !c->tuned_lanes.count(key)) {
              // time each shape on scratch accumulators (results discarded)
              prepare_tune();
              cqx::EriLaunch T = E;
              T.planes = ps_tune;
              const int nc = cqx::ncart(E.la) * cqx::ncart(E.lb) * cqx::ncart(E.lc) * cqx::ncart(E.ld);
              // lane counts of the lane-group kernel (a value 1-16) and quartets
              // per wave of the wave kernel (kWaveTag + QT for a class built in
              // both families, QT otherwise)
              std::vector<int> shapes;
              if (rt == 0 || both) {
                if (c->fast_lanes) shapes.push_back(c->fast_lanes);
                else
                  for (int g : {1, 2, 4, 8, 16})
                    if (!(g == 1 && nc > 9)) shapes.push_back(g);  // one lane per quartet: LDS tile too large
              }
              if (rt == 1 || both)
                for (int qt : {1, 2, 4, 8})
                  if (cqx::eri_wave_qt_built(E.la, E.lb, E.lc, E.ld, E.fp32, qt))
                    shapes.push_back(both ? kWaveTag + qt : qt);
              c->tuned_lanes[key] = time_shapes(shapes, [&](int v, hipStream_t st) {
                const bool wave = both ? v >= kWaveTag : rt == 1;
                if (wave) T.force_qt = both ? v - kWaveTag : v;
                else T.force_g = v;
                launch_family(T, wave ? 1 : 0, st);
              });
            }
            apply_tuned(E);
          }
          launch_eri(E, stream_of(nlaunched++, nst));
          const double fl = class_pair_flops(E.la, E.lb, E.lc, E.ld, pure_d, psum[kk], (double)counts[kk],
                                             is_complex ? 2 : 1, nJ, nK);
          c->stats.n_quartets += (long long)counts[kk];
          c->stats.n_prim_quartets_max += (long long)psum[kk];
          c->stats.flops_model += fl;
          if (f32) {
            c->stats.n_quartets_f32 += (long long)counts[kk];
            c->stats.flops_model_f32 += fl;
          }
          c->stats.n_launch_eri++;
        }
        join_side(nst);
        HIPCHK(hipEventRecord(c->ev1, s));
        HIPCHK(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        eri_ms_total += ms;
      }
      if (!tiles_done) {  // no quartet-list batch: the tile lists' launches alone
        HIPCHK(hipEventRecord(c->ev0, s));
        ensure_side(nst_all);
        fork_side(nst_all);
        int nl = 0;
        launch_tiles(nst_all, nl);
        join_side(nst_all);
        HIPCHK(hipEventRecord(c->ev1, s));
        HIPCHK(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        eri_ms_total += ms;
      }
      ensure_side(nst_all);
      back(nst_all);
      if (!scr) {
        const hipMemcpyKind dk = mem == CQX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        for (int m = 0; m < nmol; ++m)
          for (int j = 0; j < nmat; ++j)
            HIPCHK(hipMemcpyAsync(list[m * nper + start + j].AX, w.AX + xoff[m] + (size_t)j * mnb2[m] * ew,
                                  sizeof(double) * mnb2[m] * ew, dk, s));
      }
      HIPCHK(hipStreamSynchronize(s));
      if (scr) scratch_peak = std::max(scratch_peak, scr->used);
      // Record the same build into a graph for the next calls (option
      // graph; one contraction group and one list batch).  Capacities are
      // graph_margin times this build's own sizes: each class pair's
      // candidates (the list kernel's range) and each list (class pair and
      // precision; the range its ERI launch covers).  The list kernel reads
      // the candidate totals and the ERI kernels the list counts from the
      // device, and a replay that exceeds a capacity is discarded.
      const size_t ncp = cps.size();
      std::vector<long long> gcap(ncp), glst(ncp + 1, 0), lcap(2 * ncp, 0), lseg(2 * ncp + 1, 0);
      for (size_t k = 0; k < ncp; ++k) {
        const long long want = (long long)std::ceil((double)cps[k].ncand * c->graph_margin) + cqx::kListChunk;
        gcap[k] = (want + cqx::kListChunk - 1) / cqx::kListChunk * cqx::kListChunk;
        glst[k + 1] = glst[k] + gcap[k] / cqx::kListChunk;
        int la_, lb_, lc_, ld_;
        class_ls(cps[k].cb, la_, lb_);
        class_ls(cps[k].ck, lc_, ld_);
        lcap[k] = (long long)std::ceil(cp_count[k] * c->graph_margin) + 1024;
        if (mixed && fp32_class(la_, lb_, lc_, ld_))
          lcap[ncp + k] = (long long)std::ceil(cp_count[ncp + k] * c->graph_margin) + 1024;
      }
      for (size_t kk = 0; kk < 2 * ncp; ++kk) lseg[kk + 1] = lseg[kk] + lcap[kk];
      if (!scr && c->use_graph && start == 0 && nmat == nper && batches.size() == 1 && ncp > 0 &&
          lseg[2 * ncp] <= c->list_cap * qmul) {
        c->d_quart.ensure((size_t)lseg[2 * ncp]);
        c->d_count.ensure(2 * ncp);
        c->d_psum.ensure(2 * ncp);
        w.quart = c->d_quart.p;
        w.count = c->d_count.p;
        w.psum = c->d_psum.p;
        std::vector<cqx::ListLaunch> gl(ncp);
        std::vector<long long> gidx(ncp);
        std::vector<cqx_ctx::GraphSlot> slots(2 * ncp, cqx_ctx::GraphSlot{-1, 0, 0, 0, 0});
        for (size_t k = 0; k < ncp; ++k) {
          const auto &cp = cps[k];
          cqx::ListLaunch &LL = gl[k];
          LL = cqx::ListLaunch{};
          LL.bra = c->d_scr[cp.cb].p;
          LL.ket = c->d_scr[cp.ck].p;
          LL.bra_pn = w.pn[cp.cb];
          LL.ket_pn = w.pn[cp.ck];
          LL.same_class = cp.cb == cp.ck;
          LL.cand_off = w.cand + cp.coff;
          LL.cand_base = 0;
          LL.ket_lo = w.ket_lo + cp.coff;
          LL.nbra = cp.nbra;
          LL.ncand = gcap[k];
          LL.ncand_dev = w.cand + cp.coff + cp.nbra;
          LL.Bg = w.Bg;
          LL.Bm = w.Bm;
          LL.mol_boff = c->d_mol_boff.p;
          LL.mol_ns = c->d_mol_ns.p;
          LL.nmat = nmat;
          LL.thresh = c->thresh;
          LL.screen = screen;
          LL.out = w.quart + lseg[k];
          LL.count = w.count + k;
          LL.prim_sum = w.psum + k;
          LL.cap = lcap[k];
          int la_, lb_, lc_, ld_;
          class_ls(cp.cb, la_, lb_);
          class_ls(cp.ck, lc_, ld_);
          slots[k] = {la_, lb_, lc_, ld_, 0};
          if (mixed && fp32_class(la_, lb_, lc_, ld_)) {
            LL.cut32 = c->fp32_cut;
            LL.out32 = w.quart + lseg[ncp + k];
            LL.cap32 = lcap[ncp + k];
            LL.count32 = w.count + ncp + k;
            LL.prim_sum32 = w.psum + ncp + k;
            slots[ncp + k] = {la_, lb_, lc_, ld_, 1};
          }
          gidx[k] = (long long)(cp.coff + cp.nbra);
        }
        c->d_glists.upload(gl, s);
        c->d_glstart.upload(glst, s);
        c->d_gidx.upload(gidx, s);
        c->d_gchk.ensure(ncp);
        const int nst = c->nstreams;
        ensure_side(nst);
        HIPCHK(hipStreamSynchronize(s));  // every upload done before recording
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        bool forked = false;
        try {
          front();
          HIPCHK(cqx::launch_zero(reinterpret_cast<double *>(w.count), 2 * ncp, s));
          HIPCHK(cqx::launch_zero(w.psum, 2 * ncp, s));
          HIPCHK(cqx::launch_quartet_lists(c->d_glists.p, c->d_glstart.p, (int)ncp, glst[ncp], s));
          fork_side(nst);
          forked = true;
          int nlaunched = 0;
          for (size_t kk = 0; kk < 2 * ncp; ++kk) {
            if (slots[kk].la < 0) continue;
            const size_t k = kk % ncp;
            const bool f32 = kk >= ncp;
            cqx::EriLaunch E = eri_base(cps[k], 0, f32);
            E.quartets = w.quart + lseg[kk];
            E.nq = lcap[kk];
            E.nq_dev = w.count + kk;
            const double n0 = cp_count[kk];
            E.mean_prims = n0 > 0.0 ? cp_psum[kk] / n0 : 1.0;
            apply_tuned(E);
            launch_eri(E, stream_of(nlaunched++, nst));
          }
          join_side(nst);
          forked = false;
          back(nst);
          HIPCHK(cqx::launch_gather(w.cand, c->d_gidx.p, (int)ncp, c->d_gchk.p, s));
        } catch (...) {
          if (forked)  // an unjoined side stream would leave the capture open
            for (int i = 0; i < nst - 1; ++i) {
              (void)hipEventRecord(c->ev_join[i], c->side[i]);
              (void)hipStreamWaitEvent(s, c->ev_join[i], 0);
            }
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        HIPCHK(hipStreamEndCapture(s, &g));
        hipGraphExec_t ge = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
          (void)hipGraphDestroy(g);
          HIPCHK(ie);
        }
        c->graph = g;
        c->graph_exec = ge;
        c->graph_key = key;
        c->g_nmat = nmat;
        c->g_nJ = nJ;
        c->g_nK = nK;
        c->g_pure_d = pure_d;
        c->g_xoff.assign(xoff.begin(), xoff.end());
        c->g_mnb2 = mnb2;
        c->g_cap = gcap;
        c->g_lcap = lcap;
        c->g_slots = slots;
        c->stats.graph = 1;
      }
      start += nmat;
    }
    c->stats.ms_eri = eri_ms_total;
    if (!scr) c->autotune_pending = false;
    c->ws_scratch_peak = scratch_peak;
    c->stats.workspace_bytes = scr ? (long long)scratch_peak : (long long)workspace_bytes(c);
  } catch (const Unsupported &e) {
    return fail(c, CQX_ERR_UNSUPPORTED, e.what());
  } catch (const std::exception &e) {
    return fail(c, CQX_ERR_HIP, e.what());
  }
  c->stats.ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return CQX_OK;
}

}  // namespace

extern "C" {

int cqx_two_body_contract(cqx_ctx *c, int screen, int ncon, const cqx_contraction *list, int is_complex,
                          int mem) {
  return contract(c, screen, ncon, list, is_complex, mem, nullptr);
}

int cqx_two_body_contract_noalloc(cqx_ctx *c, int screen, int ncon, const cqx_contraction *list, int is_complex) {
  if (!c || ncon <= 0 || !list) return contract(c, screen, ncon, list, is_complex, CQX_DEVICE, nullptr);
  if (!c->scratch) return fail(c, CQX_ERR_HIP, "cqx_two_body_contract_noalloc: no scratch region on this device");
  if (ncon % c->nmol)
    return fail(c, CQX_ERR_ARG, "a batched context takes the same number of contractions for every molecule");
  const int nper = ncon / c->nmol;
  // every group's fixed workspace (the pieces contract() carves before its
  // quartet lists, 256-byte aligned) plus a reserve and one list entry
  try {
    HIPCHK(hipSetDevice(c->device));
    const int qmul = c->fp32_cut > 0.0 ? 2 : 1;
    int start = 0;
    while (start < nper) {
      const int nmat = group_size(list + start, nper - start, is_complex, nullptr);
      const GroupPlan g = plan_group(c, list + start, nmat, is_complex, c->stream);
      size_t need = 0;
      auto take = [&](size_t bytes) { need = ((need + 255) & ~(size_t)255) + std::max<size_t>(bytes, 1); };
      take(g.x * sizeof(double));
      take(nmat * sizeof(int));
      take(g.bm * sizeof(double));
      take(g.bg * sizeof(double));
      take(c->nmol * sizeof(double));
      for (int k = 0; k < kNumClass; ++k)
        if (!c->cls[k].empty()) take(g.pn[k] * sizeof(double));
      take(g.planes * sizeof(double));
      take(g.planes * sizeof(double));
      take(c->nmol * sizeof(cqx::MolPlanes));
      take(g.cand * sizeof(long long));
      take(g.cand * sizeof(int));
      take(g.scan_bytes);
      if (g.ncp) {
        take(g.ncp * sizeof(long long));
        take(g.ncp * sizeof(long long));
      }
      need += 4096 + (g.ncp + 64) * (2 * sizeof(cqx::ListLaunch) + 64) + 256 + qmul * sizeof(Quartet);
      if (need > CQX_NOALLOC_SCRATCH_BYTES)
        return fail(c, CQX_ERR_UNSUPPORTED,
                    "cqx_two_body_contract_noalloc: the contractions need " + std::to_string(need) +
                        " bytes of fixed workspace; the scratch region holds " +
                        std::to_string((size_t)CQX_NOALLOC_SCRATCH_BYTES));
      start += nmat;
    }
  } catch (const std::exception &e) {
    return fail(c, CQX_ERR_HIP, e.what());
  }
  Arena a;
  a.base = static_cast<char *>(c->scratch);
  a.cap = CQX_NOALLOC_SCRATCH_BYTES;
  return contract(c, screen, ncon, list, is_complex, CQX_DEVICE, &a);
}

int cqx_scratch_size(cqx_ctx *c, int ncon, const cqx_contraction *list, int is_complex, size_t *bytes) {
  if (!c || ncon <= 0 || !list || !bytes) return fail(c, CQX_ERR_ARG, "cqx_scratch_size: bad arguments");
  const int nmol = c->nmol;
  if (ncon % nmol)
    return fail(c, CQX_ERR_ARG, "a batched context takes the same number of contractions for every molecule");
  const int nper = ncon / nmol;
  try {
    HIPCHK(hipSetDevice(c->device));
    // the buffers grow to the largest group's need; every group is sized alike
    size_t x = 0, bm = 0, bg = 0, planes = 0, cand = 0, scan = 0, ncp = 0, nmatm = 0, pn[kNumClass] = {};
    long long quart = 1;
    const bool mixed = c->fp32_cut > 0.0;
    const int qmul = mixed ? 2 : 1;
    int start = 0;
    while (start < nper) {
      const int nmat = group_size(list + start, nper - start, is_complex, nullptr);
      const GroupPlan g = plan_group(c, list + start, nmat, is_complex, c->stream);
      x = std::max(x, g.x);
      bm = std::max(bm, g.bm);
      bg = std::max(bg, g.bg);
      planes = std::max(planes, g.planes);
      cand = std::max(cand, g.cand);
      scan = std::max(scan, g.scan_bytes);
      ncp = std::max(ncp, g.ncp);
      nmatm = std::max(nmatm, (size_t)nmat);
      for (int k = 0; k < kNumClass; ++k) pn[k] = std::max(pn[k], g.pn[k]);
      long long q = list_entries_bound(g, c->list_cap) * qmul;
      if (c->use_graph) q = std::max(q, c->list_cap * qmul);  // a recorded build's lists
      quart = std::max(quart, q);
      start += nmat;
    }
    // pieces of a batch: each class pair whole, or bra ranges of one
    size_t nbra_all = 0;
    for (int cb = 0; cb < kNumClass; ++cb) nbra_all += c->cls[cb].size() * (size_t)(cb + 1);
    const size_t npiece = ncp + nbra_all;
    size_t b = 0;
    b += 2 * x * sizeof(double);                     // staged densities and outputs
    b += nmatm * sizeof(int);                        // HER flags
    b += (bm + bg + nmol) * sizeof(double);          // block norms, molecule maxima
    for (int k = 0; k < kNumClass; ++k) b += pn[k] * sizeof(double);
    b += 2 * planes * sizeof(double);                // Cartesian planes and accumulators
    if (c->autotune_pending) b += planes * sizeof(double) + nmol * sizeof(cqx::MolPlanes);
    b += 2 * nmol * sizeof(cqx::MolPlanes);
    b += cand * (sizeof(long long) + sizeof(int)) + scan;
    b += 2 * ncp * sizeof(long long);                // gathered candidate totals
    b += (size_t)quart * sizeof(Quartet);            // quartet lists
    b += npiece * (2 * sizeof(unsigned long long) + 2 * sizeof(double) + sizeof(cqx::ListLaunch) +
                   sizeof(long long)) + sizeof(long long);
    if (c->use_graph) b += ncp * (sizeof(cqx::ListLaunch) + 4 * sizeof(long long)) + sizeof(long long);
    // tile lists: at most one tile per sub-block and precision, when the
    // build takes the tile path (Hermitian contractions, not recorded)
    bool her_all = true;
    for (int i = 0; i < ncon; ++i) her_all = her_all && list[i].her;
    if (c->use_tiles && !c->use_graph && her_all) {
      size_t nsub = 0, nt = 0;
      for (const auto &tc : c->tile_cps) {
        int la, lb, lc, ld;
        class_ls(tc.cb, la, lb);
        class_ls(tc.ck, lc, ld);
        if (cqx::ncart(la) * cqx::ncart(lb) * cqx::ncart(lc) * cqx::ncart(ld) > c->fast_max_nc) continue;
        nsub += (size_t)tc.n;
        ++nt;
      }
      b += nsub * qmul * sizeof(cqx::Tile) + nt * (3 * sizeof(unsigned long long) + 2 * sizeof(double) +
                                                   sizeof(cqx::TileListLaunch) + sizeof(long long)) +
           sizeof(long long);
    }
    *bytes = b;
  } catch (const std::exception &e) {
    return fail(c, CQX_ERR_HIP, e.what());
  }
  return CQX_OK;
}

int cqx_grad_two_body_contract(cqx_ctx *c, int ngrad, int nmat, const cqx_contraction *list, int mem) {
  if (!c || ngrad <= 0 || ngrad % 3 || nmat <= 0 || !list)
    return fail(c, CQX_ERR_ARG, "cqx_grad_two_body_contract: bad arguments");
  if (c->nmol != 1) return fail(c, CQX_ERR_UNSUPPORTED, "gradient contractions take a one-molecule context");
  const cqx::Basis &BS = c->basis;
  int natom = 0;
  for (const auto &sh : BS.shells) {
    natom = std::max(natom, sh.atom + 1);
    if (sh.l > cqx::kMaxLDev) return fail(c, CQX_ERR_UNSUPPORTED, "gradient contractions cover s through g shells");
  }
  if (ngrad < 3 * natom) return fail(c, CQX_ERR_ARG, "fewer gradient components than 3 x atoms of the basis");
  unsigned kmask = 0;
  for (int g = 0; g < ngrad; ++g)
    for (int m = 0; m < nmat; ++m) {
      const cqx_contraction &cc = list[(size_t)g * nmat + m];
      if (!cc.X || !cc.AX || (cc.type != CQX_COULOMB && cc.type != CQX_EXCHANGE))
        return fail(c, CQX_ERR_ARG, "invalid gradient contraction " + std::to_string(g * nmat + m));
      // direct.hpp:2498-2500
      if (!cc.her) return fail(c, CQX_ERR_UNSUPPORTED, "Nonhermetian NYI!");
      // one density per matrix, shared by every coordinate (FockBuilder::getGDGrad)
      if (cc.X != list[m].X || cc.type != list[m].type)
        return fail(c, CQX_ERR_ARG, "matrix " + std::to_string(m) + " must use one density and type for every coordinate");
      if (g == 0 && cc.type == CQX_EXCHANGE) kmask |= 1u << m;
    }
  if (nmat > 32) return fail(c, CQX_ERR_ARG, "at most 32 matrices per gradient contraction");
  const auto t_start = std::chrono::steady_clock::now();
  c->stats = cqx_stats{};
  try {
    HIPCHK(hipSetDevice(c->device));
    drop_graph(c);  // the workspace below may be reallocated
    hipStream_t s = c->stream;
    const int nb = c->nb, nbc = c->nbc;
    const size_t nb2 = (size_t)nb * nb, nbc2 = (size_t)nbc * nbc;
    std::vector<int> atom(BS.shells.size());
    for (size_t k = 0; k < BS.shells.size(); ++k) atom[k] = BS.shells[k].atom;
    c->d_shell_atom.upload(atom, s);
    // densities to Cartesian planes
    const hipMemcpyKind hk = mem == CQX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    c->d_X.ensure((size_t)nmat * nb2);
    c->d_gX.ensure((size_t)nmat * nbc2);
    for (int m = 0; m < nmat; ++m) {
      HIPCHK(hipMemcpyAsync(c->d_X.p + (size_t)m * nb2, list[m].X, sizeof(double) * nb2, hk, s));
      HIPCHK(cqx::launch_to_cart(c->cmap, c->d_X.p + (size_t)m * nb2, 0, c->d_gX.p + (size_t)m * nbc2, nullptr, 1, s));
    }
    const size_t nacc = (size_t)ngrad * nmat * nbc2;
    c->d_gacc.ensure(nacc);
    HIPCHK(cqx::launch_zero(c->d_gacc.p, nacc, s));
    // every quartet of every class pair (bra class >= ket class)
    HIPCHK(hipEventRecord(c->ev0, s));
    for (int cb = 0; cb < kNumClass; ++cb)
      for (int ck = 0; ck <= cb; ++ck) {
        const long long nbra = (long long)c->cls[cb].size(), nket = (long long)c->cls[ck].size();
        if (!nbra || !nket) continue;
        cqx::GradLaunch G{};
        class_ls(cb, G.la, G.lb);
        class_ls(ck, G.lc, G.ld);
        // the largest LDS a block may take (dynamic LDS up to the device's 160 KB)
        G.chunk = cqx::grad_chunk(G.la, G.lb, G.lc, G.ld, std::max<size_t>((size_t)c->lds_max, 160 * 1024), &G.rpp);
        if (G.chunk <= 0) throw Unsupported("gradient kernel LDS beyond the device limit");
        G.bra = c->d_pairs[cb].p;
        G.ket = c->d_pairs[ck].p;
        G.nbra = (int)nbra;
        G.nket = (int)nket;
        G.same_class = cb == ck;
        G.nq = cb == ck ? nbra * (nbra + 1) / 2 : nbra * nket;
        G.prims = c->d_prims.p;
        G.prim_alpha = c->d_prim_alpha.p;
        G.rys = c->d_rys.p;
        G.shell_atom = c->d_shell_atom.p;
        G.X = c->d_gX.p;
        G.acc = c->d_gacc.p;
        G.nmat = nmat;
        G.nbc = nbc;
        G.kmask = kmask;
        HIPCHK(cqx::launch_eri_grad(G, s));
        c->stats.n_quartets += G.nq;
        c->stats.n_launch_eri++;
      }
    HIPCHK(hipEventRecord(c->ev1, s));
    // back to pure functions, Hermitian symmetrisation into AX
    // (AX += 1/2 (A + A^T), direct.hpp:2548-2570), per coordinate and matrix
    const size_t nout = (size_t)ngrad * nmat;
    c->d_gAX.ensure(nout * nb2);
    for (size_t k = 0; k < nout; ++k) {
      HIPCHK(hipMemcpyAsync(c->d_gAX.p + k * nb2, list[k].AX, sizeof(double) * nb2, hk, s));
      HIPCHK(cqx::launch_from_cart(c->cmap, c->d_gacc.p + k * nbc2, nullptr, 1, 1, 0, c->d_gAX.p + k * nb2, s));
    }
    const hipMemcpyKind dk = mem == CQX_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    for (size_t k = 0; k < nout; ++k)
      HIPCHK(hipMemcpyAsync(list[k].AX, c->d_gAX.p + k * nb2, sizeof(double) * nb2, dk, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.ms_eri = ms;
  } catch (const std::exception &e) {
    return fail(c, CQX_ERR_HIP, e.what());
  }
  c->stats.ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  return CQX_OK;
}

int cqx_device_alloc(cqx_ctx *c, size_t bytes, void **ptr) {
  if (!c || !ptr) return fail(c, CQX_ERR_ARG, "cqx_device_alloc: bad arguments");
  if (hipSetDevice(c->device) != hipSuccess || hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess)
    return fail(c, CQX_ERR_MEM, "hipMalloc failed");
  return CQX_OK;
}

int cqx_device_free(cqx_ctx *c, void *ptr) {
  if (!c) return fail(c, CQX_ERR_ARG, "cqx_device_free: null context");
  if (ptr && hipFree(ptr) != hipSuccess) return fail(c, CQX_ERR_HIP, "hipFree failed");
  return CQX_OK;
}

int cqx_memcpy(cqx_ctx *c, void *dst, const void *src, size_t bytes, int kind) {
  if (!c || (bytes && (!dst || !src)) || kind < 0 || kind > 2)
    return fail(c, CQX_ERR_ARG, "cqx_memcpy: bad arguments");
  const hipMemcpyKind k[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice};
  if (hipSetDevice(c->device) != hipSuccess || hipMemcpyAsync(dst, src, bytes, k[kind], c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(c, CQX_ERR_HIP, "hipMemcpy failed");
  return CQX_OK;
}

int cqx_device_zero(cqx_ctx *c, void *ptr, size_t bytes) {
  if (!c || (bytes && !ptr) || bytes % sizeof(double))
    return fail(c, CQX_ERR_ARG, "cqx_device_zero: bad arguments (size must be a multiple of 8)");
  if (hipSetDevice(c->device) != hipSuccess ||
      cqx::launch_zero(static_cast<double *>(ptr), bytes / sizeof(double), c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(c, CQX_ERR_HIP, "device zero failed");
  return CQX_OK;
}

}  // extern "C"
namespace cqx {
hipStream_t ctx_stream(const cqx_ctx *c) { return c->stream; }
int ctx_device(const cqx_ctx *c) { return c->device; }
double ctx_thresh(const cqx_ctx *c) { return c->thresh; }
const Basis &ctx_basis(const cqx_ctx *c) { return c->basis; }
int ctx_fail(cqx_ctx *c, int code, const std::string &msg) { return fail(c, code, msg); }
}  // namespace cqx

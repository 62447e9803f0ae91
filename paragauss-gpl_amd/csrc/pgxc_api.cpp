// C-ABI layer of the MI355X XC build (declared in include/pgxc.h).
//
// A build is a plan: for every super-step, the units (grid chunks) of every
// molecule taking part, and per stage the list of work items over those units.
// Stage kernels walk their lists persistently; one molecule is a batch of one.
// Per super-step:  basis -> density contraction -> functional -> W ->
// potential contraction; then per molecule a fold of the split partials (H + H^T)
// and one packing into the lower triangle.
// Chunks are whole numbers of reference batches, so the sums are those of
// calc_xcks_nl (ParaGauss modules/xc_hamiltonian.f90:1348-1575), regrouped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pgxc.h"

#include "pgxc_internal.h"

using namespace pgxc;

namespace {

struct Plan;

long round_up(long a, long b) { return (a + b - 1) / b * b; }

// schedule of the contraction stages (PGXC_SCHED_*) and item logging
struct Sched {
  int mode = PGXC_SCHED_ANALYTIC;  // no slower than the static walk in any measurement
  pgxc_costmodel *model = nullptr;
  int log = 0;
};

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

// Device buffers freed when the scope ends, on every return path.
struct ScopedDev {
  std::vector<void *> bufs;
  ~ScopedDev() {
    for (void *b : bufs) dfree(b);
  }
  template <typename T>
  T *keep(T *p) {
    bufs.push_back((void *)p);
    return p;
  }
};

}  // namespace

struct pgxc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // chunks alternate between two lanes (this stream and stream2, each with its
  // own work buffers) so one chunk's stages overlap the next chunk's
  hipStream_t stream2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  int nlanes = 1, nlanes_cap = 1;  // lanes in use / with buffers
  std::string err;
  // sizes
  int n_atoms = 0, n_shells = 0, nbas = 0, NP = 0, veclen = 128, tf = 128;
  int lmax = 0;  // largest angular momentum of the basis
  int var[2] = {0, 1};  // contraction variant for FP64 and for FP32 builds
  long npts = 0, npad = 0, Mc = 0;
  int nchunks = 0, nsplit = 1, ntiles = 0;
  long nblk_total = 0;
  // device: geometry and basis
  double *grid = nullptr;  // [4][npad] x, y, z, w
  AtomD *atoms = nullptr;
  ShellD *shells = nullptr;
  RangeD *ranges = nullptr;  // atom ranges of the basis stage
  int *ratoms = nullptr;     // atoms in shell order
  int n_ranges = 0;
  std::vector<RangeD> h_ranges;
  double *exps = nullptr, *norms = nullptr, *contr = nullptr;
  // device: density (NP x NP); D holds the operand in the build's precision
  double *P = nullptr, *Pref = nullptr;
  void *D = nullptr;
  // device: work buffers of one chunk
  void *phi = nullptr, *W = nullptr;
  size_t phi_bytes = 0, w_bytes = 0, phi_elem = 0;
  size_t w_lane = 0, phi_lane = 0;  // bytes of one lane's W / basis chunk buffer
  // basis values of every chunk kept across builds, per element type
  // ([0] FP64, [1] FP32): built once, reused while the geometry lives
  int cache_off = 0;
  void *phic[2] = {nullptr, nullptr};
  int phic_valid[2] = {0, 0}, phic_failed[2] = {0, 0};
  int phic_grad[2] = {0, 0};  // the cache also holds the gradient planes (a GGA basis pass ran)
  size_t phic_bytes[2] = {0, 0};
  void *phi_base = nullptr;  // Phi of the current build: chunk k at phi_base + k * phi_stride
  size_t phi_stride = 0;
  int skip_basis = 0;        // the current build reads cached basis values
  double *part = nullptr, *coef = nullptr, *Hpart = nullptr;
  // open shell (ispin = 2): the fields above and below serve the alpha spin,
  // these the beta spin (allocated with the first open-shell density; the
  // basis values and the work tables are shared)
  int ispin = 1;
  struct Beta {
    double *P = nullptr, *Pref = nullptr, *part = nullptr, *coef = nullptr, *Hpart = nullptr;
    double *state = nullptr, *Vstate = nullptr, *Vout = nullptr;
    void *D = nullptr, *W = nullptr;
    size_t w_bytes = 0;
  } beta;
  // device: state and results
  double *state = nullptr;                    // [S_NSLOT][npad]
  double *Vstate = nullptr, *Vout = nullptr;  // packed
  double *epart = nullptr;
  // host
  int have_density = 0, density_mode = 0, have_state = 0;
  uint32_t state_functional = 0;
  uint64_t geom = 0;  // hash of the inputs that fix the layout (state files)
  double last_exc = 0, last_charge = 0;
  double timing[6] = {0, 0, 0, 0, 0, 0};
  double stats[5] = {0, 0, 0, 0, 0};
  std::vector<hipEvent_t> ev;
  Plan *plan = nullptr;
  Plan *plan_e = nullptr;  // energy-only builds (pgxc_energy)
  double *eterm = nullptr;  // [nblk_total][5] energy-term block sums
  double last_terms[5] = {0, 0, 0, 0, 0};
  double exps_per_point = 0;
  // screening and work tables over the padded grid (setup_screen)
  int screen = 0, n8 = 0, nw8 = 0, CH = 0;
  int *nbt = nullptr;
  uint32_t *sig8 = nullptr, *wbits = nullptr;
  ScreenShell *sshells = nullptr;
  double flops_screened = 0, flops_dense = 0;  // algorithmic work of one build
  double sum_bn = 0, sum_n2 = 0;  // sums over 128-point tiles of B n_b and n_b^2 (screened)
  // host and device copies: per tile run lists, per split run unions and
  // [group][chunk] flags, per chunk density and potential items
  std::vector<int> h_tl_off, h_su_off, h_tf_off;
  std::vector<uint16_t> h_tl, h_su, h_tfl;
  std::vector<uint8_t> h_tlm;
  std::vector<long> h_gcf_off, h_di_off, h_pi_off;
  std::vector<uint8_t> h_gcf;
  std::vector<uint32_t> h_di, h_pi;
  int *tl_off = nullptr, *su_off = nullptr, *tf_off = nullptr;
  uint16_t *tl = nullptr, *su = nullptr, *tfl = nullptr;
  uint8_t *tlm = nullptr;
  uint32_t *tfb = nullptr;  // per tile, its functions as a bitmap of nwf words
  int nwf = 0;
  // per entry of a split's function list: bit t set when the function is
  // in the list of the split's tile t (every bit when a split has more than
  // 32 tiles)
  std::vector<uint32_t> h_sumask;
  uint32_t *sumask = nullptr;
  // nuclear gradient work space (first gradient on): Hessian planes and
  // products of one chunk, per-(tile, function) sums
  double *ghess = nullptr, *gXY = nullptr, *gpart = nullptr;
  long gpart_n = 0;
  long *gcf_off = nullptr;
  uint8_t *gcf = nullptr;
  uint32_t *di = nullptr, *pi = nullptr;
  // blocks of the basis values some stage reads (see setup_screen): per
  // chunk the basis items (256-point block << 16 | atom range) that touch
  // them and their write flags
  uint32_t *bi = nullptr, *bfl = nullptr;
  std::vector<long> h_bi_off;
  std::vector<int> fpos;          // internal position of each caller function
  std::vector<int> fatom;         // atom of each internal function position
  std::vector<long> perm;         // caller index of each internal point
  std::vector<double> axyz;       // atom positions (bohr), caller order
  int *d_fpos = nullptr;
  Sched sched;
};

namespace {

int fail(pgxc_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                              \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, e_ == hipErrorOutOfMemory ? PGXC_ERR_NOMEM : PGXC_ERR_HIP,       \
                  std::string(#call) + ": " + hipGetErrorString(e_));                  \
  } while (0)

template <typename T>
int dalloc(pgxc_ctx *c, T **p, size_t count) {
  if (*p) return PGXC_OK;
  hipError_t e = hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, PGXC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return PGXC_OK;
}

const uint32_t kSupported = PGXC_XC_PBE;
const int kTiles[3] = {64, 96, 128};

// ------------------------------------------------------------------- plans
// density / potential lists per (function tile, variant): index 3 * t + v
constexpr int kLists = 3 * kNumVariants;

struct StepPlan {
  int nunits = 0;
  StageList basis, func, func2, wmat, wmat2;  // func2, wmat2: open-shell units
  StageList dens[kLists], pot[kLists];
};

struct Plan {
  uint32_t functional = 0;
  int precision = 0;
  std::vector<int> incr;  // per molecule
  long order_version = -1;
  int sched_mode = 0, log = 0;
  int energy = 0;  // energy terms only: no W, potential, fold or packing
  int wbasis = 0;  // split-BF16 W from the basis evaluated again (k_wbasis)
  const pgxc_costmodel *model = nullptr;
  long model_version = -1;
  int *d_ctr = nullptr;  // queue counters of every contraction list
  long nctr = 0;
  // H partials, cleared at the start: (buffer, ints, lane); lane 1's on the
  // second stream, overlapping the first lane's first stages
  struct HZero {
    int *p;
    long n;
    int lane;
  };
  std::vector<HZero> hzero;
  // item log: features of every contraction item (row-major), and per list
  // the device log and the list's first feature row
  std::vector<double> feat;
  std::vector<std::pair<uint32_t *, long>> logs;  // (device ticks, items)
  double tick_us = 0.0;
  std::vector<StepPlan> steps;
  int nlanes = 1;  // 2: chunk k runs on lane k % 2 (two streams)
  std::vector<int> lanes;  // lanes in use per molecule when the plan was made
  std::vector<int> spins;  // ispin per molecule when the plan was made
  PackD *d_pack = nullptr;
  long *d_pack_base = nullptr, *d_prep_base = nullptr, *d_fold_base = nullptr;
  long pack_total = 0, prep_total = 0, fold_total = 0;
  int nmol = 0;
  char *blob = nullptr;  // device memory holding every list
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_src = nullptr;
  ~Plan() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (graph_src) (void)hipGraphDestroy(graph_src);
    dfree(blob);
  }
};

void free_plan(Plan *&p) {
  delete p;
  p = nullptr;
}

// Schedule of one stage list: a unit's items stay together, units in
// descending priority (largest predicted cost first).  Position L of the
// persistent walk runs on XCD L % 8, so contiguous runs of the sequence are
// spread to keep neighbouring items on one XCD.  Without priorities the kernels
// use the natural XCD-blocked order.
void make_order(const std::vector<long> &base, const std::vector<double> &prio,
                std::vector<int> &order) {
  order.clear();
  if (prio.empty()) return;
  const int nu = (int)base.size() - 1;
  std::vector<int> us(nu);
  std::iota(us.begin(), us.end(), 0);
  std::stable_sort(us.begin(), us.end(), [&](int a, int b) { return prio[a] > prio[b]; });
  std::vector<int> seq;
  seq.reserve(base[nu]);
  for (int u : us)
    for (long i = base[u]; i < base[u + 1]; ++i) seq.push_back((int)i);
  const long total = base[nu];
  order.resize(total);
  const long per = total / 8;
  for (long L = 0; L < total; ++L) {
    long r = L;
    if (L < per * 8) r = (L % 8) * per + L / 8;
    order[L] = seq[r];
  }
}

enum ItemKind { K_BASIS, K_FUNC, K_WMAT, K_DENS, K_POT, K_WBASIS };

long unit_items(const UnitD &U, ItemKind k, int tf) {
  switch (k) {
    case K_BASIS:
    case K_WBASIS: return U.bscreen ? U.nbi : items_basis(U.Mc, U.n_ranges);
    case K_FUNC: return items_functional(U.Mc);
    case K_WMAT: return items_wmat(U.Mc, U.NP);
    case K_DENS: return U.ndi;
    default: return U.npi;
  }
}

}  // namespace

// Features (PGXC_NFEATURES) and counted flops of the contraction items of
// one unit, in the items' order.  kact: K length left after screening
// (gathered functions of a density item, points of a potential item);
// kdense: the dense K length.
static void unit_item_features(const pgxc_ctx *c, int step, ItemKind kind, int precision,
                               std::vector<double> &feat, std::vector<double> &flops) {
  const int tf = c->tf;
  const int var = precision == PGXC_PREC_FP64 ? c->var[0] : precision == PGXC_PREC_FP32 ? c->var[1] : 0;
  auto push = [&](int kd, double kact, double kdense, bool active, double fl) {
    const double f[PGXC_NFEATURES] = {(double)kd,
                                      precision == PGXC_PREC_FP64 ? 1.0 : 0.0,
                                      precision == PGXC_PREC_FP32 ? 1.0 : 0.0,
                                      precision == PGXC_PREC_BF16X3 ? 1.0 : 0.0,
                                      tf / 128.0,
                                      std::log2(1.0 + kact) / 16.0,
                                      std::log2(1.0 + kdense) / 16.0,
                                      active ? 1.0 : 0.0,
                                      var / 2.0};
    feat.insert(feat.end(), f, f + PGXC_NFEATURES);
    flops.push_back(fl);
  };
  if (kind == K_DENS) {
    const long T0 = (long)step * c->Mc / kScreenTile;
    for (long i = c->h_di_off[step]; i < c->h_di_off[step + 1]; ++i) {
      const uint32_t e = c->h_di[(size_t)i];
      const long T = T0 + (e >> 8);
      const int pr = (int)(e & 255), c8 = c->h_tl_off[(size_t)T + 1] - c->h_tl_off[(size_t)T];
      const int nf = c->h_tf_off[(size_t)T + 1] - c->h_tf_off[(size_t)T];
      const double cols = (double)std::min(2 * tf, nf - 2 * pr * tf);
      push(0, (double)c8 * kRun, c->NP, true, 2.0 * kScreenTile * cols * c8 * kRun);
    }
  } else {
    const long spts = c->Mc / c->nsplit, G = spts / kGroup;
    for (long i = c->h_pi_off[step]; i < c->h_pi_off[step + 1]; ++i) {
      const uint32_t e = c->h_pi[(size_t)i];
      const long S = (long)step * c->nsplit + (e >> 20);
      const int ci = (int)((e >> 10) & 1023), cj = (int)(e & 1023);
      const int nu = c->h_su_off[(size_t)S + 1] - c->h_su_off[(size_t)S], nch = (nu + tf - 1) / tf;
      const uint8_t *fl = c->h_gcf.data() + c->h_gcf_off[(size_t)S];
      long cnt = 0;
      for (long g = 0; g < G; ++g) cnt += (fl[g * nch + ci] & fl[g * nch + cj]) != 0;
      const double rows = (double)std::min(tf, nu - ci * tf);
      const double cols = (double)std::min(tf, nu - cj * tf);
      push(1, (double)cnt * kGroup, (double)spts, cnt > 0, 2.0 * rows * cols * cnt * kGroup);
    }
  }
}

// Build (or reuse) the plan of a set of molecules for one configuration.
static int get_plan(std::vector<pgxc_ctx *> &mols, uint32_t functional, int precision,
                    const std::vector<int> &incr, const std::vector<double> &unit_prio,
                    long order_version, const Sched &sc, Plan *&plan, pgxc_ctx *errc,
                    int energy = 0, int dense_basis = 0) {
  const int smode = sc.model ? sc.mode : std::min(sc.mode, (int)PGXC_SCHED_ANALYTIC);
  std::vector<int> lanes, spins;
  for (auto *c : mols) {
    lanes.push_back(c->nlanes);
    spins.push_back(c->ispin);
  }
  if (plan && plan->functional == functional && plan->precision == precision &&
      plan->incr == incr && plan->lanes == lanes && plan->spins == spins &&
      plan->energy == energy &&
      plan->order_version == order_version &&
      plan->sched_mode == smode &&
      plan->log == sc.log && plan->model == sc.model &&
      plan->model_version == costmodel_version(sc.model))
    return PGXC_OK;
  free_plan(plan);
  plan = new Plan();
  plan->functional = functional;
  plan->precision = precision;
  plan->incr = incr;
  plan->lanes = lanes;
  plan->spins = spins;
  plan->energy = energy;
  plan->order_version = order_version;
  plan->sched_mode = smode;
  plan->log = sc.log;
  plan->model = sc.model;
  plan->model_version = costmodel_version(sc.model);
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, errc->device) != hipSuccess ||
        khz <= 0)
      khz = 100000;
    plan->tick_us = 1000.0 / khz;
  }
  const bool queues = smode != PGXC_SCHED_STATIC;
  const int nm = (int)mols.size();
  int nsteps = 0;
  for (auto *c : mols) nsteps = std::max(nsteps, c->nchunks);
  // Two lanes (step k on stream k & 1) only when every member of more than
  // one chunk asked for two.  A member's chunk k then takes the work buffers
  // of the lane whose stream runs it, so two chunks that share buffers never
  // run at the same time; a member of one chunk only has lane 0.
  int plan_lanes = 0;
  for (auto *c : mols)
    if (c->nchunks > 1) plan_lanes = plan_lanes ? std::min(plan_lanes, c->nlanes) : c->nlanes;
  if (plan_lanes < 1 || nsteps < 2) plan_lanes = 1;
  auto lanes_of = [plan_lanes](const pgxc_ctx *c) { return c->nchunks > 1 ? plan_lanes : 1; };

  // host image of everything that goes to the device
  std::vector<char> img;
  auto put = [&](const void *src, size_t bytes) {  // src == nullptr: zeros
    const size_t off = (size_t)round_up((long)img.size(), 256);
    img.resize(off + std::max<size_t>(bytes, 1));
    if (bytes && src) memcpy(img.data() + off, src, bytes);
    return off;
  };
  struct Pending {
    size_t units_off = 0, base_off = 0, order_off = 0, qbeg_off = 0, ctr_off = 0, tlog_off = 0;
    int nu = 0;
    long total = 0;
    bool has_order = false, has_queues = false, has_log = false;
    size_t frow = 0;  // first feature row of the list's items
    int lmax = 0;     // basis lists: the largest angular momentum of their molecules
  };
  std::vector<double> flops;  // counted flops per feature row
  size_t plan_ctr_off_tmp = 0;
  auto stage = [&](const std::vector<UnitD> &us, const std::vector<int> &um, ItemKind kind,
                   int tf, int step) {
    Pending P;
    P.nu = (int)us.size();
    std::vector<long> base(us.size() + 1, 0);
    for (size_t i = 0; i < us.size(); ++i) base[i + 1] = base[i] + unit_items(us[i], kind, tf);
    P.total = base.back();
    P.units_off = put(us.data(), sizeof(UnitD) * us.size());
    P.base_off = put(base.data(), sizeof(long) * base.size());
    const bool contr = kind == K_DENS || kind == K_POT;
    if (contr && (queues || sc.log) && P.total > 0) {
      P.frow = plan->feat.size() / PGXC_NFEATURES;
      for (int m : um) unit_item_features(mols[m], step, kind, precision, plan->feat, flops);
    }
    if (contr && queues && P.total > 0) {
      P.has_queues = true;  // queue bounds filled in once every cost is known
    } else {
      std::vector<double> pr;
      if (!unit_prio.empty())
        for (int m : um) pr.push_back(unit_prio[m]);
      std::vector<int> order;
      make_order(base, pr, order);
      P.has_order = !order.empty();
      if (P.has_order) P.order_off = put(order.data(), sizeof(int) * order.size());
    }
    if (contr && sc.log && P.total > 0) {
      P.has_log = true;
      P.tlog_off = put(nullptr, sizeof(uint32_t) * P.total);
    }
    return P;
  };
  // per step: basis, func, wmat, dens lists, pot lists, func2, wmat2
  std::vector<std::vector<Pending>> pend(nsteps);
  for (int k = 0; k < nsteps; ++k) {
    std::vector<UnitD> units;
    std::vector<int> umol, uspin;
    for (int m = 0; m < nm; ++m) {
      pgxc_ctx *c = mols[m];
      if (k >= c->nchunks) continue;
      const long p0 = (long)k * c->Mc;
      UnitD U;
      memset(&U, 0, sizeof(U));
      U.gx = c->grid + p0;
      U.gy = c->grid + c->npad + p0;
      U.gz = c->grid + 2 * c->npad + p0;
      U.gw = c->grid + 3 * c->npad + p0;
      U.atoms = c->atoms;
      U.shells = c->shells;
      U.ranges = c->ranges;
      U.ratoms = c->ratoms;
      U.n_ranges = c->n_ranges;
      U.exps = c->exps;
      U.norms = c->norms;
      U.contr = c->contr;
      const int lane = k % lanes_of(c);  // work buffers of chunk k: its stream's lane
      U.phi = (char *)c->phi_base + (c->phi_stride ? (size_t)k * c->phi_stride : (size_t)lane * c->phi_lane);
      U.W = (char *)c->W + (size_t)lane * c->w_lane;
      U.D = c->D;
      U.part = c->part + (size_t)lane * c->ntiles * 4 * c->Mc;
      U.coef = c->coef + (size_t)lane * 4 * c->Mc;
      U.state = c->state + p0;
      U.epart = c->epart + 2 * (p0 / kThreads);
      U.eterm = c->eterm + 5 * (p0 / kThreads);
      U.Hpart = c->Hpart + (size_t)lane * c->nsplit * c->NP * c->NP;
      U.Mc = c->Mc;
      U.npad = c->npad;
      U.n_atoms = c->n_atoms;
      U.NP = c->NP;
      U.ntiles = c->ntiles;
      U.nsplit = c->nsplit;
      U.incremental = incr[m];
      U.mask = functional;
      U.CH = c->CH;
      U.nw8 = c->nw8;
      U.ditems = c->di + c->h_di_off[k];
      U.pitems = c->pi + c->h_pi_off[k];
      U.ndi = c->h_di_off[k + 1] - c->h_di_off[k];
      U.npi = c->h_pi_off[k + 1] - c->h_pi_off[k];
      U.tl_off = c->tl_off + p0 / kScreenTile;
      U.tl = c->tl;
      U.tf_off = c->tf_off + p0 / kScreenTile;
      U.tfl = c->tfl;
      U.tfb = c->tfb + (p0 / kScreenTile) * c->nwf;
      U.tlm = c->tlm;
      U.nwf = c->nwf;
      U.su_off = c->su_off + (long)k * c->nsplit;
      U.su = c->su;
      U.sumask = c->sumask;
      U.gcf_off = c->gcf_off + (long)k * c->nsplit;
      U.gcf = c->gcf;
      U.sig8 = c->sig8 + (p0 / kGroup) * c->nw8;
      U.wbits = c->wbits + (p0 / kGroup) * c->nw8;
      U.bscreen = dense_basis ? 0 : 1;  // the basis stage writes the blocks read later
      U.bitems = c->bi + c->h_bi_off[k];
      U.bflags = c->bfl + 16 * c->h_bi_off[k];
      U.nbi = c->h_bi_off[k + 1] - c->h_bi_off[k];
      if (c->ispin == 2) {  // the beta spin's unit: same points and basis values
        UnitD B = U;
        B.W = (char *)c->beta.W + (size_t)lane * c->w_lane;
        B.D = c->beta.D;
        B.part = c->beta.part + (size_t)lane * c->ntiles * 4 * c->Mc;
        B.coef = c->beta.coef + (size_t)lane * 4 * c->Mc;
        B.state = c->beta.state + p0;
        B.Hpart = c->beta.Hpart + (size_t)lane * c->nsplit * c->NP * c->NP;
        U.part2 = B.part;
        U.coef2 = B.coef;
        U.state2 = B.state;
        U.W2 = B.W;
        units.push_back(U);
        umol.push_back(m);
        uspin.push_back(0);
        units.push_back(B);
        umol.push_back(m);
        uspin.push_back(1);
      } else {
        units.push_back(U);
        umol.push_back(m);
        uspin.push_back(0);
      }
    }
    // the units of one spin (0: closed-shell and alpha units, 1: beta)
    // meeting a condition on the molecule
    auto pick = [&](int spin, auto cond, std::vector<UnitD> &out, std::vector<int> &om) {
      for (size_t i = 0; i < units.size(); ++i)
        if (uspin[i] == spin && cond(mols[umol[i]])) {
          out.push_back(units[i]);
          om.push_back(umol[i]);
        }
    };
    {  // molecules whose basis values are cached skip the basis stage
      std::vector<UnitD> bu;
      std::vector<int> bm;
      for (size_t i = 0; i < units.size(); ++i)
        if (uspin[i] == 0 && (dense_basis || !mols[umol[i]]->skip_basis)) {
          bu.push_back(units[i]);
          bm.push_back(umol[i]);
        }
      pend[k].push_back(stage(bu, bm, K_BASIS, 0, k));
      for (int m : bm) pend[k].back().lmax = std::max(pend[k].back().lmax, mols[m]->lmax);
    }
    {
      std::vector<UnitD> fu;
      std::vector<int> fm;
      pick(0, [](const pgxc_ctx *c) { return c->ispin == 1; }, fu, fm);
      pend[k].push_back(stage(fu, fm, K_FUNC, 0, k));
    }
    // W: split-BF16 builds evaluate the basis again over the basis items
    // (test switch PGXC_W_READ=1: they read the stored values back instead,
    // the same bits)
    const bool wb = precision == PGXC_PREC_BF16X3 && !getenv("PGXC_W_READ");
    plan->wbasis = wb ? 1 : 0;
    const ItemKind wkind = wb ? K_WBASIS : K_WMAT;
    {  // closed-shell units; the open-shell ones in wmat2 below
      std::vector<UnitD> wu;
      std::vector<int> wm;
      pick(0, [](const pgxc_ctx *c) { return c->ispin == 1; }, wu, wm);
      pend[k].push_back(stage(wu, wm, wkind, 0, k));
      for (int m : wm) pend[k].back().lmax = std::max(pend[k].back().lmax, mols[m]->lmax);
    }
    const int cls = precision == PGXC_PREC_FP64 ? 0 : 1;
    for (int kind = 0; kind < 2; ++kind)
      for (int li = 0; li < kLists; ++li) {
        const int t = li / kNumVariants, v = li % kNumVariants;
        std::vector<UnitD> sub;
        std::vector<int> subm;
        for (size_t i = 0; i < units.size(); ++i) {
          const pgxc_ctx *c = mols[umol[i]];
          if (c->tf == kTiles[t] && c->var[cls] == v) {
            sub.push_back(units[i]);
            subm.push_back(umol[i]);
          }
        }
        pend[k].push_back(stage(sub, subm, kind == 0 ? K_DENS : K_POT, kTiles[t], k));
      }
    {  // spin functional over the alpha units of open-shell molecules
      std::vector<UnitD> fu;
      std::vector<int> fm;
      pick(0, [](const pgxc_ctx *c) { return c->ispin == 2; }, fu, fm);
      pend[k].push_back(stage(fu, fm, K_FUNC, 0, k));
      pend[k].push_back(stage(fu, fm, wkind, 0, k));  // both spins' W per alpha unit
      for (int m : fm) pend[k].back().lmax = std::max(pend[k].back().lmax, mols[m]->lmax);
    }
  }
  // queue bounds: each list's positions (natural order, so neighbouring items
  // keep sharing operand panels) cut into kQueues runs of equal predicted cost
  {
    std::vector<double> cost = flops;
    const size_t nrow = flops.size();
    if (smode == PGXC_SCHED_MODEL && nrow) {
      if (costmodel_features(sc.model) != PGXC_NFEATURES)
        return fail(errc, PGXC_ERR_ARG, "cost model does not take PGXC_NFEATURES features");
      std::vector<double> pred(nrow);
      int rc = pgxc_costmodel_predict(sc.model, plan->feat.data(), (int64_t)nrow, pred.data());
      if (rc) return fail(errc, rc, std::string("cost model: ") + pgxc_costmodel_last_error(sc.model));
      for (size_t i = 0; i < nrow; ++i) cost[i] = std::exp(std::min(pred[i], 50.0));
    } else {
      for (size_t i = 0; i < nrow; ++i) cost[i] = flops[i] + 1.0;
    }
    long nq = 0;
    for (auto &pk : pend)
      for (auto &P : pk) nq += P.has_queues;
    std::vector<int> zeros((size_t)std::max<long>(1, nq * kQueues), 0);
    const size_t ctr0 = put(zeros.data(), sizeof(int) * zeros.size());
    plan->nctr = nq * kQueues;
    long qi = 0;
    for (auto &pk : pend)
      for (auto &P : pk) {
        if (!P.has_queues) continue;
        double tot = 0.0;
        for (long i = 0; i < P.total; ++i) tot += cost[P.frow + i];
        long qb[kQueues + 1];
        qb[0] = 0;
        double acc = 0.0;
        int q = 1;
        for (long i = 0; i < P.total && q < kQueues; ++i) {
          acc += cost[P.frow + i];
          while (q < kQueues && acc >= tot * q / kQueues) qb[q++] = i + 1;
        }
        while (q <= kQueues) qb[q++] = P.total;
        P.qbeg_off = put(qb, sizeof(qb));
        // within a queue the costliest items go first, so its last items,
        // which set when the stage ends, are the cheapest; equal costs keep
        // their natural order (the items of one density tile stay together)
        std::vector<int> order((size_t)P.total);
        std::iota(order.begin(), order.end(), 0);
        for (int k = 0; k < kQueues; ++k)
          std::stable_sort(order.begin() + qb[k], order.begin() + qb[k + 1], [&](int a, int b) {
            return cost[P.frow + a] > cost[P.frow + b];
          });
        P.has_order = true;
        P.order_off = put(order.data(), sizeof(int) * order.size());
        P.ctr_off = ctr0 + sizeof(int) * kQueues * qi++;
      }
    plan_ctr_off_tmp = ctr0;
  }
  // operand preparation and packing per molecule and spin
  std::vector<PackD> packs;
  for (int m = 0; m < nm; ++m) {
    pgxc_ctx *c = mols[m];
    for (int sp = 0; sp < c->ispin; ++sp) {
      PackD Q;
      memset(&Q, 0, sizeof(Q));
      Q.Hpart = sp ? c->beta.Hpart : c->Hpart;
      Q.Vout = sp ? c->beta.Vout : c->Vout;
      Q.Vstate = sp ? c->beta.Vstate : c->Vstate;
      Q.P = sp ? c->beta.P : c->P;
      Q.Pref = sp ? c->beta.Pref : c->Pref;
      Q.D = sp ? c->beta.D : c->D;
      Q.fpos = c->d_fpos;
      Q.n = c->nbas;
      Q.NP = c->NP;
      Q.nsplit = c->nsplit * lanes_of(c);  // the fold sums the partials of every lane
      Q.incremental = incr[m];
      packs.push_back(Q);
      for (int l = 0; l < lanes_of(c); ++l)
        plan->hzero.push_back({(int *)(Q.Hpart + (size_t)l * c->nsplit * c->NP * c->NP),
                               2L * c->nsplit * c->NP * c->NP, l});
    }
  }
  const int npack = (int)packs.size();
  std::vector<long> pbase(npack + 1, 0), qbase(npack + 1, 0), fbase(npack + 1, 0);
  for (int q = 0; q < npack; ++q) {
    const PackD &Q = packs[q];
    const long npk = Q.n * (Q.n + 1) / 2;
    pbase[q + 1] = pbase[q] + (npk + kThreads - 1) / kThreads;
    qbase[q + 1] = qbase[q] + (Q.NP * Q.NP + kThreads - 1) / kThreads;
    const long nb = Q.NP / 32;
    fbase[q + 1] = fbase[q] + nb * (nb + 1) / 2;
  }
  plan->nmol = npack;
  const size_t pack_off = put(packs.data(), sizeof(PackD) * npack);
  const size_t pb_off = put(pbase.data(), sizeof(long) * pbase.size());
  const size_t qb_off = put(qbase.data(), sizeof(long) * qbase.size());
  const size_t fb_off = put(fbase.data(), sizeof(long) * fbase.size());
  HIPCHK(errc, hipMalloc((void **)&plan->blob, std::max<size_t>(img.size(), 256)));
  HIPCHK(errc, hipMemcpy(plan->blob, img.data(), img.size(), hipMemcpyHostToDevice));
  char *B = plan->blob;
  plan->d_ctr = (int *)(B + plan_ctr_off_tmp);
  auto mk = [&](const Pending &P) {
    StageList L{};
    L.units = (const UnitD *)(B + P.units_off);
    L.nu = P.nu;
    L.base = (const long *)(B + P.base_off);
    L.order = P.has_order ? (const int *)(B + P.order_off) : nullptr;
    L.total = P.total;
    L.ctr = P.has_queues ? (int *)(B + P.ctr_off) : nullptr;
    L.qbeg = P.has_queues ? (const long *)(B + P.qbeg_off) : nullptr;
    L.tlog = P.has_log ? (uint32_t *)(B + P.tlog_off) : nullptr;
    L.lmax = P.lmax;
    if (P.has_log) plan->logs.push_back({L.tlog, P.total});
    return L;
  };
  plan->steps.resize(nsteps);
  for (int k = 0; k < nsteps; ++k) {
    StepPlan &S = plan->steps[k];
    S.basis = mk(pend[k][0]);
    S.func = mk(pend[k][1]);
    S.wmat = mk(pend[k][2]);
    for (int li = 0; li < kLists; ++li) S.dens[li] = mk(pend[k][3 + li]);
    for (int li = 0; li < kLists; ++li) S.pot[li] = mk(pend[k][3 + kLists + li]);
    S.func2 = mk(pend[k][3 + 2 * kLists]);
    S.wmat2 = mk(pend[k][4 + 2 * kLists]);
    S.nunits = pend[k][0].nu;
  }
  plan->nlanes = plan_lanes;
  plan->d_pack = (PackD *)(B + pack_off);
  plan->d_pack_base = (long *)(B + pb_off);
  plan->d_prep_base = (long *)(B + qb_off);
  plan->d_fold_base = (long *)(B + fb_off);
  plan->fold_total = fbase[npack];
  plan->pack_total = pbase[npack];
  plan->prep_total = qbase[npack];
  return PGXC_OK;
}

// Item log of a plan's last build: features and microseconds per item.
static int64_t read_item_log(const Plan *pl, hipStream_t s, double *features, double *us,
                             int64_t cap) {
  if (!pl || !pl->log) return 0;
  int64_t n = 0;
  for (auto &lg : pl->logs) n += lg.second;
  if (!features || !us || cap < n) return n;
  if (hipStreamSynchronize(s) != hipSuccess) return -PGXC_ERR_HIP;
  int64_t row = 0;
  std::vector<uint32_t> t;
  for (auto &lg : pl->logs) {
    t.resize((size_t)lg.second);
    if (hipMemcpy(t.data(), lg.first, sizeof(uint32_t) * lg.second, hipMemcpyDeviceToHost) != hipSuccess)
      return -PGXC_ERR_HIP;
    for (long i = 0; i < lg.second; ++i, ++row) us[row] = t[(size_t)i] * pl->tick_us;
  }
  memcpy(features, pl->feat.data(), sizeof(double) * PGXC_NFEATURES * n);
  return n;
}

static int check_sched(int32_t mode, const pgxc_costmodel *model) {
  if (mode < PGXC_SCHED_STATIC || mode > PGXC_SCHED_MODEL) return PGXC_ERR_ARG;
  if (mode == PGXC_SCHED_MODEL && !model) return PGXC_ERR_ARG;
  return PGXC_OK;
}

// Events of a timed build: kEvStart / kEvPack / kEvEnd bracket the build and
// its packing; step k of the plan records five more from kEvStep + 5 k on
// (before basis, after basis, after the density contraction, after the
// functional + W, after the potential contraction).  Nothing waits on them
// until the build has ended, so timing adds no host synchronisation.
constexpr int kEvStart = 0, kEvPack = 1, kEvEnd = 2, kEvStep = 3;

// Launch a plan, recording the events above when ev is given.  Chunk k runs
// on lane k % 2 of a two-lane plan: stream s or s2, each lane with its own
// work buffers, so one chunk's stages (and the ends of its launches) overlap
// the next chunk's; s2 starts after the operand preparation (fork) and the
// fold waits for its last chunk (join).
template <typename T>
static void launch_plan(hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join,
                        const Plan &pl, bool gga, std::vector<hipEvent_t> *ev) {
  const bool x3 = pl.precision == PGXC_PREC_BF16X3;
  const bool two = pl.nlanes > 1 && pl.steps.size() > 1;
  if (pl.nctr) launch_zero_ints(s, pl.d_ctr, pl.nctr);
  // potential items add into H: each lane's partials cleared on its stream
  if (!pl.energy)
    for (auto &z : pl.hzero)
      if (z.lane == 0 || !two) launch_zero_ints(s, z.p, z.n);
  launch_prep(s, pl.d_pack, pl.nmol, pl.d_prep_base, pl.prep_total,
              x3 ? 2 : std::is_same<T, double>::value ? 0 : 1, false);
  if (two) {
    (void)hipEventRecord(fork, s);
    (void)hipStreamWaitEvent(s2, fork, 0);
    if (!pl.energy)
      for (auto &z : pl.hzero)
        if (z.lane != 0) launch_zero_ints(s2, z.p, z.n);
  }
  for (size_t k = 0; k < pl.steps.size(); ++k) {
    const StepPlan &S = pl.steps[k];
    const hipStream_t t = two && (k & 1) ? s2 : s;
    hipEvent_t *E = ev ? ev->data() + kEvStep + 5 * k : nullptr;
    if (E) (void)hipEventRecord(E[0], t);
    launch_basis<T>(t, gga, S.basis);
    if (E) (void)hipEventRecord(E[1], t);
    for (int li = 0; li < kLists; ++li) {
      const int tf = kTiles[li / kNumVariants], v = li % kNumVariants;
      if constexpr (std::is_same<T, float>::value) {
        if (x3) {
          launch_density_bf16x3(t, gga, tf, S.dens[li]);
          continue;
        }
      }
      launch_density<T>(t, gga, tf, v, S.dens[li]);
    }
    if (E) (void)hipEventRecord(E[2], t);
    if (pl.energy) {
      launch_energy(t, gga, false, S.func);
      launch_energy(t, gga, true, S.func2);
      continue;
    }
    launch_functional(t, gga, S.func);
    launch_functional_spin(t, gga, S.func2);
    if (pl.wbasis) {
      launch_wbasis(t, gga, false, S.wmat);
      launch_wbasis(t, gga, true, S.wmat2);
    } else {
      launch_wmat<T>(t, gga, x3, false, S.wmat);
      launch_wmat<T>(t, gga, x3, true, S.wmat2);
    }
    if (E) (void)hipEventRecord(E[3], t);
    for (int li = 0; li < kLists; ++li) {
      const int tf = kTiles[li / kNumVariants], v = li % kNumVariants;
      if constexpr (std::is_same<T, float>::value) {
        if (x3) {
          launch_potential_bf16x3(t, tf, S.pot[li]);
          continue;
        }
      }
      launch_potential<T>(t, tf, v, S.pot[li]);
    }
    if (E) (void)hipEventRecord(E[4], t);
  }
  if (two) {
    (void)hipEventRecord(join, s2);
    (void)hipStreamWaitEvent(s, join, 0);
  }
  if (pl.energy) return;
  if (ev) (void)hipEventRecord((*ev)[kEvPack], s);
  launch_fold(s, pl.d_pack, pl.nmol, pl.d_fold_base, pl.fold_total);
  launch_pack(s, pl.d_pack, pl.nmol, pl.d_pack_base, pl.pack_total);
  launch_prep(s, pl.d_pack, pl.nmol, pl.d_prep_base, pl.prep_total, 0, true);
  if (ev) (void)hipEventRecord((*ev)[kEvEnd], s);
}

// Bytes of basis caches held per device by all contexts of the process.
static std::mutex g_cache_mu;
static std::map<int, size_t> g_cache_bytes;

// bytes of basis values per (function, point): value and gradient planes
static size_t phi_elem_bytes(size_t elem) { return 4 * elem; }

// Test switch PGXC_POISON_BASIS=1: basis buffers, when they get their first
// contents (allocation, or the zeroing of the lane buffers at a change of
// precision), are filled with NaN bytes wherever a build writes basis values
// (rows of the real functions), so a stage
// reading a block the basis stage left unwritten turns the results into NaN.
static hipError_t poison_basis(const pgxc_ctx *c, void *base, size_t elem, size_t per_buf,
                               size_t nbuf, hipStream_t s) {
  const char *pz = getenv("PGXC_POISON_BASIS");
  if (!pz || pz[0] != '1') return hipSuccess;
  const size_t rows = (size_t)c->nbas * c->Mc, plane = (size_t)c->NP * c->Mc;
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < nbuf && e == hipSuccess; ++k) {
    char *b = (char *)base + k * per_buf;
    for (int pl = 0; pl < 4 && e == hipSuccess; ++pl)
      e = hipMemsetAsync(b + pl * plane * elem, 0xff, rows * elem, s);
  }
  return e;
}

// grad: the build reads the gradient planes of the basis values (GGA)
static int ensure_work(pgxc_ctx *c, size_t elem, int precision, hipStream_t s, bool grad) {
  const size_t need_phi = phi_elem_bytes(elem) * (size_t)c->NP * c->Mc;
  const size_t need_w = (size_t)c->NP * c->Mc * elem * c->nlanes_cap;
  c->w_lane = (size_t)c->NP * c->Mc * elem;
  bool moved = false;
  if (c->w_bytes < need_w) {
    dfree(c->W);
    c->W = nullptr;
    c->w_bytes = 0;
    HIPCHK(c, hipMalloc(&c->W, need_w));
    c->w_bytes = need_w;
    moved = true;
  }
  if (c->ispin == 2 && c->beta.w_bytes < need_w) {
    dfree(c->beta.W);
    c->beta.W = nullptr;
    c->beta.w_bytes = 0;
    HIPCHK(c, hipMalloc(&c->beta.W, need_w));
    c->beta.w_bytes = need_w;
    moved = true;
  }
  // basis cache: every chunk's values, when they take at most a quarter of
  // the device, leave 32 GiB free, and the caches of all contexts on the
  // device stay within half of it (the members of a batch keep room for their
  // own work buffers)
  const int e = elem == 8 ? 0 : 1;
  if (!c->cache_off && !c->phic[e] && !c->phic_failed[e]) {
    const size_t need = need_phi * (size_t)c->nchunks;
    size_t fr = 0, tot = 0;
    c->phic_failed[e] = 1;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    size_t &held = g_cache_bytes[c->device];
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && need <= tot / 4 && need + (32ull << 30) <= fr &&
        held + need <= tot / 2 && hipMalloc(&c->phic[e], need) == hipSuccess) {
      c->phic_failed[e] = 0;
      c->phic_bytes[e] = need;
      held += need;
      HIPCHK(c, hipMemsetAsync(c->phic[e], 0, need, s));  // padding rows read as zero
      HIPCHK(c, poison_basis(c, c->phic[e], elem, need_phi, c->nchunks, s));
      c->phic_valid[e] = 0;
      c->phic_grad[e] = 0;
    } else {
      c->phic[e] = nullptr;
      (void)hipGetLastError();
    }
  }
  c->phi_lane = need_phi;
  if (!c->phic[e]) {  // one chunk buffer per lane, rewritten by every build
    if (c->phi_bytes < need_phi * c->nlanes_cap) {
      dfree(c->phi);
      c->phi = nullptr;
      c->phi_bytes = 0;
      HIPCHK(c, hipMalloc(&c->phi, need_phi * c->nlanes_cap));
      c->phi_bytes = need_phi * c->nlanes_cap;
      c->phi_elem = 0;
    }
    // function rows nbas..NP-1 are never written and must read as zero in the
    // element type of this build (a buffer last used at another precision
    // holds that precision's bytes there)
    if (c->phi_elem != elem) {
      HIPCHK(c, hipMemsetAsync(c->phi, 0, c->phi_bytes, s));
      c->phi_elem = elem;
      HIPCHK(c, poison_basis(c, c->phi, elem, need_phi, c->nlanes_cap, s));
    }
  }
  void *base = c->phic[e] ? c->phic[e] : c->phi;
  const size_t stride = c->phic[e] ? need_phi : 0;
  // an LDA pass writes only the value planes: a GGA build after it evaluates again
  const int skip = c->phic[e] && c->phic_valid[e] && (c->phic_grad[e] || !grad);
  if (base != c->phi_base || stride != c->phi_stride || skip != c->skip_basis) moved = true;
  c->phi_base = base;
  c->phi_stride = stride;
  c->skip_basis = skip;
  if (moved) {  // their unit descriptors hold the old pointers
    free_plan(c->plan);
    free_plan(c->plan_e);
  }
  return PGXC_OK;
}

// after a build: the cache of this element type now holds valid values
static void mark_basis_cached(pgxc_ctx *c, size_t elem, int precision, bool grad) {
  const int e = elem == 8 ? 0 : 1;
  if (c->phic[e]) {
    c->phic_valid[e] = 1;
    if (grad) c->phic_grad[e] = 1;
  }
}

static int validate_build(pgxc_ctx *c, uint32_t functional, int32_t precision, int &incr) {
  if (!c) return PGXC_ERR_ARG;
  if (functional == 0 || (functional & ~kSupported))
    return fail(c, PGXC_ERR_UNSUPPORTED, "functional mask has terms other than XAlpha, PWLDAc, PBEx, PBEc");
  if ((functional & PGXC_XC_PBEC) && !(functional & PGXC_XC_PWLDAC))
    return fail(c, PGXC_ERR_ARG, "PBEc needs the PW92 LDA correlation it is built on (PWLDAc)");
  if (precision != PGXC_PREC_FP64 && precision != PGXC_PREC_FP32 && precision != PGXC_PREC_BF16X3)
    return fail(c, PGXC_ERR_UNSUPPORTED, "precision mode " + std::to_string(precision) + " not available");
  if (!c->have_density) return fail(c, PGXC_ERR_STATE, "pgxc_build before pgxc_set_density/orbitals");
  incr = 0;
  if (c->density_mode == PGXC_DENSITY_INCREMENTAL) {
    if (!c->have_state)
      return fail(c, PGXC_ERR_STATE, "incremental density but no previous build on this context");
    if (c->state_functional != functional)
      return fail(c, PGXC_ERR_STATE, "incremental build with a functional different from the previous build");
    incr = 1;
  }
  return PGXC_OK;
}

static void set_stats(pgxc_ctx *c, int precision) {
  const double B = (double)c->npts;
  c->stats[0] = B;
  c->stats[1] = std::ceil(B / c->veclen);
  // sum over 128-point tiles of 4 B n_b^2 + 16 B n_b, n_b = functions of the
  // shells not cut off on the tile (dense: n_b = nbas)
  c->stats[2] = (c->screen ? c->flops_screened : c->flops_dense) * c->ispin;
  c->stats[3] = B * c->exps_per_point;
  // sum over the same tiles of 32 B (x, y, z, w) + 16 n_b^2 per spin (P
  // sub-block read, V partial written) + 4 s B n_b (basis values and
  // gradients materialised at s bytes per element)
  const double sb = precision == PGXC_PREC_FP64 ? 8.0 : 4.0;
  c->stats[4] = 32.0 * B + 16.0 * c->sum_n2 * c->ispin + 4.0 * sb * c->sum_bn;
}

static int collect(pgxc_ctx *c, hipStream_t s, uint32_t functional, int precision) {
  std::vector<double> ep(2 * (size_t)c->nblk_total);
  HIPCHK(c, hipMemcpyAsync(ep.data(), c->epart, ep.size() * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  double e = 0.0, q = 0.0;
  for (long b = 0; b < c->nblk_total; ++b) {
    e += ep[2 * b];
    q += ep[2 * b + 1];
  }
  c->last_exc = e;
  c->last_charge = q;
  c->have_state = 1;
  c->state_functional = functional;
  set_stats(c, precision);
  return PGXC_OK;
}

// FNV-1a style hash over 8-byte words (a shorter tail byte by byte).
static uint64_t hash_words(uint64_t h, const void *p, size_t bytes) {
  const unsigned char *b = (const unsigned char *)p;
  size_t i = 0;
  for (; i + 8 <= bytes; i += 8) {
    uint64_t w;
    memcpy(&w, b + i, 8);
    h = (h ^ w) * 1099511628211ull;
  }
  for (; i < bytes; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

// Order of n points (coordinate k of point p at xyz[k * stride + p * step])
// along a Morton curve over their bounding box (10 bits per axis), ties in
// input order.
static void morton_order(const double *xyz, long n, long stride, long step,
                         std::vector<long> &perm) {
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = hi[k] = xyz[(size_t)k * stride];
    for (long p = 1; p < n; ++p) {
      lo[k] = std::min(lo[k], xyz[(size_t)k * stride + (size_t)p * step]);
      hi[k] = std::max(hi[k], xyz[(size_t)k * stride + (size_t)p * step]);
    }
  }
  std::vector<std::pair<uint32_t, long>> key((size_t)n);
  for (long p = 0; p < n; ++p) {
    uint32_t code = 0;
    for (int k = 0; k < 3; ++k) {
      const double span = hi[k] - lo[k];
      const double f = span > 0 ? (xyz[(size_t)k * stride + (size_t)p * step] - lo[k]) / span : 0.0;
      const uint32_t q = (uint32_t)std::min(1023.0, std::max(0.0, f * 1023.0));
      for (int b = 0; b < 10; ++b) code |= ((q >> b) & 1u) << (3 * b + k);
    }
    key[(size_t)p] = {code, p};
  }
  std::sort(key.begin(), key.end());
  perm.resize((size_t)n);
  for (long p = 0; p < n; ++p) perm[(size_t)p] = key[(size_t)p].second;
}

// Internal order of the basis functions: shells by reach class (the radius
// sqrt(50 / alpha_min) beyond which the shell is cut off, in steps of 10 % on
// a log scale, longest reach first), then by their atom's place on a Morton
// curve, then in caller order.  Shells of one element share their exponents,
// so a class holds the like shells of neighbouring atoms: the functions that
// can be nonzero on a compact group of points fill whole runs of 8 and few
// 32-function blocks.  A shell keeps its functions together and in caller
// order.  Sets each shell's internal func_off and fpos[caller function] =
// internal position.
static void function_order(std::vector<ShellD> &sh, const double *exps, const double *atom_xyz,
                           int n_atoms, std::vector<int> &fpos) {
  const size_t ns = sh.size();
  std::vector<double> reach(ns);
  for (size_t i = 0; i < ns; ++i) {
    double amin = exps[sh[i].exp_off];
    for (int e = 1; e < sh[i].n_exp; ++e) amin = std::min(amin, exps[sh[i].exp_off + e]);
    reach[i] = std::sqrt(50.0 / amin);
  }
  std::vector<long> aperm;
  morton_order(atom_xyz, n_atoms, 1, 3, aperm);
  std::vector<int> arank(n_atoms);
  for (int r = 0; r < n_atoms; ++r) arank[aperm[r]] = r;
  std::vector<int> idx(ns);
  std::iota(idx.begin(), idx.end(), 0);
  auto cls = [&](int i) { return -(int)std::floor(std::log(reach[i]) / std::log(1.1)); };
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (cls(a) != cls(b)) return cls(a) < cls(b);
    return arank[sh[a].atom] < arank[sh[b].atom];
  });
  int nbas = 0;
  for (auto &d : sh) nbas += (2 * d.l + 1) * (d.n_unc + d.n_contr);
  fpos.assign(nbas, 0);
  int off = 0;
  for (int i : idx) {
    const int nf = (2 * sh[i].l + 1) * (sh[i].n_unc + sh[i].n_contr);
    for (int f = 0; f < nf; ++f) fpos[sh[i].func_off + f] = off + f;
    sh[i].func_off = off;
    off += nf;
  }
}

// Screening and work tables of the padded grid and the algorithmic work of one
// build.  k_screen marks, per 32-point group, the runs of 8 functions that can
// be nonzero there (all runs when screening is off); the host turns the marks
// into run lists per 128-point tile and per point split, [group][chunk]
// flags per split, and the work items of every chunk.
template <typename V>
static int upload(pgxc_ctx *c, V **dst, const std::vector<V> &src) {
  int rc = dalloc(c, dst, src.size());
  if (rc) return rc;
  if (!src.empty()) HIPCHK(c, hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
  return PGXC_OK;
}

static int setup_screen(pgxc_ctx *c, const std::vector<ShellD> &sh, const double *exps, bool on) {
  const long ntp = c->npad / kScreenTile, ngr = c->npad / kGroup;
  const double n = (double)c->nbas;
  c->flops_dense = 4.0 * c->npts * n * n + 16.0 * c->npts * n;
  c->n8 = c->NP / kRun;
  c->CH = c->tf / kRun;
  c->screen = on;
  c->nw8 = (c->n8 + 31) / 32;
  const int n8 = c->n8, nw8 = c->nw8;
  const long G = c->Mc / c->nsplit / kGroup;
  if (n8 > kMaxRuns) return fail(c, PGXC_ERR_UNSUPPORTED, "more than 8192 basis functions (padded)");
  if (G > kMaxSplitGroups) return fail(c, PGXC_ERR_UNSUPPORTED, "point split above 65536 points");
  std::vector<ScreenShell> ss(sh.size());
  for (size_t i = 0; i < sh.size(); ++i) {
    double amin = exps[sh[i].exp_off];
    for (int e = 1; e < sh[i].n_exp; ++e) amin = std::min(amin, exps[sh[i].exp_off + e]);
    ss[i].amin = amin;
    ss[i].atom = sh[i].atom;
    ss[i].func_off = sh[i].func_off;
    ss[i].nfunc = (2 * sh[i].l + 1) * (sh[i].n_unc + sh[i].n_contr);
    ss[i].pad = 0;
  }
  int rc;
  if ((rc = upload(c, &c->sshells, ss))) return rc;
  if ((rc = dalloc(c, &c->nbt, ntp))) return rc;
  if ((rc = dalloc(c, &c->sig8, (size_t)ngr * nw8))) return rc;
  const int nws = (c->n_shells + 31) / 32;
  ScopedDev tmp;
  uint32_t *d_smark = nullptr;
  HIPCHK(c, hipMalloc((void **)&d_smark, sizeof(uint32_t) * (size_t)ntp * nws));
  tmp.keep(d_smark);
  launch_screen(c->stream, ntp, c->grid, c->grid + c->npad, c->grid + 2 * c->npad, c->n_atoms,
                c->atoms, c->n_shells, c->sshells, n8, c->nbt, c->sig8, d_smark);
  HIPCHK(c, hipGetLastError());
  std::vector<int> nb((size_t)ntp);
  std::vector<uint32_t> sig((size_t)ngr * nw8), smark((size_t)ntp * nws);
  HIPCHK(c, hipMemcpyAsync(nb.data(), c->nbt, sizeof(int) * ntp, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(sig.data(), c->sig8, sizeof(uint32_t) * sig.size(), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(smark.data(), d_smark, sizeof(uint32_t) * smark.size(),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (on) {
    double f = 0.0;
    for (long t = 0; t < ntp; ++t) {
      const double B = (double)std::max<long>(0, std::min<long>(kScreenTile, c->npts - t * kScreenTile));
      f += 4.0 * B * nb[(size_t)t] * nb[(size_t)t] + 16.0 * B * nb[(size_t)t];
      if (B > 0) {
        c->sum_bn += B * nb[(size_t)t];
        c->sum_n2 += (double)nb[(size_t)t] * nb[(size_t)t];
      }
    }
    c->flops_screened = f;
  } else {  // every run counts on every group
    c->flops_screened = c->flops_dense;
    c->sum_bn = (double)c->npts * n;
    c->sum_n2 = std::ceil((double)c->npts / kScreenTile) * n * n;
    for (long g = 0; g < ngr; ++g)
      for (int r = 0; r < n8; ++r) sig[(size_t)g * nw8 + r / 32] |= 1u << (r % 32);
    HIPCHK(c, hipMemcpy(c->sig8, sig.data(), sizeof(uint32_t) * sig.size(), hipMemcpyHostToDevice));
    for (long t = 0; t < ntp; ++t)
      for (int si = 0; si < c->n_shells; ++si) smark[(size_t)t * nws + si / 32] |= 1u << (si % 32);
  }
  // per 128-point tile: the functions of its marked shells, ascending (the
  // density contraction's output columns)
  // (and as a bitmap of nwf words per tile)
  c->nwf = (c->NP + 31) / 32;
  std::vector<uint32_t> tfb((size_t)ntp * c->nwf, 0u);
  {
    c->h_tf_off.assign(1, 0);
    c->h_tfl.clear();
    for (long t = 0; t < ntp; ++t) {
      uint32_t *fm = tfb.data() + (size_t)t * c->nwf;
      for (int si = 0; si < c->n_shells; ++si)
        if ((smark[(size_t)t * nws + si / 32] >> (si % 32)) & 1u)
          for (int f = ss[(size_t)si].func_off; f < ss[(size_t)si].func_off + ss[(size_t)si].nfunc; ++f)
            fm[(size_t)f / 32] |= 1u << (f % 32);
      for (int f = 0; f < c->NP; ++f)
        if ((fm[(size_t)f / 32] >> (f % 32)) & 1u) c->h_tfl.push_back((uint16_t)f);
      c->h_tf_off.push_back((int)c->h_tfl.size());
    }
  }
  if ((rc = upload(c, &c->tfb, tfb))) return rc;
  auto bit = [&](long g, int r) { return (sig[(size_t)g * nw8 + r / 32] >> (r % 32)) & 1u; };
  // per 128-point tile: its runs (union over its four groups), and per run
  // its functions in the tile's list
  c->h_tl_off.assign(1, 0);
  c->h_tl.clear();
  c->h_tlm.clear();
  for (long t = 0; t < ntp; ++t) {
    for (int r = 0; r < n8; ++r)
      if (bit(4 * t, r) | bit(4 * t + 1, r) | bit(4 * t + 2, r) | bit(4 * t + 3, r)) {
        c->h_tl.push_back((uint16_t)r);
        c->h_tlm.push_back((uint8_t)(tfb[(size_t)t * c->nwf + (size_t)r / 4] >> (8 * (r % 4))));
      }
    c->h_tl_off.push_back((int)c->h_tl.size());
  }
  // per split: the union of its tiles' functions (the potential's rows and
  // columns, in chunks of tf functions), and [group][chunk] flags: whether a
  // run holding a function of the chunk is live on the group
  const long NS = (long)c->nchunks * c->nsplit;
  std::vector<uint32_t> wneed((size_t)ngr * nw8, 0u);
  c->h_su_off.assign(1, 0);
  c->h_su.clear();
  c->h_sumask.clear();
  c->h_gcf_off.clear();
  c->h_gcf.clear();
  {
    const int tf = c->tf;
    std::vector<uint32_t> fm((size_t)(c->NP + 31) / 32);
    std::vector<int> cr;  // the distinct runs of one chunk
    for (long S = 0; S < NS; ++S) {
      const long g0 = S * G;
      std::fill(fm.begin(), fm.end(), 0u);
      for (long t = g0 / 4; t < (g0 + G) / 4; ++t)
        for (int i = c->h_tf_off[(size_t)t]; i < c->h_tf_off[(size_t)t + 1]; ++i) {
          const int f = c->h_tfl[(size_t)i];
          fm[(size_t)f / 32] |= 1u << (f % 32);
        }
      const size_t u0 = c->h_su.size();
      for (int f = 0; f < c->NP; ++f)
        if ((fm[(size_t)f / 32] >> (f % 32)) & 1u) c->h_su.push_back((uint16_t)f);
      c->h_su_off.push_back((int)c->h_su.size());
      // per entry: the split's tiles whose list holds the function
      const long st0 = g0 / 4, nst = G / 4;
      for (size_t e = u0; e < c->h_su.size(); ++e) {
        const int f = c->h_su[e];
        uint32_t m = 0;
        if (nst > 32)
          m = ~0u;
        else
          for (long t = 0; t < nst; ++t)
            m |= ((tfb[(size_t)(st0 + t) * c->nwf + (size_t)f / 32] >> (f % 32)) & 1u) << t;
        c->h_sumask.push_back(m);
      }
      const int nu = (int)(c->h_su.size() - u0), nch = (nu + tf - 1) / tf;
      c->h_gcf_off.push_back((long)c->h_gcf.size());
      const size_t gf0 = c->h_gcf.size();
      c->h_gcf.resize(gf0 + (size_t)G * nch, 0);
      for (int ch = 0; ch < nch; ++ch) {
        cr.clear();
        for (int e = ch * tf; e < std::min(nu, (ch + 1) * tf); ++e) {
          const int r = c->h_su[u0 + e] / kRun;
          if (cr.empty() || cr.back() != r) cr.push_back(r);
        }
        for (long g = g0; g < g0 + G; ++g) {
          uint8_t f = 0;
          for (size_t i = 0; i < cr.size() && !f; ++i) f = (uint8_t)bit(g, cr[i]);
          c->h_gcf[gf0 + (size_t)(g - g0) * nch + ch] = f;
          // the potential contraction reads W on this group for every function of a flagged chunk
          if (f)
            for (int r : cr) wneed[(size_t)g * nw8 + r / 32] |= 1u << (r % 32);
        }
      }
    }
  }
  if ((rc = upload(c, &c->wbits, wneed))) return rc;
  // What the basis stage writes, per (shell, 32-point group): values and
  // gradients of the shells marked on the group's tile, the functions of the
  // tile's list (the density's operand and epilogue, the W pass and the
  // potentials read no others: they stage zeros for the rest).  When a point
  // split spans more than 32 tiles, the potentials' tile masks are off, and
  // values are also written on the runs of the tile (the density's K) and of
  // the chunks the potentials read there (wneed).  No stage reads any other
  // block of a basis-value buffer.
  {
    std::vector<uint32_t> need((size_t)ngr * 2 * nw8, 0u), tw((size_t)nw8);
    for (long t = 0; t < ntp; ++t) {
      std::fill(tw.begin(), tw.end(), 0u);
      for (int i = c->h_tl_off[(size_t)t]; i < c->h_tl_off[(size_t)t + 1]; ++i)
        tw[c->h_tl[(size_t)i] >> 5] |= 1u << (c->h_tl[(size_t)i] & 31);
      for (long g = 4 * t; g < 4 * t + 4; ++g)
        for (int w = 0; w < nw8; ++w) {
          need[(size_t)g * 2 * nw8 + w] = tw[(size_t)w] | wneed[(size_t)g * nw8 + w];
          need[(size_t)g * 2 * nw8 + nw8 + w] = tw[(size_t)w];
        }
    }
    // runs holding each atom range's functions
    const int nr = c->n_ranges;
    std::vector<uint32_t> rmask((size_t)nr * nw8, 0u);
    for (int ri = 0; ri < nr; ++ri) {
      const RangeD &R = c->h_ranges[(size_t)ri];
      for (int si = R.shell_begin; si < R.shell_begin + R.n_shells; ++si) {
        const ShellD &d = sh[(size_t)si];
        const int nf = (2 * d.l + 1) * (d.n_unc + d.n_contr);
        for (int r = d.func_off / kRun; r <= (d.func_off + nf - 1) / kRun; ++r)
          rmask[(size_t)ri * nw8 + r / 32] |= 1u << (r % 32);
      }
    }
    std::vector<uint32_t> bi, bfl, bm((size_t)nw8);
    const bool masked = G / 4 <= 32;
    c->h_bi_off.assign(1, 0);
    const long gpb = kThreads / kGroup;  // groups per basis block
    for (int k = 0; k < c->nchunks; ++k) {
      const long g0 = (long)k * c->Mc / kGroup;
      for (long b = 0; b < c->Mc / kThreads; ++b) {
        std::fill(bm.begin(), bm.end(), 0u);
        for (long g = g0 + b * gpb; g < g0 + (b + 1) * gpb; ++g)
          for (int w = 0; w < nw8; ++w) bm[(size_t)w] |= need[(size_t)g * 2 * nw8 + w];
        for (int ri = 0; ri < nr; ++ri) {
          uint32_t any = 0;
          for (int w = 0; w < nw8 && !any; ++w) any = bm[(size_t)w] & rmask[(size_t)ri * nw8 + w];
          if (!any) continue;
          // per (shell j of the range, group g of the block): values, gradients read
          uint32_t fw[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
          const RangeD &R = c->h_ranges[(size_t)ri];
          for (int j = 0; j < R.n_shells; ++j) {
            const int sg = R.shell_begin + j;
            const ShellD &d = sh[(size_t)sg];
            const int nf = (2 * d.l + 1) * (d.n_unc + d.n_contr);
            for (long gg = 0; gg < gpb; ++gg) {
              const uint32_t *nd = need.data() + (size_t)(g0 + b * gpb + gg) * 2 * nw8;
              uint32_t v = 0;
              for (int r = d.func_off / kRun; r <= (d.func_off + nf - 1) / kRun; ++r)
                v |= nd[r >> 5] >> (r & 31);
              // gradients: read only for the tile's functions (the density
              // epilogue's columns, and the W pass's live functions)
              const long tg = (g0 + b * gpb + gg) / 4;
              const uint32_t gr = smark[(size_t)tg * nws + sg / 32] >> (sg % 32);
              const int bit = 8 * j + (int)gg;
              // values: with the contractions skipping the functions outside
              // a tile's list (always in the density; in the potentials when a
              // split spans at most 32 tiles), only those of the shells marked
              // on the tile
              fw[bit >> 5] |= ((masked ? gr : v) & 1u) << (bit & 31);
              fw[8 + (bit >> 5)] |= (gr & 1u) << (bit & 31);
            }
          }
          if (!(fw[0] | fw[1] | fw[2] | fw[3] | fw[4] | fw[5] | fw[6] | fw[7])) continue;
          bi.push_back((uint32_t)(b << 16 | ri));
          bfl.insert(bfl.end(), fw, fw + 16);
        }
      }
      c->h_bi_off.push_back((long)bi.size());
    }
    if ((rc = upload(c, &c->bi, bi))) return rc;
    if ((rc = upload(c, &c->bfl, bfl))) return rc;
    if (getenv("PGXC_DEBUG")) {  // sizes of the screened basis work
      long live = 0, nv = 0, ng = 0;
      for (long g = 0; g < ngr; ++g)
        for (int w = 0; w < nw8; ++w) {
          live += __builtin_popcount(sig[(size_t)g * nw8 + w]);
          nv += __builtin_popcount(need[(size_t)g * 2 * nw8 + w]);
          ng += __builtin_popcount(need[(size_t)g * 2 * nw8 + nw8 + w]);
        }
      fprintf(stderr, "pgxc: %ld groups x %d runs; blocks live %ld, values read %ld, gradients read %ld; "
                      "basis items %zu (of %ld)\n",
              ngr, n8, live, nv, ng, bi.size(), (long)c->nchunks * (c->Mc / kThreads) * c->n_ranges);
      // elements (function x point): basis values and gradients written,
      // W written (runs live or read by the potential) and of those the
      // ones in the tile's list
      double ev = 0, eg = 0, ew = 0, ewl = 0;
      for (size_t i = 0; i < bi.size(); ++i) {
        const RangeD &R = c->h_ranges[bi[i] & 0xffffu];
        const uint32_t *fw = bfl.data() + 16 * i;
        for (int j = 0; j < R.n_shells; ++j) {
          const ShellD &d = sh[(size_t)(R.shell_begin + j)];
          const int nf = (2 * d.l + 1) * (d.n_unc + d.n_contr);
          for (int gg = 0; gg < 8; ++gg) {
            const int bit = 8 * j + gg;
            ev += ((fw[bit >> 5] >> (bit & 31)) & 1u) * nf * 32.0;
            eg += ((fw[8 + (bit >> 5)] >> (bit & 31)) & 1u) * nf * 32.0;
          }
        }
      }
      for (long g = 0; g < ngr; ++g)
        for (int r = 0; r < n8; ++r)
          if (((sig[(size_t)g * nw8 + r / 32] | wneed[(size_t)g * nw8 + r / 32]) >> (r % 32)) & 1u) {
            ew += kRun * 32.0;
            for (int f = r * kRun; f < std::min(c->NP, (r + 1) * kRun); ++f)
              ewl += ((tfb[(size_t)(g / 4) * c->nwf + f / 32] >> (f % 32)) & 1u) * 32.0;
          }
      // values if only the tiles' runs were read (no potential chunks)
      double evt = 0;
      for (long g = 0; g < ngr; ++g) {
        const long t = g / 4;
        std::vector<uint32_t> tw2((size_t)nw8, 0u);
        for (int i = c->h_tl_off[(size_t)t]; i < c->h_tl_off[(size_t)t + 1]; ++i)
          tw2[c->h_tl[(size_t)i] >> 5] |= 1u << (c->h_tl[(size_t)i] & 31);
        for (const ShellD &d : sh) {
          const int nf = (2 * d.l + 1) * (d.n_unc + d.n_contr);
          uint32_t v = 0;
          for (int r = d.func_off / kRun; r <= (d.func_off + nf - 1) / kRun; ++r) v |= tw2[r >> 5] >> (r & 31);
          evt += (v & 1u) * nf * 32.0;
        }
      }
      fprintf(stderr, "pgxc: elements per build: basis values %.4g (tile runs only %.4g), gradients %.4g; "
                      "W %.4g, of them in the tiles' lists %.4g\n", ev, evt, eg, ew, ewl);
    }
  }
  // work items per chunk
  const long tpc = c->Mc / kScreenTile;
  c->h_di_off.assign(1, 0);
  c->h_pi_off.assign(1, 0);
  c->h_di.clear();
  c->h_pi.clear();
  for (int k = 0; k < c->nchunks; ++k) {
    for (long t = 0; t < tpc; ++t) {
      const long T = k * tpc + t;
      const int nf = c->h_tf_off[(size_t)T + 1] - c->h_tf_off[(size_t)T];
      // density items: pairs of column chunks (tf functions) of the tile's functions
      for (int pr = 0; pr < (nf + 2 * c->tf - 1) / (2 * c->tf); ++pr)
        c->h_di.push_back((uint32_t)(t << 8 | pr));
    }
    c->h_di_off.push_back((long)c->h_di.size());
    for (int sl = 0; sl < c->nsplit; ++sl) {
      const long S = (long)k * c->nsplit + sl;
      const int nu = c->h_su_off[(size_t)S + 1] - c->h_su_off[(size_t)S], nch = (nu + c->tf - 1) / c->tf;
      for (int ci = 0; ci < nch; ++ci)
        for (int cj = 0; cj < nch; ++cj)
          c->h_pi.push_back((uint32_t)sl << 20 | (uint32_t)ci << 10 | (uint32_t)cj);
    }
    c->h_pi_off.push_back((long)c->h_pi.size());
  }
  if ((rc = upload(c, &c->tl_off, c->h_tl_off))) return rc;
  if ((rc = upload(c, &c->tl, c->h_tl))) return rc;
  if ((rc = upload(c, &c->tlm, c->h_tlm))) return rc;
  if ((rc = upload(c, &c->tf_off, c->h_tf_off))) return rc;
  if ((rc = upload(c, &c->tfl, c->h_tfl))) return rc;
  if ((rc = upload(c, &c->su_off, c->h_su_off))) return rc;
  if ((rc = upload(c, &c->su, c->h_su))) return rc;
  if ((rc = upload(c, &c->sumask, c->h_sumask))) return rc;
  if ((rc = upload(c, &c->gcf_off, c->h_gcf_off))) return rc;
  if ((rc = upload(c, &c->gcf, c->h_gcf))) return rc;
  if ((rc = upload(c, &c->di, c->h_di))) return rc;
  if ((rc = upload(c, &c->pi, c->h_pi))) return rc;
  return PGXC_OK;
}

extern "C" {

static char g_free_err[256];  // message of the last failed context-free call

const char *pgxc_last_error(const pgxc_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  return g_free_err[0] ? g_free_err : "null context";
}

int64_t pgxc_nbas(const pgxc_ctx *ctx) { return ctx ? ctx->nbas : -1; }

int pgxc_create_opt(pgxc_ctx **out, int32_t n_atoms, const double *atom_xyz, int32_t n_shells,
                    const int32_t *shell_info, const double *exps, const double *norms,
                    const double *contr, const double *grid, int64_t npts, int32_t veclen,
                    const pgxc_options *opt) {
  if (!out) return PGXC_ERR_ARG;
  *out = nullptr;
  pgxc_ctx *c = new pgxc_ctx();
  *out = c;
  if (n_atoms <= 0 || n_shells <= 0 || !atom_xyz || !shell_info || !exps || !norms ||
      (!contr && n_shells > 0) || !grid || npts <= 0 || veclen <= 0)
    return fail(c, PGXC_ERR_ARG, "pgxc_create: null pointer or non-positive size");
  if (opt && opt->tile != 0 && opt->tile != 64 && opt->tile != 96 && opt->tile != 128)
    return fail(c, PGXC_ERR_ARG, "pgxc_create: tile must be 0, 64, 96 or 128");
  if (opt && (opt->chunk_points < 0 || opt->chunk_points % kSplitPts != 0))
    return fail(c, PGXC_ERR_ARG, "pgxc_create: chunk_points must be a non-negative multiple of 2048");
  if (opt && (opt->variant_fp64 < -1 || opt->variant_fp64 >= kNumVariants || opt->variant_fp32 < -1 ||
              opt->variant_fp32 >= kNumVariants))
    return fail(c, PGXC_ERR_ARG, "pgxc_create: variant must be -1 (default) or 0.." +
                                     std::to_string(kNumVariants - 1));
  if (opt && (opt->no_screen < 0 || opt->no_screen > 1 || opt->no_basis_cache < 0 ||
              opt->no_basis_cache > 1))
    return fail(c, PGXC_ERR_ARG, "pgxc_create: no_screen and no_basis_cache must be 0 or 1");
  c->cache_off = opt ? opt->no_basis_cache : 0;
  if (opt && opt->variant_fp64 >= 0) c->var[0] = opt->variant_fp64;
  if (opt && opt->variant_fp32 >= 0) c->var[1] = opt->variant_fp32;
  if (n_shells > kMaxShells)
    return fail(c, PGXC_ERR_UNSUPPORTED, "pgxc_create: more than " + std::to_string(kMaxShells) + " shells");
  c->n_atoms = n_atoms;
  c->n_shells = n_shells;
  c->npts = npts;
  c->veclen = veclen;

  // ---- basis tables (shells must be grouped by atom) ---------------------
  std::vector<ShellD> sh(n_shells);
  std::vector<AtomD> at(n_atoms);
  for (int a = 0; a < n_atoms; ++a) {
    at[a].x = atom_xyz[3 * a];
    at[a].y = atom_xyz[3 * a + 1];
    at[a].z = atom_xyz[3 * a + 2];
    at[a].shell_begin = at[a].shell_end = 0;
    at[a].lmax = 0;
    at[a].pad = 0;
    at[a].exp_begin = at[a].n_exp = at[a].contr_begin = at[a].n_contr = 0;
  }
  long eoff = 0, coff = 0, foff = 0;
  int prev_atom = -1;
  std::vector<int> seen(n_atoms, 0);
  for (int s = 0; s < n_shells; ++s) {
    const int32_t *si = shell_info + 5 * s;
    ShellD &d = sh[s];
    d.atom = si[0];
    d.l = si[1];
    d.n_exp = si[2];
    d.n_unc = si[3];
    d.n_contr = si[4];
    if (d.atom < 0 || d.atom >= n_atoms)
      return fail(c, PGXC_ERR_ARG, "pgxc_create: shell " + std::to_string(s) + " atom out of range");
    if (d.l < 0 || d.l > kLMax)
      return fail(c, PGXC_ERR_UNSUPPORTED, "pgxc_create: l = " + std::to_string(d.l) +
                                               " above the supported maximum " + std::to_string(kLMax));
    if (d.n_exp <= 0 || d.n_unc < 0 || d.n_contr < 0 || d.n_unc > d.n_exp ||
        d.n_unc + d.n_contr <= 0)
      return fail(c, PGXC_ERR_ARG, "pgxc_create: shell " + std::to_string(s) + " has inconsistent counts");
    if (d.n_contr > kMaxContr)
      return fail(c, PGXC_ERR_UNSUPPORTED, "pgxc_create: more than " + std::to_string(kMaxContr) +
                                               " contracted functions in one shell");
    if (d.atom != prev_atom) {
      if (seen[d.atom])
        return fail(c, PGXC_ERR_ARG, "pgxc_create: shells of one atom must be contiguous");
      seen[d.atom] = 1;
      at[d.atom].shell_begin = s;
      prev_atom = d.atom;
    }
    at[d.atom].shell_end = s + 1;
    at[d.atom].lmax = std::max(at[d.atom].lmax, d.l);
    c->lmax = std::max(c->lmax, d.l);
    d.exp_off = (int)eoff;
    d.contr_off = (int)coff;
    d.func_off = (int)foff;
    eoff += d.n_exp;
    coff += (long)d.n_exp * d.n_contr;
    foff += (long)(2 * d.l + 1) * (d.n_unc + d.n_contr);
  }
  for (int a = 0; a < n_atoms; ++a) {
    long np = 0, nc = 0;
    for (int s = at[a].shell_begin; s < at[a].shell_end; ++s) {
      np += sh[s].n_exp;
      nc += (long)sh[s].n_exp * sh[s].n_contr;
    }
    if (at[a].shell_end > at[a].shell_begin) {
      const ShellD &f = sh[at[a].shell_begin];
      at[a].exp_begin = f.exp_off;
      at[a].contr_begin = f.contr_off;
    }
    at[a].n_exp = (int)np;
    at[a].n_contr = (int)nc;
    if (at[a].shell_end - at[a].shell_begin > kAtomShells || np > kAtomPrims || nc > kAtomContr)
      return fail(c, PGXC_ERR_UNSUPPORTED,
                  "pgxc_create: atom " + std::to_string(a) + " has more than " +
                      std::to_string(kAtomShells) + " shells, " + std::to_string(kAtomPrims) +
                      " primitives or " + std::to_string(kAtomContr) + " contraction coefficients");
  }
  for (long e = 0; e < eoff; ++e)
    if (!(exps[e] > 0.0) || !std::isfinite(norms[e]))
      return fail(c, PGXC_ERR_ARG, "pgxc_create: exponents must be positive and norms finite");
  c->nbas = (int)foff;
  c->exps_per_point = (double)eoff;
  function_order(sh, exps, atom_xyz, n_atoms, c->fpos);
  c->fatom.assign(c->nbas, 0);
  for (const ShellD &d : sh)
    for (int j = 0; j < (2 * d.l + 1) * (d.n_unc + d.n_contr); ++j) c->fatom[d.func_off + j] = d.atom;
  // function tile: requested, or the one of 128 / 96 / 64 that pads nbas least
  if (opt && opt->tile) {
    c->tf = opt->tile;
    c->NP = (int)round_up(c->nbas, c->tf);
  } else {
    c->tf = 128;
    c->NP = (int)round_up(c->nbas, 128);
    for (int t : {96, 64})
      if (round_up(c->nbas, t) < c->NP) {
        c->NP = (int)round_up(c->nbas, t);
        c->tf = t;
      }
  }
  c->ntiles = c->NP / c->tf;

  // ---- chunking of the grid ----------------------------------------------
  // a chunk holds at most 256 Ki points and 8 GiB of FP64 basis values; a
  // point split of the potential contraction is a whole number m of
  // 2048-point blocks, the fewest that keep the splits of a chunk at 64 or
  // less, and the chunk is rounded up to whole splits (one more chunk when
  // that would pass the limit)
  long mc_max = std::max<long>(kSplitPts, ((8L << 30) / (32L * c->NP)) / kSplitPts * kSplitPts);
  mc_max = std::min<long>(mc_max, 1L << 18);
  if (opt && opt->chunk_points) mc_max = opt->chunk_points;
  for (c->nchunks = (int)((npts + mc_max - 1) / mc_max);; ++c->nchunks) {
    const long q = round_up((npts + c->nchunks - 1) / c->nchunks, kSplitPts) / kSplitPts;
    const long m = (q + 63) / 64;
    c->nsplit = (int)((q + m - 1) / m);
    c->Mc = (long)c->nsplit * m * kSplitPts;
    if (c->Mc <= mc_max) break;
  }
  c->npad = c->Mc * c->nchunks;
  c->nblk_total = c->npad / kThreads;

  HIPCHK(c, hipGetDevice(&c->device));
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->nlanes = c->nlanes_cap = c->nchunks > 1 ? 2 : 1;
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
  HIPCHK(c, hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
  HIPCHK(c, hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
  // atom ranges of the basis stage: consecutive atoms in shell order, as
  // many as the kernel's LDS image holds
  std::vector<int> ratoms;
  for (int a = 0; a < n_atoms; ++a)
    if (at[a].shell_end > at[a].shell_begin) ratoms.push_back(a);
  std::sort(ratoms.begin(), ratoms.end(),
            [&](int p, int q) { return at[p].shell_begin < at[q].shell_begin; });
  c->h_ranges.clear();
  for (size_t i = 0; i < ratoms.size(); ++i) {
    const AtomD &a = at[ratoms[i]];
    const int ns = a.shell_end - a.shell_begin;
    if (c->h_ranges.empty() || c->h_ranges.back().n_atoms == kRangeAtoms ||
        c->h_ranges.back().n_shells + ns > kRangeShells ||
        c->h_ranges.back().n_exp + a.n_exp > kRangePrims ||
        c->h_ranges.back().n_contr + a.n_contr > kRangeContr) {
      RangeD r{};
      r.atom_begin = (int)i;
      r.shell_begin = a.shell_begin;
      r.exp_begin = a.exp_begin;
      r.contr_begin = a.contr_begin;
      c->h_ranges.push_back(r);
    }
    RangeD &r = c->h_ranges.back();
    ++r.n_atoms;
    r.n_shells += ns;
    r.n_exp += a.n_exp;
    r.n_contr += a.n_contr;
  }
  c->n_ranges = (int)c->h_ranges.size();
  int rc;
  if ((rc = dalloc(c, &c->grid, 4 * (size_t)c->npad))) return rc;
  if ((rc = dalloc(c, &c->atoms, n_atoms))) return rc;
  if ((rc = dalloc(c, &c->shells, n_shells))) return rc;
  if ((rc = dalloc(c, &c->ranges, std::max<size_t>(c->h_ranges.size(), 1)))) return rc;
  if ((rc = dalloc(c, &c->ratoms, std::max<size_t>(ratoms.size(), 1)))) return rc;
  if ((rc = dalloc(c, &c->exps, eoff))) return rc;
  if ((rc = dalloc(c, &c->norms, eoff))) return rc;
  if ((rc = dalloc(c, &c->contr, std::max<long>(coff, 1)))) return rc;
  // grid in space-filling-curve order, padded with far points of zero weight
  std::vector<long> perm;
  morton_order(grid, npts, npts, 1, perm);
  c->perm = perm;
  c->axyz.assign(atom_xyz, atom_xyz + 3 * (size_t)n_atoms);
  std::vector<double> g(4 * (size_t)c->npad);
  for (int k = 0; k < 4; ++k)
    for (long p = 0; p < c->npad; ++p)
      g[(size_t)k * c->npad + p] =
          p < npts ? grid[(size_t)k * npts + perm[p]] : (k < 3 ? kFar : 0.0);
  HIPCHK(c, hipMemcpy(c->grid, g.data(), g.size() * 8, hipMemcpyHostToDevice));
  {
    uint64_t h = 1469598103934665603ull;
    h = hash_words(h, atom_xyz, 3 * (size_t)n_atoms * 8);
    h = hash_words(h, shell_info, 5 * (size_t)n_shells * 4);
    h = hash_words(h, exps, (size_t)eoff * 8);
    h = hash_words(h, norms, (size_t)eoff * 8);
    h = hash_words(h, contr, (size_t)coff * 8);
    h = hash_words(h, grid, 4 * (size_t)npts * 8);
    const int64_t lay[2] = {c->NP, c->npad};
    c->geom = hash_words(h, lay, sizeof(lay));
  }
  HIPCHK(c, hipMemcpy(c->atoms, at.data(), sizeof(AtomD) * n_atoms, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->shells, sh.data(), sizeof(ShellD) * n_shells, hipMemcpyHostToDevice));
  if (!ratoms.empty()) {
    HIPCHK(c, hipMemcpy(c->ranges, c->h_ranges.data(), sizeof(RangeD) * c->h_ranges.size(),
                        hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->ratoms, ratoms.data(), sizeof(int) * ratoms.size(), hipMemcpyHostToDevice));
  }
  if ((rc = dalloc(c, &c->d_fpos, c->nbas))) return rc;
  HIPCHK(c, hipMemcpy(c->d_fpos, c->fpos.data(), sizeof(int) * c->nbas, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->exps, exps, 8 * eoff, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->norms, norms, 8 * eoff, hipMemcpyHostToDevice));
  if (coff > 0) HIPCHK(c, hipMemcpy(c->contr, contr, 8 * coff, hipMemcpyHostToDevice));
  if ((rc = setup_screen(c, sh, exps, !(opt && opt->no_screen)))) return rc;
  const size_t np2 = (size_t)c->NP * c->NP;
  if ((rc = dalloc(c, &c->P, np2))) return rc;
  if ((rc = dalloc(c, &c->Pref, np2))) return rc;
  HIPCHK(c, hipMalloc(&c->D, np2 * 8));
  HIPCHK(c, hipMemset(c->P, 0, np2 * 8));
  HIPCHK(c, hipMemset(c->Pref, 0, np2 * 8));
  if ((rc = dalloc(c, &c->state, (size_t)S_NSLOT * c->npad))) return rc;
  HIPCHK(c, hipMemset(c->state, 0, (size_t)S_NSLOT * c->npad * 8));
  const size_t npk = (size_t)c->nbas * (c->nbas + 1) / 2;
  if ((rc = dalloc(c, &c->Vstate, npk))) return rc;
  if ((rc = dalloc(c, &c->Vout, npk))) return rc;
  if ((rc = dalloc(c, &c->epart, 2 * (size_t)c->nblk_total))) return rc;
  if ((rc = dalloc(c, &c->eterm, 5 * (size_t)c->nblk_total))) return rc;
  const size_t L = (size_t)c->nlanes_cap;
  if ((rc = dalloc(c, &c->part, L * c->ntiles * 4 * c->Mc))) return rc;
  if ((rc = dalloc(c, &c->coef, L * 4 * c->Mc))) return rc;
  if ((rc = dalloc(c, &c->Hpart, L * c->nsplit * np2))) return rc;
  HIPCHK(c, hipMemset(c->part, 0, L * c->ntiles * 4 * c->Mc * 8));
  HIPCHK(c, hipMemset(c->coef, 0, L * 4 * c->Mc * 8));
  return PGXC_OK;
}

int pgxc_create(pgxc_ctx **out, int32_t n_atoms, const double *atom_xyz, int32_t n_shells,
                const int32_t *shell_info, const double *exps, const double *norms,
                const double *contr, const double *grid, int64_t npts, int32_t veclen) {
  return pgxc_create_opt(out, n_atoms, atom_xyz, n_shells, shell_info, exps, norms, contr, grid,
                         npts, veclen, nullptr);
}

int pgxc_layout(const pgxc_ctx *c, int64_t *out8) {
  if (!c || !out8) return PGXC_ERR_ARG;
  out8[0] = c->NP;
  out8[1] = c->tf;
  out8[2] = c->Mc;
  out8[3] = c->nchunks;
  out8[4] = c->nsplit;
  out8[5] = c->npad;
  out8[6] = c->var[0];
  out8[7] = c->var[1];
  return PGXC_OK;
}

}  // extern "C"

// Buffers of the beta spin, zeroed like the alpha spin's at creation.
static int alloc_beta(pgxc_ctx *c) {
  auto &B = c->beta;
  if (B.P) return PGXC_OK;
  const size_t np2 = (size_t)c->NP * c->NP, npk = (size_t)c->nbas * (c->nbas + 1) / 2;
  const size_t L = (size_t)c->nlanes_cap;
  int rc;
  if ((rc = dalloc(c, &B.P, np2)) || (rc = dalloc(c, &B.Pref, np2))) return rc;
  HIPCHK(c, hipMalloc(&B.D, np2 * 8));
  if ((rc = dalloc(c, &B.state, (size_t)S_NSLOT * c->npad)) || (rc = dalloc(c, &B.Vstate, npk)) ||
      (rc = dalloc(c, &B.Vout, npk)) || (rc = dalloc(c, &B.part, L * c->ntiles * 4 * c->Mc)) ||
      (rc = dalloc(c, &B.coef, L * 4 * c->Mc)) || (rc = dalloc(c, &B.Hpart, L * c->nsplit * np2)))
    return rc;
  HIPCHK(c, hipMemset(B.P, 0, np2 * 8));
  HIPCHK(c, hipMemset(B.Pref, 0, np2 * 8));
  HIPCHK(c, hipMemset(B.state, 0, (size_t)S_NSLOT * c->npad * 8));
  HIPCHK(c, hipMemset(B.part, 0, L * c->ntiles * 4 * c->Mc * 8));
  HIPCHK(c, hipMemset(B.coef, 0, L * 4 * c->Mc * 8));
  return PGXC_OK;
}

// Switch the context to ispin spins: a build state of the other spin count
// is dropped, and the plan (whose units depend on it) rebuilt.
static int set_spins(pgxc_ctx *c, int ispin) {
  if (ispin == c->ispin) return PGXC_OK;
  if (ispin == 2) {
    int rc = alloc_beta(c);
    if (rc) return rc;
  }
  c->ispin = ispin;
  c->have_state = 0;
  c->have_density = 0;
  free_plan(c->plan);
  free_plan(c->plan_e);
  return PGXC_OK;
}

static int check_density_args(pgxc_ctx *c, int32_t ispin, int32_t nbas, int32_t mode) {
  if (!c) return PGXC_ERR_ARG;
  if (ispin != 1 && ispin != 2)
    return fail(c, PGXC_ERR_ARG, "ispin = " + std::to_string(ispin) + ": must be 1 or 2");
  if (nbas != c->nbas)
    return fail(c, PGXC_ERR_ARG, "nbas = " + std::to_string(nbas) + " but the basis has " +
                                     std::to_string(c->nbas) + " functions");
  if (mode != PGXC_DENSITY_FULL && mode != PGXC_DENSITY_INCREMENTAL)
    return fail(c, PGXC_ERR_ARG, "unknown density mode " + std::to_string(mode));
  if (mode == PGXC_DENSITY_INCREMENTAL && c->have_state && ispin != c->ispin)
    return fail(c, PGXC_ERR_STATE, "incremental density with ispin = " + std::to_string(ispin) +
                                       " after a build with ispin = " + std::to_string(c->ispin));
  return set_spins(c, ispin);
}

extern "C" int pgxc_set_density(pgxc_ctx *c, int32_t ispin, int32_t nbas, const double *P,
                                int32_t mode) {
  int rc = check_density_args(c, ispin, nbas, mode);
  if (rc) return rc;
  if (!P) return fail(c, PGXC_ERR_ARG, "pgxc_set_density: P is null");
  const int n = c->nbas, NP = c->NP;
  std::vector<double> h((size_t)NP * NP, 0.0);
  const int *pos = c->fpos.data();
  for (int sp = 0; sp < ispin; ++sp) {
    const double *Ps = P + (size_t)sp * n * n;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) h[(size_t)pos[i] * NP + pos[j]] = Ps[(size_t)i + (size_t)j * n];
    HIPCHK(c, hipMemcpy(sp ? c->beta.P : c->P, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  }
  c->have_density = 1;
  c->density_mode = mode;
  return PGXC_OK;
}

// Temporary device buffers of a call, freed when the call returns on any path.

extern "C" int pgxc_set_orbitals(pgxc_ctx *c, int32_t ispin, int32_t nbas, int32_t nocc,
                                 const double *C, const double *occ, int32_t mode) {
  int rc = check_density_args(c, ispin, nbas, mode);
  if (rc) return rc;
  if (nocc < 0 || (nocc > 0 && (!C || !occ)))
    return fail(c, PGXC_ERR_ARG, "pgxc_set_orbitals: bad nocc or null C/occ");
  const int n = c->nbas;
  double *dC = nullptr, *docc = nullptr;
  ScopedDev tmp;  // dC and docc, freed on every return below
  for (int sp = 0; sp < ispin; ++sp)
    HIPCHK(c, hipMemsetAsync(sp ? c->beta.P : c->P, 0, (size_t)c->NP * c->NP * 8, c->stream));
  if (nocc > 0) {
    HIPCHK(c, hipMalloc((void **)&dC, (size_t)ispin * n * nocc * 8));
    tmp.keep(dC);
    HIPCHK(c, hipMalloc((void **)&docc, (size_t)ispin * nocc * 8));
    tmp.keep(docc);
    std::vector<double> Cp((size_t)ispin * n * nocc);  // rows in the internal function order
    for (int sp = 0; sp < ispin; ++sp)
      for (int k = 0; k < nocc; ++k)
        for (int i = 0; i < n; ++i)
          Cp[(size_t)c->fpos[i] + ((size_t)sp * nocc + k) * n] = C[(size_t)i + ((size_t)sp * nocc + k) * n];
    HIPCHK(c, hipMemcpy(dC, Cp.data(), Cp.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(docc, occ, (size_t)ispin * nocc * 8, hipMemcpyHostToDevice));
    for (int sp = 0; sp < ispin; ++sp)
      launch_form_density(c->stream, dC + (size_t)sp * n * nocc, docc + (size_t)sp * nocc, n, nocc,
                          c->NP, sp ? c->beta.P : c->P);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_density = 1;
  c->density_mode = mode;
  return PGXC_OK;
}

extern "C" {

int pgxc_build(pgxc_ctx *c, uint32_t functional, int32_t precision, double *ham, double *exc,
               double *charge) {
  int incr = 0;
  int rc = validate_build(c, functional, precision, incr);
  if (rc) return rc;
  const bool gga = (functional & (PGXC_XC_PBEX | PGXC_XC_PBEC)) != 0;
  hipStream_t s = c->stream;
  const size_t elem = precision == PGXC_PREC_FP64 ? 8 : 4;
  if ((rc = ensure_work(c, elem, precision, s, gga))) return rc;
  std::vector<pgxc_ctx *> mols{c};
  std::vector<int> incrv{incr};
  if ((rc = get_plan(mols, functional, precision, incrv, {}, -1, c->sched, c->plan, c))) return rc;
  const size_t nev = kEvStep + 5 * c->plan->steps.size();
  while (c->ev.size() < nev) {
    hipEvent_t e = nullptr;
    HIPCHK(c, hipEventCreate(&e));
    c->ev.push_back(e);
  }
  HIPCHK(c, hipEventRecord(c->ev[kEvStart], s));
  if (elem == 8)
    launch_plan<double>(s, c->stream2, c->fork, c->join, *c->plan, gga, &c->ev);
  else
    launch_plan<float>(s, c->stream2, c->fork, c->join, *c->plan, gga, &c->ev);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(s));
  for (int k = 0; k < 4; ++k) c->timing[k] = 0.0;
  float t = 0;
  for (size_t k = 0; k < c->plan->steps.size(); ++k) {
    const hipEvent_t *E = c->ev.data() + kEvStep + 5 * k;
    for (int j = 0; j < 4; ++j) {
      HIPCHK(c, hipEventElapsedTime(&t, E[j], E[j + 1]));
      c->timing[j] += t;
    }
  }
  HIPCHK(c, hipEventElapsedTime(&t, c->ev[kEvPack], c->ev[kEvEnd]));
  c->timing[4] = t;
  HIPCHK(c, hipEventElapsedTime(&t, c->ev[kEvStart], c->ev[kEvEnd]));
  c->timing[5] = t;
  mark_basis_cached(c, elem, precision, gga);
  if ((rc = collect(c, s, functional, precision))) return rc;
  return pgxc_get_result(c, ham, exc, charge);
}

int pgxc_get_result(pgxc_ctx *c, double *ham, double *exc, double *charge) {
  if (!c) return PGXC_ERR_ARG;
  if (!c->have_state) return fail(c, PGXC_ERR_STATE, "no build has run on this context");
  if (ham) {  // ispin = 2: alpha / beta at even / odd positions
    const size_t npk = (size_t)c->nbas * (c->nbas + 1) / 2, is = (size_t)c->ispin;
    std::vector<double> v(npk);
    for (size_t sp = 0; sp < is; ++sp) {
      HIPCHK(c, hipMemcpy(v.data(), sp ? c->beta.Vout : c->Vout, npk * 8, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < npk; ++i) ham[i * is + sp] += v[i];
    }
  }
  if (exc) *exc += c->last_exc;
  if (charge) *charge += c->last_charge;
  return PGXC_OK;
}

int pgxc_energy(pgxc_ctx *c, uint32_t functional, int32_t precision, double *terms4,
                double *charge) {
  int incr = 0;
  int rc = validate_build(c, functional, precision, incr);
  if (rc == PGXC_ERR_STATE && c->have_density) rc = PGXC_OK;  // the full density is always there
  if (rc) return rc;
  const bool gga = (functional & (PGXC_XC_PBEX | PGXC_XC_PBEC)) != 0;
  hipStream_t s = c->stream;
  const size_t elem = precision == PGXC_PREC_FP64 ? 8 : 4;
  if ((rc = ensure_work(c, elem, precision, s, gga))) return rc;
  std::vector<pgxc_ctx *> mols{c};
  std::vector<int> full{0};
  if ((rc = get_plan(mols, functional, precision, full, {}, -1, c->sched, c->plan_e, c, 1))) return rc;
  if (elem == 8)
    launch_plan<double>(s, c->stream2, c->fork, c->join, *c->plan_e, gga, nullptr);
  else
    launch_plan<float>(s, c->stream2, c->fork, c->join, *c->plan_e, gga, nullptr);
  HIPCHK(c, hipGetLastError());
  std::vector<double> et(5 * (size_t)c->nblk_total);
  HIPCHK(c, hipMemcpyAsync(et.data(), c->eterm, et.size() * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  double t[5] = {0, 0, 0, 0, 0};
  for (long b = 0; b < c->nblk_total; ++b)
    for (int k = 0; k < 5; ++k) t[k] += et[5 * b + k];
  memcpy(c->last_terms, t, sizeof(t));
  mark_basis_cached(c, elem, precision, gga);
  if (terms4) memcpy(terms4, t, 4 * sizeof(double));
  if (charge) *charge = t[4];
  return PGXC_OK;
}

// The basis-centre term into grad; with pm (device [3][npad], zeroed) also
// w grad f at every point, summed over spins.
static int gradient_impl(pgxc_ctx *c, double *grad, double *pm) {
  if (!c || !grad) return PGXC_ERR_ARG;
  if (!c->have_state) return fail(c, PGXC_ERR_STATE, "pgxc_gradient before a build on this context");
  const uint32_t functional = c->state_functional;
  const bool gga = (functional & (PGXC_XC_PBEX | PGXC_XC_PBEC)) != 0;
  hipStream_t s = c->stream;
  int rc;
  if ((rc = ensure_work(c, 8, PGXC_PREC_FP64, s, true))) return rc;
  // an FP64 plan for its basis stage: chunk k's values land where its unit
  // points (the FP64 cache, or the lane's work buffer when there is none)
  Plan *gp = nullptr;
  std::vector<pgxc_ctx *> mols{c};
  std::vector<int> full{0};
  // every block of the values: the products below read all functions on all points
  if ((rc = get_plan(mols, functional, PGXC_PREC_FP64, full, {}, -1, c->sched, gp, c, 0, 1)))
    return rc;
  const long Mc = c->Mc, NP = c->NP, cs = NP * Mc, tpc = Mc / kScreenTile;
  auto done = [&](int code) {
    free_plan(gp);
    return code;
  };
  // the chunk's Hessian planes, products and per-(tile, function) sums stay
  // with the context for later gradients
  long nlist = 1;  // the most list entries of one chunk
  for (int k = 0; k < c->nchunks; ++k)
    nlist = std::max<long>(nlist, c->h_tf_off[(size_t)(k + 1) * tpc] - c->h_tf_off[(size_t)k * tpc]);
  if (!c->ghess && hipMalloc((void **)&c->ghess, 6 * (size_t)cs * 8) != hipSuccess) c->ghess = nullptr;
  if (!c->gXY && hipMalloc((void **)&c->gXY, 4 * (size_t)cs * 8) != hipSuccess) c->gXY = nullptr;
  if (c->gpart_n < 3 * nlist) {
    dfree(c->gpart);
    c->gpart_n = 0;
    if (hipMalloc((void **)&c->gpart, 3 * (size_t)nlist * 8) == hipSuccess) c->gpart_n = 3 * nlist;
    else c->gpart = nullptr;
  }
  if (!c->ghess || !c->gXY || !c->gpart) {
    (void)hipGetLastError();
    return done(fail(c, PGXC_ERR_NOMEM, "pgxc_gradient: no device memory for the chunk's Hessian and products"));
  }
  double *hess = c->ghess, *XY = c->gXY, *part = c->gpart;
  std::vector<double> gfn(3 * (size_t)NP, 0.0), hp(3 * (size_t)nlist);
  for (int k = 0; k < c->nchunks; ++k) {
    const long p0 = (long)k * Mc;
    const int lane = k % c->nlanes;
    const double *phi = (const double *)((const char *)c->phi_base +
                                         (c->phi_stride ? (size_t)k * c->phi_stride : (size_t)lane * c->phi_lane));
    launch_basis<double>(s, true, gp->steps[k].basis);
    launch_basis_hess(s, c->grid + p0, c->grid + c->npad + p0, c->grid + 2 * c->npad + p0, c->atoms,
                      c->n_atoms, c->shells, c->exps, c->norms, c->contr, (int)NP, Mc, hess);
    const long l0 = c->h_tf_off[(size_t)k * tpc], nl = c->h_tf_off[(size_t)(k + 1) * tpc] - l0;
    for (int sp = 0; sp < c->ispin; ++sp) {
      // X = Phi P, Y_c = (d/dc Phi) P with P of the last build (P_ref),
      // over the density items of the chunk, for each tile's functions
      const double *P = sp ? c->beta.Pref : c->Pref;
      for (int li = 0; li < kLists; ++li)
        launch_xy(s, kTiles[li / kNumVariants], li % kNumVariants, gp->steps[k].dens[li], P, XY,
                  gga ? 4 : 1, c->ispin, sp);
      const double *st = (sp ? c->beta.state : c->state) + p0;
      const long np = c->npad;
      launch_grad_fn(s, gga, tpc, c->tf_off + p0 / kScreenTile, c->tfl, (int)NP, Mc, XY, phi, hess,
                     st + S_A * np, st + S_BX * np, st + S_BY * np, st + S_BZ * np, part);
      if (pm)
        launch_grad_pt(s, gga, c->tf_off + p0 / kScreenTile, c->tfl, (int)NP, Mc, XY, phi, hess,
                       st + S_A * np, st + S_BX * np, st + S_BY * np, st + S_BZ * np, pm + p0, np);
      if (nl > 0 && (hipMemcpyAsync(hp.data(), part, (size_t)nl * 3 * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
                     hipStreamSynchronize(s) != hipSuccess))
        return done(fail(c, PGXC_ERR_HIP, "pgxc_gradient: kernel or copy failed"));
      for (long e = 0; e < nl; ++e)  // tiles in order, each tile's functions in order
        for (int d = 0; d < 3; ++d) gfn[3 * (size_t)c->h_tfl[(size_t)(l0 + e)] + d] += hp[(size_t)e * 3 + d];
    }
  }
  if (hipGetLastError() != hipSuccess) return done(fail(c, PGXC_ERR_HIP, "pgxc_gradient: launch failed"));
  mark_basis_cached(c, 8, PGXC_PREC_FP64, true);
  for (int a = 0; a < 3 * c->n_atoms; ++a) grad[a] = 0.0;
  for (int i = 0; i < c->nbas; ++i)
    for (int d = 0; d < 3; ++d) grad[3 * c->fatom[i] + d] += -2.0 * gfn[3 * (size_t)i + d];
  return done(PGXC_OK);
}

int pgxc_gradient(pgxc_ctx *c, double *grad) { return gradient_impl(c, grad, nullptr); }

int pgxc_gradient_grid(pgxc_ctx *c, const int32_t *owner, double *grad) {
  if (!c || !owner || !grad) return PGXC_ERR_ARG;
  const long npts = (long)c->perm.size(), np = c->npad, Mc = c->Mc;
  const int na = c->n_atoms;
  std::vector<int> own((size_t)np, -1);
  for (long p = 0; p < npts; ++p) {
    const int a = owner[c->perm[p]];
    if (a < 0 || a >= na) return fail(c, PGXC_ERR_ARG, "pgxc_gradient_grid: owner out of range");
    own[p] = a;
  }
  std::vector<double> inv((size_t)na * na, 0.0);
  for (int i = 0; i < na; ++i)
    for (int k = 0; k < na; ++k)
      if (i != k) {
        const double dx = c->axyz[3 * i] - c->axyz[3 * k], dy = c->axyz[3 * i + 1] - c->axyz[3 * k + 1],
                     dz = c->axyz[3 * i + 2] - c->axyz[3 * k + 2];
        const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (!(d > 0.0)) return fail(c, PGXC_ERR_ARG, "pgxc_gradient_grid: two atoms at one position");
        inv[(size_t)i * na + k] = 1.0 / d;
      }
  double *pm = nullptr, *d_inv = nullptr, *fp = nullptr, *q = nullptr, *rm = nullptr, *part = nullptr;
  int *d_own = nullptr;
  const long nblk = Mc / kThreads;
  auto done = [&](int code) {
    dfree(pm);
    dfree(d_inv);
    dfree(fp);
    dfree(q);
    dfree(rm);
    dfree(part);
    dfree(d_own);
    return code;
  };
  if (hipMalloc((void **)&pm, 3 * (size_t)np * 8) != hipSuccess ||
      hipMalloc((void **)&d_inv, inv.size() * 8) != hipSuccess ||
      hipMalloc((void **)&fp, (size_t)Mc * 8) != hipSuccess ||
      hipMalloc((void **)&q, (size_t)na * Mc * 8) != hipSuccess ||
      hipMalloc((void **)&rm, 3 * (size_t)Mc * 8) != hipSuccess ||
      hipMalloc((void **)&part, (size_t)nblk * na * 3 * 8) != hipSuccess ||
      hipMalloc((void **)&d_own, (size_t)np * 4) != hipSuccess)
    return done(fail(c, PGXC_ERR_NOMEM, "pgxc_gradient_grid: no device memory"));
  hipStream_t s = c->stream;
  if (hipMemsetAsync(pm, 0, 3 * (size_t)np * 8, s) != hipSuccess ||
      hipMemcpy(d_inv, inv.data(), inv.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_own, own.data(), own.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    return done(fail(c, PGXC_ERR_HIP, "pgxc_gradient_grid: upload"));
  // basis-centre term, and w grad f per point (the points move with their atoms)
  int rc = gradient_impl(c, grad, pm);
  if (rc) return done(rc);
  std::vector<double> hpm(3 * (size_t)np);
  if (hipMemcpy(hpm.data(), pm, hpm.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return done(fail(c, PGXC_ERR_HIP, "pgxc_gradient_grid: copy"));
  for (long p = 0; p < np; ++p)
    if (own[p] >= 0)
      for (int d = 0; d < 3; ++d) grad[3 * own[p] + d] += hpm[(size_t)d * np + p];
  // Becke weight derivatives times f
  const uint32_t functional = c->state_functional;
  const bool gga = (functional & (PGXC_XC_PBEX | PGXC_XC_PBEC)) != 0;
  std::vector<double> hp((size_t)nblk * na * 3);
  for (int k = 0; k < c->nchunks; ++k) {
    const long p0 = (long)k * Mc;
    const double *st = c->state + p0, *st2 = c->ispin == 2 ? c->beta.state + p0 : st;
    launch_point_f(s, gga, c->ispin == 2, functional, Mc, np, st, st2, fp);
    launch_becke_grad(s, na, c->atoms, d_inv, Mc, c->grid + p0, c->grid + np + p0,
                      c->grid + 2 * np + p0, c->grid + 3 * np + p0, fp, d_own + p0, q, rm, part);
    if (hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return done(fail(c, PGXC_ERR_HIP, "pgxc_gradient_grid: weight derivatives failed"));
    for (long b = 0; b < nblk; ++b)
      for (int a = 0; a < na; ++a)
        for (int d = 0; d < 3; ++d) grad[3 * a + d] += hp[((size_t)b * na + a) * 3 + d];
  }
  return done(PGXC_OK);
}

int pgxc_last_stats(const pgxc_ctx *c, double *s5) {
  if (!c || !s5) return PGXC_ERR_ARG;
  for (int i = 0; i < 5; ++i) s5[i] = c->stats[i];
  return PGXC_OK;
}

int pgxc_last_work(const pgxc_ctx *c, double *w3) {
  if (!c || !w3) return PGXC_ERR_ARG;
  w3[0] = c->sum_bn;
  w3[1] = c->sum_n2;
  w3[2] = std::ceil((double)c->npts / kScreenTile);
  return PGXC_OK;
}

int pgxc_last_timing(const pgxc_ctx *c, double *ms6) {
  if (!c || !ms6) return PGXC_ERR_ARG;
  for (int i = 0; i < 6; ++i) ms6[i] = c->timing[i];
  return PGXC_OK;
}

int pgxc_set_schedule(pgxc_ctx *c, int32_t mode, pgxc_costmodel *model) {
  if (!c) return PGXC_ERR_ARG;
  if (check_sched(mode, model))
    return fail(c, PGXC_ERR_ARG, "pgxc_set_schedule: mode 0..2, and a model for PGXC_SCHED_MODEL");
  c->sched.mode = mode;
  c->sched.model = mode == PGXC_SCHED_MODEL ? model : nullptr;
  return PGXC_OK;
}

// ---- state of incremental builds, for checkpoint / restart ----------------
struct StateHeader {
  char magic[8];
  uint64_t geom;
  int64_t nbas, NP, npad, npk, nslot;
  uint32_t functional;
  int32_t ispin;
  double exc, charge;
};
static const char kStateMagic[8] = {'P', 'G', 'X', 'C', 'S', 'T', '0', '1'};

// one block per spin: Pref, Vstate, state
static int64_t state_block(const pgxc_ctx *c) {
  const int64_t npk = (int64_t)c->nbas * (c->nbas + 1) / 2;
  return 8 * ((int64_t)c->NP * c->NP + npk + (int64_t)S_NSLOT * c->npad);
}

static int64_t state_bytes(const pgxc_ctx *c) {
  return (int64_t)sizeof(StateHeader) + c->ispin * state_block(c);
}

int64_t pgxc_state_bytes(const pgxc_ctx *c) { return c ? state_bytes(c) : -PGXC_ERR_ARG; }

int pgxc_save_state(pgxc_ctx *c, void *buf, int64_t bytes) {
  if (!c || !buf) return PGXC_ERR_ARG;
  if (!c->have_state) return fail(c, PGXC_ERR_STATE, "pgxc_save_state: no build has run on this context");
  if (bytes < state_bytes(c)) return fail(c, PGXC_ERR_ARG, "pgxc_save_state: buffer smaller than pgxc_state_bytes");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  StateHeader h{};
  memcpy(h.magic, kStateMagic, 8);
  h.geom = c->geom;
  h.nbas = c->nbas;
  h.NP = c->NP;
  h.npad = c->npad;
  h.npk = (int64_t)c->nbas * (c->nbas + 1) / 2;
  h.nslot = S_NSLOT;
  h.functional = c->state_functional;
  h.ispin = c->ispin;
  h.exc = c->last_exc;
  h.charge = c->last_charge;
  char *o = (char *)buf;
  memcpy(o, &h, sizeof(h));
  o += sizeof(h);
  for (int sp = 0; sp < c->ispin; ++sp) {
    HIPCHK(c, hipMemcpy(o, sp ? c->beta.Pref : c->Pref, 8 * (size_t)c->NP * c->NP, hipMemcpyDeviceToHost));
    o += 8 * (size_t)c->NP * c->NP;
    HIPCHK(c, hipMemcpy(o, sp ? c->beta.Vstate : c->Vstate, 8 * (size_t)h.npk, hipMemcpyDeviceToHost));
    o += 8 * (size_t)h.npk;
    HIPCHK(c, hipMemcpy(o, sp ? c->beta.state : c->state, 8 * (size_t)S_NSLOT * c->npad,
                        hipMemcpyDeviceToHost));
    o += 8 * (size_t)S_NSLOT * c->npad;
  }
  return PGXC_OK;
}

int pgxc_load_state(pgxc_ctx *c, const void *buf, int64_t bytes) {
  if (!c || !buf) return PGXC_ERR_ARG;
  if (bytes < (int64_t)sizeof(StateHeader)) return fail(c, PGXC_ERR_ARG, "pgxc_load_state: buffer too short");
  StateHeader h;
  memcpy(&h, buf, sizeof(h));
  if (memcmp(h.magic, kStateMagic, 8) != 0) return fail(c, PGXC_ERR_ARG, "pgxc_load_state: not a state record");
  if (h.geom != c->geom || h.nbas != c->nbas || h.NP != c->NP || h.npad != c->npad ||
      h.nslot != S_NSLOT || h.npk != (int64_t)c->nbas * (c->nbas + 1) / 2 ||
      (h.ispin != 1 && h.ispin != 2))
    return fail(c, PGXC_ERR_ARG, "pgxc_load_state: the state belongs to another geometry or layout");
  if (bytes < (int64_t)sizeof(StateHeader) + h.ispin * state_block(c))
    return fail(c, PGXC_ERR_ARG, "pgxc_load_state: buffer too short");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  {  // the next density must be set with the state's spin count
    int rc = set_spins(c, h.ispin);
    if (rc) return rc;
  }
  const char *o = (const char *)buf + sizeof(h);
  for (int sp = 0; sp < c->ispin; ++sp) {
    HIPCHK(c, hipMemcpy(sp ? c->beta.Pref : c->Pref, o, 8 * (size_t)c->NP * c->NP, hipMemcpyHostToDevice));
    o += 8 * (size_t)c->NP * c->NP;
    HIPCHK(c, hipMemcpy(sp ? c->beta.Vstate : c->Vstate, o, 8 * (size_t)h.npk, hipMemcpyHostToDevice));
    o += 8 * (size_t)h.npk;
    HIPCHK(c, hipMemcpy(sp ? c->beta.state : c->state, o, 8 * (size_t)S_NSLOT * c->npad,
                        hipMemcpyHostToDevice));
    o += 8 * (size_t)S_NSLOT * c->npad;
  }
  c->state_functional = h.functional;
  c->last_exc = h.exc;
  c->last_charge = h.charge;
  c->have_state = 1;
  return PGXC_OK;
}

int pgxc_set_lanes(pgxc_ctx *c, int32_t lanes) {
  if (!c) return PGXC_ERR_ARG;
  if (lanes < 1 || lanes > 2) return fail(c, PGXC_ERR_ARG, "pgxc_set_lanes: lanes must be 1 or 2");
  c->nlanes = std::min<int>(lanes, c->nlanes_cap);
  return PGXC_OK;
}

int pgxc_set_item_log(pgxc_ctx *c, int32_t on) {
  if (!c) return PGXC_ERR_ARG;
  c->sched.log = on ? 1 : 0;
  return PGXC_OK;
}

int64_t pgxc_item_log(pgxc_ctx *c, double *features, double *us, int64_t cap) {
  if (!c) return -PGXC_ERR_ARG;
  return read_item_log(c->plan, c->stream, features, us, cap);
}

int pgxc_destroy(pgxc_ctx *c) {
  if (!c) return PGXC_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  free_plan(c->plan);
  free_plan(c->plan_e);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache_bytes[c->device] -= c->phic_bytes[0] + c->phic_bytes[1];
  }
  void *bufs[] = {c->grid, c->atoms, c->shells, c->ranges, c->ratoms, c->exps, c->norms, c->contr, c->P, c->Pref,
                  c->D, c->phi, c->W, c->part, c->coef, c->Hpart, c->state,
                  c->Vstate, c->Vout, c->epart, c->nbt, c->sig8, c->sshells, c->d_fpos,
                  c->phic[0], c->phic[1], c->wbits, c->tl_off, c->tl, c->tf_off, c->tfl, c->tfb, c->tlm, c->su_off, c->su, c->sumask,
                  c->gcf_off, c->gcf, c->di, c->pi, c->bi, c->bfl, c->ghess, c->gXY, c->gpart, c->beta.P, c->beta.Pref, c->beta.D,
                  c->beta.W, c->beta.part, c->beta.coef, c->beta.Hpart, c->beta.state,
                  c->beta.Vstate, c->beta.Vout, c->eterm};
  for (void *b : bufs) dfree(b);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  delete c;
  return PGXC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches
struct pgxc_batch {
  std::vector<pgxc_ctx *> mols;
  hipStream_t stream = nullptr, stream2 = nullptr;  // the two lanes of its plans
  hipEvent_t fork = nullptr, join = nullptr;
  std::string err;
  Plan *plan = nullptr;
  std::vector<double> prio;
  long order_version = -1;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0;
  Sched sched;
};

extern "C" {

const char *pgxc_batch_last_error(const pgxc_batch *b) { return b ? b->err.c_str() : "null batch"; }

int pgxc_batch_create(pgxc_batch **out, int32_t n, pgxc_ctx *const *ctxs) {
  if (!out) return PGXC_ERR_ARG;
  *out = nullptr;
  if (n <= 0 || !ctxs) return PGXC_ERR_ARG;
  pgxc_batch *b = new pgxc_batch();
  *out = b;
  int dev = -1;
  for (int i = 0; i < n; ++i) {
    if (!ctxs[i]) {
      b->err = "pgxc_batch_create: null context " + std::to_string(i);
      return PGXC_ERR_ARG;
    }
    if (dev >= 0 && ctxs[i]->device != dev) {
      b->err = "pgxc_batch_create: contexts on different devices";
      return PGXC_ERR_ARG;
    }
    dev = ctxs[i]->device;
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) {
        b->err = "pgxc_batch_create: context listed twice";
        return PGXC_ERR_ARG;
      }
    b->mols.push_back(ctxs[i]);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) {
    b->err = "pgxc_batch_create: stream/event creation failed";
    return PGXC_ERR_HIP;
  }
  return PGXC_OK;
}

int pgxc_batch_set_priority(pgxc_batch *b, const double *cost) {
  if (!b) return PGXC_ERR_ARG;
  if (cost)
    b->prio.assign(cost, cost + b->mols.size());
  else
    b->prio.clear();
  ++b->order_version;
  return PGXC_OK;
}

int pgxc_batch_build(pgxc_batch *b, uint32_t functional, int32_t precision, int32_t use_graph) {
  if (!b) return PGXC_ERR_ARG;
  const int nm = (int)b->mols.size();
  std::vector<int> incr(nm, 0);
  for (int m = 0; m < nm; ++m) {
    int rc = validate_build(b->mols[m], functional, precision, incr[m]);
    if (rc) {
      b->err = "molecule " + std::to_string(m) + ": " + b->mols[m]->err;
      return rc;
    }
  }
  const bool gga = (functional & (PGXC_XC_PBEX | PGXC_XC_PBEC)) != 0;
  const size_t elem = precision == PGXC_PREC_FP64 ? 8 : 4;
  hipStream_t s = b->stream;
  bool moved = false;
  for (int m = 0; m < nm; ++m) {
    pgxc_ctx *c = b->mols[m];
    const void *phi0 = c->phi_base, *w0 = c->W, *wb0 = c->beta.W;
    const int skip0 = c->skip_basis;
    int rc = ensure_work(c, elem, precision, s, gga);
    if (rc) {
      b->err = "molecule " + std::to_string(m) + ": " + c->err;
      return rc;
    }
    moved = moved || phi0 != c->phi_base || w0 != c->W || wb0 != c->beta.W || skip0 != c->skip_basis;
  }
  if (moved) free_plan(b->plan);
  int rc = get_plan(b->mols, functional, precision, incr, b->prio, b->order_version, b->sched,
                    b->plan, b->mols[0]);
  if (rc) {
    b->err = b->mols[0]->err;
    return rc;
  }
  Plan &pl = *b->plan;
  if (use_graph && !pl.graph) {
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      if (elem == 8)
        launch_plan<double>(s, b->stream2, b->fork, b->join, pl, gga, nullptr);
      else
        launch_plan<float>(s, b->stream2, b->fork, b->join, pl, gga, nullptr);
      e = hipStreamEndCapture(s, &pl.graph_src);
      if (e == hipSuccess) e = hipGraphInstantiate(&pl.graph, pl.graph_src, nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
      b->err = std::string("hipGraph capture: ") + hipGetErrorString(e);
      return PGXC_ERR_HIP;
    }
  }
  if (hipEventRecord(b->e0, s) != hipSuccess) return PGXC_ERR_HIP;
  if (use_graph) {
    if (hipGraphLaunch(pl.graph, s) != hipSuccess) {
      b->err = "hipGraphLaunch failed";
      return PGXC_ERR_HIP;
    }
  } else if (elem == 8) {
    launch_plan<double>(s, b->stream2, b->fork, b->join, pl, gga, nullptr);
  } else {
    launch_plan<float>(s, b->stream2, b->fork, b->join, pl, gga, nullptr);
  }
  (void)hipEventRecord(b->e1, s);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    b->err = std::string("batch build: ") + hipGetErrorString(e);
    return PGXC_ERR_HIP;
  }
  float t = 0;
  (void)hipEventElapsedTime(&t, b->e0, b->e1);
  b->last_ms = t;
  for (int m = 0; m < nm; ++m) {
    pgxc_ctx *c = b->mols[m];
    rc = collect(c, s, functional, precision);
    if (rc) {
      b->err = c->err;
      return rc;
    }
    c->timing[5] = t;
    mark_basis_cached(c, elem, precision, gga);
  }
  return PGXC_OK;
}

int pgxc_batch_set_schedule(pgxc_batch *b, int32_t mode, pgxc_costmodel *model) {
  if (!b) return PGXC_ERR_ARG;
  if (check_sched(mode, model)) {
    b->err = "pgxc_batch_set_schedule: mode 0..2, and a model for PGXC_SCHED_MODEL";
    return PGXC_ERR_ARG;
  }
  b->sched.mode = mode;
  b->sched.model = mode == PGXC_SCHED_MODEL ? model : nullptr;
  return PGXC_OK;
}

int pgxc_batch_set_item_log(pgxc_batch *b, int32_t on) {
  if (!b) return PGXC_ERR_ARG;
  b->sched.log = on ? 1 : 0;
  return PGXC_OK;
}

int64_t pgxc_batch_item_log(pgxc_batch *b, double *features, double *us, int64_t cap) {
  if (!b) return -PGXC_ERR_ARG;
  return read_item_log(b->plan, b->stream, features, us, cap);
}

int pgxc_batch_last_ms(const pgxc_batch *b, double *ms) {
  if (!b || !ms) return PGXC_ERR_ARG;
  *ms = b->last_ms;
  return PGXC_OK;
}

int pgxc_batch_destroy(pgxc_batch *b) {
  if (!b) return PGXC_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  free_plan(b->plan);
  if (b->e0) (void)hipEventDestroy(b->e0);
  if (b->e1) (void)hipEventDestroy(b->e1);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  delete b;
  return PGXC_OK;
}

int pgxc_becke_weights(int32_t n_atoms, const double *atom_xyz, int64_t npts, const double *pts,
                       const int32_t *owner, const double *w_atomic, double *w_out) {
  if (n_atoms <= 0 || npts < 0 || !atom_xyz ||
      (npts > 0 && (!pts || !owner || !w_atomic || !w_out)))
    return PGXC_ERR_ARG;
  if (npts == 0) return PGXC_OK;
  for (int64_t p = 0; p < npts; ++p)
    if (owner[p] < 0 || owner[p] >= n_atoms) return PGXC_ERR_ARG;
  for (int i = 0; i < n_atoms; ++i)
    for (int k = 0; k < i; ++k) {
      const double dx = atom_xyz[3 * i] - atom_xyz[3 * k];
      const double dy = atom_xyz[3 * i + 1] - atom_xyz[3 * k + 1];
      const double dz = atom_xyz[3 * i + 2] - atom_xyz[3 * k + 2];
      if (!(dx * dx + dy * dy + dz * dz > 0.0)) return PGXC_ERR_ARG;  // coincident centres
    }
  const size_t ninv = (size_t)n_atoms * n_atoms;
  double *d_a = nullptr, *d_inv = nullptr, *d_p = nullptr, *d_w = nullptr, *d_o = nullptr;
  int *d_own = nullptr;
  hipStream_t s = nullptr;
  int rc = PGXC_OK;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == PGXC_OK)
      rc = e == hipErrorOutOfMemory ? PGXC_ERR_NOMEM : PGXC_ERR_HIP;
  };
  chk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  chk(hipMalloc((void **)&d_a, 24 * (size_t)n_atoms));
  chk(hipMalloc((void **)&d_inv, ninv * 8));
  chk(hipMalloc((void **)&d_p, 24 * (size_t)npts));
  chk(hipMalloc((void **)&d_w, 8 * (size_t)npts));
  chk(hipMalloc((void **)&d_o, 8 * (size_t)npts));
  chk(hipMalloc((void **)&d_own, 4 * (size_t)npts));
  if (rc == PGXC_OK) {
    chk(hipMemcpyAsync(d_a, atom_xyz, 24 * (size_t)n_atoms, hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_p, pts, 24 * (size_t)npts, hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_w, w_atomic, 8 * (size_t)npts, hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_own, owner, 4 * (size_t)npts, hipMemcpyHostToDevice, s));
  }
  if (rc == PGXC_OK) {
    launch_inv_dist(s, n_atoms, d_a, d_inv);
    launch_becke(s, n_atoms, d_a, d_inv, (long)npts, d_p, d_own, d_w, d_o);
    chk(hipGetLastError());
    chk(hipMemcpyAsync(w_out, d_o, 8 * (size_t)npts, hipMemcpyDeviceToHost, s));
    chk(hipStreamSynchronize(s));
  }
  dfree(d_a);
  dfree(d_inv);
  dfree(d_p);
  dfree(d_w);
  dfree(d_o);
  dfree(d_own);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

// ---- Becke weights without allocation: file-scope scratch -----------------
// The table of inverse pair distances lives in `scratch`, one copy per
// device, held until the process ends.  Host code never passes the array's
// name to a kernel (it is a host-side stand-in there): each device's address
// is taken once with hipGetSymbolAddress and kept in g_scratch_ptr.
constexpr int kNoAllocMaxAtoms = 1024;  // 8 MiB of scratch per device
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[(size_t)kNoAllocMaxAtoms * kNoAllocMaxAtoms];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kNoAllocMaxDevices];
static int free_fail(int code, const char *msg) {
  snprintf(g_free_err, sizeof(g_free_err), "%s", msg);
  return code;
}

int pgxc_becke_weights_noalloc(int32_t n_atoms, const double *atom_xyz, int64_t npts,
                               const double *pts, const int32_t *owner, const double *w_atomic,
                               double *w_out, void *stream) {
  if (n_atoms <= 0 || npts < 0 || !atom_xyz ||
      (npts > 0 && (!pts || !owner || !w_atomic || !w_out)))
    return free_fail(PGXC_ERR_ARG, "pgxc_becke_weights_noalloc: null pointer or negative size");
  if (n_atoms > kNoAllocMaxAtoms)
    return free_fail(PGXC_ERR_ARG, "pgxc_becke_weights_noalloc: more than 1024 atoms (the scratch "
                                   "table holds 1024 x 1024 inverse distances)");
  if (npts == 0) return PGXC_OK;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
    return free_fail(PGXC_ERR_HIP, "pgxc_becke_weights_noalloc: no usable current device");
  double *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void *a = nullptr;
      if (hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)) != hipSuccess || !a)
        return free_fail(PGXC_ERR_HIP, "pgxc_becke_weights_noalloc: hipGetSymbolAddress(scratch)");
      g_scratch_ptr[dev] = (double *)a;
    }
    ptr = g_scratch_ptr[dev];
  }
  hipStream_t s = (hipStream_t)stream;
  launch_inv_dist(s, n_atoms, atom_xyz, ptr);  // every entry rewritten: nothing to clear
  launch_becke(s, n_atoms, atom_xyz, ptr, (long)npts, pts, owner, w_atomic, w_out);
  if (hipGetLastError() != hipSuccess)
    return free_fail(PGXC_ERR_HIP, "pgxc_becke_weights_noalloc: kernel launch failed");
  return PGXC_OK;
}

}  // extern "C"

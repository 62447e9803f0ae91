// C-ABI layer of the MI355X XC build (declared in include/pgxc.h).
//
// Orchestration of one build, chunk by chunk over the grid (a chunk is a
// whole number of reference batches, so the sums are those of calc_xcks_nl,
// ParaGauss modules/xc_hamiltonian.f90:1348-1575, regrouped):
//   basis -> density contraction -> functional -> W -> potential contraction
// followed by one symmetrisation into the packed lower triangle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/pgxc.h"
#include "pgxc_internal.h"

using namespace pgxc;

struct pgxc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // sizes
  int n_atoms = 0, n_shells = 0, nbas = 0, NP = 0, veclen = 128;
  long npts = 0, npad = 0, Mc = 0;
  int nchunks = 0, nsplit = 1, ntiles = 0, tf = 128;
  // device: geometry and basis
  double *grid = nullptr;  // [4][npad] x, y, z, w
  AtomD *atoms = nullptr;
  ShellD *shells = nullptr;
  double *exps = nullptr, *norms = nullptr, *contr = nullptr;
  // device: density
  double *P = nullptr, *Pref = nullptr, *Dop = nullptr;  // NP x NP
  float *Dlow = nullptr;
  // device: work buffers (sized for the largest precision used so far)
  void *phi = nullptr;
  void *W = nullptr;
  size_t phi_bytes = 0, w_bytes = 0;
  size_t phi_elem = 0;  // element size the padding rows of phi were last cleared for
  double *part = nullptr, *coef = nullptr, *Hpart = nullptr;
  // device: state and results
  double *state = nullptr;  // [S_NSLOT][npad]
  double *Vstate = nullptr, *Vout = nullptr;  // packed
  double *epart = nullptr;
  long nblk_total = 0;
  // host
  int have_density = 0, density_mode = 0, have_state = 0;
  uint32_t state_functional = 0;
  double last_exc = 0, last_charge = 0;
  double timing[6] = {0, 0, 0, 0, 0, 0};
  double stats[5] = {0, 0, 0, 0, 0};
  std::vector<hipEvent_t> ev;
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

long round_up(long a, long b) { return (a + b - 1) / b * b; }

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

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

const uint32_t kSupported = PGXC_XC_PBE;

}  // namespace

extern "C" {

const char *pgxc_last_error(const pgxc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t pgxc_nbas(const pgxc_ctx *ctx) { return ctx ? ctx->nbas : -1; }

int pgxc_create(pgxc_ctx **out, int32_t n_atoms, const double *atom_xyz, int32_t n_shells,
                const int32_t *shell_info, const double *exps, const double *norms,
                const double *contr, const double *grid, int64_t npts, int32_t veclen) {
  if (!out) return PGXC_ERR_ARG;
  *out = nullptr;
  pgxc_ctx *c = new pgxc_ctx();
  *out = c;
  if (n_atoms <= 0 || n_shells <= 0 || !atom_xyz || !shell_info || !exps || !norms ||
      (!contr && n_shells > 0) || !grid || npts <= 0 || veclen <= 0)
    return fail(c, PGXC_ERR_ARG, "pgxc_create: null pointer or non-positive size");
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
    d.exp_off = (int)eoff;
    d.contr_off = (int)coff;
    d.func_off = (int)foff;
    eoff += d.n_exp;
    coff += (long)d.n_exp * d.n_contr;
    foff += (long)(2 * d.l + 1) * (d.n_unc + d.n_contr);
  }
  for (long e = 0; e < eoff; ++e)
    if (!(exps[e] > 0.0) || !std::isfinite(norms[e]))
      return fail(c, PGXC_ERR_ARG, "pgxc_create: exponents must be positive and norms finite");
  c->nbas = (int)foff;
  // function tile: the one of 128 / 96 / 64 that pads nbas least (larger on ties)
  c->tf = 128;
  c->NP = (int)round_up(c->nbas, 128);
  for (int t : {96, 64})
    if (round_up(c->nbas, t) < c->NP) {
      c->NP = (int)round_up(c->nbas, t);
      c->tf = t;
    }
  c->ntiles = density_tiles(c->NP, c->tf);

  // ---- chunking of the grid ----------------------------------------------
  // phi chunk = 4 x NP x Mc doubles, kept under ~2 GiB
  const long cap = std::max<long>(256, ((1L << 31) / (4L * c->NP * 8)) / 256 * 256);
  const long mc_max = std::min<long>(cap, 65536);
  c->nchunks = (int)((npts + mc_max - 1) / mc_max);
  c->Mc = round_up((npts + c->nchunks - 1) / c->nchunks, 256);
  c->npad = c->Mc * c->nchunks;
  // split of the point range in the potential contraction: enough
  // workgroups for 256 CUs, each split a whole number of 16-point steps
  const int ntile_pot = (c->NP / c->tf) * (c->NP / c->tf);
  c->nsplit = 1;
  while (c->nsplit < 16 && ntile_pot * c->nsplit < 512 && (c->Mc / 32) % (c->nsplit * 2) == 0)
    c->nsplit *= 2;
  c->nblk_total = c->npad / kThreads;

  HIPCHK(c, hipGetDevice(&c->device));
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  int rc;
  if ((rc = dalloc(c, &c->grid, 4 * (size_t)c->npad))) return rc;
  if ((rc = dalloc(c, &c->atoms, n_atoms))) return rc;
  if ((rc = dalloc(c, &c->shells, n_shells))) return rc;
  if ((rc = dalloc(c, &c->exps, eoff))) return rc;
  if ((rc = dalloc(c, &c->norms, eoff))) return rc;
  if ((rc = dalloc(c, &c->contr, std::max<long>(coff, 1)))) return rc;
  // grid with padding points far away and zero weight
  std::vector<double> g(4 * (size_t)c->npad);
  for (int k = 0; k < 4; ++k)
    for (long p = 0; p < c->npad; ++p)
      g[(size_t)k * c->npad + p] = p < npts ? grid[(size_t)k * npts + p] : (k < 3 ? kFar : 0.0);
  HIPCHK(c, hipMemcpy(c->grid, g.data(), g.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->atoms, at.data(), sizeof(AtomD) * n_atoms, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->shells, sh.data(), sizeof(ShellD) * n_shells, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->exps, exps, 8 * eoff, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->norms, norms, 8 * eoff, hipMemcpyHostToDevice));
  if (coff > 0) HIPCHK(c, hipMemcpy(c->contr, contr, 8 * coff, hipMemcpyHostToDevice));
  const size_t np2 = (size_t)c->NP * c->NP;
  if ((rc = dalloc(c, &c->P, np2))) return rc;
  if ((rc = dalloc(c, &c->Pref, np2))) return rc;
  if ((rc = dalloc(c, &c->Dop, np2))) return rc;
  HIPCHK(c, hipMemset(c->P, 0, np2 * 8));
  HIPCHK(c, hipMemset(c->Pref, 0, np2 * 8));
  if ((rc = dalloc(c, &c->state, (size_t)S_NSLOT * c->npad))) return rc;
  HIPCHK(c, hipMemset(c->state, 0, (size_t)S_NSLOT * c->npad * 8));
  const size_t npk = (size_t)c->nbas * (c->nbas + 1) / 2;
  if ((rc = dalloc(c, &c->Vstate, npk))) return rc;
  if ((rc = dalloc(c, &c->Vout, npk))) return rc;
  if ((rc = dalloc(c, &c->epart, 2 * (size_t)c->nblk_total))) return rc;
  if ((rc = dalloc(c, &c->part, (size_t)c->ntiles * 4 * c->Mc))) return rc;
  if ((rc = dalloc(c, &c->coef, 4 * (size_t)c->Mc))) return rc;
  if ((rc = dalloc(c, &c->Hpart, (size_t)c->nsplit * np2))) return rc;
  HIPCHK(c, hipMemset(c->part, 0, (size_t)c->ntiles * 4 * c->Mc * 8));
  HIPCHK(c, hipMemset(c->coef, 0, 4 * (size_t)c->Mc * 8));
  c->ev.resize(12);
  for (auto &e : c->ev) HIPCHK(c, hipEventCreate(&e));
  return PGXC_OK;
}

}  // extern "C"

static int check_density_args(pgxc_ctx *c, int32_t ispin, int32_t nbas, int32_t mode) {
  if (!c) return PGXC_ERR_ARG;
  if (ispin != 1)
    return fail(c, PGXC_ERR_UNSUPPORTED, "ispin = " + std::to_string(ispin) +
                                             ": only the closed-shell build (ispin = 1) is implemented");
  if (nbas != c->nbas)
    return fail(c, PGXC_ERR_ARG, "nbas = " + std::to_string(nbas) + " but the basis has " +
                                     std::to_string(c->nbas) + " functions");
  if (mode != PGXC_DENSITY_FULL && mode != PGXC_DENSITY_INCREMENTAL)
    return fail(c, PGXC_ERR_ARG, "unknown density mode " + std::to_string(mode));
  return PGXC_OK;
}

extern "C" int pgxc_set_density(pgxc_ctx *c, int32_t ispin, int32_t nbas, const double *P, int32_t mode) {
  int rc = check_density_args(c, ispin, nbas, mode);
  if (rc) return rc;
  if (!P) return fail(c, PGXC_ERR_ARG, "pgxc_set_density: P is null");
  const int n = c->nbas, NP = c->NP;
  std::vector<double> h((size_t)NP * NP, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) h[(size_t)i * NP + j] = P[(size_t)i + (size_t)j * n];
  HIPCHK(c, hipMemcpyAsync(c->P, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_density = 1;
  c->density_mode = mode;
  return PGXC_OK;
}

extern "C" int pgxc_set_orbitals(pgxc_ctx *c, int32_t ispin, int32_t nbas, int32_t nocc, const double *C,
                      const double *occ, int32_t mode) {
  int rc = check_density_args(c, ispin, nbas, mode);
  if (rc) return rc;
  if (nocc < 0 || (nocc > 0 && (!C || !occ)))
    return fail(c, PGXC_ERR_ARG, "pgxc_set_orbitals: bad nocc or null C/occ");
  const int n = c->nbas;
  double *dC = nullptr, *docc = nullptr;
  HIPCHK(c, hipMemsetAsync(c->P, 0, (size_t)c->NP * c->NP * 8, c->stream));
  if (nocc > 0) {
    HIPCHK(c, hipMalloc((void **)&dC, (size_t)n * nocc * 8));
    HIPCHK(c, hipMalloc((void **)&docc, (size_t)nocc * 8));
    HIPCHK(c, hipMemcpyAsync(dC, C, (size_t)n * nocc * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(docc, occ, (size_t)nocc * 8, hipMemcpyHostToDevice, c->stream));
    launch_form_density(c->stream, dC, docc, n, nocc, c->NP, c->P);
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(dC);
  dfree(docc);
  c->have_density = 1;
  c->density_mode = mode;
  return PGXC_OK;
}

static int ensure_work(pgxc_ctx *c, size_t elem) {
  const size_t need_phi = 4 * (size_t)c->NP * c->Mc * elem;
  const size_t need_w = (size_t)c->NP * c->Mc * elem;
  if (c->phi_bytes < need_phi) {
    dfree(c->phi);
    c->phi = nullptr;
    c->phi_bytes = 0;
    HIPCHK(c, hipMalloc(&c->phi, need_phi));
    c->phi_bytes = need_phi;
    c->phi_elem = 0;
  }
  // function rows nbas..NP-1 are never written and must read as zero in the
  // element type of this build (a buffer last used at another precision
  // holds that precision's bytes there)
  if (c->phi_elem != elem) {
    HIPCHK(c, hipMemsetAsync(c->phi, 0, c->phi_bytes, c->stream));
    c->phi_elem = elem;
  }
  if (c->w_bytes < need_w) {
    dfree(c->W);
    c->W = nullptr;
    c->w_bytes = 0;
    HIPCHK(c, hipMalloc(&c->W, need_w));
    c->w_bytes = need_w;
  }
  return PGXC_OK;
}

template <typename T>
static int run_chunks(pgxc_ctx *c, bool gga, bool incremental, uint32_t functional,
                      const T *D, bool bf16x3) {
  hipStream_t s = c->stream;
  T *phi = (T *)c->phi, *W = (T *)c->W;
  float ms[4] = {0, 0, 0, 0};
  for (int ch = 0; ch < c->nchunks; ++ch) {
    const long p0 = (long)ch * c->Mc;
    HIPCHK(c, hipEventRecord(c->ev[0], s));
    launch_basis<T>(s, gga, c->grid, c->grid + c->npad, c->grid + 2 * c->npad, p0, c->Mc,
                    c->n_atoms, c->atoms, c->shells, c->exps, c->norms, c->contr, phi, c->NP);
    HIPCHK(c, hipEventRecord(c->ev[1], s));
    if constexpr (std::is_same<T, float>::value) {
      if (bf16x3)
        launch_density_bf16x3(s, gga, phi, D, c->Mc, c->NP, c->tf, c->part);
      else
        launch_density<T>(s, gga, phi, D, c->Mc, c->NP, c->tf, c->part);
    } else {
      launch_density<T>(s, gga, phi, D, c->Mc, c->NP, c->tf, c->part);
    }
    HIPCHK(c, hipEventRecord(c->ev[2], s));
    launch_functional(s, gga, incremental, functional, c->part, c->ntiles, c->Mc,
                      c->grid + 3 * c->npad, p0, c->state, c->npad, c->coef, c->epart,
                      p0 / kThreads);
    launch_wmat<T>(s, gga, phi, c->coef, c->Mc, c->NP, W);
    HIPCHK(c, hipEventRecord(c->ev[3], s));
    if constexpr (std::is_same<T, float>::value) {
      if (bf16x3)
        launch_potential_bf16x3(s, W, phi, c->Mc, c->NP, c->tf, c->nsplit, c->Hpart, ch == 0);
      else
        launch_potential<T>(s, W, phi, c->Mc, c->NP, c->tf, c->nsplit, c->Hpart, ch == 0);
    } else {
      launch_potential<T>(s, W, phi, c->Mc, c->NP, c->tf, c->nsplit, c->Hpart, ch == 0);
    }
    HIPCHK(c, hipEventRecord(c->ev[4], s));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(c->ev[4]));
    for (int k = 0; k < 4; ++k) {
      float t = 0;
      HIPCHK(c, hipEventElapsedTime(&t, c->ev[k], c->ev[k + 1]));
      ms[k] += t;
    }
  }
  for (int k = 0; k < 4; ++k) c->timing[k] = ms[k];
  return PGXC_OK;
}

extern "C" {

int pgxc_build(pgxc_ctx *c, uint32_t functional, int32_t precision, double *ham, double *exc,
               double *charge) {
  if (!c) return PGXC_ERR_ARG;
  if (functional == 0 || (functional & ~kSupported))
    return fail(c, PGXC_ERR_UNSUPPORTED, "functional mask has terms other than XAlpha, PWLDAc, PBEx, PBEc");
  if ((functional & PGXC_XC_PBEC) && !(functional & PGXC_XC_PWLDAC))
    return fail(c, PGXC_ERR_ARG, "PBEc needs the PW92 LDA correlation it is built on (PWLDAc)");
  if (precision != PGXC_PREC_FP64 && precision != PGXC_PREC_FP32 && precision != PGXC_PREC_BF16X3)
    return fail(c, PGXC_ERR_UNSUPPORTED, "precision mode " + std::to_string(precision) + " not available");
  if (!c->have_density) return fail(c, PGXC_ERR_STATE, "pgxc_build before pgxc_set_density/orbitals");
  const bool gga = (functional & (PGXC_XC_PBEX | PGXC_XC_PBEC)) != 0;
  bool incremental = false;
  if (c->density_mode == PGXC_DENSITY_INCREMENTAL) {
    if (!c->have_state)
      return fail(c, PGXC_ERR_STATE, "incremental density but no previous build on this context");
    if (c->state_functional != functional)
      return fail(c, PGXC_ERR_STATE, "incremental build with a functional different from the previous build");
    incremental = true;
  }
  hipStream_t s = c->stream;
  const long np2 = (long)c->NP * c->NP;
  HIPCHK(c, hipEventRecord(c->ev[10], s));
  if (incremental)
    launch_sub(s, c->P, c->Pref, np2, c->Dop);
  else
    HIPCHK(c, hipMemcpyAsync(c->Dop, c->P, np2 * 8, hipMemcpyDeviceToDevice, s));
  int rc;
  if (precision == PGXC_PREC_FP64) {
    if ((rc = ensure_work(c, 8))) return rc;
    if ((rc = run_chunks<double>(c, gga, incremental, functional, c->Dop, false))) return rc;
  } else {
    if ((rc = dalloc(c, &c->Dlow, np2))) return rc;
    launch_cast<float>(s, c->Dop, np2, c->Dlow);
    if ((rc = ensure_work(c, 4))) return rc;
    if ((rc = run_chunks<float>(c, gga, incremental, functional, c->Dlow,
                                 precision == PGXC_PREC_BF16X3)))
      return rc;
  }
  HIPCHK(c, hipEventRecord(c->ev[5], s));
  launch_pack(s, c->Hpart, c->nsplit, c->NP, c->nbas, incremental ? c->Vstate : nullptr, c->Vout);
  const size_t npk = (size_t)c->nbas * (c->nbas + 1) / 2;
  HIPCHK(c, hipMemcpyAsync(c->Vstate, c->Vout, npk * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->Pref, c->P, np2 * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(c, hipEventRecord(c->ev[6], s));
  HIPCHK(c, hipGetLastError());
  std::vector<double> ep(2 * (size_t)c->nblk_total);
  HIPCHK(c, hipMemcpyAsync(ep.data(), c->epart, ep.size() * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  float t = 0;
  HIPCHK(c, hipEventElapsedTime(&t, c->ev[5], c->ev[6]));
  c->timing[4] = t;
  HIPCHK(c, hipEventElapsedTime(&t, c->ev[10], c->ev[6]));
  c->timing[5] = t;
  double e = 0.0, q = 0.0;
  for (long b = 0; b < c->nblk_total; ++b) {
    e += ep[2 * b];
    q += ep[2 * b + 1];
  }
  c->last_exc = e;
  c->last_charge = q;
  c->have_state = 1;
  c->state_functional = functional;
  // dense-equivalent work model (no screening on this path)
  const double B = (double)c->npts, n = (double)c->nbas;
  c->stats[0] = B;
  c->stats[1] = std::ceil(B / c->veclen);
  c->stats[2] = 4.0 * B * n * n + 16.0 * B * n;
  c->stats[3] = 0;
  c->stats[4] = 32.0 * B + 16.0 * n * n;
  return pgxc_get_result(c, ham, exc, charge);
}

int pgxc_get_result(pgxc_ctx *c, double *ham, double *exc, double *charge) {
  if (!c) return PGXC_ERR_ARG;
  if (!c->have_state) return fail(c, PGXC_ERR_STATE, "no build has run on this context");
  if (ham) {
    const size_t npk = (size_t)c->nbas * (c->nbas + 1) / 2;
    std::vector<double> v(npk);
    HIPCHK(c, hipMemcpy(v.data(), c->Vout, npk * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < npk; ++i) ham[i] += v[i];
  }
  if (exc) *exc += c->last_exc;
  if (charge) *charge += c->last_charge;
  return PGXC_OK;
}

int pgxc_last_stats(const pgxc_ctx *c, double *s5) {
  if (!c || !s5) return PGXC_ERR_ARG;
  for (int i = 0; i < 5; ++i) s5[i] = c->stats[i];
  return PGXC_OK;
}

int pgxc_last_timing(const pgxc_ctx *c, double *ms6) {
  if (!c || !ms6) return PGXC_ERR_ARG;
  for (int i = 0; i < 6; ++i) ms6[i] = c->timing[i];
  return PGXC_OK;
}

int pgxc_destroy(pgxc_ctx *c) {
  if (!c) return PGXC_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *bufs[] = {c->grid, c->atoms, c->shells, c->exps, c->norms, c->contr, c->P, c->Pref,
                  c->Dop, c->Dlow, c->phi, c->W, c->part, c->coef, c->Hpart, c->state,
                  c->Vstate, c->Vout, c->epart};
  for (void *b : bufs) dfree(b);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PGXC_OK;
}

}  // extern "C"

extern "C" int pgxc_becke_weights(int32_t n_atoms, const double *atom_xyz, int64_t npts,
                                  const double *pts, const int32_t *owner,
                                  const double *w_atomic, double *w_out) {
  if (n_atoms <= 0 || npts < 0 || !atom_xyz || (npts > 0 && (!pts || !owner || !w_atomic || !w_out)))
    return PGXC_ERR_ARG;
  if (npts == 0) return PGXC_OK;
  for (int64_t p = 0; p < npts; ++p)
    if (owner[p] < 0 || owner[p] >= n_atoms) return PGXC_ERR_ARG;
  std::vector<double> inv((size_t)n_atoms * n_atoms, 0.0);
  for (int i = 0; i < n_atoms; ++i)
    for (int k = 0; k < n_atoms; ++k) {
      if (i == k) continue;
      const double dx = atom_xyz[3 * i] - atom_xyz[3 * k], dy = atom_xyz[3 * i + 1] - atom_xyz[3 * k + 1],
                   dz = atom_xyz[3 * i + 2] - atom_xyz[3 * k + 2];
      const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (!(d > 0.0)) return PGXC_ERR_ARG;  // coincident centres
      inv[(size_t)i * n_atoms + k] = 1.0 / d;
    }
  double *d_a = nullptr, *d_inv = nullptr, *d_p = nullptr, *d_w = nullptr, *d_o = nullptr;
  int *d_own = nullptr;
  hipStream_t s = nullptr;
  int rc = PGXC_OK;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == PGXC_OK) rc = e == hipErrorOutOfMemory ? PGXC_ERR_NOMEM : PGXC_ERR_HIP;
  };
  chk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  chk(hipMalloc((void **)&d_a, 24 * (size_t)n_atoms));
  chk(hipMalloc((void **)&d_inv, inv.size() * 8));
  chk(hipMalloc((void **)&d_p, 24 * (size_t)npts));
  chk(hipMalloc((void **)&d_w, 8 * (size_t)npts));
  chk(hipMalloc((void **)&d_o, 8 * (size_t)npts));
  chk(hipMalloc((void **)&d_own, 4 * (size_t)npts));
  if (rc == PGXC_OK) {
    chk(hipMemcpyAsync(d_a, atom_xyz, 24 * (size_t)n_atoms, hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_inv, inv.data(), inv.size() * 8, hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_p, pts, 24 * (size_t)npts, hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_w, w_atomic, 8 * (size_t)npts, hipMemcpyHostToDevice, s));
    chk(hipMemcpyAsync(d_own, owner, 4 * (size_t)npts, hipMemcpyHostToDevice, s));
  }
  if (rc == PGXC_OK) {
    launch_becke(s, n_atoms, d_a, d_inv, (long)npts, d_p, d_own, d_w, d_o);
    chk(hipGetLastError());
    chk(hipMemcpyAsync(w_out, d_o, 8 * (size_t)npts, hipMemcpyDeviceToHost, s));
    chk(hipStreamSynchronize(s));
  }
  dfree(d_a); dfree(d_inv); dfree(d_p); dfree(d_w); dfree(d_o); dfree(d_own);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

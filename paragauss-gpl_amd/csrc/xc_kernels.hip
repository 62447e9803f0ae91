// HIP kernels of the XC build for gfx950 (CDNA4, wave64).  All stage kernels
// run over work lists of units (a unit = one grid chunk of one molecule), so one
// launch per stage serves a single molecule or a batch of many.
//
//   k_basis       basis values and gradients  (ParaGauss modules/orbital_module.f90:1742-1932,
//                 2421-2619; solid harmonics modules/solid_harmonics_module.f90:369-568)
//   k_density     X = Phi D on MFMA, fused rho = sum Phi X, grad rho = 2 sum dPhi X
//                 (modules/density_calc_module.f90:1478-1577)
//   k_functional  Slater + PW92 + PBE-x + PBE-c in FP64, closed-form derivatives
//                 (modules/xc_func.f90:187-305 and the term modules)
//   k_wmat        W = a Phi + b . grad Phi  (modules/xc_hamiltonian.f90:1429-1483)
//   k_potential   H = W^T Phi on MFMA, split over points  (:1528-1531)
//   k_pack        V = H + H^T into the packed lower triangle (:1561-1567)
#include <type_traits>

#include "pgxc_internal.h"
#include "xc_device.h"

namespace pgxc {

// Widest vector (elements, <= 16 bytes) that splits a BM x BK operand tile
// evenly over the workgroup and along its contiguous dimension.
constexpr int pick_vec(int elems, int contig, int maxv) {
  for (int v = maxv; v > 1; v /= 2)
    if (elems % (kThreads * v) == 0 && contig % v == 0) return v;
  return 1;
}

template <typename T, int N>
struct VecT {
  typedef T type __attribute__((ext_vector_type(N)));
  __device__ static __forceinline__ T get(const type &v, int t) { return v[t]; }
};
template <typename T>
struct VecT<T, 1> {
  typedef T type;
  __device__ static __forceinline__ T get(const T &v, int) { return v; }
};

// Block tile BM x BN over K in steps of BK; 4 waves arranged 2 x 2, wave tile
// (BM/2) x (BN/2) made of 16x16 MFMA blocks.  ST = 2: two LDS stages, the next
// slab prefetched into registers while the current one feeds the MFMAs (one
// barrier per step).  ST = 1: one LDS stage, the next slab's global loads are
// in flight during the MFMAs (two barriers per step, half the LDS).
//   AKC: A(m,k) = A[m*sA + k] (k contiguous) else A[m + k*sA]
//   BKC: B(k,n) = B[n*sB + k] (k contiguous) else B[k*sB + n]
// LDS layout follows the operands.  Row-contiguous operands (both !KC) are
// kept [k][row]: the staging writes whole vectors and an MFMA operand is one
// element per lane.  k-contiguous operands (both KC) are kept [row][k] with
// rows padded by 16 bytes: the staging again writes the loaded vectors as
// they are, and each lane reads E = 16 / sizeof(T) consecutive k with one
// 16-byte read that feeds E MFMAs (MFMA e of a k chunk takes k = 4 E c + E g
// + e in lane group g, the same for A and B).  Rows 16 bytes apart put the
// 16-byte reads of 16 lanes on distinct banks.
template <typename T>
constexpr int kc_elems() { return 16 / (int)sizeof(T); }
template <typename T, int BK>
constexpr int kc_row() { return BK + kc_elems<T>(); }
template <typename T, int R, int BK, int ST, bool KC>
constexpr int lds_elems() { return ST * R * (KC ? kc_row<T, BK>() : BK); }

// Operand addressing (gathered rows and K runs): aaddr(i, kt, kk) is the
// element offset in A of the vector that starts at tile row / column i and
// local k index kk of K step kt (vectors run along the contiguous dimension),
// or -1 for a vector of zeros; baddr likewise for B.  Each operand is kept
// in LDS in its own layout ([row][k] padded when k-contiguous, else
// [k][row]); with both k-contiguous the MFMAs take E k per 16-byte read,
// otherwise one element per lane and k.
template <typename T, int BM, int BN, int BK, int ST, bool AKC, bool BKC, class AAddr,
          class BAddr>
__device__ __forceinline__ void gemm_tile(const T *__restrict__ A, const T *__restrict__ B,
                                          int ksteps, AAddr aaddr, BAddr baddr, T *As, T *Bs,
                                          typename MF<T>::acc (&acc)[BM / 32][BN / 32]) {
  constexpr bool KC = AKC && BKC;
  constexpr int E = kc_elems<T>(), BKP = kc_row<T, BK>();
  static_assert(!KC || BK % (4 * E) == 0, "k chunk");
  // elements of one LDS stage of each operand
  constexpr int SA = AKC ? BM * BKP : BK * BM, SB = BKC ? BN * BKP : BK * BN;
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int MAXV = 16 / (int)sizeof(T);
  constexpr int VA = pick_vec(BM * BK, AKC ? BK : BM, MAXV);
  constexpr int VB = pick_vec(BN * BK, BKC ? BK : BN, MAXV);
  constexpr int LA = BM * BK / (kThreads * VA), LB = BN * BK / (kThreads * VB);
  static_assert(LA * kThreads * VA == BM * BK && LB * kThreads * VB == BN * BK, "tile split");
  typedef VecT<T, VA> VA_t;
  typedef VecT<T, VB> VB_t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  typename VA_t::type ra[LA];
  typename VB_t::type rb[LB];
  unsigned oka = 0, okb = 0;  // bit e: vector e is real (else zeros)
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = typename MF<T>::acc{0, 0, 0, 0};

  if (ksteps <= 0) return;
  // Loads are unconditional global loads (a vector of zeros reads the
  // operand's first vector instead, and the zeros are applied when it is
  // written to LDS), and a step past the last one loads zeros: the pipeline
  // is straight-line code, so the compiler's wait counts keep the next
  // step's loads in flight during the MFMAs.
  auto load = [&](int kt) {
    const bool live = kt < ksteps;
    unsigned m = 0;
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      const int idx = tid + kThreads * e;
      long o;
      if (AKC) {
        const int mm = idx / (BK / VA), kv = idx % (BK / VA);
        o = aaddr(mm, kt, kv * VA);
      } else {
        const int mv = idx % (BM / VA), kk = idx / (BM / VA);
        o = aaddr(mv * VA, kt, kk);
      }
      const bool g = live && o >= 0;
      ra[e] = gld((const typename VA_t::type *)(A + (g ? o : 0)));
      m |= (unsigned)g << e;
    }
    oka = m;
    m = 0;
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      const int idx = tid + kThreads * e;
      long o;
      if (BKC) {
        const int nn = idx / (BK / VB), kv = idx % (BK / VB);
        o = baddr(nn, kt, kv * VB);
      } else {
        const int nv = idx % (BN / VB), kk = idx / (BN / VB);
        o = baddr(nv * VB, kt, kk);
      }
      const bool g = live && o >= 0;
      rb[e] = gld((const typename VB_t::type *)(B + (g ? o : 0)));
      m |= (unsigned)g << e;
    }
    okb = m;
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < LA; ++e)
      if (!((oka >> e) & 1u)) ra[e] = typename VA_t::type{};
#pragma unroll
    for (int e = 0; e < LB; ++e)
      if (!((okb >> e) & 1u)) rb[e] = typename VB_t::type{};
    T *as = As + buf * SA, *bs = Bs + buf * SB;
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      const int idx = tid + kThreads * e;
      if constexpr (AKC) {  // the vector as loaded, row-major [row][k]
        const int mm = idx / (BK / VA), kv = idx % (BK / VA);
        *(typename VA_t::type *)(as + mm * BKP + kv * VA) = ra[e];
      } else {
#pragma unroll
        for (int t = 0; t < VA; ++t) {
          const int mm = (idx % (BM / VA)) * VA + t, kk = idx / (BM / VA);
          as[kk * BM + mm] = VA_t::get(ra[e], t);
        }
      }
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      const int idx = tid + kThreads * e;
      if constexpr (BKC) {
        const int nn = idx / (BK / VB), kv = idx % (BK / VB);
        *(typename VB_t::type *)(bs + nn * BKP + kv * VB) = rb[e];
      } else {
#pragma unroll
        for (int t = 0; t < VB; ++t) {
          const int nn = (idx % (BN / VB)) * VB + t, kk = idx / (BN / VB);
          bs[kk * BN + nn] = VB_t::get(rb[e], t);
        }
      }
    }
  };
  auto compute = [&](int buf) {
    if constexpr (KC) {
      typedef typename VecT<T, E>::type VE;
      const T *as = As + buf * BM * BKP, *bs = Bs + buf * BN * BKP;
      const int ko = E * (lane >> 4);
#pragma unroll
      for (int c = 0; c < BK / (4 * E); ++c) {
        VE a[TM], b[TN];
#pragma unroll
        for (int i = 0; i < TM; ++i)
          a[i] = *(const VE *)(as + (wm * WM + i * 16 + (lane & 15)) * BKP + 4 * E * c + ko);
#pragma unroll
        for (int j = 0; j < TN; ++j)
          b[j] = *(const VE *)(bs + (wn * WN + j * 16 + (lane & 15)) * BKP + 4 * E * c + ko);
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = MF<T>::mma(a[i][e], b[j][e], acc[i][j]);
      }
      return;
    }
    const T *as = As + buf * SA, *bs = Bs + buf * SB;
#pragma unroll
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      const int kr = k4 * 4 + (lane >> 4);
      T a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int r = wm * WM + i * 16 + (lane & 15);
        a[i] = AKC ? as[r * BKP + kr] : as[kr * BM + r];
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int r = wn * WN + j * 16 + (lane & 15);
        b[j] = BKC ? bs[r * BKP + kr] : bs[kr * BN + r];
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = MF<T>::mma(a[i], b[j], acc[i][j]);
    }
  };

  if constexpr (ST == 2) {
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < ksteps; ++kt) {
      const int cur = kt & 1;
      load(kt + 1);
      compute(cur);
      store(cur ^ 1);
      __syncthreads();
    }
  } else {
    load(0);
    for (int kt = 0; kt < ksteps; ++kt) {
      store(0);
      __syncthreads();
      load(kt + 1);
      compute(0);
      __syncthreads();
    }
  }
}

// --------------------------------------------------------- solid harmonics
// Racah-normalised real solid harmonics up to L and their gradients, by the
// reference's recursion in l (m = 1 is z-like, then (2m, 2m+1) pairs, top
// pair by the (x, y) rotation); gradients by differentiating the recursion.
template <int L, typename R>
__device__ __forceinline__ void sh_eval(R x, R y, R z, R *C, R *Cx, R *Cy, R *Cz) {
  const R r2 = x * x + y * y + z * z;
  C[0] = R(1);
  Cx[0] = Cy[0] = Cz[0] = R(0);
  if constexpr (L >= 1) {
    C[1] = z; C[2] = x; C[3] = y;
    Cx[1] = R(0); Cy[1] = R(0); Cz[1] = R(1);
    Cx[2] = R(1); Cy[2] = R(0); Cz[2] = R(0);
    Cx[3] = R(0); Cy[3] = R(1); Cz[3] = R(0);
  }
  const R dr2x = R(2) * x, dr2y = R(2) * y, dr2z = R(2) * z;
#pragma unroll
  for (int l = 2; l <= L; ++l) {
    const int o = l * l, o1 = (l - 1) * (l - 1), o2 = (l - 2) * (l - 2);
    const R lr = l, l1r = l - 1, ll1r = 2 * l - 1;
    C[o] = (ll1r * z * C[o1] - l1r * r2 * C[o2]) / lr;
    Cx[o] = (ll1r * z * Cx[o1] - l1r * (dr2x * C[o2] + r2 * Cx[o2])) / lr;
    Cy[o] = (ll1r * z * Cy[o1] - l1r * (dr2y * C[o2] + r2 * Cy[o2])) / lr;
    Cz[o] = (ll1r * (C[o1] + z * Cz[o1]) - l1r * (dr2z * C[o2] + r2 * Cz[o2])) / lr;
#pragma unroll
    for (int im = 1; im <= l - 2; ++im) {
      const R sim = (R)(im * im);
      const R rlm = R(1) / sqrt(lr * lr - sim);
      const R rl1m = sqrt(l1r * l1r - sim);
      const R a = ll1r * z * rlm, b = rl1m * r2 * rlm;
      const R daz = ll1r * rlm, db = rl1m * rlm;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int m = 2 * im + k - 1;
        C[o + m] = a * C[o1 + m] - b * C[o2 + m];
        Cx[o + m] = a * Cx[o1 + m] - b * Cx[o2 + m] - db * dr2x * C[o2 + m];
        Cy[o + m] = a * Cy[o1 + m] - b * Cy[o2 + m] - db * dr2y * C[o2 + m];
        Cz[o + m] = a * Cz[o1 + m] + daz * C[o1 + m] - b * Cz[o2 + m] - db * dr2z * C[o2 + m];
      }
    }
    {
      const R sim = (R)((l - 1) * (l - 1));
      const R rlm = R(1) / sqrt(lr * lr - sim);
      const R a = ll1r * z * rlm, daz = ll1r * rlm;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int m = 2 * l - 3 + k;
        C[o + m] = a * C[o1 + m];
        Cx[o + m] = a * Cx[o1 + m];
        Cy[o + m] = a * Cy[o1 + m];
        Cz[o + m] = a * Cz[o1 + m] + daz * C[o1 + m];
      }
    }
    {
      const R rll = sqrt(ll1r / (R(2) * lr));
      const int i2 = o1 + 2 * l - 3, i1 = o1 + 2 * l - 2;
      C[o + 2 * l - 1] = (x * C[i2] - y * C[i1]) * rll;
      C[o + 2 * l] = (y * C[i2] + x * C[i1]) * rll;
      Cx[o + 2 * l - 1] = (C[i2] + x * Cx[i2] - y * Cx[i1]) * rll;
      Cy[o + 2 * l - 1] = (x * Cy[i2] - C[i1] - y * Cy[i1]) * rll;
      Cz[o + 2 * l - 1] = (x * Cz[i2] - y * Cz[i1]) * rll;
      Cx[o + 2 * l] = (y * Cx[i2] + C[i1] + x * Cx[i1]) * rll;
      Cy[o + 2 * l] = (C[i2] + y * Cy[i2] + x * Cy[i1]) * rll;
      Cz[o + 2 * l] = (y * Cz[i2] + x * Cz[i1]) * rll;
    }
  }
}

// ------------------------------------------------------------------- basis
// Evaluation in the output precision T: FP64 for the plain build, FP32 for
// the mixed builds (whose low-precision error is proportional to ||dP||).
// Value and gradient of function m of a shell of angular momentum L at one
// point (r0, r1: the contracted radial factor and its derivative factor);
// every consumer of basis values forms them with this one expression.
template <int L, bool GRAD, typename T>
__device__ __forceinline__ void fn_values(int m, T r0, T r1, T x, T y, T z, const T *C,
                                          const T *Cx, const T *Cy, const T *Cz, T &v, T &gx,
                                          T &gy, T &gz) {
  constexpr int o = L * L;
  const T sv = C[o + m];
  v = sv * r0;
  if (GRAD) {
    const T hv = sv * r1;
    gx = x * hv + Cx[o + m] * r0;
    gy = y * hv + Cy[o + m] * r0;
    gz = z * hv + Cz[o + m] * r0;
  }
}

// Emitter of the basis stage: stores a shell's values (and gradients, when
// wg) at point p of the [4][NP][Mc] planes.  Global (not flat) stores: a flat
// store also counts against the LDS wait counter, so every wait for the
// staged shell data would wait for the stores issued before it.
template <typename T>
struct StorePhi {
  T *phi;
  long Mc, cs;
  int p;
  template <int L, bool GRAD>
  __device__ __forceinline__ void emit(long f0, int nf, T r0, T r1, T x, T y, T z, const T *C,
                                       const T *Cx, const T *Cy, const T *Cz, bool wg) const {
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) {
      const long q = (f0 + (long)m * nf) * Mc + p;
      T v, gx = T(0), gy = T(0), gz = T(0);
      fn_values<L, false>(m, r0, r1, x, y, z, C, Cx, Cy, Cz, v, gx, gy, gz);
      gst(phi + q, v);
      if (GRAD && wg) {
        fn_values<L, true>(m, r0, r1, x, y, z, C, Cx, Cy, Cz, v, gx, gy, gz);
        gst(phi + cs + q, gx);
        gst(phi + 2 * cs + q, gy);
        gst(phi + 3 * cs + q, gz);
      }
    }
  }
};

// A range of atoms' basis data staged in LDS for a basis item: the shells,
// their exponents, norms and contraction coefficients in the evaluation
// precision, per atom its centre, shell range (relative to the range's
// first shell) and largest l, and per (shell, 32-point group of the item)
// what to write: bit 0 the values, bit 1 the gradients.
template <typename T>
struct RangeLds {
  ShellD sh[kRangeShells];
  T ex[kRangePrims], nm[kRangePrims], ct[kRangeContr];
  double ax[kRangeAtoms], ay[kRangeAtoms], az[kRangeAtoms];
  int s0[kRangeAtoms], s1[kRangeAtoms], lm[kRangeAtoms];
  uint8_t fl[kRangeShells][kThreads / kGroup];
};

// ex, nm: the shell's primitives; ct: its n_exp x n_contr coefficients.
// Primitives go in blocks of kPB: the block's exponents, norms and
// coefficients are read from LDS together (one wait, not one per read), its
// radial factors formed without a branch per primitive (a block whose
// primitives are all cut off on the wave's points skips the exponentials),
// then added into the contractions in primitive order.
constexpr int kPB = 4;
template <int L, bool GRAD, typename T, class Put>
__device__ __forceinline__ void shell_emit(const ShellD &sh, const T *ex, const T *nm, const T *ct,
                                           T x, T y, T z, T r2, const T *C, const T *Cx,
                                           const T *Cy, const T *Cz, const Put &put, bool wg) {
  const int ne = sh.n_exp, nu = sh.n_unc, nc = sh.n_contr, nf = nu + nc;
  T rc0[kMaxContr], rc1[kMaxContr];
#pragma unroll
  for (int k = 0; k < kMaxContr; ++k) rc0[k] = rc1[k] = T(0);
  for (int e0 = 0; e0 < ne; e0 += kPB) {
    T al[kPB], ar2[kPB], nv[kPB], r0[kPB], r1[kPB];
    bool any = false;
#pragma unroll
    for (int i = 0; i < kPB; ++i) {
      const int e = min(e0 + i, ne - 1);  // past the shell: a repeat, never used
      al[i] = ex[e];
      nv[i] = nm[e];
      ar2[i] = al[i] * r2;
      any = any || (e0 + i < ne && ar2[i] < T(50));
    }
    if (__ballot(any)) {
#pragma unroll
      for (int i = 0; i < kPB; ++i) {
        const T v = nv[i] * exp(-ar2[i]);
        r0[i] = (e0 + i < ne && ar2[i] < T(50)) ? v : T(0);  // primitive cut-off of radial1()
        r1[i] = (T(-2) * al[i]) * r0[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < kPB; ++i) r0[i] = r1[i] = T(0);
    }
#pragma unroll
    for (int i = 0; i < kPB; ++i)
      if (e0 + i < nu)
        put.template emit<L, GRAD>(sh.func_off + e0 + i, nf, r0[i], r1[i], x, y, z, C, Cx, Cy, Cz, wg);
#pragma unroll
    for (int k = 0; k < kMaxContr; ++k)
      if (k < nc) {
        T c[kPB];
#pragma unroll
        for (int i = 0; i < kPB; ++i) c[i] = ct[k * ne + min(e0 + i, ne - 1)];
#pragma unroll
        for (int i = 0; i < kPB; ++i)
          if (e0 + i < ne) {
            rc0[k] += r0[i] * c[i];
            rc1[k] += r1[i] * c[i];
          }
      }
  }
#pragma unroll
  for (int k = 0; k < kMaxContr; ++k)
    if (k < nc)
      put.template emit<L, GRAD>(sh.func_off + nu + k, nf, rc0[k], rc1[k], x, y, z, C, Cx, Cy, Cz, wg);
}

// One atom's shells sb .. se-1 of the LDS image A at one point (e0, c0: the
// range's first primitive and coefficient); grp: the point's group in the
// item.
template <int LA, bool GRAD, bool ALLG = false, typename T, class Put>
__device__ __forceinline__ void basis_atom(const RangeLds<T> &A, int sb, int se, int e0, int c0,
                                           T x, T y, T z, const Put &put, int grp) {
  constexpr int NLM = (LA + 1) * (LA + 1);
  T C[NLM], Cx[NLM], Cy[NLM], Cz[NLM];
  sh_eval<LA, T>(x, y, z, C, Cx, Cy, Cz);
  const T r2 = x * x + y * y + z * z;
  for (int s = sb; s < se; ++s) {
    const int f = A.fl[s][grp];
    if (!(f & 1)) continue;  // no stage reads the shell's values on this group
    const bool wg = GRAD && (ALLG || (f & 2));  // ALLG: gradients wherever values are formed
    // every lane reads the same shell: its fields are uniform over the wave,
    // so the loops over its primitives and functions run on scalar counters
    ShellD sh = A.sh[s];
    sh.l = __builtin_amdgcn_readfirstlane(sh.l);
    sh.n_exp = __builtin_amdgcn_readfirstlane(sh.n_exp);
    sh.n_unc = __builtin_amdgcn_readfirstlane(sh.n_unc);
    sh.n_contr = __builtin_amdgcn_readfirstlane(sh.n_contr);
    sh.exp_off = __builtin_amdgcn_readfirstlane(sh.exp_off);
    sh.contr_off = __builtin_amdgcn_readfirstlane(sh.contr_off);
    sh.func_off = __builtin_amdgcn_readfirstlane(sh.func_off);
    const T *ex = A.ex + (sh.exp_off - e0), *nm = A.nm + (sh.exp_off - e0);
    const T *ct = A.ct + (sh.contr_off - c0);
    switch (sh.l) {
      case 0: shell_emit<0, GRAD>(sh, ex, nm, ct, x, y, z, r2, C, Cx, Cy, Cz, put, wg); break;
      case 1: if constexpr (LA >= 1) shell_emit<1, GRAD>(sh, ex, nm, ct, x, y, z, r2, C, Cx, Cy, Cz, put, wg); break;
      case 2: if constexpr (LA >= 2) shell_emit<2, GRAD>(sh, ex, nm, ct, x, y, z, r2, C, Cx, Cy, Cz, put, wg); break;
      case 3: if constexpr (LA >= 3) shell_emit<3, GRAD>(sh, ex, nm, ct, x, y, z, r2, C, Cx, Cy, Cz, put, wg); break;
      case 4: if constexpr (LA >= 4) shell_emit<4, GRAD>(sh, ex, nm, ct, x, y, z, r2, C, Cx, Cy, Cz, put, wg); break;
    }
  }
}

// -------------------------------------------------------------- functional
// Closed-shell terms; each returns F, dF/drho, dF/dgamma added into (f, fr, fg).
namespace xcf {
constexpr double PI = 3.14159265358979324;
constexpr double DEPS = 1.0e-30;

// Slater exchange: -(3/4)(3/pi)^(1/3) rho^(4/3), zero for rho <= 0
__device__ __forceinline__ void xalpha(double rho, double &f, double &fr) {
  const double xa = -(3.0 / 4.0) * cbrt(3.0 / PI);
  if (rho > 0.0) {
    const double c = cbrt(rho);
    f += xa * (rho * c);
    fr += xa * ((4.0 / 3.0) * c);
  }
}

// PW92 paramagnetic G(rs) and dG/drs
__device__ __forceinline__ void pw92_g(double rs, double &E, double &dE) {
  const double a = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382,
               b4 = 0.49294;
  const double s = sqrt(rs);
  const double Q0 = -(2.0 * a * (1.0 + a1 * rs));
  const double Q1 = 2.0 * a * (b1 * s + b2 * rs + b3 * rs * s + b4 * rs * rs);
  const double Q2 = log(1.0 + 1.0 / Q1);
  const double dQ1 = 2.0 * a * (b1 * 0.5 / s + b2 + 1.5 * b3 * s + 2.0 * b4 * rs);
  const double dQ2 = (-dQ1 / (Q1 * Q1)) / (1.0 + 1.0 / Q1);
  E = Q0 * Q2;
  dE = (-2.0 * a * a1) * Q2 + Q0 * dQ2;
}

// n * eps_c^PW92(n, zeta = 0) with the n < DEPS clamp; value and dn
__device__ __forceinline__ void pw92_unpol(double n, double &F, double &dF) {
  const double crs = cbrt(3.0 / (4.0 * PI));
  double nn = n;
  bool fixed = false;
  if (nn < DEPS) { nn = DEPS; fixed = true; }
  const double rs = crs / cbrt(nn);
  double E, dE;
  pw92_g(rs, E, dE);
  F = nn * E;
  dF = fixed ? 0.0 : E + nn * dE * (-rs / (3.0 * nn));
}

__device__ __forceinline__ void pwldac(double rho, double &f, double &fr) {
  double F, dF;
  pw92_unpol(0.5 * rho + 0.5 * rho, F, dF);
  f += F;
  fr += dF;
}

// PBE exchange excess e_x G(s^2), G = kappa mu s2 / (kappa + mu s2)
__device__ __forceinline__ void pbex(double rho, double gam, double &f, double &fr,
                                     double &fg) {
  const double xa = -(3.0 / 4.0) * cbrt(3.0 / PI);
  const double kappa = 0.804, mu = 0.066725 * (PI * PI / 3.0);
  const double s2_fact = 1.0 / (4.0 * cbrt(3.0 * PI * PI) * cbrt(3.0 * PI * PI));
  double ex = 0.0, dex = 0.0;
  if (rho > 0.0) {
    const double c = cbrt(rho);
    ex = xa * (rho * c);
    dex = xa * ((4.0 / 3.0) * c);
  }
  double s2 = 0.0, s2r = 0.0, s2g = 0.0;
  if (rho >= 1.0e-20) {
    const double c = cbrt(rho);
    s2g = s2_fact / (rho * rho * c * c);
    s2 = gam * s2g;
    s2r = -(8.0 / 3.0) * s2 / rho;
  }
  const double den = kappa + mu * s2;
  const double G = kappa * mu * s2 / den;
  const double dG = (kappa * kappa * mu) / (den * den);
  f += ex * G;
  fr += dex * G + ex * dG * s2r;
  fg += ex * dG * s2g;
}

// PBE correlation gradient term n H(n, gamma), zeta = 0 (phi = 1)
__device__ __forceinline__ void pbec(double rho, double g2, double &f, double &fr,
                                     double &fg) {
  const double beta = 0.066725, gc = 0.03109069086965489503494086371;
  const double bg = beta / gc;
  const double K = PI / (16.0 * cbrt(3.0 * PI * PI));
  const double nv = 0.5 * rho + 0.5 * rho;
  if (nv <= 0.0) return;
  double Fl, dFl;
  pw92_unpol(nv, Fl, dFl);
  const double El = Fl / nv, dEl = dFl / nv - Fl / (nv * nv);
  // phi = (za^(2/3) + zb^(2/3)) / 2^(1/3) = 1 for za = zb = 1/2
  const double fi3 = 1.0;
  const double Ex = exp(-El / (gc * fi3));
  const double dd = Ex - 1.0 + DEPS;
  const double A = bg / dd;
  const double dA = (bg / (dd * dd)) * Ex * (dEl / (gc * fi3));
  const double c = cbrt(nv);
  const double n73 = nv * nv * c;
  const double D = n73 + DEPS;
  const double t2 = g2 / D * K;
  const double dt2n = -(g2 * K / (D * D)) * (7.0 / 3.0) * (nv * c);
  const double dt2g = K / D;
  const double x = A * t2;
  const double dxn = dA * t2 + A * dt2n, dxg = A * dt2g;
  const double q = 1.0 + x + x * x;
  const double rat = (1.0 + x) / q;
  const double drat = -x * (2.0 + x) / (q * q);
  const double arg = 1.0 + bg * t2 * rat;
  const double H = gc * fi3 * log(arg);
  const double pre = gc * fi3 / arg * bg;
  f += nv * H;
  fr += H + nv * pre * (dt2n * rat + t2 * drat * dxn);
  fg += nv * pre * (dt2g * rat + t2 * drat * dxg);
}

// The terms of mask in enum order (XAlpha, PWLDAc, PBEx, PBEc): each term's
// F into e[0..3] (zeroed by the caller), dF/drho and dF/dgamma summed over
// them; the functional's F is ((e0 + e1) + e2) + e3.
template <bool GGA>
__device__ __forceinline__ void terms(uint32_t mask, double rho, double gam, double (&e)[4],
                                      double &fr, double &fg) {
  if (mask & (1u << 1)) xalpha(rho, e[0], fr);
  if (mask & (1u << 3)) pwldac(rho, e[1], fr);
  if (GGA) {
    if (mask & (1u << 10)) pbex(rho, gam, e[2], fr, fg);
    if (mask & (1u << 18)) pbec(rho, gam, e[3], fr, fg);
  }
}
__device__ __forceinline__ double sum4(const double (&e)[4]) { return ((e[0] + e[1]) + e[2]) + e[3]; }
}  // namespace xcf

// ============================================================ work lists
// Every stage kernel is persistent: a grid of at most a few waves of
// workgroups walks the stage's item list (item = unit-local tile), in the
// order given by `order` (cost-model schedule) or, without one, in an
// XCD-blocked order so that neighbouring tiles share an XCD's L2.  Items write
// disjoint outputs, so the order never changes a result bit.

__device__ __forceinline__ int find_unit(const long *__restrict__ base, int nu, long it) {
  int lo = 0, hi = nu - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= it) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ long item_at(const int *__restrict__ order, long L, long total) {
  return order ? (long)order[L] : (long)xcd_remap((int)L, (int)total);
}

// ------------------------------------------------------------------- basis
// item = (256-point block, range of atoms) of a unit; one thread per point,
// walking the range's atoms.  The workgroup first stages the range's shells,
// primitives, coefficients and atom centres and the per-(shell, group) write
// flags in LDS, all loads in flight together, so the evaluation itself waits
// on no global load.  Screened units (bscreen) walk their list of the
// (block, range) pairs whose values a later stage reads, and write only the
// shells read on each point's group; dense units write every shell.
// A basis item staged: its first point, the range's first primitive and
// coefficient, its atom count, and the point's coordinates.
struct BasisItem {
  long p;
  int e0, c0, n_atoms;
  double gx, gy, gz;
};

// Stages basis item `local` of unit U in A (every thread calls it; starts
// and ends with a barrier).
template <bool GRAD, typename T>
__device__ __forceinline__ BasisItem stage_basis_item(const UnitD &U, long local, RangeLds<T> &A) {
  const int tid = threadIdx.x;
  long blk;
  int rng;
  if (U.bscreen) {
    const uint32_t e = U.bitems[local];
    blk = e >> 16;
    rng = (int)(e & 0xffffu);
  } else {
    rng = (int)(local % U.n_ranges);
    blk = local / U.n_ranges;
  }
  const RangeD R = U.ranges[rng];
  BasisItem it;
  it.e0 = R.exp_begin;
  it.c0 = R.contr_begin;
  it.n_atoms = R.n_atoms;
  it.p = blk * kThreads + tid;
  // the point's coordinates in FP64, loads in flight across the staging
  it.gx = gld(U.gx + it.p);
  it.gy = gld(U.gy + it.p);
  it.gz = gld(U.gz + it.p);
  __syncthreads();  // the previous item's reads of A are done
  for (int i = tid; i < R.n_shells; i += kThreads) A.sh[i] = U.shells[R.shell_begin + i];
  for (int i = tid; i < R.n_exp; i += kThreads) {
    A.ex[i] = (T)gld(U.exps + it.e0 + i);
    A.nm[i] = (T)gld(U.norms + it.e0 + i);
  }
  for (int i = tid; i < R.n_contr; i += kThreads) A.ct[i] = (T)gld(U.contr + it.c0 + i);
  if (tid < R.n_atoms) {
    const AtomD at = U.atoms[gld(U.ratoms + R.atom_begin + tid)];
    A.ax[tid] = at.x;
    A.ay[tid] = at.y;
    A.az[tid] = at.z;
    A.s0[tid] = at.shell_begin - R.shell_begin;
    A.s1[tid] = at.shell_end - R.shell_begin;
    A.lm[tid] = at.lmax;
  }
  if (tid < R.n_shells * (kThreads / kGroup)) {  // flag of (shell tid / 8, group tid % 8)
    uint32_t f = GRAD ? 3u : 1u;
    if (U.bscreen) {
      const uint32_t *fw = U.bflags + 16 * local;
      f = ((fw[tid >> 5] >> (tid & 31)) & 1u) |
          (GRAD ? ((fw[8 + (tid >> 5)] >> (tid & 31)) & 1u) << 1 : 0u);
    }
    A.fl[tid / (kThreads / kGroup)][tid % (kThreads / kGroup)] = (uint8_t)f;
  }
  __syncthreads();
  return it;
}

// The staged range's atoms in turn at the thread's point, each shell written
// on the point's group handed to put (ALLG: with its gradients wherever its
// values are).
template <bool GRAD, bool ALLG, int LMAX, typename T, class Put>
__device__ __forceinline__ void eval_basis_item(const RangeLds<T> &A, const BasisItem &it,
                                                const Put &put) {
  const int grp = threadIdx.x / kGroup;
  const int na = __builtin_amdgcn_readfirstlane(it.n_atoms);
  for (int a = 0; a < na; ++a) {
    const int sb = __builtin_amdgcn_readfirstlane(A.s0[a]), se = __builtin_amdgcn_readfirstlane(A.s1[a]);
    bool any = false;
    for (int s = sb; s < se; ++s) any = any || (A.fl[s][grp] & 1);
    if (!any) continue;  // no stage reads this atom's values on the point's group
    const T x = (T)(it.gx - A.ax[a]), y = (T)(it.gy - A.ay[a]), z = (T)(it.gz - A.az[a]);
    // the kernel holds the registers of its largest case, so it is built
    // for the largest angular momentum of the list (every atom's lmax <= LMAX)
    switch (__builtin_amdgcn_readfirstlane(A.lm[a])) {
      case 0: basis_atom<0, GRAD, ALLG>(A, sb, se, it.e0, it.c0, x, y, z, put, grp); break;
      case 1: if constexpr (LMAX >= 1) basis_atom<1, GRAD, ALLG>(A, sb, se, it.e0, it.c0, x, y, z, put, grp); break;
      case 2: if constexpr (LMAX >= 2) basis_atom<2, GRAD, ALLG>(A, sb, se, it.e0, it.c0, x, y, z, put, grp); break;
      case 3: if constexpr (LMAX >= 3) basis_atom<3, GRAD, ALLG>(A, sb, se, it.e0, it.c0, x, y, z, put, grp); break;
      default: if constexpr (LMAX >= 4) basis_atom<4, GRAD, ALLG>(A, sb, se, it.e0, it.c0, x, y, z, put, grp); break;
    }
  }
}

template <bool GRAD, typename T, int LMAX>
__global__ __launch_bounds__(kThreads) void k_basis(const UnitD *__restrict__ units, int nu,
                                                    const long *__restrict__ base,
                                                    const int *__restrict__ order, long total) {
  __shared__ RangeLds<T> A;
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const long item = item_at(order, L, total);
    const int u = find_unit(base, nu, item);
    const UnitD &U = units[u];
    const BasisItem it = stage_basis_item<GRAD>(U, item - base[u], A);
    const StorePhi<T> put{(T *)U.phi, U.Mc, (long)U.NP * U.Mc, (int)it.p};
    eval_basis_item<GRAD, false, LMAX>(A, it, put);
  }
}

// Emitter of the split-BF16 W pass: W = a Phi + b . grad Phi of a shell's
// functions at point p (FP64 from the FP32 values, as a pass reading stored
// values forms it), stored as bf16 hi / lo planes for NS spins.
template <int NS>
struct StoreWSplit {
  uint16_t *wh[2], *wl[2];
  long Mc;
  int p;
  double a[2], b0[2], b1[2], b2[2];
  template <int L, bool GRAD>
  __device__ __forceinline__ void emit(long f0, int nf, float r0, float r1, float x, float y,
                                       float z, const float *C, const float *Cx, const float *Cy,
                                       const float *Cz, bool) const {
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) {
      const long q = (f0 + (long)m * nf) * Mc + p;
      float v, gx = 0.0f, gy = 0.0f, gz = 0.0f;
      fn_values<L, GRAD>(m, r0, r1, x, y, z, C, Cx, Cy, Cz, v, gx, gy, gz);
#pragma unroll
      for (int sp = 0; sp < NS; ++sp) {
        const double w = GRAD ? a[sp] * (double)v + b0[sp] * (double)gx + b1[sp] * (double)gy +
                                    b2[sp] * (double)gz
                              : a[sp] * (double)v;
        __bf16 h, l;
        split_bf16((float)w, h, l);
        gst(wh[sp] + q, __builtin_bit_cast(uint16_t, h));
        gst(wl[sp] + q, __builtin_bit_cast(uint16_t, l));
      }
    }
  }
};

// Split-BF16 W pass over the basis items: the basis functions evaluated again
// (FP32, the same expressions as the basis stage, so the same values) instead
// of read back, W written for the shells whose values the basis stage writes
// (TWO: both spins of an open-shell unit).  The split-BF16 potential reads W
// only for the functions of a tile's list (with wide point splits, also the
// zeros of the other functions of its chunks, which are written here too).
template <bool GRAD, bool TWO, int LMAX>
__global__ __launch_bounds__(kThreads) void k_wbasis(const UnitD *__restrict__ units, int nu,
                                                     const long *__restrict__ base,
                                                     const int *__restrict__ order, long total) {
  __shared__ RangeLds<float> A;
  constexpr int NS = TWO ? 2 : 1;
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const long item = item_at(order, L, total);
    const int u = find_unit(base, nu, item);
    const UnitD &U = units[u];
    const BasisItem it = stage_basis_item<GRAD>(U, item - base[u], A);
    const long Mc = U.Mc, cs = (long)U.NP * Mc;
    StoreWSplit<NS> put;
    put.Mc = Mc;
    put.p = (int)it.p;
#pragma unroll
    for (int sp = 0; sp < NS; ++sp) {
      uint16_t *w = (uint16_t *)(sp ? U.W2 : U.W);
      const double *cf = sp ? U.coef2 : U.coef;
      put.wh[sp] = w;
      put.wl[sp] = w + cs;
      put.a[sp] = gld(cf + it.p);
      put.b0[sp] = GRAD ? gld(cf + Mc + it.p) : 0.0;
      put.b1[sp] = GRAD ? gld(cf + 2 * Mc + it.p) : 0.0;
      put.b2[sp] = GRAD ? gld(cf + 3 * Mc + it.p) : 0.0;
    }
    eval_basis_item<GRAD, true, LMAX>(A, it, put);
  }
}

// ----------------------------------------------------------------- density
constexpr int kDBM = 128;

// K step and LDS stages of the contraction variants the layout can select
template <int V> struct Var;
template <> struct Var<0> { static constexpr int BK = 16, ST = 2; };
template <> struct Var<1> { static constexpr int BK = 32, ST = 1; };
template <> struct Var<2> { static constexpr int BK = 32, ST = 2; };

// item = (128-point tile, pair of function tiles of its gathered runs) of
// X = Phi D, then the tile's partial contractions per point, one slot per
// function tile
template <typename T, int BN, int VAR, bool GRAD>
__global__ __launch_bounds__(kThreads)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 || VAR != 2 ? 2 : 1))) void k_density(const StageList S) {
  constexpr int BM = kDBM, BK = Var<VAR>::BK, ST = Var<VAR>::ST, TM = BM / 32, TN = BN / 32;
  __shared__ T As[lds_elems<T, BM, BK, ST, false>()];
  __shared__ __attribute__((aligned(16))) T Bs[lds_elems<T, BN, BK, ST, true>()];
  __shared__ double red[2][4][BM];
  __shared__ uint16_t kl[kMaxRuns];
  __shared__ uint8_t km[kMaxRuns];  // per run of kl: its functions in the tile's list
  __shared__ uint16_t cl[2 * BN];
  __shared__ long slot;
  for (long L = walk_next(S, -1, &slot); L < S.total;) {
    const uint64_t t0 = S.tlog ? wall_clock64() : 0;
    const long it = walk_item(S, L);
    const int u = find_unit(S.base, S.nu, it);
    const UnitD &U = S.units[u];
    const uint32_t e = U.ditems[it - S.base[u]];
    const int tile = (int)(e >> 8), pair = (int)(e & 255);
    const long m0 = (long)tile * BM, Mc = U.Mc, NP = U.NP;
    const int off = U.tl_off[tile], c8 = U.tl_off[tile + 1] - off;
    const int foff = U.tf_off[tile] + 2 * pair * BN, nf = U.tf_off[tile + 1] - foff;
    __syncthreads();  // the previous item's readers of kl and cl are done
    for (int i = threadIdx.x; i < min(2 * BN, nf); i += kThreads) cl[i] = gld(U.tfl + foff + i);
    for (int i = threadIdx.x; i < c8; i += kThreads) km[i] = gld(U.tlm + off + i);
    lds_list(U.tl + off, c8, kl);
    const int ksteps = (c8 * kRun + BK - 1) / BK;
    const T *phi = (const T *)U.phi, *D = (const T *)U.D;
    int ahead = 0;
    // the item's two column chunks in turn (the second may be absent)
    for (int h = 0; h < 2 && h * BN < nf; ++h) {
      const uint16_t *cfn = cl + h * BN;
      const int ncol = min(BN, nf - h * BN);
      // A(m, k) = Phi[f(k)][m0 + m], B(k, n) = D[f(k)][col(n)] = D[col(n)][f(k)]
      // (D symmetric, so B is k-contiguous along runs): K runs over the
      // tile's gathered runs, the columns over its functions; A is 0 for the
      // functions of the runs outside the tile's list (left unwritten by the
      // basis stage)
      auto aaddr = [&](int m, int kt, int kk) -> long {
        const int k = kt * BK + kk;
        const long f = run_fn(kl, c8, k);
        return (f < 0 || !((km[k >> 3] >> (k & 7)) & 1u)) ? -1 : f * Mc + m;
      };
      auto baddr = [&](int n, int kt, int kk) -> long {
        const long f = run_fn(kl, c8, kt * BK + kk);
        return (f < 0 || n >= ncol) ? -1 : (long)cfn[min(n, ncol - 1)] * NP + f;
      };
      typename MF<T>::acc acc[TM][TN];
      gemm_tile<T, BM, BN, BK, ST, false, true>(phi + m0, D, ksteps, aaddr, baddr, As, Bs, acc);
      if (h == 1 || (h + 1) * BN >= nf) ahead = walk_ahead(S);  // the last chunk
      // FP64 variants 0 and 1 run two waves per SIMD: their epilogue reads
      // one component plane at a time to fit the registers
      constexpr bool PW = sizeof(T) == 8 && VAR != 2;
      density_epilogue<T, typename MF<T>::acc, BM, BN, GRAD, 1, kThreads, 2, PW>(
          acc, phi, Mc, U.NP, m0, cfn, ncol, 2 * pair + h, U.part, red);
    }
    item_log(S, it, t0);
    L = walk_take(S, L, ahead, &slot);
  }
}

// item = 256-point block of a unit
template <bool GGA>
__global__ __launch_bounds__(kThreads) void k_functional(const UnitD *__restrict__ units, int nu,
                                                         const long *__restrict__ base,
                                                         const int *__restrict__ order,
                                                         long total) {
  __shared__ double se[kThreads], sq[kThreads];
  const int tid = threadIdx.x;
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const long it = item_at(order, L, total);
    const int u = find_unit(base, nu, it);
    const UnitD &U = units[u];
    const long blk = it - base[u];
    const long p = blk * kThreads + tid, Mc = U.Mc, npad = U.npad;
    const double *part = U.part;
    double *state = U.state, *coef = U.coef;  // state is offset to the unit's first point
    double rho = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    // partial sums of the density items of this point's tile, one per chunk
    const int tile = (int)(p / kScreenTile);
    const int tf = U.CH * kRun, nch = (U.tf_off[tile + 1] - U.tf_off[tile] + tf - 1) / tf;
    for (int t = 0; t < nch; ++t) {
      rho += part[((long)t * 4 + 0) * Mc + p];
      if (GGA) {
        g0 += part[((long)t * 4 + 1) * Mc + p];
        g1 += part[((long)t * 4 + 2) * Mc + p];
        g2 += part[((long)t * 4 + 3) * Mc + p];
      }
    }
    if (U.incremental) {
      rho = state[S_RHO * npad + p] + rho;
      if (GGA) {
        g0 = state[S_GX * npad + p] + g0;
        g1 = state[S_GY * npad + p] + g1;
        g2 = state[S_GZ * npad + p] + g2;
      }
    }
    const double gam = g0 * g0 + g1 * g1 + g2 * g2;
    double e[4] = {0.0, 0.0, 0.0, 0.0}, fr = 0.0, fg = 0.0;
    xcf::terms<GGA>(U.mask, rho, gam, e, fr, fg);
    const double f = xcf::sum4(e);
    const double ww = U.gw[p];
    const double a = 0.5 * fr * ww;
    const double bq = ww * 2.0 * fg;
    const double b0 = bq * g0, b1 = bq * g1, b2 = bq * g2;
    if (U.incremental) {
      coef[0 * Mc + p] = a - state[S_A * npad + p];
      if (GGA) {
        coef[1 * Mc + p] = b0 - state[S_BX * npad + p];
        coef[2 * Mc + p] = b1 - state[S_BY * npad + p];
        coef[3 * Mc + p] = b2 - state[S_BZ * npad + p];
      }
    } else {
      coef[0 * Mc + p] = a;
      if (GGA) {
        coef[1 * Mc + p] = b0;
        coef[2 * Mc + p] = b1;
        coef[3 * Mc + p] = b2;
      }
    }
    state[S_RHO * npad + p] = rho;
    state[S_A * npad + p] = a;
    if (GGA) {
      state[S_GX * npad + p] = g0;
      state[S_GY * npad + p] = g1;
      state[S_GZ * npad + p] = g2;
      state[S_BX * npad + p] = b0;
      state[S_BY * npad + p] = b1;
      state[S_BZ * npad + p] = b2;
    }
    se[tid] = f * ww;
    sq[tid] = rho * ww;
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
      if (tid < st) {
        se[tid] += se[tid + st];
        sq[tid] += sq[tid + st];
      }
      __syncthreads();
    }
    if (tid == 0) {
      U.epart[2 * blk + 0] = se[0];
      U.epart[2 * blk + 1] = sq[0];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------- functional, ispin = 2
// Spin-polarised correlation on values with first derivatives in
// (n_a, n_b, g2): PW92 c_lda (modules/pw_ldac_module.f90:127-195) and PBE
// c_gga (modules/pbe_ggcxc_module.f90:96-140); exchange by spin scaling
// (modules/exchange.f90:253-288, :511-601) through the closed-shell terms.
namespace xcs {
using xcf::PI;
using xcf::DEPS;
struct Dv {
  double v, d0, d1, d2;
};
__device__ __forceinline__ Dv cst(double v) { return {v, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ Dv operator+(Dv a, Dv b) { return {a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2}; }
__device__ __forceinline__ Dv operator-(Dv a, Dv b) { return {a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2}; }
__device__ __forceinline__ Dv operator*(Dv a, Dv b) {
  return {a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2};
}
__device__ __forceinline__ Dv operator*(Dv a, double c) { return {a.v * c, a.d0 * c, a.d1 * c, a.d2 * c}; }
__device__ __forceinline__ Dv operator/(Dv a, Dv b) {
  const double q = a.v / b.v, r = 1.0 / b.v;
  return {q, (a.d0 - q * b.d0) * r, (a.d1 - q * b.d1) * r, (a.d2 - q * b.d2) * r};
}
__device__ __forceinline__ Dv chain(Dv a, double v, double g) { return {v, g * a.d0, g * a.d1, g * a.d2}; }
// the powers the formulas take, through sqrt / cbrt (arguments > 0 here)
__device__ __forceinline__ Dv sqrtd(Dv a) {
  const double t = sqrt(a.v);
  return chain(a, t, 0.5 / t);
}
__device__ __forceinline__ Dv cbrtd(Dv a) {  // a^(1/3)
  const double t = cbrt(a.v);
  return chain(a, t, t / (3.0 * a.v));
}
__device__ __forceinline__ Dv p23(Dv a) {  // a^(2/3)
  const double t = cbrt(a.v);
  return chain(a, t * t, (2.0 / 3.0) / t);
}
__device__ __forceinline__ Dv p43(Dv a) {  // a^(4/3)
  const double t = cbrt(a.v);
  return chain(a, a.v * t, (4.0 / 3.0) * t);
}
__device__ __forceinline__ Dv p73(Dv a) {  // a^(7/3)
  const double t = cbrt(a.v);
  return chain(a, a.v * a.v * t, (7.0 / 3.0) * a.v * t);
}
__device__ __forceinline__ Dv logd(Dv a) { return chain(a, log(a.v), 1.0 / a.v); }
__device__ __forceinline__ Dv expd(Dv a) {
  const double e = exp(a.v);
  return chain(a, e, e);
}

// PW92 G(rs) for the parameter sets of eps_c(zeta = 0), eps_c(zeta = 1) and
// -alpha_c (p = 1)
__device__ __forceinline__ Dv pw92_gs(Dv rs, int set) {
  const double A[3] = {0.031091, 0.015545, 0.016887};
  const double A1[3] = {0.21370, 0.20548, 0.11125};
  const double B1[3] = {7.5957, 14.1189, 10.357};
  const double B2[3] = {3.5876, 6.1977, 3.6231};
  const double B3[3] = {1.6382, 3.3662, 0.88026};
  const double B4[3] = {0.49294, 0.62517, 0.49671};
  const double a = A[set];
  const Dv s = sqrtd(rs);
  const Dv Q0 = (cst(1.0) + rs * A1[set]) * (-2.0 * a);
  const Dv Q1 = (s * B1[set] + rs * B2[set] + rs * s * B3[set] + rs * rs * B4[set]) * (2.0 * a);
  const Dv Q2 = logd(cst(1.0) + cst(1.0) / Q1);
  return Q0 * Q2;
}

// n eps_c^PW92(n_a, n_b)
__device__ __forceinline__ Dv pw92_spin(Dv na, Dv nb) {
  const double c13 = 1.0 / 3.0, c43 = 4.0 / 3.0;
  const double crs = cbrt(3.0 / (4.0 * PI));
  const double cf = 1.0 / (2.0 * cbrt(2.0) - 2.0);
  const double cfzz = 2.0 * c43 * c13 * cf;
  Dv n = na + nb;
  if (n.v < DEPS) n = cst(DEPS);
  const Dv rs = cst(crs) / cbrtd(n);
  const Dv eu = pw92_gs(rs, 0), ep = pw92_gs(rs, 1), ma = pw92_gs(rs, 2);
  const Dv z = (na - nb) / n, z2 = z * z, z4 = z2 * z2;
  const Dv fz = (p43(cst(1.0) + z + cst(DEPS)) + p43(cst(1.0) - z + cst(DEPS)) - cst(2.0)) * cf;
  const Dv ec = eu + (ep - eu) * fz * z4 - ma * fz * (cst(1.0) - z4) * (1.0 / cfzz);
  return n * ec;
}

// n H(n_a, n_b, g2), g2 = |grad n|^2, on lda = pw92_spin(n_a, n_b) (the
// PW92 term PBE-c is built on, computed once for both)
__device__ __forceinline__ Dv pbec_spin(Dv na, Dv nb, Dv g2, Dv lda) {
  const double beta = 0.066725, gc = 0.03109069086965489503494086371;
  const Dv n = na + nb;
  if (n.v <= 0.0) return cst(0.0);
  const Dv el = lda / n;
  Dv za = na / n, zb = nb / n;
  if (za.v < DEPS) za = cst(DEPS);
  if (zb.v < DEPS) zb = cst(DEPS);
  const Dv fi = (p23(za) + p23(zb)) * (1.0 / cbrt(2.0));
  const Dv fi2 = fi * fi, fi3 = fi2 * fi;
  const Dv A = cst(beta / gc) / (expd(el * (-1.0) / (fi3 * gc)) - cst(1.0) + cst(DEPS));
  const Dv t2 = g2 / (fi2 * p73(n) + cst(DEPS)) * (PI / (16.0 * cbrt(3.0 * PI * PI)));
  const Dv x = A * t2;
  const Dv rat = (cst(1.0) + x) / (cst(1.0) + x + x * x);
  const Dv H = fi3 * gc * logd(cst(1.0) + t2 * (beta / gc) * rat);
  return n * H;
}

// The spin-polarised terms of mask in enum order: exchange by spin scaling,
// PW92 once (its own term and the base of PBE-c, which needs it); each
// term's F into e[0..3], dF/drho_s into fr[s], dF/dg_ss into fgs[s] and
// dF/dg_ab into fab (all zeroed by the caller).
template <bool GGA>
__device__ __forceinline__ void terms(uint32_t mask, const double (&rho)[2], const double (&gss)[2],
                                      double gab, double (&e)[4], double (&fr)[2],
                                      double (&fgs)[2], double &fab) {
  if (mask & (1u << 1))
    for (int sp = 0; sp < 2; ++sp) {
      double x = 0.0, xr = 0.0;
      xcf::xalpha(2.0 * rho[sp], x, xr);
      e[0] += 0.5 * x;
      fr[sp] += xr;
    }
  Dv lda = cst(0.0);
  if (mask & (1u << 3)) {
    lda = pw92_spin({rho[0], 1.0, 0.0, 0.0}, {rho[1], 0.0, 1.0, 0.0});
    e[1] += lda.v;
    fr[0] += lda.d0;
    fr[1] += lda.d1;
  }
  if (GGA) {
    if (mask & (1u << 10))
      for (int sp = 0; sp < 2; ++sp) {
        double x = 0.0, xr = 0.0, xg = 0.0;
        xcf::pbex(2.0 * rho[sp], 4.0 * gss[sp], x, xr, xg);
        e[2] += 0.5 * x;
        fr[sp] += xr;
        fgs[sp] += 2.0 * xg;
      }
    if (mask & (1u << 18)) {
      const Dv F = pbec_spin({rho[0], 1.0, 0.0, 0.0}, {rho[1], 0.0, 1.0, 0.0},
                             {gss[0] + gss[1] + 2.0 * gab, 0.0, 0.0, 1.0}, lda);
      e[3] += F.v;
      fr[0] += F.d0;
      fr[1] += F.d1;
      fgs[0] += F.d2;
      fgs[1] += F.d2;
      fab += 2.0 * F.d2;
    }
  }
}
}  // namespace xcs

// item = 256-point block of an open-shell unit (the alpha spin's unit, which
// carries the beta spin's partials, coefficients and state in part2, coef2,
// state2).  Per point: rho_s, grad rho_s from the partials, f, dF/drho_s,
// dF/dg_aa, dF/dg_bb, dF/dg_ab, then a_s = w dF/drho_s / 2 and
// b_s = w (2 dF/dg_ss grad rho_s + dF/dg_ab grad rho_s')
// (modules/xc_hamiltonian.f90:1428-1445).
template <bool GGA>
__global__ __launch_bounds__(kThreads) void k_functional_spin(const UnitD *__restrict__ units, int nu,
                                                              const long *__restrict__ base,
                                                              const int *__restrict__ order,
                                                              long total) {
  __shared__ double se[kThreads], sq[kThreads];
  const int tid = threadIdx.x;
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const long it = item_at(order, L, total);
    const int u = find_unit(base, nu, it);
    const UnitD &U = units[u];
    const long blk = it - base[u];
    const long p = blk * kThreads + tid, Mc = U.Mc, npad = U.npad;
    const double *part[2] = {U.part, U.part2};
    double *state[2] = {U.state, U.state2};
    double *coef[2] = {U.coef, U.coef2};
    double rho[2], g[2][3];
    const int tile = (int)(p / kScreenTile);
    const int tf = U.CH * kRun, nch = (U.tf_off[tile + 1] - U.tf_off[tile] + tf - 1) / tf;
    for (int sp = 0; sp < 2; ++sp) {
      double r = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
      for (int t = 0; t < nch; ++t) {
        r += part[sp][((long)t * 4 + 0) * Mc + p];
        if (GGA) {
          g0 += part[sp][((long)t * 4 + 1) * Mc + p];
          g1 += part[sp][((long)t * 4 + 2) * Mc + p];
          g2 += part[sp][((long)t * 4 + 3) * Mc + p];
        }
      }
      if (U.incremental) {
        r = state[sp][S_RHO * npad + p] + r;
        if (GGA) {
          g0 = state[sp][S_GX * npad + p] + g0;
          g1 = state[sp][S_GY * npad + p] + g1;
          g2 = state[sp][S_GZ * npad + p] + g2;
        }
      }
      rho[sp] = r;
      g[sp][0] = g0;
      g[sp][1] = g1;
      g[sp][2] = g2;
    }
    const double gaa = g[0][0] * g[0][0] + g[0][1] * g[0][1] + g[0][2] * g[0][2];
    const double gbb = g[1][0] * g[1][0] + g[1][1] * g[1][1] + g[1][2] * g[1][2];
    const double gab = g[0][0] * g[1][0] + g[0][1] * g[1][1] + g[0][2] * g[1][2];
    const double gss[2] = {gaa, gbb};
    double e[4] = {0.0, 0.0, 0.0, 0.0}, fr[2] = {0.0, 0.0}, fgs[2] = {0.0, 0.0}, fab = 0.0;
    xcs::terms<GGA>(U.mask, rho, gss, gab, e, fr, fgs, fab);
    const double f = xcf::sum4(e);
    const double ww = U.gw[p];
    for (int sp = 0; sp < 2; ++sp) {
      const double a = 0.5 * fr[sp] * ww;
      double b[3];
      for (int c = 0; c < 3; ++c) b[c] = ww * (2.0 * fgs[sp] * g[sp][c] + fab * g[1 - sp][c]);
      double *cf = coef[sp], *st = state[sp];
      if (U.incremental) {
        cf[0 * Mc + p] = a - st[S_A * npad + p];
        if (GGA) {
          cf[1 * Mc + p] = b[0] - st[S_BX * npad + p];
          cf[2 * Mc + p] = b[1] - st[S_BY * npad + p];
          cf[3 * Mc + p] = b[2] - st[S_BZ * npad + p];
        }
      } else {
        cf[0 * Mc + p] = a;
        if (GGA) {
          cf[1 * Mc + p] = b[0];
          cf[2 * Mc + p] = b[1];
          cf[3 * Mc + p] = b[2];
        }
      }
      st[S_RHO * npad + p] = rho[sp];
      st[S_A * npad + p] = a;
      if (GGA) {
        st[S_GX * npad + p] = g[sp][0];
        st[S_GY * npad + p] = g[sp][1];
        st[S_GZ * npad + p] = g[sp][2];
        st[S_BX * npad + p] = b[0];
        st[S_BY * npad + p] = b[1];
        st[S_BZ * npad + p] = b[2];
      }
    }
    se[tid] = f * ww;
    sq[tid] = (rho[0] + rho[1]) * ww;
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
      if (tid < st) {
        se[tid] += se[tid + st];
        sq[tid] += sq[tid + st];
      }
      __syncthreads();
    }
    if (tid == 0) {
      U.epart[2 * blk + 0] = se[0];
      U.epart[2 * blk + 1] = sq[0];
    }
    __syncthreads();
  }
}

// Each term's F at a point from rho and grad rho (per spin when SPIN)
template <bool GGA, bool SPIN>
__device__ __forceinline__ void point_terms(uint32_t mask, const double (&rho)[2],
                                            const double (&g)[2][3], double (&e)[4]) {
  if (!SPIN) {
    const double gam = g[0][0] * g[0][0] + g[0][1] * g[0][1] + g[0][2] * g[0][2];
    double fr = 0.0, fg = 0.0;
    xcf::terms<GGA>(mask, rho[0], gam, e, fr, fg);
  } else {
    const double gss[2] = {g[0][0] * g[0][0] + g[0][1] * g[0][1] + g[0][2] * g[0][2],
                           g[1][0] * g[1][0] + g[1][1] * g[1][1] + g[1][2] * g[1][2]};
    const double gab = g[0][0] * g[1][0] + g[0][1] * g[1][1] + g[0][2] * g[1][2];
    double fr[2] = {0.0, 0.0}, fgs[2] = {0.0, 0.0}, fab = 0.0;
    xcs::terms<GGA>(mask, rho, gss, gab, e, fr, fgs, fab);
  }
}

// ------------------------------------------------------- energy terms
// item = 256-point block of a unit (SPIN: an open-shell molecule's alpha unit):
// the XC energy of each functional term and the charge on the block, without
// potential coefficients and without touching the per-point state
// (post_scf_calc_xc_energy, modules/post_scf_module.f90:2204-2420).  Block
// sums go to eterm[blk][5]: XAlpha, PWLDAc, PBEx, PBEc, charge.
template <bool GGA, bool SPIN>
__global__ __launch_bounds__(kThreads) void k_energy(const UnitD *__restrict__ units, int nu,
                                                     const long *__restrict__ base,
                                                     const int *__restrict__ order, long total) {
  __shared__ double sh[5][kThreads];
  const int tid = threadIdx.x;
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const long it = item_at(order, L, total);
    const int u = find_unit(base, nu, it);
    const UnitD &U = units[u];
    const long blk = it - base[u];
    const long p = blk * kThreads + tid, Mc = U.Mc;
    const int tile = (int)(p / kScreenTile);
    const int tf = U.CH * kRun, nch = (U.tf_off[tile + 1] - U.tf_off[tile] + tf - 1) / tf;
    double rho[2] = {0.0, 0.0}, g[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int sp = 0; sp < (SPIN ? 2 : 1); ++sp) {
      const double *part = sp ? U.part2 : U.part;
      for (int t = 0; t < nch; ++t) {
        rho[sp] += part[((long)t * 4 + 0) * Mc + p];
        if (GGA)
          for (int c = 0; c < 3; ++c) g[sp][c] += part[((long)t * 4 + 1 + c) * Mc + p];
      }
    }
    const double ww = U.gw[p];
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    point_terms<GGA, SPIN>(U.mask, rho, g, e);
    for (int k = 0; k < 4; ++k) sh[k][tid] = e[k] * ww;
    sh[4][tid] = (rho[0] + rho[1]) * ww;
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
      if (tid < st)
        for (int k = 0; k < 5; ++k) sh[k][tid] += sh[k][tid + st];
      __syncthreads();
    }
    if (tid < 5) U.eterm[5 * blk + tid] = sh[tid][0];
    __syncthreads();
  }
}

// ------------------------------------------------ energy density per point
// f at each point of a chunk from the per-point state of the last build
// (rho, grad rho per spin): the weight-derivative term of the nuclear
// gradient multiplies it by dw/dR.
template <bool GGA, bool SPIN>
__global__ __launch_bounds__(kThreads) void k_point_f(long Mc, long npad, uint32_t mask,
                                                      const double *__restrict__ st,
                                                      const double *__restrict__ st2,
                                                      double *__restrict__ fp) {
  for (long p = (long)blockIdx.x * kThreads + threadIdx.x; p < Mc; p += (long)gridDim.x * kThreads) {
    double rho[2] = {0.0, 0.0}, g[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int sp = 0; sp < (SPIN ? 2 : 1); ++sp) {
      const double *S = sp ? st2 : st;
      rho[sp] = S[S_RHO * npad + p];
      if (GGA)
        for (int c = 0; c < 3; ++c) g[sp][c] = S[(S_GX + c) * npad + p];
    }
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    point_terms<GGA, SPIN>(mask, rho, g, e);
    const double f = xcf::sum4(e);
    fp[p] = f;
  }
}

// -------------------------------------------------------------------- W
// item = (1024-point block, 16-function group), 4 consecutive points of one
// 32-point group per thread: W = a Phi + b . grad Phi.  SPLIT (split-BF16
// builds, T = float): W stored as bf16 hi / lo planes.  A run of 8 functions
// cut off on the group is 0 without reading Phi, and is not written at all
// where the potential contraction never reads it; inside a live run, a
// function outside the tile's function list (its shell cut off on the whole
// tile) is 0 without reading Phi either.
template <typename T, bool GRAD, bool SPLIT, bool TWO>
__global__ __launch_bounds__(kThreads) void k_wmat(const UnitD *__restrict__ units, int nu,
                                                   const long *__restrict__ base,
                                                   const int *__restrict__ order, long total) {
  typedef T t4 __attribute__((ext_vector_type(4)));
  typedef uint16_t h4 __attribute__((ext_vector_type(4)));
  constexpr int NS = TWO ? 2 : 1;  // spins written from one read of the basis values
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const long it = item_at(order, L, total);
    const int u = find_unit(base, nu, it);
    const UnitD &U = units[u];
    const long local = it - base[u];
    const int ng = U.NP / 16;
    const long p = (local / ng) * (4 * kThreads) + 4 * threadIdx.x;
    const long f0 = (local % ng) * 16;
    const long Mc = U.Mc, cs = (long)U.NP * Mc;
    const T *phi = (const T *)U.phi;
    void *Wout[2] = {U.W, TWO ? U.W2 : U.W};
    const double *cf[2] = {U.coef, TWO ? U.coef2 : U.coef};
    auto put = [&](int sp, long q, const double (&v)[4]) {
      if constexpr (SPLIT) {
        uint16_t *Wh = (uint16_t *)Wout[sp], *Wl = Wh + cs;
        h4 hi, lo;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          __bf16 h, l;
          split_bf16((float)v[j], h, l);
          hi[j] = __builtin_bit_cast(uint16_t, h);
          lo[j] = __builtin_bit_cast(uint16_t, l);
        }
        gst((h4 *)(Wh + q), hi);
        gst((h4 *)(Wl + q), lo);
      } else {
        t4 w;
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] = (T)v[j];
        gst((t4 *)((T *)Wout[sp] + q), w);
      }
    };
    const uint32_t *sg = U.sig8 + (p / kGroup) * U.nw8, *wn = U.wbits + (p / kGroup) * U.nw8;
    const uint32_t *fb = U.tfb + (p / kScreenTile) * U.nwf;
    // split-BF16 W is read only by the split-BF16 potential, which skips the
    // functions outside a tile's list (when a split spans at most 32 tiles):
    // no zeros are written for them
    const bool skip = SPLIT && Mc / U.nsplit <= 32 * kScreenTile;
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // the item's two runs of 8 functions
      const long r = f0 / kRun + h, fr = r * kRun;
      const uint32_t bit = 1u << (r & 31);
      const uint32_t fm = gld(fb + (fr >> 5)) >> (fr & 31);  // the run's functions in the tile's list
      const uint32_t wm = skip ? fm : 0xffu;                 // the run's functions written
      if (!(gld(sg + (r >> 5)) & bit)) {
        if (gld(wn + (r >> 5)) & bit) {
          const double z[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
          for (int k = 0; k < kRun; ++k)
            if ((wm >> k) & 1u)
#pragma unroll
              for (int sp = 0; sp < NS; ++sp) put(sp, (fr + k) * Mc + p, z);
        }
        continue;
      }
      double a[NS][4], b0[NS][4] = {}, b1[NS][4] = {}, b2[NS][4] = {};
#pragma unroll
      for (int sp = 0; sp < NS; ++sp)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          a[sp][j] = gld(cf[sp] + p + j);
          if (GRAD) {
            b0[sp][j] = gld(cf[sp] + Mc + p + j);
            b1[sp][j] = gld(cf[sp] + 2 * Mc + p + j);
            b2[sp][j] = gld(cf[sp] + 3 * Mc + p + j);
          }
        }
#pragma unroll 4
      for (int k = 0; k < kRun; ++k) {
        if (!((wm >> k) & 1u)) continue;
        const long q = (fr + k) * Mc + p;
        t4 v0 = {}, v1 = {}, v2 = {}, v3 = {};
        if ((fm >> k) & 1u) {
          v0 = gld((const t4 *)(phi + q));
          if (GRAD) {
            v1 = gld((const t4 *)(phi + cs + q));
            v2 = gld((const t4 *)(phi + 2 * cs + q));
            v3 = gld((const t4 *)(phi + 3 * cs + q));
          }
        }
#pragma unroll
        for (int sp = 0; sp < NS; ++sp) {
          double v[4];
#pragma unroll
          for (int j = 0; j < 4; ++j)
            v[j] = GRAD ? a[sp][j] * (double)v0[j] + b0[sp][j] * (double)v1[j] +
                              b1[sp][j] * (double)v2[j] + b2[sp][j] * (double)v3[j]
                        : a[sp][j] * (double)v0[j];
          put(sp, q, v);
        }
      }
    }
  }
}

// --------------------------------------------------------------- potential
// item = (point split, function tile pair) of H[split] (+)= W^T Phi
// Variants 0 and 1 keep one LDS stage set under 64 KB: built for two waves
// per SIMD, two workgroups share a CU and hide each other's waits.
template <typename T, int BT, int VAR>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(VAR == 2 ? 1 : 2))) void k_potential(const StageList S) {
  constexpr int BK = Var<VAR>::BK, ST = Var<VAR>::ST, TT = BT / 32;
  constexpr int R = kGroup / BK;  // K steps per 32-point group
  __shared__ __attribute__((aligned(16))) T As[lds_elems<T, BT, BK, ST, true>()];
  __shared__ __attribute__((aligned(16))) T Bs[lds_elems<T, BT, BK, ST, true>()];
  __shared__ uint16_t ul[2 * BT];  // the functions of the row chunk, then of the column chunk
  __shared__ uint32_t cm[BT];      // per column: the split's tiles whose list holds its function
  __shared__ uint16_t gl[kMaxSplitGroups];
  __shared__ int wtot[4];
  __shared__ long slot;
  for (long L = walk_next(S, -1, &slot); L < S.total;) {
    const uint64_t t0 = S.tlog ? wall_clock64() : 0;
    const long it = walk_item(S, L);
    const int u = find_unit(S.base, S.nu, it);
    const UnitD &U = S.units[u];
    const uint32_t e = U.pitems[it - S.base[u]];
    const int sl = (int)(e >> 20), ci = (int)((e >> 10) & 1023), cj = (int)(e & 1023);
    const long Mc = U.Mc;
    const int so = U.su_off[sl], nu = U.su_off[sl + 1] - so;
    const int nch = (nu + BT - 1) / BT, G = (int)(Mc / U.nsplit) / kGroup;
    const long kb = (long)sl * G * kGroup;  // first point of the split
    const int nri = min(BT, nu - ci * BT), nrj = min(BT, nu - cj * BT);
    __syncthreads();  // the previous item's readers of ul / gl are done
    for (int i = threadIdx.x; i < 2 * BT; i += kThreads)
      if (i < BT ? i < nri : i - BT < nrj) ul[i] = gld(U.su + so + (i < BT ? ci * BT + i : cj * BT + i - BT));
    for (int i = threadIdx.x; i < nrj; i += kThreads) cm[i] = gld(U.sumask + so + cj * BT + i);
    __syncthreads();
    const int ng = potential_groups(U.gcf + U.gcf_off[sl], G, nch, ci, cj, gl, wtot);
    const uint16_t *ri = ul, *rj = ul + BT;
    // A(m, k) = W[f_i(m)][kb + p(k)], B(k, n) = Phi[f_j(n)][kb + p(k)]: rows from
    // the split's functions, K over the groups where both chunks can be
    // nonzero; B is 0 on the tiles whose list lacks the column's function
    // (left unwritten by the basis stage)
    auto kpt = [&](int kt, int kk) -> long { return (long)gl[kt / R] * kGroup + (kt % R) * BK + kk; };
    auto aaddr = [&](int m, int kt, int kk) -> long {
      return m < nri ? (long)ri[min(m, nri - 1)] * Mc + kb + kpt(kt, kk) : -1;
    };
    auto baddr = [&](int n, int kt, int kk) -> long {
      const long k = kpt(kt, kk);
      return (n < nrj && ((cm[min(n, nrj - 1)] >> ((int)(k >> 7) & 31)) & 1u))
                 ? (long)rj[min(n, nrj - 1)] * Mc + kb + k : -1;
    };
    const T *W = (const T *)U.W, *phi = (const T *)U.phi;
    typename MF<T>::acc acc[TT][TT];
    gemm_tile<T, BT, BT, BK, ST, true, true>(W, phi, ng * R, aaddr, baddr, As, Bs, acc);
    const int ahead = walk_ahead(S);
    potential_epilogue<T, typename MF<T>::acc, BT, BT>(acc, U.Hpart + (long)sl * U.NP * U.NP,
                                                       U.NP, ri, nri, rj, nrj);
    item_log(S, it, t0);
    L = walk_take(S, L, ahead, &slot);
  }
}

// -------------------------------------------------------------------- pack
// Fold, item = 32 x 32 tile (rb >= cb) of one molecule's internal matrix:
// S(r, c) = sum_split H(r, c) + sum_split H(c, r) for both the tile and its
// mirror, read along rows; the mirror's sums change places through LDS.  S
// replaces split 0's partial on the tile (only this item reads that tile and
// its mirror, and it writes after every read).
__global__ __launch_bounds__(kThreads) void k_fold(const PackD *__restrict__ mols, int nm,
                                                   const long *__restrict__ base, long total) {
  __shared__ double t[32][33];
  const int tid = threadIdx.x;
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const int m = find_unit(base, nm, L);
    const PackD &M = mols[m];
    long e = L - base[m], rb = (long)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
    while (rb * (rb + 1) / 2 > e) --rb;
    while ((rb + 1) * (rb + 2) / 2 <= e) ++rb;
    const long cb = e - rb * (rb + 1) / 2, NP = M.NP, np2 = NP * NP;
    double sa[4] = {}, sb[4] = {};
    for (int s = 0; s < M.nsplit; ++s) {
      const double *H = M.Hpart + (long)s * np2;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = tid + kThreads * i, r = x >> 5, c = x & 31;
        sa[i] += H[(rb * 32 + r) * NP + cb * 32 + c];
        sb[i] += H[(cb * 32 + r) * NP + rb * 32 + c];
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = tid + kThreads * i;
      t[x >> 5][x & 31] = sb[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = tid + kThreads * i, r = x >> 5, c = x & 31;
      M.Hpart[(rb * 32 + r) * NP + cb * 32 + c] = sa[i] + t[c][r];
    }
    __syncthreads();
  }
}

// item = 256 packed elements of one molecule (caller's function order):
// V = base + S at the functions' internal rows (S folded above, symmetric,
// read from its lower tiles); then the molecule's state for the next
// incremental build: V_state = V.
__global__ __launch_bounds__(kThreads) void k_pack(const PackD *__restrict__ mols, int nm,
                                                   const long *__restrict__ base, long total) {
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const int m = find_unit(base, nm, L);
    const PackD &M = mols[m];
    const long idx = (L - base[m]) * kThreads + threadIdx.x;
    const long n = M.n, NP = M.NP, npk = n * (n + 1) / 2;
    if (idx >= npk) continue;
    long j = (long)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while (j * (j + 1) / 2 > idx) --j;
    while ((j + 1) * (j + 2) / 2 <= idx) ++j;
    const long k = idx - j * (j + 1) / 2;
    const long pj = M.fpos[j], pk = M.fpos[k];  // internal rows of caller functions j, k
    const long r = pj > pk ? pj : pk, c = pj > pk ? pk : pj;
    const double v = M.Hpart[r * NP + c];
    const double out = M.incremental ? M.Vstate[idx] + v : v;
    M.Vout[idx] = out;
    M.Vstate[idx] = out;
  }
}

// Density operand of one molecule: D = P - P_ref (incremental) or P, in the
// contraction precision; afterwards P_ref = P.  item = 256 elements.
__global__ __launch_bounds__(kThreads) void k_prep(const PackD *__restrict__ mols, int nm,
                                                   const long *__restrict__ base, long total,
                                                   int mode, int finish) {
  for (long L = blockIdx.x; L < total; L += gridDim.x) {
    const int m = find_unit(base, nm, L);
    const PackD &M = mols[m];
    const long i = (L - base[m]) * kThreads + threadIdx.x;
    if (i >= (long)M.NP * M.NP) continue;
    if (finish) {
      M.Pref[i] = M.P[i];
    } else {
      const double d = M.incremental ? M.P[i] - M.Pref[i] : M.P[i];
      if (mode == 2) {  // split planes of D^T: element (r, c) goes to (c, r)
        const long np2 = M.NP * M.NP, r = i / M.NP, c = i % M.NP;
        __bf16 hi, lo;
        split_bf16((float)d, hi, lo);
        ((__bf16 *)M.D)[c * M.NP + r] = hi;
        ((__bf16 *)M.D)[np2 + c * M.NP + r] = lo;
      } else if (mode == 1) {
        ((float *)M.D)[i] = (float)d;
      } else {
        ((double *)M.D)[i] = d;
      }
    }
  }
}

// ---------------------------------------------------------------- launches
static unsigned persistent_grid(long total, int per_cu) {
  const long cap = 256L * per_cu;
  return (unsigned)(total < cap ? (total > 0 ? total : 1) : cap);
}

// grid of a contraction launch: with queues, whole workgroups per queue
unsigned contraction_grid(const StageList &L, int per_cu) {
  long g = persistent_grid(L.total, per_cu);
  if (L.ctr) g = std::max<long>(kQueues, (g + kQueues - 1) / kQueues * kQueues);
  return (unsigned)g;
}

template <typename T, int LMAX>
static void basis_l(hipStream_t s, bool grad, const StageList &L) {
  dim3 g(persistent_grid(L.total, 8));
  if (grad)
    hipLaunchKernelGGL((k_basis<true, T, LMAX>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
  else
    hipLaunchKernelGGL((k_basis<false, T, LMAX>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
}

template <int LMAX>
static void wbasis_l(hipStream_t s, bool grad, bool two, const StageList &L) {
  dim3 g(persistent_grid(L.total, 8));
  if (grad && two)
    hipLaunchKernelGGL((k_wbasis<true, true, LMAX>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
  else if (grad)
    hipLaunchKernelGGL((k_wbasis<true, false, LMAX>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
  else if (two)
    hipLaunchKernelGGL((k_wbasis<false, true, LMAX>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
  else
    hipLaunchKernelGGL((k_wbasis<false, false, LMAX>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
}

void launch_wbasis(hipStream_t s, bool grad, bool two, const StageList &L) {
  if (!L.total) return;
  switch (L.lmax) {
    case 0: case 1: wbasis_l<1>(s, grad, two, L); break;
    case 2: wbasis_l<2>(s, grad, two, L); break;
    case 3: wbasis_l<3>(s, grad, two, L); break;
    default: wbasis_l<4>(s, grad, two, L); break;
  }
}

template <typename T>
void launch_basis(hipStream_t s, bool grad, const StageList &L) {
  if (!L.total) return;
  switch (L.lmax) {
    case 0: case 1: basis_l<T, 1>(s, grad, L); break;
    case 2: basis_l<T, 2>(s, grad, L); break;
    case 3: basis_l<T, 3>(s, grad, L); break;
    default: basis_l<T, 4>(s, grad, L); break;
  }
}

// ---------------------------------------------------- gradient products
// The products of the nuclear gradient, X = Phi P and Y_c = (d/dc Phi) P
// (P the density matrix of the last build), over the density contraction's
// items: 128-point tile x pair of chunks of the tile's functions, K over the
// tile's runs, the same GEMM with P for D.  Written to XY[plane][f][p] for
// the tile's functions f only: the gradient's per-point terms take X_f, Y_f
// where phi_f is not cut off.  nplanes: 1 (X, LDA) or 4.  Units u with
// u % nspin == spin only (a plan lists a molecule's spins as consecutive
// units; their items are the same tiles).
template <int BN, int VAR>
__global__ __launch_bounds__(kThreads) void k_xy(const StageList S, const double *__restrict__ P,
                                                 double *__restrict__ XY, int nplanes, int nspin,
                                                 int spin) {
  typedef double T;
  constexpr int BM = kDBM, BK = Var<VAR>::BK, ST = Var<VAR>::ST, TM = BM / 32, TN = BN / 32;
  __shared__ T As[lds_elems<T, BM, BK, ST, false>()];
  __shared__ __attribute__((aligned(16))) T Bs[lds_elems<T, BN, BK, ST, true>()];
  __shared__ uint16_t kl[kMaxRuns];
  __shared__ uint16_t cl[2 * BN];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, wm = wave >> 1, wn = wave & 1;
  for (long L = blockIdx.x; L < S.total; L += gridDim.x) {
    const long it = item_at(S.order, L, S.total);
    const int u = find_unit(S.base, S.nu, it);
    if (u % nspin != spin) continue;  // uniform over the workgroup
    const UnitD &U = S.units[u];
    const uint32_t e = U.ditems[it - S.base[u]];
    const int tile = (int)(e >> 8), pair = (int)(e & 255);
    const long m0 = (long)tile * BM, Mc = U.Mc, NP = U.NP, cs = NP * Mc;
    const int off = U.tl_off[tile], c8 = U.tl_off[tile + 1] - off;
    const int foff = U.tf_off[tile] + 2 * pair * BN, nf = U.tf_off[tile + 1] - foff;
    __syncthreads();  // the previous item's readers of kl and cl are done
    for (int i = threadIdx.x; i < min(2 * BN, nf); i += kThreads) cl[i] = gld(U.tfl + foff + i);
    lds_list(U.tl + off, c8, kl);
    const int ksteps = (c8 * kRun + BK - 1) / BK;
    const T *phi = (const T *)U.phi;
    for (int h = 0; h < 2 && h * BN < nf; ++h) {
      const uint16_t *cfn = cl + h * BN;
      const int ncol = min(BN, nf - h * BN);
      auto baddr = [&](int n, int kt, int kk) -> long {
        const long f = run_fn(kl, c8, kt * BK + kk);
        return (f < 0 || n >= ncol) ? -1 : (long)cfn[min(n, ncol - 1)] * NP + f;
      };
      for (int c = 0; c < nplanes; ++c) {
        auto aaddr = [&](int m, int kt, int kk) -> long {
          const long f = run_fn(kl, c8, kt * BK + kk);
          return f < 0 ? -1 : c * cs + f * Mc + m;
        };
        typename MF<T>::acc acc[TM][TN];
        gemm_tile<T, BM, BN, BK, ST, false, true>(phi + m0, P, ksteps, aaddr, baddr, As, Bs, acc);
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int n = wn * (BN / 2) + j * 16 + mf_col(lane);
          if (n >= ncol) continue;
          double *xy = XY + c * cs + (long)cfn[n] * Mc + m0;
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) gst(xy + wm * (BM / 2) + i * 16 + MF<T>::row(lane, r), acc[i][j][r]);
        }
      }
    }
  }
}

template <int BN>
static void xy_bn(hipStream_t s, int var, const StageList &L, const double *P, double *XY, int np,
                  int ns, int sp) {
  dim3 g(persistent_grid(L.total, 2));
  switch (var) {
    case 1: hipLaunchKernelGGL((k_xy<BN, 1>), g, dim3(kThreads), 0, s, L, P, XY, np, ns, sp); break;
    case 2: hipLaunchKernelGGL((k_xy<BN, 2>), g, dim3(kThreads), 0, s, L, P, XY, np, ns, sp); break;
    default: hipLaunchKernelGGL((k_xy<BN, 0>), g, dim3(kThreads), 0, s, L, P, XY, np, ns, sp); break;
  }
}

void launch_xy(hipStream_t s, int tf, int var, const StageList &L, const double *P, double *XY,
               int nplanes, int nspin, int spin) {
  if (!L.total) return;
  StageList W = L;  // a plain static walk (no queues, no schedule order, no log)
  W.ctr = nullptr;
  W.qbeg = nullptr;
  W.order = nullptr;
  W.tlog = nullptr;
  switch (tf) {
    case 64: xy_bn<64>(s, var, W, P, XY, nplanes, nspin, spin); break;
    case 96: xy_bn<96>(s, var, W, P, XY, nplanes, nspin, spin); break;
    default: xy_bn<128>(s, var, W, P, XY, nplanes, nspin, spin); break;
  }
}

template <typename T, int BN, int VAR>
static void density_bn(hipStream_t s, bool grad, const StageList &L) {
  dim3 g(contraction_grid(L, 2));
  if (grad)
    hipLaunchKernelGGL((k_density<T, BN, VAR, true>), g, dim3(kThreads), 0, s, L);
  else
    hipLaunchKernelGGL((k_density<T, BN, VAR, false>), g, dim3(kThreads), 0, s, L);
}

template <typename T, int VAR>
static void density_var(hipStream_t s, bool grad, int tf, const StageList &L) {
  switch (tf) {
    case 64: density_bn<T, 64, VAR>(s, grad, L); break;
    case 96: density_bn<T, 96, VAR>(s, grad, L); break;
    default: density_bn<T, 128, VAR>(s, grad, L); break;
  }
}

template <typename T>
void launch_density(hipStream_t s, bool grad, int tf, int var, const StageList &L) {
  if (!L.total) return;
  switch (var) {
    case 1: density_var<T, 1>(s, grad, tf, L); break;
    case 2: density_var<T, 2>(s, grad, tf, L); break;
    default: density_var<T, 0>(s, grad, tf, L); break;
  }
}

void launch_functional(hipStream_t s, bool gga, const StageList &L) {
  if (!L.total) return;
  dim3 g(persistent_grid(L.total, 8));
  if (gga)
    hipLaunchKernelGGL((k_functional<true>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
  else
    hipLaunchKernelGGL((k_functional<false>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
}

void launch_point_f(hipStream_t s, bool gga, bool spin, uint32_t mask, long Mc, long npad,
                    const double *st, const double *st2, double *fp) {
  const unsigned g = (unsigned)std::min<long>((Mc + kThreads - 1) / kThreads, 256L * 8);
  auto k = spin ? (gga ? k_point_f<true, true> : k_point_f<false, true>)
                : (gga ? k_point_f<true, false> : k_point_f<false, false>);
  hipLaunchKernelGGL(k, dim3(g), dim3(kThreads), 0, s, Mc, npad, mask, st, st2, fp);
}

void launch_energy(hipStream_t s, bool gga, bool spin, const StageList &L) {
  if (!L.total) return;
  dim3 g(persistent_grid(L.total, 8));
  auto k = spin ? (gga ? k_energy<true, true> : k_energy<false, true>)
                : (gga ? k_energy<true, false> : k_energy<false, false>);
  hipLaunchKernelGGL(k, g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
}

void launch_functional_spin(hipStream_t s, bool gga, const StageList &L) {
  if (!L.total) return;
  dim3 g(persistent_grid(L.total, 8));
  if (gga)
    hipLaunchKernelGGL((k_functional_spin<true>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
  else
    hipLaunchKernelGGL((k_functional_spin<false>), g, dim3(kThreads), 0, s, L.units, L.nu, L.base, L.order, L.total);
}

template <typename T>
void launch_wmat(hipStream_t s, bool grad, bool split, bool two, const StageList &L) {
  if (!L.total) return;
  dim3 g(persistent_grid(L.total, 8));
#define PGXC_WMAT(SPLIT)                                                                           \
  {                                                                                                \
    if (two) {                                                                                     \
      if (grad)                                                                                    \
        hipLaunchKernelGGL((k_wmat<T, true, SPLIT, true>), g, dim3(kThreads), 0, s, L.units, L.nu, \
                           L.base, L.order, L.total);                                              \
      else                                                                                         \
        hipLaunchKernelGGL((k_wmat<T, false, SPLIT, true>), g, dim3(kThreads), 0, s, L.units,     \
                           L.nu, L.base, L.order, L.total);                                        \
    } else if (grad) {                                                                             \
      hipLaunchKernelGGL((k_wmat<T, true, SPLIT, false>), g, dim3(kThreads), 0, s, L.units, L.nu,  \
                         L.base, L.order, L.total);                                                \
    } else {                                                                                       \
      hipLaunchKernelGGL((k_wmat<T, false, SPLIT, false>), g, dim3(kThreads), 0, s, L.units, L.nu, \
                         L.base, L.order, L.total);                                                \
    }                                                                                              \
  }
  if constexpr (std::is_same<T, float>::value) {
    if (split) {
      PGXC_WMAT(true)
      return;
    }
  }
  PGXC_WMAT(false)
#undef PGXC_WMAT
}

template <typename T, int VAR>
static void potential_var(hipStream_t s, int tf, const StageList &L) {
  dim3 g(contraction_grid(L, 2));
  switch (tf) {
    case 64: hipLaunchKernelGGL((k_potential<T, 64, VAR>), g, dim3(kThreads), 0, s, L); break;
    case 96: hipLaunchKernelGGL((k_potential<T, 96, VAR>), g, dim3(kThreads), 0, s, L); break;
    default: hipLaunchKernelGGL((k_potential<T, 128, VAR>), g, dim3(kThreads), 0, s, L); break;
  }
}

template <typename T>
void launch_potential(hipStream_t s, int tf, int var, const StageList &L) {
  if (!L.total) return;
  switch (var) {
    case 1: potential_var<T, 1>(s, tf, L); break;
    case 2: potential_var<T, 2>(s, tf, L); break;
    default: potential_var<T, 0>(s, tf, L); break;
  }
}

__global__ __launch_bounds__(kThreads) void k_zero_ints(int *p, long n) {
  for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads)
    p[i] = 0;
}

void launch_zero_ints(hipStream_t s, int *p, long n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_ints, dim3(persistent_grid((n + kThreads - 1) / kThreads, 1)),
                     dim3(kThreads), 0, s, p, n);
}

void launch_fold(hipStream_t s, const PackD *mols, int nm, const long *base, long total) {
  if (!total) return;
  hipLaunchKernelGGL(k_fold, dim3(persistent_grid(total, 8)), dim3(kThreads), 0, s, mols, nm, base, total);
}

void launch_pack(hipStream_t s, const PackD *mols, int nm, const long *base, long total) {
  if (!total) return;
  hipLaunchKernelGGL(k_pack, dim3(persistent_grid(total, 8)), dim3(kThreads), 0, s, mols, nm, base, total);
}

void launch_prep(hipStream_t s, const PackD *mols, int nm, const long *base, long total,
                 int mode, bool finish) {
  if (!total) return;
  hipLaunchKernelGGL(k_prep, dim3(persistent_grid(total, 8)), dim3(kThreads), 0, s, mols, nm, base,
                     total, mode, (int)finish);
}

#define PGXC_INST(T)                                                                  \
  template void launch_basis<T>(hipStream_t, bool, const StageList &);               \
  template void launch_density<T>(hipStream_t, bool, int, int, const StageList &);   \
  template void launch_wmat<T>(hipStream_t, bool, bool, bool, const StageList &);    \
  template void launch_potential<T>(hipStream_t, int, int, const StageList &);
PGXC_INST(double)
PGXC_INST(float)
#undef PGXC_INST

// ---------------------------------------------------------------- screening
// item = one 128-point tile (four 32-point groups).  The smallest r^2 of each
// group to each atom (FP64, reduced over the group's 32 lanes), then per
// shell and group alpha_min r^2_min < cut-off marks the shell's runs of 8
// functions.  Outputs per group: the run flags (nw8 words, sig8); per tile:
// n_b, the functions of the shells marked on any of its groups, and those
// shells (smark, one bit per shell).
__global__ __launch_bounds__(kThreads) void k_screen(long ntiles, const double *__restrict__ gx,
                                                     const double *__restrict__ gy,
                                                     const double *__restrict__ gz, int n_atoms,
                                                     const AtomD *__restrict__ atoms, int n_shells,
                                                     const ScreenShell *__restrict__ shells,
                                                     int n8, int *nb, uint32_t *sig8,
                                                     uint32_t *smark) {
  constexpr int AC = 512;  // atoms per pass
  constexpr int W8 = kMaxRuns / 32;
  constexpr int WS = kMaxShells / 32;
  __shared__ double r2g[4][AC];
  __shared__ unsigned bits[4][W8];
  __shared__ unsigned sb[WS];
  __shared__ int nbs;
  const int tid = threadIdx.x;
  const int nw8 = (n8 + 31) / 32, nws = (n_shells + 31) / 32;
  for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    for (int i = tid; i < 4 * W8; i += kThreads) (&bits[0][0])[i] = 0u;
    for (int i = tid; i < WS; i += kThreads) sb[i] = 0u;
    if (tid == 0) nbs = 0;
    const long p = t * kScreenTile + (tid & 127);
    const double x = gx[p], y = gy[p], z = gz[p];
    for (int a0 = 0; a0 < n_atoms; a0 += AC) {
      const int na = min(AC, n_atoms - a0);
      __syncthreads();
      // threads 0..127 take even atoms, 128..255 odd ones (whole waves each)
      for (int a = tid >> 7; a < na; a += 2) {
        const AtomD A = atoms[a0 + a];
        const double dx = x - A.x, dy = y - A.y, dz = z - A.z;
        double r2 = dx * dx + dy * dy + dz * dz;
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) r2 = fmin(r2, __shfl_xor(r2, off));
        if ((tid & 31) == 0) r2g[(tid & 127) >> 5][a] = r2;
      }
      __syncthreads();
      for (int si = tid; si < n_shells; si += kThreads) {
        const ScreenShell S = shells[si];
        if (S.atom < a0 || S.atom >= a0 + na) continue;
        bool any = false;
        const int b0 = S.func_off / kRun, b1 = (S.func_off + S.nfunc - 1) / kRun;
        for (int q = 0; q < 4; ++q) {
          if (S.amin * r2g[q][S.atom - a0] < kCutoff) {
            any = true;
            for (int b = b0; b <= b1; ++b) atomicOr(&bits[q][b >> 5], 1u << (b & 31));
          }
        }
        if (any) {
          atomicAdd(&nbs, S.nfunc);
          atomicOr(&sb[si >> 5], 1u << (si & 31));
        }
      }
    }
    __syncthreads();
    if (tid == 0) nb[t] = nbs;
    for (int i = tid; i < 4 * nw8; i += kThreads) {
      const int q = i / nw8, w = i % nw8;
      sig8[(t * 4 + q) * nw8 + w] = bits[q][w];
    }
    for (int i = tid; i < nws; i += kThreads) smark[t * nws + i] = sb[i];
    __syncthreads();
  }
}

void launch_screen(hipStream_t s, long ntiles_pts, const double *gx, const double *gy,
                   const double *gz, int n_atoms, const AtomD *atoms, int n_shells,
                   const ScreenShell *shells, int n8, int *nb, uint32_t *sig8, uint32_t *smark) {
  if (ntiles_pts <= 0) return;
  hipLaunchKernelGGL(k_screen, dim3(persistent_grid(ntiles_pts, 8)), dim3(kThreads), 0, s,
                     ntiles_pts, gx, gy, gz, n_atoms, atoms, n_shells, shells, n8, nb, sig8, smark);
}

// ------------------------------------------------------------ small kernels
__global__ __launch_bounds__(kThreads) void k_form_density(const double *__restrict__ C,
                                                           const double *__restrict__ occ, int n,
                                                           int nocc, int NP, double *P) {
  const int j = blockIdx.x * kThreads + threadIdx.x, i = blockIdx.y;
  if (j >= n) return;
  double s = 0.0;
  for (int k = 0; k < nocc; ++k) s += C[i + (long)k * n] * occ[k] * C[j + (long)k * n];
  P[(long)i * NP + j] = s;
}

void launch_form_density(hipStream_t s, const double *C, const double *occ, int n, int nocc,
                         int NP, double *P) {
  dim3 grid((unsigned)((n + kThreads - 1) / kThreads), (unsigned)n);
  hipLaunchKernelGGL(k_form_density, grid, dim3(kThreads), 0, s, C, occ, n, nocc, NP, P);
}



}  // namespace pgxc

// ------------------------------------------------------------------ Becke
namespace pgxc {
// Becke partition weight of the owning atom (J. Chem. Phys. 88, 2547 (1988)),
// as becke0 (ParaGauss modules/grid_module.f90:2442-2497) with the step
// function mu3 = p(p(p(mu))), p(m) = m (3 - m^2) / 2
// (autodiff/becke_step_func.f90:84-124), no atomic-size adjustment.
__device__ __forceinline__ double becke_step(double mu) {
#pragma unroll
  for (int k = 0; k < 3; ++k) mu = mu * (3.0 - mu * mu) / 2.0;
  return mu;
}

__global__ __launch_bounds__(kThreads) void k_becke(int na, const double *__restrict__ axyz,
                                                    const double *__restrict__ inv_dist,
                                                    long npts, const double *__restrict__ pts,
                                                    const int *__restrict__ owner,
                                                    const double *__restrict__ wat,
                                                    double *__restrict__ wout) {
  const long p = (long)blockIdx.x * kThreads + threadIdx.x;
  if (p >= npts) return;
  const double x = pts[p], y = pts[npts + p], z = pts[2 * npts + p];
  const int own = owner[p];
  double sum = 0.0, pown = 0.0;
  for (int i = 0; i < na; ++i) {
    const double dxi = axyz[3 * i] - x, dyi = axyz[3 * i + 1] - y, dzi = axyz[3 * i + 2] - z;
    const double ri = sqrt(dxi * dxi + dyi * dyi + dzi * dzi);
    double pi = 1.0;
    for (int k = 0; k < na && pi != 0.0; ++k) {
      if (k == i) continue;
      const double dxk = axyz[3 * k] - x, dyk = axyz[3 * k + 1] - y, dzk = axyz[3 * k + 2] - z;
      const double rk = sqrt(dxk * dxk + dyk * dyk + dzk * dzk);
      const double mu = (ri - rk) * inv_dist[(long)i * na + k];
      pi *= (1.0 - becke_step(mu)) / 2.0;
    }
    sum += pi;
    if (i == own) pown = pi;
  }
  wout[p] = wat[p] * (pown / sum);
}

// Inverse distances of every pair of centres, inv[i][k] = 1/|R_i - R_k|
// (0 on the diagonal), one thread per pair; the table k_becke reads.
__global__ __launch_bounds__(kThreads) void k_inv_dist(int na, const double *__restrict__ axyz,
                                                       double *__restrict__ inv) {
  const long q = (long)blockIdx.x * kThreads + threadIdx.x;
  if (q >= (long)na * na) return;
  const int i = (int)(q / na), k = (int)(q % na);
  double v = 0.0;
  if (i != k) {
    const double dx = axyz[3 * i] - axyz[3 * k], dy = axyz[3 * i + 1] - axyz[3 * k + 1],
                 dz = axyz[3 * i + 2] - axyz[3 * k + 2];
    v = 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
  }
  inv[q] = v;
}

void launch_inv_dist(hipStream_t s, int na, const double *axyz, double *inv) {
  const long n = (long)na * na;
  hipLaunchKernelGGL(k_inv_dist, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     s, na, axyz, inv);
}

void launch_becke(hipStream_t s, int na, const double *axyz, const double *inv_dist, long npts,
                  const double *pts, const int *owner, const double *wat, double *wout) {
  hipLaunchKernelGGL(k_becke, dim3((unsigned)((npts + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, s, na, axyz, inv_dist, npts, pts, owner, wat, wout);
}
}  // namespace pgxc

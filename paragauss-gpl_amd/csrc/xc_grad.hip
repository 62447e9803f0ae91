// Nuclear gradient of E_xc, the basis-centre term: ParaGauss's F1
// (modules/post_scf_module.f90:150-185, stored in grad_xc when weight_grads
// is off), on a fixed grid:
//
//   dE/dR_A = -2 sum_p sum_{i on A} [ grad_d phi_i (2 a X_i + b . Y_i)
//                                     + X_i (b . grad) grad_d phi_i ]
//
// per spin, with X = Phi P, Y_c = (d/dc Phi) P (plain DGEMMs), a = w dF/drho / 2
// and b = w (2 dF/dg_ss grad rho_s + dF/dg_ab grad rho_s') the potential
// coefficients of the build (the second term only for GGA).  This file holds
// the basis Hessian and the per-function reduction over points.
#include "pgxc_internal.h"
#include "xc_device.h"

namespace pgxc {
namespace {

// value, gradient and Hessian (xx, xy, xz, yy, yz, zz) of a function of r
struct J2 {
  double v, g[3], h[6];
};
__device__ constexpr int kHc[6] = {0, 0, 0, 1, 1, 2}, kHd[6] = {0, 1, 2, 1, 2, 2};
// position of the symmetric pair (c, d) in that order
__device__ __forceinline__ int hidx(int c, int d) {
  const int lo = c < d ? c : d, hi = c < d ? d : c;
  return lo == 0 ? hi : lo == 1 ? 2 + hi : 5;
}

__device__ __forceinline__ J2 jconst(double v) {
  J2 r;
  r.v = v;
  for (int c = 0; c < 3; ++c) r.g[c] = 0.0;
  for (int k = 0; k < 6; ++k) r.h[k] = 0.0;
  return r;
}
__device__ __forceinline__ J2 jcoord(double x, int c) {  // the coordinate x_c
  J2 r = jconst(x);
  r.g[c] = 1.0;
  return r;
}
__device__ __forceinline__ J2 jmul(const J2 &a, const J2 &b) {
  J2 r;
  r.v = a.v * b.v;
  for (int c = 0; c < 3; ++c) r.g[c] = a.g[c] * b.v + a.v * b.g[c];
  for (int k = 0; k < 6; ++k) {
    const int c = kHc[k], d = kHd[k];
    r.h[k] = a.h[k] * b.v + a.g[c] * b.g[d] + a.g[d] * b.g[c] + a.v * b.h[k];
  }
  return r;
}
__device__ __forceinline__ J2 jaxpby(double s, const J2 &a, double t, const J2 &b) {
  J2 r;
  r.v = s * a.v + t * b.v;
  for (int c = 0; c < 3; ++c) r.g[c] = s * a.g[c] + t * b.g[c];
  for (int k = 0; k < 6; ++k) r.h[k] = s * a.h[k] + t * b.h[k];
  return r;
}

// Solid harmonics up to L as jets: the recursion of the value evaluation
// (m = 1 z-like, (2m, 2m+1) pairs, top pair by the (x, y) rotation) carried
// out on value + first + second derivatives.
template <int L>
__device__ __forceinline__ void sh_jets(double x, double y, double z, J2 *C) {
  const J2 X = jcoord(x, 0), Y = jcoord(y, 1), Z = jcoord(z, 2);
  J2 R2 = jconst(x * x + y * y + z * z);
  R2.g[0] = 2.0 * x;
  R2.g[1] = 2.0 * y;
  R2.g[2] = 2.0 * z;
  R2.h[0] = R2.h[3] = R2.h[5] = 2.0;
  C[0] = jconst(1.0);
  if constexpr (L >= 1) {
    C[1] = Z;
    C[2] = X;
    C[3] = Y;
  }
#pragma unroll
  for (int l = 2; l <= L; ++l) {
    const int o = l * l, o1 = (l - 1) * (l - 1), o2 = (l - 2) * (l - 2);
    const double lr = l, l1r = l - 1, ll1r = 2 * l - 1;
    C[o] = jaxpby(ll1r / lr, jmul(Z, C[o1]), -l1r / lr, jmul(R2, C[o2]));
#pragma unroll
    for (int im = 1; im <= l - 2; ++im) {
      const double sim = (double)(im * im);
      const double rlm = 1.0 / sqrt(lr * lr - sim), rl1m = sqrt(l1r * l1r - sim);
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int m = 2 * im + k - 1;
        C[o + m] = jaxpby(ll1r * rlm, jmul(Z, C[o1 + m]), -rl1m * rlm, jmul(R2, C[o2 + m]));
      }
    }
    {
      const double rlm = 1.0 / sqrt(lr * lr - (double)((l - 1) * (l - 1)));
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int m = 2 * l - 3 + k;
        C[o + m] = jaxpby(ll1r * rlm, jmul(Z, C[o1 + m]), 0.0, C[o1 + m]);
      }
    }
    {
      const double rll = sqrt(ll1r / (2.0 * lr));
      const int i2 = o1 + 2 * l - 3, i1 = o1 + 2 * l - 2;
      C[o + 2 * l - 1] = jaxpby(rll, jmul(X, C[i2]), -rll, jmul(Y, C[i1]));
      C[o + 2 * l] = jaxpby(rll, jmul(Y, C[i2]), rll, jmul(X, C[i1]));
    }
  }
}

// Hessian planes of the functions of one shell at one point: phi = C_lm R(r)
// with R a contraction of norm exp(-alpha r^2) (zero where alpha r^2 >= 50,
// as the value evaluation)
template <int L>
__device__ __forceinline__ void shell_hess(const ShellD &sh, const double *__restrict__ exps,
                                           const double *__restrict__ norms,
                                           const double *__restrict__ contr, double x, double y,
                                           double z, const J2 *C, double *hess, long Mc, long cs,
                                           long p) {
  const int ne = sh.n_exp, nu = sh.n_unc, nc = sh.n_contr, nf = nu + nc;
  const double r2 = x * x + y * y + z * z, xs[3] = {x, y, z};
  // radial jet of R = sum_e c_e r0_e: grad = R1 x, Hessian = R2 x x^T + R1 I,
  // R1 = sum c (-2 alpha) r0, R2 = sum c (4 alpha^2) r0
  auto emit = [&](int f, double R0, double R1, double R2) {
    J2 R = jconst(R0);
    for (int c = 0; c < 3; ++c) R.g[c] = R1 * xs[c];
    for (int k = 0; k < 6; ++k)
      R.h[k] = R2 * xs[kHc[k]] * xs[kHd[k]] + (kHc[k] == kHd[k] ? R1 : 0.0);
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) {
      const J2 F = jmul(C[L * L + m], R);
      const long q = (long)(f + m * nf) * Mc + p;
      for (int k = 0; k < 6; ++k) hess[k * cs + q] = F.h[k];
    }
  };
  double s0[kMaxContr], s1[kMaxContr], s2[kMaxContr];
  for (int k = 0; k < kMaxContr; ++k) s0[k] = s1[k] = s2[k] = 0.0;
  for (int e = 0; e < ne; ++e) {
    const double alpha = exps[sh.exp_off + e], ar2 = alpha * r2;
    double r0 = 0.0;
    if (ar2 < 50.0) r0 = norms[sh.exp_off + e] * exp(-ar2);
    const double r1 = -2.0 * alpha * r0, rr = 4.0 * alpha * alpha * r0;
    if (e < nu) emit(sh.func_off + e, r0, r1, rr);
    for (int k = 0; k < nc; ++k) {
      const double cf = contr[sh.contr_off + k * ne + e];
      s0[k] += r0 * cf;
      s1[k] += r1 * cf;
      s2[k] += rr * cf;
    }
  }
  for (int k = 0; k < nc; ++k) emit(sh.func_off + nu + k, s0[k], s1[k], s2[k]);
}

template <int LA>
__device__ __forceinline__ void atom_hess(const AtomD &at, const ShellD *__restrict__ shells,
                                          const double *exps, const double *norms,
                                          const double *contr, double x, double y, double z,
                                          double *hess, long Mc, long cs, long p) {
  J2 C[(LA + 1) * (LA + 1)];
  sh_jets<LA>(x, y, z, C);
  for (int s = at.shell_begin; s < at.shell_end; ++s) {
    const ShellD sh = shells[s];
    switch (sh.l) {
      case 0: shell_hess<0>(sh, exps, norms, contr, x, y, z, C, hess, Mc, cs, p); break;
      case 1: if constexpr (LA >= 1) shell_hess<1>(sh, exps, norms, contr, x, y, z, C, hess, Mc, cs, p); break;
      case 2: if constexpr (LA >= 2) shell_hess<2>(sh, exps, norms, contr, x, y, z, C, hess, Mc, cs, p); break;
      case 3: if constexpr (LA >= 3) shell_hess<3>(sh, exps, norms, contr, x, y, z, C, hess, Mc, cs, p); break;
      case 4: if constexpr (LA >= 4) shell_hess<4>(sh, exps, norms, contr, x, y, z, C, hess, Mc, cs, p); break;
    }
  }
}

// item = (256-point block, atom); hess = [6][NP][Mc] of one chunk
__global__ __launch_bounds__(kThreads) void k_basis_hess(
    const double *__restrict__ gx, const double *__restrict__ gy, const double *__restrict__ gz,
    const AtomD *__restrict__ atoms, int n_atoms, const ShellD *__restrict__ shells,
    const double *__restrict__ exps, const double *__restrict__ norms,
    const double *__restrict__ contr, int NP, long Mc, double *__restrict__ hess) {
  const long total = (Mc / kThreads) * n_atoms, cs = (long)NP * Mc;
  for (long it = blockIdx.x; it < total; it += gridDim.x) {
    const int atom = (int)(it % n_atoms);
    const long p = (it / n_atoms) * kThreads + threadIdx.x;
    const AtomD at = atoms[atom];
    const double x = gx[p] - at.x, y = gy[p] - at.y, z = gz[p] - at.z;
    switch (at.lmax) {
      case 0: atom_hess<0>(at, shells, exps, norms, contr, x, y, z, hess, Mc, cs, p); break;
      case 1: atom_hess<1>(at, shells, exps, norms, contr, x, y, z, hess, Mc, cs, p); break;
      case 2: atom_hess<2>(at, shells, exps, norms, contr, x, y, z, hess, Mc, cs, p); break;
      case 3: atom_hess<3>(at, shells, exps, norms, contr, x, y, z, hess, Mc, cs, p); break;
      default: atom_hess<4>(at, shells, exps, norms, contr, x, y, z, hess, Mc, cs, p); break;
    }
  }
}

// The gradient bracket of function i at point p (q = i Mc + p) added into t:
//   t_d += grad_d phi_i (a2 X_i + b . Y_i) + X_i sum_c b_c H_dc phi_i
template <bool GGA>
__device__ __forceinline__ void grad_bracket(const double *__restrict__ XY,
                                             const double *__restrict__ phi,
                                             const double *__restrict__ hess, long cs, long q,
                                             double a2, const double (&b)[3], double (&t)[3]) {
  const double X = XY[q];
  double s = a2 * X;
  if (GGA) s += b[0] * XY[cs + q] + b[1] * XY[2 * cs + q] + b[2] * XY[3 * cs + q];
  for (int d = 0; d < 3; ++d) {
    double u = phi[(1 + d) * cs + q] * s;
    if (GGA) {
      double hb = 0.0;
      for (int c = 0; c < 3; ++c) hb += b[c] * hess[hidx(c, d) * cs + q];
      u += X * hb;
    }
    t[d] += u;
  }
}

// item = one 128-point tile of the chunk: for each function i of its list
// (tfl[tf_off[t] ..], the functions not cut off on the tile; the others add
// exact zeros) the tile's sum of
//   t_d = grad_d phi_i (2 a X_i + b . Y_i) + X_i sum_c b_c H_dc phi_i
// into part[(list position - pbase) * 3 + d], over its points in a fixed
// order.  Two functions at a time, one per half of the workgroup (a point
// per thread).  XY = [4][NP][Mc]: X, then Y_x, Y_y, Y_z (written for the
// listed functions); phi the chunk's [4][NP][Mc] basis values; a, b the
// build's per-point coefficients.
template <bool GGA>
__global__ __launch_bounds__(kThreads) void k_grad_fn(long ntiles, const int *__restrict__ tf_off,
                                                      const uint16_t *__restrict__ tfl, int NP,
                                                      long Mc, const double *__restrict__ XY,
                                                      const double *__restrict__ phi,
                                                      const double *__restrict__ hess,
                                                      const double *__restrict__ a,
                                                      const double *__restrict__ bx,
                                                      const double *__restrict__ by,
                                                      const double *__restrict__ bz,
                                                      double *__restrict__ part) {
  static_assert(kThreads == 2 * kScreenTile, "two halves of one tile's points");
  __shared__ double sh[3][kThreads];
  const long cs = (long)NP * Mc, pbase = tf_off[0];
  const int tid = threadIdx.x, half = tid / kScreenTile, pt = tid % kScreenTile;
  for (long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const long f0 = tf_off[t];
    const int nf = tf_off[t + 1] - (int)f0;
    const long p = t * kScreenTile + pt;
    const double a2 = 2.0 * a[p];
    const double b[3] = {GGA ? bx[p] : 0.0, GGA ? by[p] : 0.0, GGA ? bz[p] : 0.0};
    for (int e0 = 0; e0 < nf; e0 += 2) {
      const int e = e0 + half;
      double tt[3] = {0.0, 0.0, 0.0};
      if (e < nf) grad_bracket<GGA>(XY, phi, hess, cs, (long)tfl[f0 + e] * Mc + p, a2, b, tt);
      for (int d = 0; d < 3; ++d) sh[d][tid] = tt[d];
      __syncthreads();
      for (int st = kScreenTile / 2; st > 0; st >>= 1) {
        if (pt < st)
          for (int d = 0; d < 3; ++d) sh[d][tid] += sh[d][tid + st];
        __syncthreads();
      }
      if (pt < 3 && e < nf) part[(f0 + e - pbase) * 3 + pt] = sh[pt][half * kScreenTile];
      __syncthreads();
    }
  }
}

// item = a point of the chunk: pm[d][p] += 2 sum_i t_d(i, p), the same terms
// as k_grad_fn summed over functions instead of points: w grad f at the point
// (the motion of the point with its atom)
template <bool GGA>
__global__ __launch_bounds__(kThreads) void k_grad_pt(const int *__restrict__ tf_off,
                                                      const uint16_t *__restrict__ tfl, int NP,
                                                      long Mc, const double *__restrict__ XY,
                                                      const double *__restrict__ phi,
                                                      const double *__restrict__ hess,
                                                      const double *__restrict__ a,
                                                      const double *__restrict__ bx,
                                                      const double *__restrict__ by,
                                                      const double *__restrict__ bz,
                                                      double *__restrict__ pm, long pstride) {
  const long cs = (long)NP * Mc;
  for (long p = (long)blockIdx.x * kThreads + threadIdx.x; p < Mc; p += (long)gridDim.x * kThreads) {
    const double a2 = 2.0 * a[p];
    double b[3] = {0.0, 0.0, 0.0};
    if (GGA) {
      b[0] = bx[p];
      b[1] = by[p];
      b[2] = bz[p];
    }
    double t[3] = {0.0, 0.0, 0.0};
    const long tile = p / kScreenTile;
    for (int e = tf_off[tile]; e < tf_off[tile + 1]; ++e)  // the functions not cut off here
      grad_bracket<GGA>(XY, phi, hess, cs, (long)tfl[e] * Mc + p, a2, b, t);
    for (int d = 0; d < 3; ++d) pm[d * pstride + p] += 2.0 * t[d];
  }
}

// ------------------------------------------------- Becke weight derivatives
// w_p = w_atomic P_A(r_p) / Z, P_C = prod_{D != C} s(mu_CD), Z = sum_C P_C,
// mu_CD = (|r - R_C| - |r - R_D|) / |R_C - R_D|, s = (1 - f3(mu)) / 2 with f3
// the three-fold iterated polynomial x (3 - x^2) / 2 (no atomic-size
// adjustment); the point moves with its owner A.  With sigma = s'/s:
//   d log P_C / dR_B = sum_{D != C} sigma(mu_CD) d mu_CD / dR_B
//   dw / dR_B = w sum_C (delta_CA - P_C / Z) d log P_C / dR_B.
// s = 0 only at mu = +-1, where s' = 0 too: sigma is taken as 0 there.
__device__ __forceinline__ void becke_s(double mu, double &s, double &sig) {
  const double f1 = mu * (3.0 - mu * mu) / 2.0, f2 = f1 * (3.0 - f1 * f1) / 2.0,
               f3 = f2 * (3.0 - f2 * f2) / 2.0;
  const double d = 1.5 * (1.0 - f2 * f2) * 1.5 * (1.0 - f1 * f1) * 1.5 * (1.0 - mu * mu);
  s = 0.5 * (1.0 - f3);
  sig = s > 0.0 ? -0.5 * d / s : 0.0;
}

struct Pt {
  double x, y, z;
};
__device__ __forceinline__ double dist(const AtomD &A, Pt r, double (&e)[3]) {
  e[0] = r.x - A.x;
  e[1] = r.y - A.y;
  e[2] = r.z - A.z;
  const double d = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  const double id = d > 0.0 ? 1.0 / d : 0.0;
  e[0] *= id;
  e[1] *= id;
  e[2] *= id;
  return d;
}

// item = a point: q[C][p] = P_C / Z, and rm[p] = sum_C (delta_CA - q_C)
// d log P_C / dr (the owner's share through the point's motion)
__global__ __launch_bounds__(kThreads) void k_becke_q(int na, const AtomD *__restrict__ atoms,
                                                      const double *__restrict__ inv_dist, long Mc,
                                                      const double *__restrict__ gx,
                                                      const double *__restrict__ gy,
                                                      const double *__restrict__ gz,
                                                      const int *__restrict__ owner,
                                                      double *__restrict__ q, double *__restrict__ rm) {
  for (long p = (long)blockIdx.x * kThreads + threadIdx.x; p < Mc; p += (long)gridDim.x * kThreads) {
    const int A = owner[p];
    if (A < 0) {
      for (int C = 0; C < na; ++C) q[C * Mc + p] = 0.0;
      for (int d = 0; d < 3; ++d) rm[d * Mc + p] = 0.0;
      continue;
    }
    const Pt r = {gx[p], gy[p], gz[p]};
    double Z = 0.0;
    for (int C = 0; C < na; ++C) {
      double eC[3];
      const double rC = dist(atoms[C], r, eC);
      double P = 1.0;
      for (int D = 0; D < na && P != 0.0; ++D) {
        if (D == C) continue;
        double eD[3], s, sig;
        const double rD = dist(atoms[D], r, eD);
        becke_s((rC - rD) * inv_dist[(long)C * na + D], s, sig);
        P *= s;
      }
      q[C * Mc + p] = P;
      Z += P;
    }
    const double iz = 1.0 / Z;
    double g[3] = {0.0, 0.0, 0.0};
    for (int C = 0; C < na; ++C) {
      const double qc = q[C * Mc + p] * iz;
      q[C * Mc + p] = qc;
      const double k = (C == A ? 1.0 : 0.0) - qc;
      if (k == 0.0) continue;
      double eC[3];
      const double rC = dist(atoms[C], r, eC);
      for (int D = 0; D < na; ++D) {
        if (D == C) continue;
        double eD[3], s, sig;
        const double rD = dist(atoms[D], r, eD);
        const double iR = inv_dist[(long)C * na + D];
        becke_s((rC - rD) * iR, s, sig);
        for (int d = 0; d < 3; ++d) g[d] += k * sig * (eC[d] - eD[d]) * iR;
      }
    }
    for (int d = 0; d < 3; ++d) rm[d * Mc + p] = g[d];
  }
}

// item = (256-point block, atom B): the block's sum over its points of
//   w f [ sum_{C != B} (delta_CA - q_C) sigma(mu_CB) d mu_CB / dR_B
//       + (delta_BA - q_B) sum_{D != B} sigma(mu_BD) d mu_BD / dR_B
//       + [B == A] rm ]
// with d mu_CB / dR_B = e_B / R + mu (R_C - R_B) / R^2 and
// d mu_BD / dR_B = -e_B / R - mu (R_B - R_D) / R^2 (R = |R_C - R_B|, ...),
// into part[blk][B][3]
__global__ __launch_bounds__(kThreads) void k_becke_atom(
    int na, const AtomD *__restrict__ atoms, const double *__restrict__ inv_dist, long Mc,
    const double *__restrict__ gx, const double *__restrict__ gy, const double *__restrict__ gz,
    const double *__restrict__ gw, const double *__restrict__ fp, const int *__restrict__ owner,
    const double *__restrict__ q, const double *__restrict__ rm, double *__restrict__ part) {
  __shared__ double sh[3][kThreads];
  const int tid = threadIdx.x;
  const long total = (Mc / kThreads) * na;
  for (long it = blockIdx.x; it < total; it += gridDim.x) {
    const int B = (int)(it % na);
    const long blk = it / na, p = blk * kThreads + tid;
    const int A = owner[p];
    double g[3] = {0.0, 0.0, 0.0};
    const double wf = gw[p] * fp[p];
    if (A >= 0 && wf != 0.0) {
      const Pt r = {gx[p], gy[p], gz[p]};
      const AtomD &AB = atoms[B];
      double eB[3];
      const double rB = dist(AB, r, eB);
      const double kB = (B == A ? 1.0 : 0.0) - q[B * Mc + p];
      for (int C = 0; C < na; ++C) {
        if (C == B) continue;
        const AtomD &AC = atoms[C];
        double eC[3], s, sig;
        const double rC = dist(AC, r, eC);
        const double iR = inv_dist[(long)C * na + B];
        const double RCB[3] = {AC.x - AB.x, AC.y - AB.y, AC.z - AB.z};
        // pair (C, B): B is the second atom
        const double kC = (C == A ? 1.0 : 0.0) - q[C * Mc + p];
        const double muCB = (rC - rB) * iR;
        if (kC != 0.0) {
          becke_s(muCB, s, sig);
          for (int d = 0; d < 3; ++d) g[d] += kC * sig * (eB[d] * iR + muCB * RCB[d] * iR * iR);
        }
        // pair (B, C): B is the first atom, mu_BC = -mu_CB, R_B - R_C = -RCB
        if (kB != 0.0) {
          becke_s(-muCB, s, sig);
          for (int d = 0; d < 3; ++d) g[d] += kB * sig * (-eB[d] * iR - muCB * RCB[d] * iR * iR);
        }
      }
      if (B == A)
        for (int d = 0; d < 3; ++d) g[d] += rm[d * Mc + p];
      for (int d = 0; d < 3; ++d) g[d] *= wf;
    }
    for (int d = 0; d < 3; ++d) sh[d][tid] = g[d];
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
      if (tid < st)
        for (int d = 0; d < 3; ++d) sh[d][tid] += sh[d][tid + st];
      __syncthreads();
    }
    if (tid < 3) part[(blk * na + B) * 3 + tid] = sh[tid][0];
    __syncthreads();
  }
}

}  // namespace

void launch_basis_hess(hipStream_t s, const double *gx, const double *gy, const double *gz,
                       const AtomD *atoms, int n_atoms, const ShellD *shells, const double *exps,
                       const double *norms, const double *contr, int NP, long Mc, double *hess) {
  const long total = (Mc / kThreads) * n_atoms;
  const unsigned g = (unsigned)std::min<long>(total, 256L * 8);
  hipLaunchKernelGGL(k_basis_hess, dim3(g), dim3(kThreads), 0, s, gx, gy, gz, atoms, n_atoms,
                     shells, exps, norms, contr, NP, Mc, hess);
}

void launch_grad_fn(hipStream_t s, bool gga, long ntiles, const int *tf_off, const uint16_t *tfl,
                    int NP, long Mc, const double *XY, const double *phi, const double *hess,
                    const double *a, const double *bx, const double *by, const double *bz,
                    double *part) {
  if (ntiles <= 0) return;
  const unsigned g = (unsigned)std::min<long>(ntiles, 256L * 8);
  if (gga)
    hipLaunchKernelGGL(k_grad_fn<true>, dim3(g), dim3(kThreads), 0, s, ntiles, tf_off, tfl, NP, Mc,
                       XY, phi, hess, a, bx, by, bz, part);
  else
    hipLaunchKernelGGL(k_grad_fn<false>, dim3(g), dim3(kThreads), 0, s, ntiles, tf_off, tfl, NP, Mc,
                       XY, phi, hess, a, bx, by, bz, part);
}

}  // namespace pgxc

namespace pgxc {
void launch_grad_pt(hipStream_t s, bool gga, const int *tf_off, const uint16_t *tfl, int NP,
                    long Mc, const double *XY, const double *phi, const double *hess,
                    const double *a, const double *bx, const double *by, const double *bz,
                    double *pm, long pstride) {
  const unsigned g = (unsigned)std::min<long>((Mc + kThreads - 1) / kThreads, 256L * 8);
  if (gga)
    hipLaunchKernelGGL(k_grad_pt<true>, dim3(g), dim3(kThreads), 0, s, tf_off, tfl, NP, Mc, XY,
                       phi, hess, a, bx, by, bz, pm, pstride);
  else
    hipLaunchKernelGGL(k_grad_pt<false>, dim3(g), dim3(kThreads), 0, s, tf_off, tfl, NP, Mc, XY,
                       phi, hess, a, bx, by, bz, pm, pstride);
}

void launch_becke_grad(hipStream_t s, int na, const AtomD *atoms, const double *inv_dist, long Mc,
                       const double *gx, const double *gy, const double *gz, const double *gw,
                       const double *fp, const int *owner, double *q, double *rm, double *part) {
  const unsigned g1 = (unsigned)std::min<long>((Mc + kThreads - 1) / kThreads, 256L * 8);
  hipLaunchKernelGGL(k_becke_q, dim3(g1), dim3(kThreads), 0, s, na, atoms, inv_dist, Mc, gx, gy,
                     gz, owner, q, rm);
  const long total = (Mc / kThreads) * na;
  const unsigned g2 = (unsigned)std::min<long>(total, 256L * 8);
  hipLaunchKernelGGL(k_becke_atom, dim3(g2), dim3(kThreads), 0, s, na, atoms, inv_dist, Mc, gx, gy,
                     gz, gw, fp, owner, q, rm, part);
}
}  // namespace pgxc

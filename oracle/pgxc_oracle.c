/*
 * pgxc_oracle.c -- CPU FP64 restatement of ParaGauss's exchange-correlation
 * grid integration (xc_build -> calc_xcks_nl / calc_xcks).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product path (libpgxc.so) never links, loads or calls it.
 *
 * Every function follows the reference lines named beside it (paths are
 * relative to the ParaGauss source root):
 *
 *   or_sh_point        modules/solid_harmonics_module.f90:369-482 (values, same
 *                      recursion and operation order).  Gradients: forward
 *                      differentiation of that recursion; the reference applies
 *                      product/differential rules instead
 *                      (modules/solid_harmonics_module.f90:488-568,
 *                      modules/solhrules_module.f90:821-880).  Both are the exact
 *                      derivative of the same polynomial.
 *   or_prim_norm       modules/unique_atom_module.f90:426-444, dfac :465-486
 *   or_basis_eval      modules/orbital_module.f90:1742-1932 (radial, radial1,
 *                      contract) and :2421-2619 (calculate_orbs_and_grads_v3,
 *                      C1 case: one equal atom, one independent function per m
 *                      with coefficient 1, order m -> radial function).
 *   or_xalpha          modules/exchange.f90:511-601 (exchange_lda, x_lda)
 *   or_pbex            modules/exchange.f90:206-292,325-439 (x_gga),
 *                      :617-639 (g_pbe), :858-867 (g_becke86_form)
 *   or_pw92            modules/pw_ldac_module.f90:127-194 (c_lda),
 *                      :255-270 (ispin=1 loop), :591-651 (set_param)
 *   or_pbec            modules/pbe_ggcxc_module.f90:96-212 (c_gga), :273-301
 *   or_functional      modules/xc_func.f90:187-305 (term loop, frac = 1),
 *                      term order by modules/xc_cntrl.f90:61-104 enum
 *   or_slater_radius   modules/grid_module.f90:1029-1079 (rslater_f, Z <= 10)
 *   or_radial_grid     modules/grid_module.f90:1246-1301 (RADIALGRID)
 *   or_lebedev         modules/grid_module.f90:1359-2100 (INIGRD, 194 and 302)
 *   or_atomic_grid     modules/grid_module.f90:407-433 (product grid)
 *   or_becke0          modules/grid_module.f90:2442-2497 (becke0, art = 0) and
 *                      autodiff/becke_step_func.f90:84-124 (mu0)
 *   or_xc_build        modules/xc_hamiltonian.f90:1280-1580 (GGA),
 *                      :812-1036 (LDA: integrate, tristore),
 *                      modules/density_calc_module.f90:1478-1577 (density),
 *                      batch source modules/grid_module.f90:1133-1181 and the
 *                      static split dlb/dlb_common.f90:516-561.
 *
 * How each is pinned (checks live in tests/):
 *   or_sh_point   compiled reference solid_harmonics_module (oracle/_ref) and
 *                 its golden vectors; closed forms l<=2; sum_m C_lm^2 = r^2l;
 *                 Euler r.grad C = l C; central differences.
 *   or_prim_norm  unit self-overlap of every primitive by radial quadrature.
 *   or_pw92       compiled reference pw_ldac_module golden vectors.
 *   or_pbec       compiled reference pbe_ggcxc_module golden vectors.
 *   or_xalpha     closed form (Dirac exchange) + central differences.
 *   or_pbex       textbook PBE enhancement 1+k-k/(1+mu s^2/k) with the
 *                 reference's mu = 0.066725 pi^2/3, + central differences.
 *                 (exchange.f90 cannot be compiled standalone: it uses vwnc,
 *                 which pulls in the SCF state modules.)
 *   or_basis_eval brute-force Python evaluation from the pinned pieces.
 *   or_radial_grid  w = r^2 dr/di of the continuous map r(i) (central
 *                 differences); Gaussian radial moments in closed form.
 *   or_lebedev    unit points, octahedral orbits, exactness on monomials to
 *                 the rule's degree, agreement with the exact rules solved
 *                 independently (pgxc/lebedev.py solve()) to the table's digits.
 *   or_becke0     partition of unity; weight 1 at the host nucleus; a
 *                 brute-force product over all ordered pairs.
 *   or_xc_build   brute-force numpy einsum of the same sums on tiny inputs;
 *                 P-form == occupied-orbital form; DGEMM arithmetic is the
 *                 textbook triple loop ("parity unpinned" against the
 *                 unversioned system BLAS the reference links).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_LMAX 6
#define OR_NLM ((OR_LMAX + 1) * (OR_LMAX + 1))

/* xc_cntrl enum values (modules/xc_cntrl.f90:61-104) used as bit positions */
#define XC_XALPHA 1
#define XC_PWLDAC 3
#define XC_PBEX 10
#define XC_PBEC 18

/* ------------------------------------------------------------------------ */
/* solid harmonics, Racah normalised; lm index = l*l + (m-1), m = 1..2l+1    */
/* ------------------------------------------------------------------------ */
void or_sh_point(int lmax, double x, double y, double z, double *C, double *Cx,
                 double *Cy, double *Cz)
{
  const double r2 = x * x + y * y + z * z;
  C[0] = 1.0;
  Cx[0] = Cy[0] = Cz[0] = 0.0;
  if (lmax == 0) return;
  C[1] = z; C[2] = x; C[3] = y;
  Cx[1] = 0; Cy[1] = 0; Cz[1] = 1;
  Cx[2] = 1; Cy[2] = 0; Cz[2] = 0;
  Cx[3] = 0; Cy[3] = 1; Cz[3] = 0;
  const double dr2[3] = {2 * x, 2 * y, 2 * z};
  for (int l = 2; l <= lmax; ++l) {
    const int o = l * l, o1 = (l - 1) * (l - 1), o2 = (l - 2) * (l - 2);
    const double lr = l, l1r = l - 1, ll1r = 2 * l - 1;
    double *D[3] = {Cx, Cy, Cz};
    /* m = 1 */
    C[o] = (ll1r * z * C[o1] - l1r * r2 * C[o2]) / lr;
    for (int q = 0; q < 3; ++q) {
      const double dz = (q == 2) ? 1.0 : 0.0;
      D[q][o] = (ll1r * (dz * C[o1] + z * D[q][o1]) -
                 l1r * (dr2[q] * C[o2] + r2 * D[q][o2])) / lr;
    }
    /* m = 2 .. 2l-3 */
    for (int im = 1; im <= l - 2; ++im) {
      const double sim = (double)(im * im);
      double rlm = sqrt(lr * lr - sim);
      rlm = 1 / rlm;
      const double rl1m = sqrt(l1r * l1r - sim);
      const double a = ll1r * z * rlm, b = rl1m * r2 * rlm;
      for (int k = 0; k < 2; ++k) {
        const int m = 2 * im + k - 1; /* zero-based offset inside shell */
        C[o + m] = a * C[o1 + m] - b * C[o2 + m];
        for (int q = 0; q < 3; ++q) {
          const double da = (q == 2) ? ll1r * rlm : 0.0;
          const double db = rl1m * dr2[q] * rlm;
          D[q][o + m] = a * D[q][o1 + m] + da * C[o1 + m] - b * D[q][o2 + m] -
                        db * C[o2 + m];
        }
      }
    }
    /* m = 2l-2, 2l-1 */
    {
      const double sim = (double)((l - 1) * (l - 1));
      double rlm = sqrt(lr * lr - sim);
      rlm = 1 / rlm;
      const double a = ll1r * z * rlm;
      for (int k = 0; k < 2; ++k) {
        const int m = 2 * l - 3 + k;
        C[o + m] = a * C[o1 + m];
        for (int q = 0; q < 3; ++q) {
          const double da = (q == 2) ? ll1r * rlm : 0.0;
          D[q][o + m] = a * D[q][o1 + m] + da * C[o1 + m];
        }
      }
    }
    /* m = 2l, 2l+1 */
    {
      const double rll = sqrt(ll1r / (2 * lr));
      const int i2 = o1 + 2 * l - 3, i1 = o1 + 2 * l - 2;
      C[o + 2 * l - 1] = (x * C[i2] - y * C[i1]) * rll;
      C[o + 2 * l] = (y * C[i2] + x * C[i1]) * rll;
      for (int q = 0; q < 3; ++q) {
        const double dx = (q == 0), dy = (q == 1);
        D[q][o + 2 * l - 1] =
            (dx * C[i2] + x * D[q][i2] - dy * C[i1] - y * D[q][i1]) * rll;
        D[q][o + 2 * l] =
            (dy * C[i2] + y * D[q][i2] + dx * C[i1] + x * D[q][i1]) * rll;
      }
    }
  }
}

/* vectorised wrapper: xyz is (npts,3) column-major relative coordinates */
void or_solid_harmonics(int lmax, int npts, const double *xyz, double *C,
                        double *G)
{
  const int nlm = (lmax + 1) * (lmax + 1);
  double c[OR_NLM], cx[OR_NLM], cy[OR_NLM], cz[OR_NLM];
  for (int p = 0; p < npts; ++p) {
    or_sh_point(lmax, xyz[p], xyz[npts + p], xyz[2 * npts + p], c, cx, cy, cz);
    for (int i = 0; i < nlm; ++i) {
      C[(size_t)i * npts + p] = c[i];
      G[((size_t)i * 3 + 0) * npts + p] = cx[i];
      G[((size_t)i * 3 + 1) * npts + p] = cy[i];
      G[((size_t)i * 3 + 2) * npts + p] = cz[i];
    }
  }
}

/* legacy primitive norm (2a/pi)^(3/4) (2 sqrt a)^L / sqrt((2L-1)!!) */
double or_prim_norm(int L, double alpha)
{
  const double pi = 3.14159265358979324;
  int n = 1;
  for (int i = 1; i <= L; ++i) n *= (2 * i - 1);
  const double df = (double)n;
  return pow(2 * alpha / pi, 0.75) * pow(2 * sqrt(alpha), L) / sqrt(df);
}

/* ------------------------------------------------------------------------ */
/* basis description shared with the product: shell_info has 5 ints/shell:  */
/*   atom, l, n_exp, n_unc, n_contr ; exps/norms concatenated per shell;    */
/*   contr concatenated per shell, each n_exp x n_contr column-major.       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n_atoms;
  const double *atom_xyz; /* 3 per atom */
  int n_shells;
  const int *shell_info;
  const double *exps, *norms, *contr;
} or_basis;

int or_nbas(int n_shells, const int *shell_info)
{
  int n = 0;
  for (int s = 0; s < n_shells; ++s) {
    const int *si = shell_info + 5 * s;
    n += (2 * si[1] + 1) * (si[3] + si[4]);
  }
  return n;
}

#define OR_MAXEXP 64

/* Basis values and gradients on npts points.  pts: column-major (ld,3).
 * phi[f*npts+p]; grad[(f*3+c)*npts+p] (the layout of orbs_grads(vl,3,dim)). */
static void basis_eval(const or_basis *b, int npts, const double *pts, long ld,
                       double *phi, double *grad)
{
  double c[OR_NLM], cx[OR_NLM], cy[OR_NLM], cz[OR_NLM];
  double rad0[OR_MAXEXP], rad1[OR_MAXEXP], rc0[OR_MAXEXP], rc1[OR_MAXEXP];
  for (int p = 0; p < npts; ++p) {
    int f = 0, eoff = 0, coff = 0, cur_atom = -1, cur_lmax = -1;
    double x = 0, y = 0, z = 0, r2 = 0;
    for (int s = 0; s < b->n_shells; ++s) {
      const int *si = b->shell_info + 5 * s;
      const int atom = si[0], l = si[1], ne = si[2], nu = si[3], nc = si[4];
      if (atom != cur_atom || l > cur_lmax) {
        /* harmonics of this centre up to the largest l it carries */
        int lm = l;
        for (int t = s + 1; t < b->n_shells && b->shell_info[5 * t] == atom; ++t)
          if (b->shell_info[5 * t + 1] > lm) lm = b->shell_info[5 * t + 1];
        const double *R = b->atom_xyz + 3 * atom;
        x = pts[p] - R[0];
        y = pts[ld + p] - R[1];
        z = pts[2 * ld + p] - R[2];
        r2 = x * x + y * y + z * z;
        or_sh_point(lm, x, y, z, c, cx, cy, cz);
        cur_atom = atom;
        cur_lmax = lm;
      }
      const double *ex = b->exps + eoff, *nm = b->norms + eoff;
      const double *ct = b->contr + coff;
      /* radial1(): primitive factors with the alpha r^2 >= 50 cut-off */
      for (int e = 0; e < ne; ++e) {
        const double alpha = ex[e], ar2 = alpha * r2;
        if (ar2 < 50.0) {
          rad0[e] = nm[e] * exp(-ar2);
          rad1[e] = (-2.0 * alpha) * rad0[e];
        } else {
          rad0[e] = 0.0;
          rad1[e] = 0.0;
        }
      }
      /* contract(): columns n_unc+1.. are contractions over all exponents */
      for (int u = 0; u < nu; ++u) { rc0[u] = rad0[u]; rc1[u] = rad1[u]; }
      for (int k = 0; k < nc; ++k) {
        double s0 = 0.0, s1 = 0.0;
        for (int e = 0; e < ne; ++e) {
          s0 += rad0[e] * ct[(size_t)k * ne + e];
          s1 += rad1[e] * ct[(size_t)k * ne + e];
        }
        rc0[nu + k] = s0;
        rc1[nu + k] = s1;
      }
      const int nf = nu + nc, o = l * l;
      for (int m = 0; m < 2 * l + 1; ++m) {
        const double sv = c[o + m], sx = cx[o + m], sy = cy[o + m], sz = cz[o + m];
        for (int k = 0; k < nf; ++k, ++f) {
          phi[(size_t)f * npts + p] = sv * rc0[k];
          const double hv = sv * rc1[k];
          grad[((size_t)f * 3 + 0) * npts + p] = x * hv + sx * rc0[k];
          grad[((size_t)f * 3 + 1) * npts + p] = y * hv + sy * rc0[k];
          grad[((size_t)f * 3 + 2) * npts + p] = z * hv + sz * rc0[k];
        }
      }
      eoff += ne;
      coff += ne * nc;
    }
  }
}

void or_basis_eval(int n_atoms, const double *atom_xyz, int n_shells,
                   const int *shell_info, const double *exps, const double *norms,
                   const double *contr, int npts, const double *pts, long ld,
                   double *phi, double *grad)
{
  or_basis b = {n_atoms, atom_xyz, n_shells, shell_info, exps, norms, contr};
  basis_eval(&b, npts, pts, ld, phi, grad);
}

/* ------------------------------------------------------------------------ */
/* functionals, closed shell (ispin = 1)                                     */
/* ------------------------------------------------------------------------ */
static const double PI = 3.14159265358979324;
static const double DEPS = 1.0e-30;

/* Slater exchange e = -(3/4)(3/pi)^(1/3) rho^(4/3), 0 for rho <= 0 */
void or_xalpha(int n, const double *rho, double *f, double *fr)
{
  const double xalpha = -(3.0 / 4.0) * pow(3.0 / PI, 1.0 / 3.0);
  for (int i = 0; i < n; ++i) {
    const double r = rho[i];
    if (r > 0.0) {
      f[i] = xalpha * pow(r, 4.0 / 3.0);
      fr[i] = xalpha * ((4.0 / 3.0) * pow(r, 1.0 / 3.0));
    } else {
      f[i] = 0.0;
      fr[i] = 0.0;
    }
  }
}

/* PBE exchange excess e_x(rho) G(s^2), G = k mu s2/(k + mu s2) */
void or_pbex(int n, const double *rho, const double *gam, double *f, double *fr,
             double *fg)
{
  const double xalpha = -(3.0 / 4.0) * pow(3.0 / PI, 1.0 / 3.0);
  const double kappa = 0.804, beta = 0.066725, mu = beta * (PI * PI / 3.0);
  const double s2_fact = 1.0 / (4.0 * pow(3.0 * PI * PI, 2.0 / 3.0));
  for (int i = 0; i < n; ++i) {
    const double r = rho[i], g = gam[i];
    double ex = 0.0, dex = 0.0;
    if (r > 0.0) {
      ex = xalpha * pow(r, 4.0 / 3.0);
      dex = xalpha * ((4.0 / 3.0) * pow(r, 1.0 / 3.0));
    }
    double s2 = 0.0, s2r = 0.0, s2g = 0.0;
    if (r >= 1.0e-20) {
      s2g = 1.0 / pow(r, 8.0 / 3.0) * s2_fact;
      s2 = g * s2g;
      s2r = -(8.0 / 3.0) * s2 / r;
    }
    const double den = kappa + mu * s2;
    const double G = kappa * mu * s2 / den;
    const double dG = (kappa * mu) / den - (kappa * mu * s2) * mu / (den * den);
    f[i] = ex * G;
    fr[i] = dex * G + ex * dG * s2r;
    fg[i] = ex * dG * s2g;
  }
}

/* PW92 G(rs) for the paramagnetic set (do_ECU, RPA = false); value + d/drs */
static void pw92_g(double rs, double *E, double *dE)
{
  const double a = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876,
               b3 = 1.6382, b4 = 0.49294;
  /* p = 1 for this set: rs**p == rs */
  const double rs1d2 = sqrt(rs), rs3d2 = rs * rs1d2, rsp = rs;
  const double Q0 = -(2.0 * a * (1.0 + a1 * rs));
  const double Q1 = 2.0 * a * (b1 * rs1d2 + b2 * rs + b3 * rs3d2 + b4 * rs * rsp);
  const double Q2 = log(1.0 + 1.0 / Q1);
  const double dQ0 = -2.0 * a * a1;
  const double dQ1 =
      2.0 * a * (b1 * 0.5 / rs1d2 + b2 + 1.5 * b3 * rs1d2 + 2.0 * b4 * rs);
  const double dQ2 = (-dQ1 / (Q1 * Q1)) / (1.0 + 1.0 / Q1);
  *E = Q0 * Q2;
  *dE = dQ0 * Q2 + Q0 * dQ2;
}

/* c_lda(rho/2, rho/2): value and d/drho, with the nn < DEPS clamp */
static void pw92_restricted(double rho, double *F, double *dF)
{
  const double crs = pow(3.0 / (4.0 * PI), 1.0 / 3.0);
  double nn = 0.5 * rho + 0.5 * rho;
  int fixed = 0;
  if (nn < DEPS) { nn = DEPS; fixed = 1; }
  const double rs = crs / pow(nn, 1.0 / 3.0);
  double E, dE;
  pw92_g(rs, &E, &dE);
  /* z = 0, so f(z) = 0 and EC = ECU exactly */
  *F = nn * E;
  *dF = fixed ? 0.0 : E + nn * dE * (-rs / (3.0 * nn));
}

void or_pw92(int n, const double *rho, double *f, double *fr)
{
  for (int i = 0; i < n; ++i) pw92_restricted(rho[i], &f[i], &fr[i]);
}

/* c_gga(rho/2, rho/2, gamma, beta = 0.066725): F = n H */
void or_pbec(int n, const double *rho, const double *gam, double *f, double *fr,
             double *fg)
{
  const double beta = 0.066725;
  const double gc = 0.03109069086965489503494086371;
  const double bg = beta / gc;
  const double K = PI / (4.0 * 4.0 * pow(3.0 * PI * PI, 1.0 / 3.0));
  const double crs = pow(3.0 / (4.0 * PI), 1.0 / 3.0);
  /* za = zb = 1/2 exactly in the closed-shell case */
  const double za = 0.5, zb = 0.5;
  const double fi = (pow(za, 2.0 / 3.0) + pow(zb, 2.0 / 3.0)) * (1.0 / pow(2.0, 1.0 / 3.0));
  const double fi2 = fi * fi, fi3 = fi2 * fi;
  for (int i = 0; i < n; ++i) {
    const double nv = 0.5 * rho[i] + 0.5 * rho[i], g2 = gam[i];
    if (nv <= 0.0) { f[i] = fr[i] = fg[i] = 0.0; continue; }
    /* ELDAC = c_lda(na, nb)/n and its n-derivative */
    double El, dEl;
    {
      double nn = nv;
      int fixed = 0;
      if (nn < DEPS) { nn = DEPS; fixed = 1; }
      const double rs = crs / pow(nn, 1.0 / 3.0);
      double E, dE;
      pw92_g(rs, &E, &dE);
      const double Flda = nn * E;
      const double dFlda = fixed ? 0.0 : E + nn * dE * (-rs / (3.0 * nn));
      El = Flda / nv;
      dEl = dFlda / nv - Flda / (nv * nv);
    }
    const double Ex = exp(-El / (gc * fi3));
    const double dden = Ex - 1.0 + DEPS;
    const double A = bg / dden;
    const double dA = (bg / (dden * dden)) * Ex * (dEl / (gc * fi3));
    const double n73 = pow(nv, 7.0 / 3.0);
    const double D = fi2 * n73 + DEPS;
    const double t2 = g2 / D * K;
    const double dt2n = -(g2 * K / (D * D)) * fi2 * (7.0 / 3.0) * (n73 / nv);
    const double dt2g = K / D;
    const double x = A * t2;
    const double dxn = dA * t2 + A * dt2n, dxg = A * dt2g;
    const double q = 1.0 + x + x * x;
    const double rat = (1.0 + x) / q;
    const double drat = -x * (2.0 + x) / (q * q);
    const double arg = 1.0 + bg * t2 * rat;
    const double H = gc * fi3 * log(arg);
    const double pre = gc * fi3 / arg * bg;
    const double dHn = pre * (dt2n * rat + t2 * drat * dxn);
    const double dHg = pre * (dt2g * rat + t2 * drat * dxg);
    f[i] = nv * H;
    fr[i] = H + nv * dHn;
    fg[i] = nv * dHg;
  }
}

/* Sum of the enabled terms in enum order: f, dF/drho, dF/dgamma */
int or_functional(int mask, int n, const double *rho, const double *gam,
                  double *f, double *fr, double *fg)
{
  const int gga = (mask & ((1 << XC_PBEX) | (1 << XC_PBEC))) != 0;
  if (mask & ~((1 << XC_XALPHA) | (1 << XC_PWLDAC) | (1 << XC_PBEX) | (1 << XC_PBEC)))
    return -1;
  if ((mask & (1 << XC_PBEC)) && !(mask & (1 << XC_PWLDAC))) return -2;
  double *x = malloc(sizeof(double) * 3 * (size_t)n);
  double *xr = x + n, *xg = x + 2 * (size_t)n;
  for (int i = 0; i < n; ++i) { f[i] = 0; fr[i] = 0; if (gga) fg[i] = 0; }
  for (int term = 1; term <= XC_PBEC; ++term) {
    if (!(mask & (1 << term))) continue;
    for (int i = 0; i < n; ++i) { x[i] = 0; xr[i] = 0; xg[i] = 0; }
    switch (term) {
      case XC_XALPHA: or_xalpha(n, rho, x, xr); break;
      case XC_PWLDAC: or_pw92(n, rho, x, xr); break;
      case XC_PBEX: or_pbex(n, rho, gam, x, xr, xg); break;
      case XC_PBEC: or_pbec(n, rho, gam, x, xr, xg); break;
    }
    for (int i = 0; i < n; ++i) {
      f[i] = f[i] + x[i] * 1.0;
      fr[i] = fr[i] + xr[i] * 1.0;
      if (gga) fg[i] = fg[i] + xg[i] * 1.0;
    }
  }
  free(x);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* functionals, spin polarised (ispin = 2)                                   */
/* ------------------------------------------------------------------------ */
/* Forward-mode values with first derivatives in three variables, as the
 * reference's ad2x3 objects (autodiff/autodiff2.f90) carry them; the
 * expressions below follow the reference's unrestricted branches line by
 * line. */
typedef struct { double v, d[3]; } dual;

static dual dv(double v, int k)
{
  dual r = {v, {0, 0, 0}};
  if (k >= 0) r.d[k] = 1.0;
  return r;
}
static dual dc(double v) { return dv(v, -1); }
static dual dadd(dual a, dual b)
{
  dual r = {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
  return r;
}
static dual dsub(dual a, dual b)
{
  dual r = {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
  return r;
}
static dual dmul(dual a, dual b)
{
  dual r;
  r.v = a.v * b.v;
  for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
static dual ddiv(dual a, dual b)
{
  dual r;
  r.v = a.v / b.v;
  for (int k = 0; k < 3; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
  return r;
}
static dual dscale(dual a, double c)
{
  dual r = {a.v * c, {a.d[0] * c, a.d[1] * c, a.d[2] * c}};
  return r;
}
static dual dpowc(dual a, double c) /* a**c, a > 0 */
{
  dual r;
  r.v = pow(a.v, c);
  const double g = c * pow(a.v, c - 1.0);
  for (int k = 0; k < 3; ++k) r.d[k] = g * a.d[k];
  return r;
}
static dual dlog(dual a)
{
  dual r;
  r.v = log(a.v);
  for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] / a.v;
  return r;
}
static dual dexp(dual a)
{
  dual r;
  r.v = exp(a.v);
  for (int k = 0; k < 3; ++k) r.d[k] = r.v * a.d[k];
  return r;
}

/* PW92 parameter sets (do_ECU, do_ECP, do_mALF; RPA = false), p = 1 */
static const double PW92_SET[3][6] = {
    {0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294},
    {0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517},
    {0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671}};

/* g(rs): pw_ldac_module.f90 c_lda's contained g */
static dual pw92_gd(dual rs, int set)
{
  const double *q = PW92_SET[set];
  const double a = q[0], a1 = q[1], b1 = q[2], b2 = q[3], b3 = q[4], b4 = q[5];
  const dual rs1d2 = dpowc(rs, 0.5), rs3d2 = dmul(rs, rs1d2), rsp = rs; /* p = 1 */
  const dual Q0 = dscale(dadd(dc(1.0), dscale(rs, a1)), -2.0 * a);
  const dual Q1 = dscale(dadd(dadd(dadd(dscale(rs1d2, b1), dscale(rs, b2)), dscale(rs3d2, b3)),
                              dscale(dmul(rs, rsp), b4)), 2.0 * a);
  const dual Q2 = dlog(dadd(dc(1.0), ddiv(dc(1.0), Q1)));
  return dmul(Q0, Q2);
}

/* c_lda(na, nb) of pw_ldac_module.f90:127-195 (same in pbe_ggcxc_module) */
static dual pw92_clda(dual na, dual nb)
{
  const double c1d3 = 1.0 / 3.0, c4d3 = 4.0 / 3.0;
  const double crs = pow(3.0 / (4.0 * PI), c1d3);
  const double cf = 1.0 / (pow(2.0, c4d3) - 2.0);
  const double cfzz = 2.0 * c4d3 * c1d3 * cf;
  dual nn = dadd(na, nb);
  if (nn.v < DEPS) nn = dc(DEPS);
  const dual rs = ddiv(dc(crs), dpowc(nn, c1d3));
  const dual ECU = pw92_gd(rs, 0), ECP = pw92_gd(rs, 1), mALF = pw92_gd(rs, 2);
  const dual z = ddiv(dsub(na, nb), nn), z2 = dmul(z, z), z4 = dmul(z2, z2);
  const dual f = dscale(dsub(dadd(dpowc(dadd(dadd(dc(1.0), z), dc(DEPS)), c4d3),
                                  dpowc(dadd(dsub(dc(1.0), z), dc(DEPS)), c4d3)), dc(2.0)), cf);
  const dual EC = dsub(dadd(ECU, dmul(dmul(dsub(ECP, ECU), f), z4)),
                       dscale(dmul(dmul(mALF, f), dsub(dc(1.0), z4)), 1.0 / cfzz));
  return dmul(nn, EC);
}

/* c_gga(na, nb, g2, beta) of pbe_ggcxc_module.f90:96-140 */
static dual pbe_cgga(dual na, dual nb, dual g2)
{
  const double beta = 0.066725, gamma = 0.03109069086965489503494086371;
  const dual n = dadd(na, nb);
  if (n.v <= 0.0) return dc(0.0);
  const dual ELDAC = ddiv(pw92_clda(na, nb), n);
  dual za = ddiv(na, n), zb = ddiv(nb, n);
  if (za.v < DEPS) za = dc(DEPS);
  if (zb.v < DEPS) zb = dc(DEPS);
  const dual fi = dscale(dadd(dpowc(za, 2.0 / 3.0), dpowc(zb, 2.0 / 3.0)), 1.0 / pow(2.0, 1.0 / 3.0));
  const dual fi2 = dmul(fi, fi), fi3 = dmul(fi2, fi);
  const dual A = ddiv(dc(beta / gamma),
                      dadd(dsub(dexp(ddiv(dscale(ELDAC, -1.0), dscale(fi3, gamma))), dc(1.0)), dc(DEPS)));
  const dual t2 = dscale(ddiv(g2, dadd(dmul(fi2, dpowc(n, 7.0 / 3.0)), dc(DEPS))),
                         PI / (4.0 * 4.0 * pow(3.0 * PI * PI, 1.0 / 3.0)));
  const dual x = dmul(A, t2);
  const dual rat = ddiv(dadd(dc(1.0), x), dadd(dadd(dc(1.0), x), dmul(x, x)));
  const dual H = dmul(dscale(fi3, gamma), dlog(dadd(dc(1.0), dmul(dscale(t2, beta / gamma), rat))));
  return dmul(n, H);
}

/* PW92 correlation, unrestricted (pw_ldac iop = PWLDAC_UNRESTR,
 * :222-245): F, dF/dna, dF/dnb */
void or_pw92_spin(int n, const double *na, const double *nb, double *f, double *fa, double *fb)
{
  for (int i = 0; i < n; ++i) {
    const dual F = pw92_clda(dv(na[i], 0), dv(nb[i], 1));
    f[i] = F.v;
    fa[i] = F.d[0];
    fb[i] = F.d[1];
  }
}

/* PBE correlation, ispin = 2 (pbe_ggcc, :301-316): F, dF/dna, dF/dnb and
 * dF/dg2 = dF/dgaa = dF/dgbb = dF/dgab / 2 (g2 = gaa + gbb + 2 gab) */
void or_pbec_spin(int n, const double *na, const double *nb, const double *gaa,
                  const double *gbb, const double *gab, double *f, double *fa, double *fb,
                  double *fg2)
{
  for (int i = 0; i < n; ++i) {
    const dual F = pbe_cgga(dv(na[i], 0), dv(nb[i], 1), dv(gaa[i] + gbb[i] + 2.0 * gab[i], 2));
    f[i] = F.v;
    fa[i] = F.d[0];
    fb[i] = F.d[1];
    fg2[i] = F.d[2];
  }
}

/* The functional for ispin = 2 (xc_functionals, xc_func.f90:187-305): per
 * point f, fr[2] = (dF/dna, dF/dnb), fg[3] = (dF/dgaa, dF/dgbb, dF/dgab).
 * Exchange by spin scaling (exchange.f90:253-288, :511-601):
 * F = 1/2 Fx(2 n_s, 4 g_ss) summed over s, dF/dn_s = Fx'(2 n_s),
 * dF/dg_ss = 2 dFx/dg(2 n_s, 4 g_ss). */
int or_functional_spin(int mask, int n, const double *na, const double *nb, const double *gaa,
                       const double *gbb, const double *gab, double *f, double *fr, double *fg)
{
  if (mask & ~((1 << XC_XALPHA) | (1 << XC_PWLDAC) | (1 << XC_PBEX) | (1 << XC_PBEC)))
    return -1;
  if ((mask & (1 << XC_PBEC)) && !(mask & (1 << XC_PWLDAC))) return -2;
  for (int i = 0; i < n; ++i) {
    f[i] = 0;
    fr[2 * i] = fr[2 * i + 1] = 0;
    fg[3 * i] = fg[3 * i + 1] = fg[3 * i + 2] = 0;
  }
  for (int term = 1; term <= XC_PBEC; ++term) {
    if (!(mask & (1 << term))) continue;
    for (int i = 0; i < n; ++i) {
      double x = 0, xa = 0, xb = 0, ga = 0, gb = 0, gab_ = 0;
      if (term == XC_XALPHA || term == XC_PBEX) {
        for (int s = 0; s < 2; ++s) {
          const double r2 = 2.0 * (s ? nb[i] : na[i]), g4 = 4.0 * (s ? gbb[i] : gaa[i]);
          double e = 0, er = 0, eg = 0;
          if (term == XC_XALPHA) or_xalpha(1, &r2, &e, &er);
          else or_pbex(1, &r2, &g4, &e, &er, &eg);
          x += 0.5 * e;
          if (s) { xb = er; gb = 2.0 * eg; } else { xa = er; ga = 2.0 * eg; }
        }
      } else if (term == XC_PWLDAC) {
        or_pw92_spin(1, na + i, nb + i, &x, &xa, &xb);
      } else {
        double g2;
        or_pbec_spin(1, na + i, nb + i, gaa + i, gbb + i, gab + i, &x, &xa, &xb, &g2);
        ga = g2;
        gb = g2;
        gab_ = 2.0 * g2;
      }
      f[i] = f[i] + x * 1.0;
      fr[2 * i] = fr[2 * i] + xa * 1.0;
      fr[2 * i + 1] = fr[2 * i + 1] + xb * 1.0;
      fg[3 * i] = fg[3 * i] + ga * 1.0;
      fg[3 * i + 1] = fg[3 * i + 1] + gb * 1.0;
      fg[3 * i + 2] = fg[3 * i + 2] + gab_ * 1.0;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* the XC build                                                              */
/* ------------------------------------------------------------------------ */

/* Y[j*ld+p] = sum_f M[f + j*n] X[f*ld+p], j < ncol  (dgemm 'n','n')        */
static void gemm_nn(int vl, int n, int ncol, const double *X, const double *M,
                    double *Y)
{
  for (int j = 0; j < ncol; ++j) {
    double *y = Y + (size_t)j * vl;
    for (int p = 0; p < vl; ++p) y[p] = 0.0;
    const double *m = M + (size_t)j * n;
    for (int f = 0; f < n; ++f) {
      const double c = m[f];
      if (c == 0.0) continue;
      const double *xf = X + (size_t)f * vl;
      for (int p = 0; p < vl; ++p) y[p] += xf[p] * c;
    }
  }
}

/* H[i + j*n] = sum_p W[i*vl+p] F[j*vl+p]   (dgemm 't','n') */
static void gemm_tn(int vl, int n, const double *W, const double *F, double *H)
{
  int i = 0;
  for (; i + 4 <= n; i += 4) {
    int j = 0;
    for (; j + 4 <= n; j += 4) {
      double acc[4][4] = {{0}};
      for (int p = 0; p < vl; ++p) {
        const double w0 = W[(size_t)(i + 0) * vl + p], w1 = W[(size_t)(i + 1) * vl + p];
        const double w2 = W[(size_t)(i + 2) * vl + p], w3 = W[(size_t)(i + 3) * vl + p];
        for (int b = 0; b < 4; ++b) {
          const double fv = F[(size_t)(j + b) * vl + p];
          acc[0][b] += w0 * fv; acc[1][b] += w1 * fv;
          acc[2][b] += w2 * fv; acc[3][b] += w3 * fv;
        }
      }
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) H[(size_t)(i + a) + (size_t)(j + b) * n] = acc[a][b];
    }
    for (; j < n; ++j)
      for (int a = 0; a < 4; ++a) {
        double s = 0;
        for (int p = 0; p < vl; ++p) s += W[(size_t)(i + a) * vl + p] * F[(size_t)j * vl + p];
        H[(size_t)(i + a) + (size_t)j * n] = s;
      }
  }
  for (; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int p = 0; p < vl; ++p) s += W[(size_t)i * vl + p] * F[(size_t)j * vl + p];
      H[(size_t)i + (size_t)j * n] = s;
    }
}

/*
 * Full XC build over contiguous batches of <= veclen points, ispin = 1 or 2.
 *   grid     (npts,4) column-major: x, y, z, w   (grdpts layout)
 *   dform    0: dens = P (n x n x ispin, column-major);
 *            1: dens = C_occ (n x nocc x ispin), occ (nocc x ispin)
 *   ham      packed lower triangle, n(n+1)/2, (j,k<=j) row order, spins
 *            interleaved (alpha / beta at even / odd positions,
 *            modules/xc_hamiltonian.f90:1455-1460); ADDED INTO
 *   batch_lo/hi   process only batches [lo, hi) (for bounded CPU samples)
 *   nthreads  contiguous static split of the batch range over threads
 */
static int xc_build_impl(int ispin, int n_atoms, const double *atom_xyz, int n_shells,
                         const int *shell_info, const double *exps, const double *norms,
                         const double *contr, const double *grid, int64_t npts, int veclen,
                         int mask, int dform, const double *dens, int nocc,
                         const double *occ, double *ham, double *exc, double *charge,
                         int64_t batch_lo, int64_t batch_hi, int nthreads)
{
  or_basis b = {n_atoms, atom_xyz, n_shells, shell_info, exps, norms, contr};
  const int n = or_nbas(n_shells, shell_info);
  const int gga = (mask & ((1 << XC_PBEX) | (1 << XC_PBEC))) != 0;
  if (mask & ~((1 << XC_XALPHA) | (1 << XC_PWLDAC) | (1 << XC_PBEX) | (1 << XC_PBEC)))
    return -1;
  if ((mask & (1 << XC_PBEC)) && !(mask & (1 << XC_PWLDAC))) return -2;
  const int64_t nbatch = (npts + veclen - 1) / veclen;
  if (batch_hi < 0 || batch_hi > nbatch) batch_hi = nbatch;
  if (batch_lo < 0) batch_lo = 0;
  const int ncol = dform == 0 ? n : nocc;
  const size_t npk = (size_t)n * (n + 1) / 2, nh = (size_t)ispin * npk;
  if (ispin != 1 && ispin != 2) return -3;
  if (nthreads < 1) nthreads = 1;
  double *hp = calloc((size_t)nthreads * nh, sizeof(double));
  double *ep = calloc((size_t)nthreads * 2, sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0, nt = 1;
#ifdef _OPENMP
    tid = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    const int64_t nb = batch_hi - batch_lo;
    const int64_t b0 = batch_lo + nb * tid / nt, b1 = batch_lo + nb * (tid + 1) / nt;
    double *phi = malloc(sizeof(double) * (size_t)n * veclen);
    double *grd = malloc(sizeof(double) * (size_t)n * 3 * veclen);
    double *gx = malloc(sizeof(double) * (size_t)n * veclen);
    double *gy = malloc(sizeof(double) * (size_t)n * veclen);
    double *gz = malloc(sizeof(double) * (size_t)n * veclen);
    double *X = malloc(sizeof(double) * (size_t)ncol * veclen);
    double *XG = malloc(sizeof(double) * (size_t)ncol * veclen * 3);
    double *help = malloc(sizeof(double) * (size_t)n * veclen);
    double *H = malloc(sizeof(double) * (size_t)n * n);
    /* per spin s: rho[s], gr[s][3]; gam[aa, bb, ab]; f; fr[2]; fg[3]; A; B[3];
       the spin functional's interleaved fr (2 rows) and fg (3 rows) */
    double *rho = malloc(sizeof(double) * veclen * 26);
    double *gr = rho + 2 * veclen, *gam = rho + 8 * veclen, *f = rho + 11 * veclen;
    double *fr = rho + 12 * veclen, *fg = rho + 14 * veclen, *A = rho + 17 * veclen;
    double *B = rho + 18 * veclen, *fri = rho + 21 * veclen, *fgi = rho + 23 * veclen;
    double *hv = hp + (size_t)tid * nh;
    double e_loc = 0.0, q_loc = 0.0;
    for (int64_t bt = b0; bt < b1; ++bt) {
      const int64_t p0 = bt * veclen;
      const int vl = (int)((npts - p0) < veclen ? (npts - p0) : veclen);
      const double *pts = grid + p0, *w = grid + 3 * npts + p0;
      basis_eval(&b, vl, pts, (long)npts, phi, grd);
      for (int fi = 0; fi < n; ++fi)
        for (int p = 0; p < vl; ++p) {
          gx[(size_t)fi * vl + p] = grd[((size_t)fi * 3 + 0) * vl + p];
          gy[(size_t)fi * vl + p] = grd[((size_t)fi * 3 + 1) * vl + p];
          gz[(size_t)fi * vl + p] = grd[((size_t)fi * 3 + 2) * vl + p];
        }
      /* density per spin (modules/density_calc_module.f90:1494-1577) */
      for (int sp = 0; sp < ispin; ++sp) {
        const double *dm = dens + (size_t)sp * n * ncol;
        const double *oc_s = occ + (size_t)sp * nocc;
        double *rs = rho + (size_t)sp * veclen, *gs = gr + (size_t)sp * 3 * veclen;
        for (int p = 0; p < vl; ++p) { rs[p] = 0; gs[p] = gs[veclen + p] = gs[2 * veclen + p] = 0; }
        gemm_nn(vl, n, ncol, phi, dm, X);
        if (dform == 1) {
          if (gga) {
            gemm_nn(vl, n, ncol, gx, dm, XG);
            gemm_nn(vl, n, ncol, gy, dm, XG + (size_t)ncol * vl);
            gemm_nn(vl, n, ncol, gz, dm, XG + (size_t)2 * ncol * vl);
          }
          for (int j = 0; j < nocc; ++j) {
            const double oc = oc_s[j];
            const double *ph = X + (size_t)j * vl;
            for (int p = 0; p < vl; ++p) rs[p] = rs[p] + ph[p] * ph[p] * oc;
            if (gga)
              for (int c = 0; c < 3; ++c) {
                const double *gp = XG + ((size_t)c * ncol + j) * vl;
                for (int p = 0; p < vl; ++p)
                  gs[c * veclen + p] = gs[c * veclen + p] + 2 * ph[p] * gp[p] * oc;
              }
          }
        } else {
          for (int k = 0; k < n; ++k) {
            const double *xk = X + (size_t)k * vl, *pk = phi + (size_t)k * vl;
            for (int p = 0; p < vl; ++p) rs[p] += pk[p] * xk[p];
            if (gga) {
              const double *g0 = gx + (size_t)k * vl, *g1 = gy + (size_t)k * vl,
                           *g2 = gz + (size_t)k * vl;
              for (int p = 0; p < vl; ++p) {
                gs[p] += 2 * g0[p] * xk[p];
                gs[veclen + p] += 2 * g1[p] * xk[p];
                gs[2 * veclen + p] += 2 * g2[p] * xk[p];
              }
            }
          }
        }
      }
      /* gamma: |grad rho|^2, or (aa, bb, ab) */
      for (int g = 0; g < 2 * ispin - 1; ++g) {
        const double *u = gr + (size_t)(g == 1 ? 3 : 0) * veclen;
        const double *v = gr + (size_t)(g == 0 ? 0 : 3) * veclen;
        for (int p = 0; p < vl; ++p)
          gam[g * veclen + p] = u[p] * v[p] + u[veclen + p] * v[veclen + p] +
                                u[2 * veclen + p] * v[2 * veclen + p];
      }
      double qs = 0.0;
      if (ispin == 1)
        for (int p = 0; p < vl; ++p) qs += rho[p] * w[p];
      else
        for (int p = 0; p < vl; ++p) qs += (rho[p] + rho[veclen + p]) * w[p];
      q_loc = q_loc + qs;
      if (ispin == 1) {
        or_functional(mask, vl, rho, gam, f, fr, fg);
      } else {
        or_functional_spin(mask, vl, rho, rho + veclen, gam, gam + veclen, gam + 2 * veclen,
                           f, fri, fgi);
        for (int p = 0; p < vl; ++p) {
          fr[p] = fri[2 * p];
          fr[veclen + p] = fri[2 * p + 1];
          for (int g = 0; g < 3; ++g) fg[g * veclen + p] = fgi[3 * p + g];
        }
      }
      double es = 0.0;
      for (int p = 0; p < vl; ++p) es += f[p] * w[p];
      e_loc = e_loc + es;
      for (int sp = 0; sp < ispin; ++sp) {
        const double *frs = fr + (size_t)sp * veclen;
        if (gga) {
          /* A_SCAL, B_VECT, orbs_help (modules/xc_hamiltonian.f90:1429-1483) */
          const double *gs = gr + (size_t)sp * 3 * veclen, *go = gr + (size_t)(1 - sp) * 3 * veclen;
          const double *fss = fg + (size_t)sp * veclen, *fab = fg + 2 * (size_t)veclen;
          for (int p = 0; p < vl; ++p) {
            A[p] = 0.5 * frs[p] * w[p];
            for (int c = 0; c < 3; ++c)
              B[c * veclen + p] = ispin == 1
                  ? w[p] * 2.0 * fss[p] * gs[c * veclen + p]
                  : w[p] * (2.0 * fss[p] * gs[c * veclen + p] + fab[p] * go[c * veclen + p]);
          }
          for (int j = 0; j < n; ++j)
            for (int p = 0; p < vl; ++p) {
              const size_t q = (size_t)j * vl + p;
              help[q] = A[p] * phi[q] + B[p] * gx[q] + B[veclen + p] * gy[q] +
                        B[2 * veclen + p] * gz[q];
            }
          gemm_tn(vl, n, help, phi, H);
          /* remap: ham_xc_arr(counter) += ham(j,k) + ham(k,j), k <= j */
          size_t cnt = 0;
          for (int j = 0; j < n; ++j)
            for (int k = 0; k <= j; ++k, ++cnt) {
              const size_t o = cnt * ispin + sp;
              hv[o] = hv[o] + (H[(size_t)j + (size_t)k * n] + H[(size_t)k + (size_t)j * n]);
            }
        } else {
          /* integrate + tristore (modules/xc_hamiltonian.f90:987-1034) */
          for (int j = 0; j < n; ++j)
            for (int p = 0; p < vl; ++p) {
              const size_t q = (size_t)j * vl + p;
              help[q] = phi[q] * (frs[p] * w[p]);
            }
          gemm_tn(vl, n, help, phi, H);
          size_t cnt = 0;
          for (int j = 0; j < n; ++j)
            for (int k = 0; k <= j; ++k, ++cnt) {
              const size_t o = cnt * ispin + sp;
              hv[o] = hv[o] + H[(size_t)k + (size_t)j * n];
            }
        }
      }
    }
    ep[2 * tid] = e_loc;
    ep[2 * tid + 1] = q_loc;
    free(phi); free(grd); free(gx); free(gy); free(gz); free(X); free(XG);
    free(help); free(H); free(rho);
  }
  /* the reduction xc_sndrcv would do, in rank order */
  for (int t = 0; t < nthreads; ++t) {
    const double *hv = hp + (size_t)t * nh;
    for (size_t q = 0; q < nh; ++q) ham[q] += hv[q];
    *exc += ep[2 * t];
    *charge += ep[2 * t + 1];
  }
  free(hp);
  free(ep);
  return 0;
}

int or_xc_build(int n_atoms, const double *atom_xyz, int n_shells,
                const int *shell_info, const double *exps, const double *norms,
                const double *contr, const double *grid, int64_t npts, int veclen,
                int mask, int dform, const double *dens, int nocc,
                const double *occ, double *ham, double *exc, double *charge,
                int64_t batch_lo, int64_t batch_hi, int nthreads)
{
  return xc_build_impl(1, n_atoms, atom_xyz, n_shells, shell_info, exps, norms, contr, grid,
                       npts, veclen, mask, dform, dens, nocc, occ, ham, exc, charge, batch_lo,
                       batch_hi, nthreads);
}

/* ispin = 2: dens and occ hold the alpha then the beta spin, ham 2 n(n+1)/2
 * interleaved; exc and charge over both spins */
int or_xc_build_spin(int n_atoms, const double *atom_xyz, int n_shells,
                     const int *shell_info, const double *exps, const double *norms,
                     const double *contr, const double *grid, int64_t npts, int veclen,
                     int mask, int dform, const double *dens, int nocc,
                     const double *occ, double *ham, double *exc, double *charge,
                     int64_t batch_lo, int64_t batch_hi, int nthreads)
{
  return xc_build_impl(2, n_atoms, atom_xyz, n_shells, shell_info, exps, norms, contr, grid,
                       npts, veclen, mask, dform, dens, nocc, occ, ham, exc, charge, batch_lo,
                       batch_hi, nthreads);
}

/* ------------------------------------------------------------------------ */
/* the integration grid (modules/grid_module.f90)                            */
/* ------------------------------------------------------------------------ */

/* Slater radius in bohr, Z = 1..10 (rslater_f :1029-1079: angstrom values of
 * :1051-1054 times angstrom = 1/bohr, modules/constants.f90:105-108) */
double or_slater_radius(int z)
{
  static const double ra[11] = {0.0, 0.53, 0.25, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 0.50};
  const double angstrom = 1.0 / 0.52917706;
  if (z < 1 || z > 10) return -1.0;
  return ra[z] * angstrom;
}

/* RADIALGRID (:1246-1301), Pople grid: the largest i <= nradpts with
 * R (i/(n+1-i))^2 < 40 bounds the shells; r, w hold nradpts entries;
 * returns the number of shells */
int or_radial_grid(int nradpts, double rcur, double *r, double *w)
{
  const double rad_limit = 40.0, rhn = nradpts + 1;
  int i;
  double rr = 0.0;
  for (i = nradpts; i >= 1; --i) {
    const double rhni = rhn - i, x = i / rhni;
    rr = rcur * (x * x);
    if (rr < rad_limit) break;
  }
  const int n = i;
  for (int k = n; k >= 1; --k) {
    const double rhni = rhn - k, x = k / rhni, x2 = x * x;
    r[k - 1] = rcur * x2;
    w[k - 1] = 2 * (rcur * rcur * rcur) * rhn * (x2 * x2 * x) / (rhni * rhni);
  }
  return n;
}

/* INIGRD orbit generators (:1826-2098); each appends its points with
 * weight wt to p (column-major n x 4: x, y, z, w) at row *np */
static void lb_add(double *p, int ntot, int *np, double x, double y, double z, double wt)
{
  p[*np] = x;
  p[ntot + *np] = y;
  p[2 * ntot + *np] = z;
  p[3 * ntot + *np] = wt;
  ++*np;
}

static void lb_a1(double *p, int nt, int *np, double wt)
{
  lb_add(p, nt, np, 0, 0, 1, wt); lb_add(p, nt, np, 0, 0, -1, wt);
  lb_add(p, nt, np, 0, 1, 0, wt); lb_add(p, nt, np, 0, -1, 0, wt);
  lb_add(p, nt, np, 1, 0, 0, wt); lb_add(p, nt, np, -1, 0, 0, wt);
}

static void lb_a2(double *p, int nt, int *np, double wt)
{
  const double s = 1.0 / sqrt(2.0);
  const double sg[4][2] = {{s, s}, {s, -s}, {-s, s}, {-s, -s}};
  for (int k = 0; k < 4; ++k) lb_add(p, nt, np, sg[k][0], sg[k][1], 0, wt);
  for (int k = 0; k < 4; ++k) lb_add(p, nt, np, sg[k][0], 0, sg[k][1], wt);
  for (int k = 0; k < 4; ++k) lb_add(p, nt, np, 0, sg[k][0], sg[k][1], wt);
}

static void lb_a3(double *p, int nt, int *np, double wt)
{
  const double t = 1.0 / sqrt(3.0);
  for (int i = -1; i <= 1; i += 2)
    for (int j = -1; j <= 1; j += 2)
      for (int k = -1; k <= 1; k += 2) lb_add(p, nt, np, i * t, j * t, k * t, wt);
}

static void lb_b(double *p, int nt, int *np, double m, double wt)
{
  const double l = sqrt((1.0 - m * m) / 2.0);
  for (int v = 0; v < 3; ++v)  /* N1, N2, N3 blocks of 8 */
    for (int i = -1; i <= 1; i += 2)
      for (int j = -1; j <= 1; j += 2)
        for (int k = -1; k <= 1; k += 2) {
          const double a = v == 2 ? m : l, b = v == 1 ? m : l, c = v == 0 ? m : l;
          lb_add(p, nt, np, i * a, j * b, k * c, wt);
        }
}

static void lb_c(double *p, int nt, int *np, double vv, double wt)
{
  const double d = sqrt(1.0 - 4.0 * vv);
  const double p1 = sqrt((1.0 + d) / 2.0), q1 = sqrt((1.0 - d) / 2.0);
  const double pr[2][2] = {{p1, q1}, {q1, p1}};
  const double sg[4][2] = {{1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
  for (int h = 0; h < 2; ++h) {
    const double a = pr[h][0], b = pr[h][1];
    for (int k = 0; k < 4; ++k) lb_add(p, nt, np, sg[k][0] * a, sg[k][1] * b, 0, wt);
    for (int k = 0; k < 4; ++k) lb_add(p, nt, np, sg[k][0] * a, 0, sg[k][1] * b, wt);
    for (int k = 0; k < 4; ++k) lb_add(p, nt, np, 0, sg[k][0] * a, sg[k][1] * b, wt);
  }
}

static void lb_d(double *p, int nt, int *np, double u, double w, double wt)
{
  const double r = sqrt(1.0 - u * u - w * w);
  const int first = *np;
  for (int v = 0; v < 3; ++v)
    for (int i = -1; i <= 1; i += 2)
      for (int j = -1; j <= 1; j += 2)
        for (int k = -1; k <= 1; k += 2) {
          const double a = v == 0 ? u : v == 1 ? w : r;
          const double b = v == 0 ? w : v == 1 ? r : u;
          const double c = v == 0 ? r : v == 1 ? u : w;
          lb_add(p, nt, np, i * a, j * b, k * c, wt);
        }
  for (int q = 0; q < 24; ++q)  /* DK(25:48) = DK(1:24) with y and z swapped */
    lb_add(p, nt, np, p[first + q], p[2 * nt + first + q], p[nt + first + q], wt);
}

/* INIGRD (:1359-2100) for ICGRD = 231 (L = 23, 194 points, :1698-1751) and
 * 291 (L = 29, 302 points, :1753-1814); returns the number of points, or
 * -1 for another ICGRD.  p: n x 4 column-major (x, y, z, weight). */
int or_lebedev(int icgrd, double *p)
{
  int np = 0;
  if (icgrd == 231) {
    const int nt = 194;
    lb_a1(p, nt, &np, 1.78234044724E-03);
    lb_a2(p, nt, &np, 5.71690594998E-03);
    lb_a3(p, nt, &np, 5.57338317884E-03);
    lb_b(p, nt, &np, 0.777493219315, 5.51877146727E-03);
    lb_b(p, nt, &np, 0.912509096867, 5.15823771181E-03);
    lb_b(p, nt, &np, 0.314196994183, 5.60870408259E-03);
    lb_b(p, nt, &np, 0.982972302707, 4.10677702817E-03);
    lb_c(p, nt, &np, 2.00 / 19.0, 5.05184606462E-03);
    lb_d(p, nt, &np, 0.159041710538, 0.525118572443, 5.53024891623E-03);
    return np;
  }
  if (icgrd == 291) {
    const int nt = 302;
    lb_a1(p, nt, &np, 8.54591172878E-04);
    lb_a3(p, nt, &np, 3.59911928502E-03);
    lb_b(p, nt, &np, 0.129238672710, 3.65004580768E-03);
    lb_b(p, nt, &np, 0.371034178385, 3.60482260142E-03);
    lb_b(p, nt, &np, 0.743452042987, 3.57672966173E-03);
    lb_b(p, nt, &np, 0.867643624544, 3.44978842429E-03);
    lb_b(p, nt, &np, 0.949454317226, 3.10895312238E-03);
    lb_b(p, nt, &np, 0.990705621379, 2.35210141366E-03);
    lb_c(p, nt, &np, 0.220093335298, 3.60082093222E-03);
    lb_c(p, nt, &np, 0.650272754659E-01, 2.98234496317E-03);
    lb_d(p, nt, &np, 0.800072749407, 0.544867737258, 3.57154055427E-03);
    lb_d(p, nt, &np, 0.902442529533, 0.412772408317, 3.39231220501E-03);
    return np;
  }
  return -1;
}

/* One atomic grid (grid_main_calc :407-433, no pruning): shells of
 * RADIALGRID times INIGRD's points, positions ANG * r + centre, weights
 * w_ang * w_rad * 16 atan(1).  xyzw: (cap x 4) column-major; returns the
 * number of points, or -1 if cap is too small or icgrd unknown. */
long or_atomic_grid(int nradpts, double rcur, int icgrd, const double *centre, long cap,
                    double *xyzw)
{
  double ang[302 * 4];
  const int na = or_lebedev(icgrd, ang);
  if (na < 0) return -1;
  double *r = malloc(sizeof(double) * (size_t)(nradpts + 1));
  double *w = malloc(sizeof(double) * (size_t)(nradpts + 1));
  const int nr = or_radial_grid(nradpts, rcur, r, w);
  const long n = (long)nr * na;
  if (n > cap) { free(r); free(w); return -1; }
  const double four_pi = 16.0 * atan(1.0);
  for (int i = 0; i < nr; ++i)
    for (int a = 0; a < na; ++a) {
      const long q = (long)i * na + a;
      for (int c = 0; c < 3; ++c) xyzw[c * cap + q] = ang[c * na + a] * r[i] + centre[c];
      xyzw[3 * cap + q] = ang[3 * na + a] * w[i] * four_pi;
    }
  free(r);
  free(w);
  return n;
}

/* becke0 (:2442-2497) with art = 0 (weight_grads = T), mu0 of
 * autodiff/becke_step_func.f90:84-124 at a = 0: the partition weight of
 * host atom a (0-based) at each of npts points x (npts x 3 column-major,
 * leading dimension ld); atm (3 x natm).  Pair distances as atomicweight2
 * precomputes them (:2326-2333). */
void or_becke0(int natm, const double *atm, int a, long npts, long ld, const double *x,
               double *wout, int nthreads)
{
  double *dst = malloc(sizeof(double) * (size_t)natm * natm);
  for (int j = 0; j < natm; ++j)
    for (int i = 0; i <= j; ++i) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) {
        const double d = atm[3 * i + c] - atm[3 * j + c];
        s += d * d;
      }
      dst[(size_t)i * natm + j] = dst[(size_t)j * natm + i] = sqrt(s);
    }
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    double *r = malloc(sizeof(double) * (size_t)natm);
    double *p = malloc(sizeof(double) * (size_t)natm);
#pragma omp for schedule(static)
    for (long q = 0; q < npts; ++q) {
      for (int i = 0; i < natm; ++i) {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) {
          const double v = atm[3 * i + c] - x[c * ld + q];
          s += v * v;
        }
        r[i] = sqrt(s);
        p[i] = 1.0;
      }
      for (int i = 0; i < natm; ++i)
        for (int k = 0; k < i; ++k) {
          double mu = (r[i] - r[k]) / dst[(size_t)k * natm + i];
          mu = (mu - 0.0) / (1.0 - 0.0 * mu);
          for (int it = 0; it < 3; ++it) mu = mu * (3.0 - mu * mu) / 2.0;
          p[i] = p[i] * ((1 - mu) / 2);
          p[k] = p[k] * ((1 + mu) / 2);
        }
      double sum = 0.0;
      for (int i = 0; i < natm; ++i) sum += p[i];
      wout[q] = p[a] / sum;
    }
    free(r);
    free(p);
  }
  free(dst);
}

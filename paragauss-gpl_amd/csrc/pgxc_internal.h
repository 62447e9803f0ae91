// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pgxc {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kLMax = 4;      // highest angular momentum of a shell
constexpr int kMaxContr = 8;  // contracted functions per shell
constexpr int kThreads = 256; // 4 wave64 per workgroup everywhere
constexpr double kFar = 1.0e6;  // coordinate of padding points (all exps cut off)

// One shell = one (atom, l) block of unique_atom_basis_type.
struct ShellD {
  int atom, l, n_exp, n_unc, n_contr, exp_off, contr_off, func_off;
};

// One centre with the contiguous shell range it carries.
struct AtomD {
  double x, y, z;
  int shell_begin, shell_end, lmax, pad;
};

// Per-point state kept across builds (incremental mode), SoA of length npad.
enum StateSlot { S_RHO = 0, S_GX, S_GY, S_GZ, S_A, S_BX, S_BY, S_BZ, S_NSLOT };

// ---- launchers (kernels live in xc_kernels.hip) -------------------------
// Basis values (+ gradients) of every function on points [p0, p0+Mc) into
// phi[c][f][p] (c = value, d/dx, d/dy, d/dz), f < nbas; rows nbas..NP-1 untouched.
template <typename T>
void launch_basis(hipStream_t s, bool grad, const double *gx, const double *gy,
                  const double *gz, long p0, long Mc, int n_atoms, const AtomD *atoms,
                  const ShellD *shells, const double *exps, const double *norms,
                  const double *contr, T *phi, long NP);

// X = Phi D (D = P or dP, NP x NP); per-point partial sums over each
// 128-function column tile: part[t][c][p], c = rho, 2 X.dPhi/dx, ...
// tf = function tile (64, 96 or 128), NP a multiple of it
template <typename T>
void launch_density(hipStream_t s, bool grad, const T *phi, const T *D, long Mc, int NP, int tf,
                    double *part);
int density_tiles(int NP, int tf);

// Functional on points of one chunk: sums the partials (in tile order), adds
// the stored state when incremental, evaluates f, f_rho, f_gamma, forms the
// potential coefficients, updates the state, writes block partials of
// (sum w f, sum w rho) to epart[blk_off + b].
void launch_functional(hipStream_t s, bool gga, bool incremental, uint32_t mask,
                       const double *part, int ntiles, long Mc, const double *w,
                       long p0, double *state, long npad, double *coef,
                       double *epart, long blk_off);
long functional_blocks(long Mc);

// W[f][p] = a phi0 + b . grad phi (T precision)
template <typename T>
void launch_wmat(hipStream_t s, bool grad, const T *phi, const double *coef, long Mc,
                 int NP, T *W);

// Hpart[sl][i][j] (+)= sum_{p in split sl} W[i][p] phi0[j][p]
template <typename T>
void launch_potential(hipStream_t s, const T *W, const T *phi, long Mc, int NP, int tf,
                      int nsplit, double *Hpart, bool first);

// split-BF16 GEMMs read FP32 operands and split while staging
void launch_density_bf16x3(hipStream_t s, bool grad, const float *phi, const float *D,
                           long Mc, int NP, int tf, double *part);
void launch_potential_bf16x3(hipStream_t s, const float *W, const float *phi, long Mc,
                             int NP, int tf, int nsplit, double *Hpart, bool first);

// V (packed, n(n+1)/2) = base + sum_sl (H[sl] + H[sl]^T) restricted to n x n
void launch_pack(hipStream_t s, const double *Hpart, int nsplit, int NP, int n,
                 const double *base, double *Vpacked);

// P = C diag(occ) C^T into an NP x NP padded array
void launch_form_density(hipStream_t s, const double *C, const double *occ, int n,
                         int nocc, int NP, double *P);

// out = a - b (NP*NP), and a cast of a double array to T
void launch_sub(hipStream_t s, const double *a, const double *b, long n, double *out);
template <typename T>
void launch_cast(hipStream_t s, const double *in, long n, T *out);

// Becke partition weights of points owned by atoms (grid generation)
void launch_becke(hipStream_t s, int na, const double *axyz, const double *inv_dist, long npts,
                  const double *pts, const int *owner, const double *wat, double *wout);

}  // namespace pgxc

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

// ---- work lists ----------------------------------------------------------
// A unit is one grid chunk of one molecule: Mc points (multiple of 256) with
// the molecule's basis, its density operand and the chunk's work buffers.
struct UnitD {
  const double *gx, *gy, *gz, *gw;  // grid of the unit's points (offset to its first point)
  const AtomD *atoms;
  const ShellD *shells;
  const double *exps, *norms, *contr;
  void *phi;        // [4][NP][Mc] values, d/dx, d/dy, d/dz (T)
  void *W;          // [NP][Mc] (T)
  const void *D;    // [NP][NP] density operand (T)
  double *part;     // [NP/tf][4][Mc] density partials
  double *coef;     // [4][Mc] potential coefficients a, b
  double *state;    // [S_NSLOT][npad] molecule state, offset to the unit's first point
  double *epart;    // [Mc/256][2] block sums of w f and w rho (offset to the unit)
  double *Hpart;    // [nsplit][NP][NP] potential partials of the molecule
  long Mc, npad;
  int n_atoms, NP, ntiles, nsplit, first, incremental;
  uint32_t mask;
  int pad;
};

// Per-molecule descriptors of the operand preparation and the final packing.
struct PackD {
  const double *Hpart;
  double *Vout, *Vstate;
  const double *P;
  double *Pref;
  void *D;
  long n, NP;
  int nsplit, incremental;
};

// One launch's items: prefix offsets base[0..nu] over units, optional order.
struct StageList {
  const UnitD *units;
  int nu;
  const long *base;
  const int *order;
  long total;
};

// ---- launchers (kernels live in xc_kernels.hip / xc_bf16x3.hip) -----------
template <typename T>
void launch_basis(hipStream_t s, bool grad, const StageList &L);
// tf = function tile (64, 96 or 128) shared by every unit of the list;
// var = contraction variant: 0 (K step 16, 2 LDS stages), 1 (32, 1), 2 (32, 2)
constexpr int kNumVariants = 3;
template <typename T>
void launch_density(hipStream_t s, bool grad, int tf, int var, const StageList &L);
void launch_functional(hipStream_t s, bool gga, const StageList &L);
template <typename T>
void launch_wmat(hipStream_t s, bool grad, const StageList &L);
template <typename T>
void launch_potential(hipStream_t s, int tf, int var, const StageList &L);
void launch_density_bf16x3(hipStream_t s, bool grad, int tf, const StageList &L);
void launch_potential_bf16x3(hipStream_t s, int tf, const StageList &L);
void launch_pack(hipStream_t s, const PackD *mols, int nm, const long *base, long total);
void launch_prep(hipStream_t s, const PackD *mols, int nm, const long *base, long total,
                 bool to_float, bool finish);

// item counts per unit, shared by the planner and the kernels' decoding
inline long items_basis(long Mc, int n_atoms) { return (Mc / kThreads) * n_atoms; }
inline long items_density(long Mc, int NP, int tf) { return (Mc / 128) * (NP / tf); }
inline long items_functional(long Mc) { return Mc / kThreads; }
inline long items_wmat(long Mc, int NP) { return (Mc / kThreads) * (NP / 16); }
inline long items_potential(int NP, int tf, int nsplit) { return (long)(NP / tf) * (NP / tf) * nsplit; }
// points per K step of the potential contraction across all variants
constexpr int kSplitStep = 32;

// P = C diag(occ) C^T into an NP x NP padded array
void launch_form_density(hipStream_t s, const double *C, const double *occ, int n,
                         int nocc, int NP, double *P);


// Becke partition weights of points owned by atoms (grid generation)
void launch_becke(hipStream_t s, int na, const double *axyz, const double *inv_dist, long npts,
                  const double *pts, const int *owner, const double *wat, double *wout);

}  // namespace pgxc

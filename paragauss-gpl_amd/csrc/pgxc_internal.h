// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct pgxc_costmodel;

namespace pgxc {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kLMax = 4;      // highest angular momentum of a shell
constexpr int kMaxContr = 8;  // contracted functions per shell
// per atom at most: shells, primitives and contraction coefficients
constexpr int kAtomShells = 16, kAtomPrims = 256, kAtomContr = 1024;
// A basis item evaluates one 256-point block against a range of atoms,
// consecutive in shell order, whose basis data the kernel stages in LDS
// together: at most kRangeAtoms atoms, kRangeShells shells, kRangePrims
// primitives and kRangeContr contraction coefficients (an atom of the
// per-atom maximum fits alone).
constexpr int kRangeAtoms = 8, kRangeShells = 32, kRangePrims = 256, kRangeContr = 1024;
constexpr int kThreads = 256; // 4 wave64 per workgroup everywhere
constexpr double kFar = 1.0e6;  // coordinate of padding points (all exps cut off)

// One shell = one (atom, l) block of unique_atom_basis_type.
struct ShellD {
  int atom, l, n_exp, n_unc, n_contr, exp_off, contr_off, func_off;
};

// One centre with the contiguous shell range it carries, and the contiguous
// ranges of its shells' primitives (exp_begin, n_exp) and contraction
// coefficients (contr_begin, n_contr).
struct AtomD {
  double x, y, z;
  int shell_begin, shell_end, lmax, pad;
  int exp_begin, n_exp, contr_begin, n_contr;
};

// A range of atoms for the basis stage: ratoms[atom_begin .. +n_atoms) are
// its atoms, in shell order; their shells (shell_begin, n_shells), primitives
// and contraction coefficients are contiguous.
struct RangeD {
  int atom_begin, n_atoms, shell_begin, n_shells;
  int exp_begin, n_exp, contr_begin, n_contr;
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
  const RangeD *ranges;  // the basis stage's atom ranges
  const int *ratoms;     // atoms in shell order (the ranges index it)
  void *phi;        // [4][NP][Mc] values, d/dx, d/dy, d/dz (T)
  void *W;          // [NP][Mc] (T); split-BF16 builds: bf16 hi / lo planes [2][NP][Mc]
  const void *D;    // [NP][NP] density operand (T); split-BF16: transposed hi / lo planes
  double *part;     // [NP/tf][4][Mc] density partials
  double *coef;     // [4][Mc] potential coefficients a, b
  double *state;    // [S_NSLOT][npad] molecule state, offset to the unit's first point
  double *epart;    // [Mc/256][2] block sums of w f and w rho (offset to the unit)
  // open-shell (ispin = 2) molecules: the alpha spin's unit also points at
  // the beta spin's partials, coefficients and state (the spin functional
  // reads both); nullptr otherwise
  const double *part2;
  double *coef2, *state2;
  void *W2;         // and the beta spin's W, written by the same W pass
  double *eterm;    // [Mc/256][5] block sums of the energy terms and the charge
  double *Hpart;    // [nsplit][NP][NP] potential partials of the molecule
  // Work of the contractions, over gathered runs of 8 functions and gathered
  // functions (all of them when screening is off).  ditems / pitems: the
  // unit's density items (tile << 8 | pair of column chunks) and potential
  // items (split << 20 | row chunk << 10 | column chunk).  tl_off / tl: per
  // 128-point tile of the unit, its ascending run list, the density's K
  // (tl_off offset to the unit's first tile, values index tl); tf_off / tfl:
  // the same for its functions (those of the shells not cut off on the
  // tile), the density's output columns in chunks of tile functions.
  // su_off / su: per point split of the unit, the union of its tiles'
  // functions, the potential's rows and columns in chunks of tile functions.
  // gcf_off / gcf: per split, [group][chunk] flags, whether the chunk can be
  // nonzero on the group (a run holding one of its functions is live there).
  // sig8: per 32-point group (from the unit's first), nw8 words of run
  // flags, the runs not cut off on the group.  wbits: the same shape, the
  // runs the potential contraction reads on the group (the FP64/FP32 W pass
  // writes W there, 0 where cut off; the split-BF16 W is written for the
  // tile's functions, the only ones its potential reads).
  const uint32_t *ditems, *pitems;
  const int *tl_off, *su_off, *tf_off;
  const uint16_t *tl, *su, *tfl;
  const uint32_t *tfb;  // the tiles' functions as bitmaps, nwf words per tile
  // per entry of tl: its run's functions in the tile's list (bit j: function
  // 8 r + j); the density contraction stages zeros for the others, which the
  // basis stage does not write
  const uint8_t *tlm;
  // per entry of su: bit t set when its function is in the list of the
  // split's tile t, so 0 there (every bit when a split has more than 32 tiles)
  const uint32_t *sumask;
  int nwf, pad_f;
  const long *gcf_off;
  const uint8_t *gcf;
  const uint32_t *sig8, *wbits;
  // Basis stage over the blocks the other stages read (bscreen = 1; 0: every
  // block).  bitems: basis items (256-point block << 16 | atom range) that
  // write some block, with bflags: 16 words per item, bit 8 j + g of words
  // 0-7 (8-15): the values (gradients) of the range's shell j are read on
  // group g of the block.
  const uint32_t *bitems, *bflags;
  long nbi;
  int bscreen, pad_b;
  long Mc, npad;
  int n_atoms, n_ranges, NP, ntiles, nsplit, incremental, pad_n;
  uint32_t mask;
  int CH, nw8;
  long ndi, npi;  // density / potential items of the unit
};

// Per-molecule descriptors of the operand preparation and the final packing.
struct PackD {
  double *Hpart;  // split partials; the fold leaves the symmetric sum in split 0
  double *Vout, *Vstate;
  const double *P;
  double *Pref;
  void *D;
  const int *fpos;  // internal position of each basis function (caller order)
  long n, NP;
  int nsplit, incremental;
};

// One launch's items: prefix offsets base[0..nu] over units, optional order.
// Queues (ctr != nullptr, contraction stages only): positions are split into
// kQueues contiguous ranges qbeg[q] .. qbeg[q+1]-1 of about equal predicted
// cost; workgroup b drains queue b % kQueues (the XCD it runs on), taking the
// next position from the counter ctr[q], then helps with the queues after
// it.  tlog (optional): per item, the
// workgroup's wall-clock ticks spent on it.
constexpr int kQueues = 8;
struct StageList {
  const UnitD *units;
  int nu;
  const long *base;
  const int *order;
  long total;
  int *ctr;
  const long *qbeg;
  uint32_t *tlog;
  int lmax;  // basis lists: the largest angular momentum of their atoms
};

// ---- launchers (kernels live in xc_kernels.hip / xc_bf16x3.hip) -----------
template <typename T>
void launch_basis(hipStream_t s, bool grad, const StageList &L);
// split-BF16 W over the basis items, the basis evaluated again (two: both
// spins of open-shell units)
void launch_wbasis(hipStream_t s, bool grad, bool two, const StageList &L);
// tf = function tile (64, 96 or 128) shared by every unit of the list;
// var = contraction variant: 0 (K step 16, 2 LDS stages), 1 (32, 1), 2 (32, 2)
constexpr int kNumVariants = 3;
template <typename T>
void launch_density(hipStream_t s, bool grad, int tf, int var, const StageList &L);
void launch_functional(hipStream_t s, bool gga, const StageList &L);
void launch_functional_spin(hipStream_t s, bool gga, const StageList &L);
// energy terms only (no coefficients, state untouched); spin: open-shell units
void launch_energy(hipStream_t s, bool gga, bool spin, const StageList &L);
// two: open-shell alpha units, whose W pass writes both spins' W (W, W2 from
// coef, coef2) from one read of the basis values
template <typename T>
void launch_wmat(hipStream_t s, bool grad, bool split, bool two, const StageList &L);
template <typename T>
void launch_potential(hipStream_t s, int tf, int var, const StageList &L);
void launch_density_bf16x3(hipStream_t s, bool grad, int tf, const StageList &L);
void launch_potential_bf16x3(hipStream_t s, int tf, const StageList &L);
void launch_fold(hipStream_t s, const PackD *mols, int nm, const long *base, long total);
void launch_pack(hipStream_t s, const PackD *mols, int nm, const long *base, long total);
void launch_zero_ints(hipStream_t s, int *p, long n);
unsigned contraction_grid(const StageList &L, int per_cu);

// cost model internals used by the planner (cost_model.hip)
long costmodel_version(const pgxc_costmodel *m);
int costmodel_features(const pgxc_costmodel *m);
// mode of the density operand: 0 FP64, 1 FP32, 2 split-BF16 planes of D^T
void launch_prep(hipStream_t s, const PackD *mols, int nm, const long *base, long total,
                 int mode, bool finish);

// item counts per unit, shared by the planner and the kernels' decoding
inline long items_basis(long Mc, int n_ranges) { return (Mc / kThreads) * n_ranges; }
inline long items_functional(long Mc) { return Mc / kThreads; }
inline long items_wmat(long Mc, int NP) { return (Mc / (4 * kThreads)) * (NP / 16); }
// point splits of the potential contraction are whole multiples of this
constexpr int kSplitPts = 2048;

// ---- screening ------------------------------------------------------------
// A radial factor is exactly zero where alpha r^2 >= 50 (the cut-off of
// ParaGauss's radial, modules/orbital_module.f90:1839-1855), so a shell is
// zero on a set of points when alpha_min r^2 >= 50 at every one of them.
// Granularity: 32-point groups x runs of 8 functions.  The contractions work
// on gathered runs: a 128-point tile contracts over the runs marked on any of
// its groups, a point split over the union of its groups' runs; the terms
// left out are exact zeros.
constexpr int kGroup = 32;
constexpr int kRun = 8;
constexpr int kScreenTile = 128;       // point tile of the density contraction
constexpr int kMaxRuns = 1024;         // runs of 8 functions per molecule (NP <= 8192)
constexpr int kMaxShells = 8192;       // shells per molecule
constexpr int kMaxSplitGroups = 2048;  // point groups per potential split
constexpr double kCutoff = 50.0 * (1.0 + 1.0e-6);  // margin: a superset in every precision
struct ScreenShell {
  double amin;                 // smallest exponent of the shell
  int atom, func_off, nfunc, pad;
};
void launch_screen(hipStream_t s, long ntiles_pts, const double *gx, const double *gy,
                   const double *gz, int n_atoms, const AtomD *atoms, int n_shells,
                   const ScreenShell *shells, int n8, int *nb, uint32_t *sig8,
                   uint32_t *smark);

// P = C diag(occ) C^T into an NP x NP padded array
void launch_form_density(hipStream_t s, const double *C, const double *occ, int n,
                         int nocc, int NP, double *P);


// Nuclear gradient, basis-centre term: the basis Hessian planes [6][NP][Mc]
// (xx, xy, xz, yy, yz, zz) of one chunk, and per 128-point tile and function
// of its list the sums of grad phi_i (2 a X_i + b . Y_i) + X_i (b . grad)
// grad phi_i into part[list position][3] (XY = [4][NP][Mc]: X = Phi P, Y_c)
void launch_basis_hess(hipStream_t s, const double *gx, const double *gy, const double *gz,
                       const AtomD *atoms, int n_atoms, const ShellD *shells, const double *exps,
                       const double *norms, const double *contr, int NP, long Mc, double *hess);
// the gradient's products X = Phi P, Y_c = (d/dc Phi) P over a density list's
// items (tile x function chunks), written for the tiles' functions only
void launch_xy(hipStream_t s, int tf, int var, const StageList &L, const double *P, double *XY,
               int nplanes, int nspin, int spin);
void launch_grad_fn(hipStream_t s, bool gga, long ntiles, const int *tf_off, const uint16_t *tfl,
                    int NP, long Mc, const double *XY, const double *phi, const double *hess,
                    const double *a, const double *bx, const double *by, const double *bz,
                    double *part);
// Weight-derivative and point-motion terms (the grid moves with the atoms):
// pm[d * pstride + p] += 2 sum_i t_d(i, p) (w grad f at each point); f per
// point from the state (st2: the beta spin's, spin builds); the Becke weight
// derivatives times w f into part[blk][atom][3] (q: [na][Mc], rm: [3][Mc]
// scratch; owner -1 for padding points)
void launch_grad_pt(hipStream_t s, bool gga, const int *tf_off, const uint16_t *tfl, int NP,
                    long Mc, const double *XY, const double *phi, const double *hess,
                    const double *a, const double *bx, const double *by, const double *bz,
                    double *pm, long pstride);
void launch_point_f(hipStream_t s, bool gga, bool spin, uint32_t mask, long Mc, long npad,
                    const double *st, const double *st2, double *fp);
void launch_becke_grad(hipStream_t s, int na, const AtomD *atoms, const double *inv_dist, long Mc,
                       const double *gx, const double *gy, const double *gz, const double *gw,
                       const double *fp, const int *owner, double *q, double *rm, double *part);

// Becke partition weights of points owned by atoms (grid generation)
void launch_inv_dist(hipStream_t s, int na, const double *axyz, double *inv);
void launch_becke(hipStream_t s, int na, const double *axyz, const double *inv_dist, long npts,
                  const double *pts, const int *owner, const double *wat, double *wout);

}  // namespace pgxc

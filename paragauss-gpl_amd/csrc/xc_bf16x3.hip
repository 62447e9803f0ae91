// Split-BF16 contractions ("bf16x3") for gfx950.
//
// Every FP32 operand x is split as
//   hi = bf16(x), lo = bf16(x - hi)
// and a product is formed from three v_mfma_f32_16x16x32_bf16 passes,
//   a*b ~ a_lo*b_hi + a_hi*b_lo + a_hi*b_hi,   FP32 accumulation,
// dropping a_lo*b_lo (relative 2^-16).  W (k_wmat) and D^T (k_prep) are
// stored split and k-contiguous, so their staging is a copy of 16-byte
// vectors into LDS; the basis values are read in FP32 (as stored, points
// contiguous) and split while they are staged, k-contiguous for the
// potential and transposed for the density.
// Same tiles and epilogues as the FP32/FP64 contractions (xc_kernels.hip);
// K advances 32 per MFMA.
#include <type_traits>

#include "pgxc_internal.h"
#include "xc_device.h"

namespace pgxc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// 16x16x32 bf16 operand placement (wave64): A(i,k) in lane i + 16 (k / 8),
// element k % 8; B(k,j) in lane j + 16 (k / 8), element k % 8.  Result as the
// FP32 16x16x4 MFMA.
__device__ __forceinline__ f4 mfma_bf16(bf16x8 a, bf16x8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

constexpr int kXBK = 32;       // K per stage
constexpr int kXBKP = 40;      // padded LDS row (bf16), 80 bytes
constexpr int kXBM = 128;
template <int BM, int BN>
struct XStage {
  static constexpr int n = 2 * BM * kXBKP + 2 * BN * kXBKP;
};

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef const u32x4 __attribute__((address_space(1))) *gvec;

// Staging of an operand stored split and k-contiguous (bf16 hi and lo planes
// at the same offsets); per stage R rows x 32 k x 2 planes, moved as 16-byte
// vectors (runs of 8).  A thread's vectors keep their rows for the whole
// item and all take k-run c = threadIdx.x & 3 of a step, so an element offset
// is row part + k part: the row pointers (plane included) are set once per
// item, and a step supplies one k offset.  Loads are unconditional global
// loads (a missing row or k run reads the plane's first vector instead) and
// the zeros are applied when the vector is written to LDS, so no load sits in
// a branch and the loads of later steps stay in flight.
template <int R, int NT = kThreads>
struct PSOperand {
  struct Src {
    const __bf16 *hi, *lo;
  };
  static constexpr int NV = R * (kXBK / 8) * 2;  // 16-byte vectors per stage
  static constexpr int PER = NV / NT;
  static_assert(PER * NT == NV, "whole vectors per thread");
  static_assert(NT % 4 == 0, "k run = threadIdx.x & 3");
  const __bf16 *row[PER];  // plane + row offset of vector e, nullptr: row absent
  const __bf16 *dummy;
  unsigned tm[PER];  // vector e's row: bit t set when it can be nonzero on tile t
  long kbase;        // k offset of tile 0
  u32x4 v[2][PER];
  unsigned ok[2];

  __device__ __forceinline__ static void place(int idx, int &plane, int &r, int &c) {
    plane = idx / (R * 4);
    const int rem = idx % (R * 4);
    r = rem >> 2;
    c = rem & 3;
  }
  // rowf(r): element offset of row r (k = 0), or -1; maskf(r): the 128-k
  // tiles from k = kb on where row r can be nonzero (bit t: tile t, at most
  // 32 tiles); a row's steps on other tiles stage zeros without reading it
  template <class RowF>
  __device__ __forceinline__ void init(const Src &S, const RowF &rowf) {
    init(S, rowf, [](int) { return ~0u; }, 0);
  }
  template <class RowF, class MaskF>
  __device__ __forceinline__ void init(const Src &S, const RowF &rowf, const MaskF &maskf, long kb) {
    dummy = S.hi;
    kbase = kb;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      int plane, r, c;
      place(threadIdx.x + NT * e, plane, r, c);
      const long o = rowf(r);
      row[e] = o >= 0 ? (plane ? S.lo : S.hi) + o : nullptr;
      tm[e] = maskf(r);
    }
  }
  // kofs(kt, c): element offset of k run c of step kt, or -1; steps past
  // ksteps load zeros
  template <int SL, class KOfs>
  __device__ __forceinline__ void load(const KOfs &kofs, int kt, int ksteps) {
    const long ko = kt < ksteps ? kofs(kt, (int)(threadIdx.x & 3)) : -1;
    const int tb = (int)((ko - kbase) >> 7) & 31;
    unsigned m = 0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const bool g = row[e] != nullptr && ko >= 0 && ((tm[e] >> tb) & 1u);
      v[SL][e] = *(gvec)(g ? row[e] + ko : dummy);
      m |= (unsigned)g << e;
    }
    ok[SL] = m;
  }
  template <int SL>
  __device__ __forceinline__ void store(__bf16 *hi, __bf16 *lo) const {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      int plane, r, c;
      place(threadIdx.x + NT * e, plane, r, c);
      const u32x4 x = ((ok[SL] >> e) & 1u) ? v[SL][e] : u32x4{0u, 0u, 0u, 0u};
      *(u32x4 *)((plane ? lo : hi) + r * kXBKP + 8 * c) = x;
    }
  }
};

// Staging of an FP32 operand stored k-contiguous, split into bf16 hi and lo
// planes on its way to LDS (split_bf16, the split every stored split operand
// gets, so the MFMA inputs are the same bits).  Per stage R rows x 32 k as
// k-quads: one 16-byte load of 4 values gives 8 bytes of hi and 8 of lo.  A
// thread's quads all take quad q = threadIdx.x & 7 of a step (k run q / 2,
// half q % 2); rows, zeros and the pipeline as PSOperand.
template <int R, int NT = kThreads>
struct PFOperand {
  static constexpr int NQ = R * (kXBK / 4);  // k-quads per stage
  static constexpr int PER = NQ / NT;
  static_assert(PER * NT == NQ, "whole quads per thread");
  static_assert(NT % 8 == 0, "k quad = threadIdx.x & 7");
  const float *row[PER];  // plane + row offset of quad e, nullptr: row absent
  const float *dummy;
  unsigned tm[PER];  // as PSOperand
  long kbase;
  f4 v[2][PER];
  unsigned ok[2];

  template <class RowF>
  __device__ __forceinline__ void init(const float *plane, const RowF &rowf) {
    init(plane, rowf, [](int) { return ~0u; }, 0);
  }
  template <class RowF, class MaskF>
  __device__ __forceinline__ void init(const float *plane, const RowF &rowf, const MaskF &maskf,
                                       long kb) {
    dummy = plane;
    kbase = kb;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int r = (int)((threadIdx.x + NT * e) >> 3);
      const long o = rowf(r);
      row[e] = o >= 0 ? plane + o : nullptr;
      tm[e] = maskf(r);
    }
  }
  template <int SL, class KOfs>
  __device__ __forceinline__ void load(const KOfs &kofs, int kt, int ksteps) {
    const int q = (int)(threadIdx.x & 7);
    const long kr = kt < ksteps ? kofs(kt, q >> 1) : -1;
    const long ko = kr >= 0 ? kr + 4 * (q & 1) : -1;
    const int tb = (int)((ko - kbase) >> 7) & 31;
    unsigned m = 0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const bool g = row[e] != nullptr && ko >= 0 && ((tm[e] >> tb) & 1u);
      v[SL][e] = gld((const f4 *)(g ? row[e] + ko : dummy));
      m |= (unsigned)g << e;
    }
    ok[SL] = m;
  }
  template <int SL>
  __device__ __forceinline__ void store(__bf16 *hi, __bf16 *lo) const {
    typedef uint16_t u4v __attribute__((ext_vector_type(4)));
    const int q = (int)(threadIdx.x & 7);
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int r = (int)((threadIdx.x + NT * e) >> 3);
      const bool g = (ok[SL] >> e) & 1u;
      u4v h, l;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        __bf16 hh, ll;
        split_bf16(g ? v[SL][e][j] : 0.0f, hh, ll);
        h[j] = __builtin_bit_cast(uint16_t, hh);
        l[j] = __builtin_bit_cast(uint16_t, ll);
      }
      *(u4v *)(hi + r * kXBKP + 4 * q) = h;
      *(u4v *)(lo + r * kXBKP + 4 * q) = l;
    }
  }
};

// Staging of an FP32 operand stored with its rows contiguous ([k][row]: the
// basis values, points contiguous) for an MFMA that takes it k-contiguous:
// per stage R rows x 4 k runs; a thread takes one (row, k run) pair, loads
// the run's 8 values of its row (8 loads of 4 bytes, k apart by ld elements;
// a wave's 64 threads hold 64 consecutive rows of one run, so each load
// reads 256 contiguous bytes), splits them with split_bf16 and writes 16
// bytes of hi and 16 of lo: the LDS image PSOperand writes from planes that
// were split and transposed beforehand, and the same bits.  The row pointers
// are set once per item; kofs(kt, c) gives the first k of run c of step kt
// (or -1 for zeros), and the loads run unconditionally as in PSOperand.
template <int R, int NT = kThreads>
struct PTOperand {
  static constexpr int NR = R * (kXBK / 8);  // (row, k run) pairs per stage
  static constexpr int PER = NR / NT;
  static_assert(PER * NT == NR, "whole runs per thread");
  static_assert(PER <= 4, "8 function flags per pair in one word");
  const float *row[PER];  // plane + row offset of pair e, nullptr: row absent
  const float *dummy;
  const uint8_t *kmk;     // per k run of the list: its k read (bit j: k 8 c + j); nullptr: all
  long ld;
  float v[2][PER][8];
  unsigned ok[2];         // 8 bits per pair: the k loaded

  template <class RowF>
  __device__ __forceinline__ void init(const float *plane, long ld_, const RowF &rowf,
                                       const uint8_t *kmask = nullptr) {
    dummy = plane;
    ld = ld_;
    kmk = kmask;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const long o = rowf((int)((threadIdx.x + NT * e) % R));
      row[e] = o >= 0 ? plane + o : nullptr;
    }
  }
  template <int SL, class KOfs>
  __device__ __forceinline__ void load(const KOfs &kofs, int kt, int ksteps) {
    unsigned m = 0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int c = (int)((threadIdx.x + NT * e) / R);
      const long k0 = kt < ksteps ? kofs(kt, c) : -1;
      const bool g = row[e] != nullptr && k0 >= 0;
      const unsigned jm = g ? (kmk ? (unsigned)kmk[kt * (kXBK / 8) + c] : 0xffu) : 0u;
      const float *p = g ? row[e] + k0 * ld : dummy;
#pragma unroll
      for (int j = 0; j < 8; ++j) v[SL][e][j] = gld(((jm >> j) & 1u) ? p + j * ld : dummy);
      m |= jm << (8 * e);
    }
    ok[SL] = m;
  }
  template <int SL>
  __device__ __forceinline__ void store(__bf16 *hi, __bf16 *lo) const {
    typedef uint16_t u8v __attribute__((ext_vector_type(8)));
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int r = (int)((threadIdx.x + NT * e) % R), c = (int)((threadIdx.x + NT * e) / R);
      const unsigned jm = ok[SL] >> (8 * e);
      u8v h, l;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        __bf16 hh, ll;
        split_bf16(((jm >> j) & 1u) ? v[SL][e][j] : 0.0f, hh, ll);
        h[j] = __builtin_bit_cast(uint16_t, hh);
        l[j] = __builtin_bit_cast(uint16_t, ll);
      }
      *(u8v *)(hi + r * kXBKP + 8 * c) = h;
      *(u8v *)(lo + r * kXBKP + 8 * c) = l;
    }
  }
};

// OA / OB: operand staging (PSOperand, PFOperand or PTOperand, rows set by
// init).  kofs(kt,
// c): element offset of k run c (8 k) of step kt, or -1 for zeros.  The NT
// threads form NT / 64 waves in a (NT / 64 / WGN) x WGN grid, each with a
// (BM / rows) x (BN / WGN) tile of 16x16 blocks; acc holds one wave's blocks.
// Pipeline: two LDS stages and two register slots; the global loads of step
// kt + 2 are issued before the MFMAs of step kt, so each load has two steps
// of MFMA work to arrive in.
template <int BM, int BN, int NT, int WGN, class OA, class OB, class KOfs>
__device__ __forceinline__ void gemm_tile_bf16x3(
    OA &A, OB &B, int ksteps, const KOfs &kofs, __bf16 *lds,
    f4 (&acc)[BM / (NT / 64 / WGN) / 16][BN / WGN / 16]) {
  constexpr int WGM = NT / 64 / WGN;
  constexpr int BKP = kXBKP, WM = BM / WGM, WN = BN / WGN, TM = WM / 16, TN = WN / 16;
  constexpr int SA = BM * BKP, SB = BN * BKP, STAGE = 2 * SA + 2 * SB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0, 0, 0, 0};

  if (ksteps <= 0) return;
  auto compute = [&](int cur) {
    const __bf16 *ah = lds + cur * STAGE, *al = ah + SA, *bh = al + SA, *bl = bh + SB;
    const int ko = 8 * (lane >> 4);
    bf16x8 ahi[TM], alo[TM], bhi[TN], blo[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int r = wm * WM + i * 16 + (lane & 15);
      ahi[i] = *(const bf16x8 *)(ah + r * BKP + ko);
      alo[i] = *(const bf16x8 *)(al + r * BKP + ko);
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int r = wn * WN + j * 16 + (lane & 15);
      bhi[j] = *(const bf16x8 *)(bh + r * BKP + ko);
      blo[j] = *(const bf16x8 *)(bl + r * BKP + ko);
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc[i][j] = mfma_bf16(alo[i], bhi[j], acc[i][j]);
        acc[i][j] = mfma_bf16(ahi[i], blo[j], acc[i][j]);
        acc[i][j] = mfma_bf16(ahi[i], bhi[j], acc[i][j]);
      }
  };
  auto stage = [&](int buf) { return lds + buf * STAGE; };
  // step kt lives in register slot kt & 1 and LDS stage kt & 1.  Loads and
  // stores run unconditionally (past the last step they move zeros), so the
  // loop body is straight-line code and the compiler's wait counts keep the
  // loads of step kt + 2 in flight while step kt + 1 is written to LDS.
  auto step = [&](auto slot, int kt) {
    constexpr int C = decltype(slot)::value, N = 1 - C;
    A.template load<C>(kofs, kt + 2, ksteps);  // slot C is free: step kt is in LDS
    B.template load<C>(kofs, kt + 2, ksteps);
    compute(C);
    __bf16 *ah = stage(N), *al = ah + SA, *bh = al + SA, *bl = bh + SB;
    A.template store<N>(ah, al);
    B.template store<N>(bh, bl);
    __syncthreads();
  };

  {
    A.template load<0>(kofs, 0, ksteps);
    B.template load<0>(kofs, 0, ksteps);
    A.template load<1>(kofs, 1, ksteps);
    B.template load<1>(kofs, 1, ksteps);
    __bf16 *ah = stage(0), *al = ah + SA, *bh = al + SA, *bl = bh + SB;
    A.template store<0>(ah, al);
    B.template store<0>(bh, bl);
  }
  __syncthreads();
  int kt = 0;
  for (; kt + 1 < ksteps; kt += 2) {
    step(std::integral_constant<int, 0>{}, kt);
    step(std::integral_constant<int, 1>{}, kt + 1);
  }
  if (kt < ksteps) step(std::integral_constant<int, 0>{}, kt);
}

__device__ __forceinline__ int find_unit_x3(const long *__restrict__ base, int nu, long it) {
  int lo = 0, hi = nu - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= it) lo = mid; else hi = mid - 1;
  }
  return lo;
}


// item = (128-point tile, pair of BN-column chunks of its gathered runs), as
// k_density: eight waves in 2 x 4, each with the 64 x 48 tile of the
// four-wave kernels, so the point panel A is staged once for 2 BN columns;
// wave columns 0-1 give the first chunk's partials, 2-3 the second's.
constexpr int kDensThreads = 512;
template <int BN, bool GRAD>
__global__ __launch_bounds__(kDensThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_density_bf16x3(const StageList S) {
  constexpr int BM = kXBM, NT = kDensThreads, WGN = 4, BN2 = 2 * BN;
  __shared__ __attribute__((aligned(16))) __bf16 lds[2 * XStage<BM, BN2>::n];
  // the epilogue's reduction buffer reuses the staging LDS (free after the
  // contraction)
  static_assert(sizeof(double) * WGN * 4 * BM <= sizeof(__bf16) * 2 * XStage<BM, BN2>::n, "alias");
  double(&red)[WGN][4][BM] = *reinterpret_cast<double(*)[WGN][4][BM]>(lds);
  __shared__ uint16_t kl[kMaxRuns];
  __shared__ uint8_t km[kMaxRuns];  // per run of kl: its functions in the tile's list
  __shared__ uint16_t cl[BN2];
  __shared__ long slot;
  static_assert(kXBK == 4 * kRun, "four runs per K step");
  for (long L = walk_next(S, -1, &slot); L < S.total;) {
    const uint64_t t0 = S.tlog ? wall_clock64() : 0;
    const long it = walk_item(S, L);
    const int u = find_unit_x3(S.base, S.nu, it);
    const UnitD &U = S.units[u];
    const uint32_t e = U.ditems[it - S.base[u]];
    const int tile = (int)(e >> 8), ch = 2 * (int)(e & 255);
    const long m0 = (long)tile * BM, NP = U.NP;
    const int off = U.tl_off[tile], c8 = U.tl_off[tile + 1] - off;
    const int foff = U.tf_off[tile] + ch * BN;
    const int ncol = min(BN2, U.tf_off[tile + 1] - foff), nslots = ncol > BN ? 2 : 1;
    __syncthreads();  // the previous item's readers of kl and cl are done
    for (int i = threadIdx.x; i < ncol; i += NT) cl[i] = gld(U.tfl + foff + i);
    for (int i = threadIdx.x; i < c8; i += NT) km[i] = gld(U.tlm + off + i);
    lds_list(U.tl + off, c8, kl);
    // A(p, k) = Phi[f(k)][m0 + p] (FP32 values, split while staged; 0 for the
    // functions of the runs outside the tile's list, which the basis stage
    // leaves unwritten), B(k, n) = D^T[col(n)][f(k)]
    const __bf16 *dh = (const __bf16 *)U.D;
    const long np2 = NP * NP;
    PTOperand<BM, NT> opa;
    PSOperand<BN2, NT> opb;
    opa.init((const float *)U.phi, U.Mc, [&](int m) -> long { return m0 + m; }, km);
    opb.init({dh, dh + np2}, [&](int n) -> long { return n < ncol ? (long)cl[n] * NP : -1; });
    auto kofs = [&](int kt, int c) -> long { return run_fn(kl, c8, kt * kXBK + 8 * c); };
    f4 acc[BM / 2 / 16][BN2 / WGN / 16];
    gemm_tile_bf16x3<BM, BN2, NT, WGN>(opa, opb, (c8 + 3) / 4, kofs, lds, acc);
    const int ahead = walk_ahead(S);
    constexpr int IB = BN2 / WGN / 16 > 3 ? 1 : 2;  // row blocks per round: within 256 registers
    density_epilogue<float, f4, BM, BN2, GRAD, IB, NT, WGN>(acc, (const float *)U.phi, U.Mc,
                                                           U.NP, m0, cl, ncol, ch, U.part, red,
                                                           nslots);
    item_log(S, it, t0);
    L = walk_take(S, L, ahead, &slot);
  }
}


template <int BN>
static void density_x3(hipStream_t s, bool grad, const StageList &L) {
  dim3 g(contraction_grid(L, 1));
  if (grad)
    hipLaunchKernelGGL((k_density_bf16x3<BN, true>), g, dim3(kDensThreads), 0, s, L);
  else
    hipLaunchKernelGGL((k_density_bf16x3<BN, false>), g, dim3(kDensThreads), 0, s, L);
}

void launch_density_bf16x3(hipStream_t s, bool grad, int tf, const StageList &L) {
  if (!L.total) return;
  switch (tf) {
    case 64: density_x3<64>(s, grad, L); break;
    case 96: density_x3<96>(s, grad, L); break;
    default: density_x3<128>(s, grad, L); break;
  }
}

// item = (point split, function tile pair), as k_potential
template <int BT>
__global__ __launch_bounds__(kThreads) void k_potential_bf16x3(const StageList S) {
  __shared__ __attribute__((aligned(16))) __bf16 lds[2 * XStage<BT, BT>::n];
  __shared__ uint16_t ul[2 * BT];  // the functions of the row chunk, then of the column chunk
  __shared__ uint16_t gl[kMaxSplitGroups];
  __shared__ int wtot[4];
  __shared__ long slot;
  static_assert(kXBK == kGroup, "one K step per 32-point group");
  for (long L = walk_next(S, -1, &slot); L < S.total;) {
    const uint64_t t0 = S.tlog ? wall_clock64() : 0;
    const long it = walk_item(S, L);
    const int u = find_unit_x3(S.base, S.nu, it);
    const UnitD &U = S.units[u];
    const uint32_t e = U.pitems[it - S.base[u]];
    const int sl = (int)(e >> 20), ci = (int)((e >> 10) & 1023), cj = (int)(e & 1023);
    const long Mc = U.Mc;
    const int so = U.su_off[sl], nu = U.su_off[sl + 1] - so;
    const int nch = (nu + BT - 1) / BT, G = (int)(Mc / U.nsplit) / kGroup;
    const long kb = (long)sl * G * kGroup;
    const int nri = min(BT, nu - ci * BT), nrj = min(BT, nu - cj * BT);
    __syncthreads();  // the previous item's readers of ul / gl are done
    for (int i = threadIdx.x; i < 2 * BT; i += kThreads)
      if (i < BT ? i < nri : i - BT < nrj) ul[i] = gld(U.su + so + (i < BT ? ci * BT + i : cj * BT + i - BT));
    __syncthreads();
    const int ng = potential_groups(U.gcf + U.gcf_off[sl], G, nch, ci, cj, gl, wtot);
    const uint16_t *ri = ul, *rj = ul + BT;
    const __bf16 *wh = (const __bf16 *)U.W;
    const long pn = (long)U.NP * Mc;
    // A(m, k) = W[f_i(m)][kb + p(k)] (stored split), B(k, n) =
    // Phi[f_j(n)][kb + p(k)] (FP32 values, split while staged)
    PSOperand<BT> opa;
    PFOperand<BT> opb;
    // rows and columns skip the tiles whose list lacks their function (W and
    // Phi are 0 there)
    const uint32_t *tmi = U.sumask + so + ci * BT, *tmj = U.sumask + so + cj * BT;
    opa.init({wh, wh + pn}, [&](int m) -> long { return m < nri ? (long)ri[m] * Mc : -1; },
             [&](int m) -> unsigned { return m < nri ? gld(tmi + m) : 0u; }, kb);
    opb.init((const float *)U.phi, [&](int n) -> long { return n < nrj ? (long)rj[n] * Mc : -1; },
             [&](int n) -> unsigned { return n < nrj ? gld(tmj + n) : 0u; }, kb);
    auto kofs = [&](int kt, int c) -> long { return kb + (long)gl[kt] * kGroup + 8 * c; };
    f4 acc[BT / 32][BT / 32];
    gemm_tile_bf16x3<BT, BT, kThreads, 2>(opa, opb, ng, kofs, lds, acc);
    const int ahead = walk_ahead(S);
    potential_epilogue<float, f4, BT, BT>(acc, U.Hpart + (long)sl * U.NP * U.NP, U.NP, ri, nri, rj,
                                          nrj);
    item_log(S, it, t0);
    L = walk_take(S, L, ahead, &slot);
  }
}

void launch_potential_bf16x3(hipStream_t s, int tf, const StageList &L) {
  if (!L.total) return;
  dim3 g(contraction_grid(L, 2));
  switch (tf) {
    case 64: hipLaunchKernelGGL((k_potential_bf16x3<64>), g, dim3(kThreads), 0, s, L); break;
    case 96: hipLaunchKernelGGL((k_potential_bf16x3<96>), g, dim3(kThreads), 0, s, L); break;
    default: hipLaunchKernelGGL((k_potential_bf16x3<128>), g, dim3(kThreads), 0, s, L); break;
  }
}

}  // namespace pgxc

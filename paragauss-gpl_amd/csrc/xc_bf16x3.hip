// Split-BF16 contractions ("bf16x3") for gfx950.
//
// Every FP32 operand x is split while it is staged into LDS as
//   hi = bf16(x), lo = bf16(x - hi)
// and a product is formed from three v_mfma_f32_16x16x32_bf16 passes,
//   a*b ~ a_lo*b_hi + a_hi*b_lo + a_hi*b_hi,   FP32 accumulation,
// dropping a_lo*b_lo (relative 2^-16).  Same tiles and epilogues as the
// FP32/FP64 contractions (xc_kernels.hip); K advances 32 per MFMA.
#include "pgxc_internal.h"
#include "xc_device.h"

namespace pgxc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float4_ __attribute__((ext_vector_type(4)));

// 16x16x32 bf16 operand placement (wave64): A(i,k) in lane i + 16 (k / 8),
// element k % 8; B(k,j) in lane j + 16 (k / 8), element k % 8.  Result as the
// FP32 16x16x4 MFMA.
__device__ __forceinline__ f4 mfma_bf16(bf16x8 a, bf16x8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

constexpr int kXBK = 32;       // K per stage
constexpr int kXBKP = 40;      // padded LDS row (bf16), 80 bytes
constexpr int kXBM = 128;
template <int BM, int BN>
struct XStage {
  static constexpr int n = 2 * BM * kXBKP + 2 * BN * kXBKP;
};

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// Staging of one operand tile of R rows x 32 K into the split LDS planes
// [row][k] (k contiguous, rows padded to kXBKP).  Work unit: a 4 x 4 block
// (4 rows x 4 k) per thread, read as four 16-byte loads along the contiguous
// dimension (KC: k contiguous, else rows contiguous, then transposed in
// registers) and written as 8-byte hi / lo vectors.
template <int R, bool KC>
struct XOperand {
  static constexpr int NB = R * kXBK / 16;              // 4 x 4 blocks
  static constexpr int PER = (NB + kThreads - 1) / kThreads;
  float4_ v[PER][4];

  // Consecutive threads take consecutive k blocks of one row block for both
  // layouts: their 8-byte LDS writes land on distinct banks (rows are 20
  // dwords apart), and each group of 8 lanes still reads whole 128-byte lines
  // (a row run when k is contiguous, a column run of 4 k columns otherwise).
  __device__ __forceinline__ void block_rc(int blk, int &r4, int &k4) const {
    k4 = blk % (kXBK / 4);
    r4 = blk / (kXBK / 4);
  }
  // X(r,k) = X[r*s + k] if KC else X[r + k*s]; kb = first k of the slab
  __device__ __forceinline__ void load(const float *__restrict__ X, long s, long kb) {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int blk = threadIdx.x + kThreads * e;
      if (NB % kThreads != 0 && blk >= NB) break;
      int r4, k4;
      block_rc(blk, r4, k4);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (KC)
          v[e][i] = *(const float4_ *)(X + (long)(4 * r4 + i) * s + kb + 4 * k4);
        else
          v[e][i] = *(const float4_ *)(X + 4 * r4 + (kb + 4 * k4 + i) * s);
      }
    }
  }
  __device__ __forceinline__ void store(__bf16 *hi, __bf16 *lo) const {
#pragma unroll
    for (int e = 0; e < PER; ++e) {
      const int blk = threadIdx.x + kThreads * e;
      if (NB % kThreads != 0 && blk >= NB) break;
      int r4, k4;
      block_rc(blk, r4, k4);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        bf16x4 h, l;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const float x = KC ? v[e][i][t] : v[e][t][i];
          __bf16 hh, ll;
          split_bf16(x, hh, ll);
          h[t] = hh;
          l[t] = ll;
        }
        const int off = (4 * r4 + i) * kXBKP + 4 * k4;
        *(bf16x4 *)(hi + off) = h;
        *(bf16x4 *)(lo + off) = l;
      }
    }
  }
};

// AKC: A(m,k) = A[m*sA + k] (k contiguous) else A[m + k*sA]; same for B(k,n).
// K step kt starts at k = kmap(kt).  Pipeline: two LDS stages and two
// register slots; the global loads of step kt + 2 are issued before the MFMAs
// of step kt, so each load has two steps of MFMA work to arrive in.
template <int BM, int BN, bool AKC, bool BKC, class KMap>
__device__ __forceinline__ void gemm_tile_bf16x3(const float *__restrict__ A, long sA,
                                                 const float *__restrict__ B, long sB,
                                                 int ksteps, KMap kmap, __bf16 *lds,
                                                 f4 (&acc)[BM / 32][BN / 32]) {
  constexpr int BKP = kXBKP, WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int SA = BM * BKP, SB = BN * BKP, STAGE = 2 * SA + 2 * SB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  XOperand<BM, AKC> ra0, ra1;
  XOperand<BN, BKC> rb0, rb1;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0, 0, 0, 0};

  if (ksteps <= 0) return;
  auto load = [&](int kt, XOperand<BM, AKC> &ra, XOperand<BN, BKC> &rb) {
    const long kb = kmap(kt);
    ra.load(A, sA, kb);
    rb.load(B, sB, kb);
  };
  auto store = [&](int buf, const XOperand<BM, AKC> &ra, const XOperand<BN, BKC> &rb) {
    __bf16 *ah = lds + buf * STAGE, *al = ah + SA, *bh = al + SA, *bl = bh + SB;
    ra.store(ah, al);
    rb.store(bh, bl);
  };
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
  // step kt lives in register slot kt & 1 and LDS stage kt & 1
  auto step = [&](int kt, XOperand<BM, AKC> &ra_cur, XOperand<BN, BKC> &rb_cur,
                  XOperand<BM, AKC> &ra_nxt, XOperand<BN, BKC> &rb_nxt) {
    if (kt + 2 < ksteps) load(kt + 2, ra_cur, rb_cur);  // slot of kt is free: kt is in LDS
    compute(kt & 1);
    if (kt + 1 < ksteps) store((kt + 1) & 1, ra_nxt, rb_nxt);
    __syncthreads();
  };

  load(0, ra0, rb0);
  if (ksteps > 1) load(1, ra1, rb1);
  store(0, ra0, rb0);
  __syncthreads();
  for (int kt = 0; kt < ksteps; kt += 2) {
    step(kt, ra0, rb0, ra1, rb1);
    if (kt + 1 < ksteps) step(kt + 1, ra1, rb1, ra0, rb0);
  }
}

__device__ __forceinline__ int find_unit_x3(const long *__restrict__ base, int nu, long it) {
  int lo = 0, hi = nu - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= it) lo = mid; else hi = mid - 1;
  }
  return lo;
}


// item = (128-point tile, function tile), as k_density
template <int BN, bool GRAD>
__global__ __launch_bounds__(kThreads) void k_density_bf16x3(const StageList S) {
  constexpr int BM = kXBM;
  __shared__ __attribute__((aligned(16))) __bf16 lds[2 * XStage<BM, BN>::n];
  __shared__ double red[2][4][BM];
  __shared__ uint16_t kl[kMaxBlocks];
  __shared__ long slot;
  static_assert(kXBK == kGroup, "one K step per 32-function block");
  for (long L = walk_next(S, -1, &slot); L < S.total; L = walk_next(S, L, &slot)) {
    const uint64_t t0 = S.tlog ? wall_clock64() : 0;
    const long it = walk_item(S, L);
    const int u = find_unit_x3(S.base, S.nu, it);
    const UnitD &U = S.units[u];
    const long local = it - S.base[u];
    const int ntn = U.NP / BN;
    const long m0 = (local / ntn) * BM;
    const int tn = (int)(local % ntn), n0 = tn * BN;
    const float *phi = (const float *)U.phi, *D = (const float *)U.D;
    bool scr = U.screen != 0;
    int ksteps = U.NP / kXBK;
    if (scr) {
      ksteps = density_screen(U, m0, tn, kl);
      if (ksteps < 0) {
        density_zero<BM, GRAD>(U.part, U.Mc, m0, tn);
        item_log(S, it, t0);
        continue;
      }
      scr = ksteps < U.nfb;  // a complete list is the dense walk
    }
    auto kmap = [&](int kt) -> long { return scr ? (long)kl[kt] * kGroup : (long)kt * kXBK; };
    f4 acc[BM / 32][BN / 32];
    gemm_tile_bf16x3<BM, BN, false, false>(phi + m0, U.Mc, D + n0, U.NP, ksteps, kmap, lds, acc);
    density_epilogue<float, f4, BM, BN, GRAD>(acc, phi, U.Mc, U.NP, m0, n0, tn, U.part, red,
                                              scr ? U.gbits : nullptr, U.nfw);
    item_log(S, it, t0);
  }
}


template <int BN>
static void density_x3(hipStream_t s, bool grad, const StageList &L) {
  dim3 g(contraction_grid(L, 2));
  if (grad)
    hipLaunchKernelGGL((k_density_bf16x3<BN, true>), g, dim3(kThreads), 0, s, L);
  else
    hipLaunchKernelGGL((k_density_bf16x3<BN, false>), g, dim3(kThreads), 0, s, L);
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
  __shared__ uint16_t gl[kMaxSplitGroups];
  __shared__ int wtot[4];
  __shared__ long slot;
  for (long L = walk_next(S, -1, &slot); L < S.total; L = walk_next(S, L, &slot)) {
    const uint64_t t0 = S.tlog ? wall_clock64() : 0;
    const long it = walk_item(S, L);
    const int u = find_unit_x3(S.base, S.nu, it);
    const UnitD &U = S.units[u];
    const long local = it - S.base[u];
    const int nt = U.NP / BT, nt2 = nt * nt;
    const int sl = (int)(local / nt2), tile = (int)(local % nt2);
    const int i0 = (tile / nt) * BT, j0 = (tile % nt) * BT;
    const int kdense = (int)(U.Mc / kXBK) / U.nsplit;
    const long kb = (long)sl * kdense * kXBK, Mc = U.Mc;
    const int ksteps =
        U.screen ? potential_screen(U, kb, kdense * kXBK, i0 / BT, j0 / BT, gl, wtot) : kdense;
    const bool scr = ksteps < kdense;  // a complete list is the dense walk
    auto kmap = [&](int kt) -> long { return scr ? (long)gl[kt] * kGroup : (long)kt * kXBK; };
    const float *W = (const float *)U.W, *phi = (const float *)U.phi;
    f4 acc[BT / 32][BT / 32];
    gemm_tile_bf16x3<BT, BT, true, true>(W + (long)i0 * Mc + kb, Mc, phi + (long)j0 * Mc + kb, Mc,
                                         ksteps, kmap, lds, acc);
    potential_epilogue<float, f4, BT, BT>(acc, U.Hpart + (long)sl * U.NP * U.NP, U.NP, i0, j0,
                                          U.first);
    item_log(S, it, t0);
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

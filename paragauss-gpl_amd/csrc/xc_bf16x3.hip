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

// AKC: A(m,k) = A[m*sA + k] (k contiguous) else A[m + k*sA]; same for B(k,n)
template <int BM, int BN, bool AKC, bool BKC>
__device__ __forceinline__ void gemm_tile_bf16x3(const float *__restrict__ A, long sA,
                                                 const float *__restrict__ B, long sB,
                                                 int ksteps, __bf16 *lds,
                                                 f4 (&acc)[BM / 32][BN / 32]) {
  constexpr int BK = kXBK, BKP = kXBKP, WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int SA = BM * BKP, SB = BN * BKP, STAGE = 2 * SA + 2 * SB;
  constexpr int LA = BM * BK / 4 / kThreads, LB = BN * BK / 4 / kThreads;  // float4 per thread
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  float4_ ra[LA], rb[LB];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f4{0, 0, 0, 0};

  auto load = [&](int kt) {
    const long kb = (long)kt * BK;
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      const int idx = tid + kThreads * e;
      if (AKC) {
        const int mm = idx / (BK / 4), k4 = idx % (BK / 4);
        ra[e] = *(const float4_ *)(A + (long)mm * sA + kb + 4 * k4);
      } else {
        const int m4 = idx % (BM / 4), kk = idx / (BM / 4);
        ra[e] = *(const float4_ *)(A + 4 * m4 + (kb + kk) * sA);
      }
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      const int idx = tid + kThreads * e;
      if (BKC) {
        const int nn = idx / (BK / 4), k4 = idx % (BK / 4);
        rb[e] = *(const float4_ *)(B + (long)nn * sB + kb + 4 * k4);
      } else {
        const int n4 = idx % (BN / 4), kk = idx / (BN / 4);
        rb[e] = *(const float4_ *)(B + (kb + kk) * sB + 4 * n4);
      }
    }
  };
  auto store = [&](int buf) {
    __bf16 *ah = lds + buf * STAGE, *al = ah + SA, *bh = al + SA, *bl = bh + SB;
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      const int idx = tid + kThreads * e;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int m, k;
        if (AKC) { m = idx / (BK / 4); k = 4 * (idx % (BK / 4)) + t; }
        else { m = 4 * (idx % (BM / 4)) + t; k = idx / (BM / 4); }
        __bf16 h, l;
        split_bf16(ra[e][t], h, l);
        ah[m * BKP + k] = h;
        al[m * BKP + k] = l;
      }
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      const int idx = tid + kThreads * e;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int n, k;
        if (BKC) { n = idx / (BK / 4); k = 4 * (idx % (BK / 4)) + t; }
        else { n = 4 * (idx % (BN / 4)) + t; k = idx / (BN / 4); }
        __bf16 h, l;
        split_bf16(rb[e][t], h, l);
        bh[n * BKP + k] = h;
        bl[n * BKP + k] = l;
      }
    }
  };

  load(0);
  store(0);
  __syncthreads();
  for (int kt = 0; kt < ksteps; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < ksteps) load(kt + 1);
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
    if (kt + 1 < ksteps) store(cur ^ 1);
    __syncthreads();
  }
}

template <int BN, bool GRAD>
__global__ __launch_bounds__(kThreads) void k_density_bf16x3(const float *__restrict__ phi,
                                                             const float *__restrict__ D, long Mc,
                                                             int NP, double *__restrict__ part,
                                                             int ntn) {
  constexpr int BM = kXBM;
  __shared__ __attribute__((aligned(16))) __bf16 lds[2 * XStage<BM, BN>::n];
  __shared__ double red[2][4][BM];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int tm = tile / ntn, tn = tile % ntn;
  const long m0 = (long)tm * BM;
  const int n0 = tn * BN;
  f4 acc[BM / 32][BN / 32];
  gemm_tile_bf16x3<BM, BN, false, false>(phi + m0, Mc, D + n0, NP, NP / kXBK, lds, acc);
  density_epilogue<float, f4, BM, BN, GRAD>(acc, phi, Mc, NP, m0, n0, tn, part, red);
}

template <int BN>
static void density_x3(hipStream_t s, bool grad, const float *phi, const float *D, long Mc, int NP,
                       double *part) {
  const int ntn = NP / BN, ntm = (int)(Mc / kXBM);
  dim3 grid((unsigned)(ntm * ntn));
  if (grad)
    hipLaunchKernelGGL((k_density_bf16x3<BN, true>), grid, dim3(kThreads), 0, s, phi, D, Mc, NP, part, ntn);
  else
    hipLaunchKernelGGL((k_density_bf16x3<BN, false>), grid, dim3(kThreads), 0, s, phi, D, Mc, NP, part, ntn);
}

void launch_density_bf16x3(hipStream_t s, bool grad, const float *phi, const float *D, long Mc,
                           int NP, int tf, double *part) {
  switch (tf) {
    case 64: density_x3<64>(s, grad, phi, D, Mc, NP, part); break;
    case 96: density_x3<96>(s, grad, phi, D, Mc, NP, part); break;
    default: density_x3<128>(s, grad, phi, D, Mc, NP, part); break;
  }
}

template <int BT>
__global__ __launch_bounds__(kThreads) void k_potential_bf16x3(const float *__restrict__ W,
                                                               const float *__restrict__ phi,
                                                               long Mc, int NP, int ksteps,
                                                               double *Hpart, int ntn, int first) {
  constexpr int BM = BT, BN = BT;
  __shared__ __attribute__((aligned(16))) __bf16 lds[2 * XStage<BM, BN>::n];
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int i0 = (tile / ntn) * BM, j0 = (tile % ntn) * BN;
  const long kb = (long)blockIdx.y * ksteps * kXBK;
  f4 acc[BM / 32][BN / 32];
  gemm_tile_bf16x3<BM, BN, true, true>(W + (long)i0 * Mc + kb, Mc, phi + (long)j0 * Mc + kb, Mc,
                                       ksteps, lds, acc);
  potential_epilogue<float, f4, BM, BN>(acc, Hpart + (long)blockIdx.y * NP * NP, NP, i0, j0, first);
}

template <int BT>
static void potential_x3(hipStream_t s, const float *W, const float *phi, long Mc, int NP,
                         int nsplit, double *Hpart, bool first) {
  const int nt = NP / BT;
  const int ksteps = (int)(Mc / kXBK) / nsplit;
  dim3 grid((unsigned)(nt * nt), (unsigned)nsplit);
  hipLaunchKernelGGL((k_potential_bf16x3<BT>), grid, dim3(kThreads), 0, s, W, phi, Mc, NP, ksteps,
                     Hpart, nt, (int)first);
}

void launch_potential_bf16x3(hipStream_t s, const float *W, const float *phi, long Mc, int NP,
                             int tf, int nsplit, double *Hpart, bool first) {
  switch (tf) {
    case 64: potential_x3<64>(s, W, phi, Mc, NP, nsplit, Hpart, first); break;
    case 96: potential_x3<96>(s, W, phi, Mc, NP, nsplit, Hpart, first); break;
    default: potential_x3<128>(s, W, phi, Mc, NP, nsplit, Hpart, first); break;
  }
}

}  // namespace pgxc

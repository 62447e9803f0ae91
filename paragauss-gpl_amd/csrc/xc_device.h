// Device helpers shared by the contraction kernels (FP64/FP32 MFMA and
// split-BF16): MFMA traits and result placement, XCD-aware tile order, and
// the two GEMM epilogues (density contraction, potential accumulation).
#pragma once
#include "pgxc_internal.h"

namespace pgxc {

// ------------------------------------------------------------------ helpers
// Workgroups are dispatched round-robin over the 8 XCDs; give every XCD a
// contiguous range of tile ids so that tiles sharing an operand panel share
// that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int per = nblk >> 3;
  if (bid >= (per << 3)) return bid;
  return (bid & 7) * per + (bid >> 3);
}

template <typename T>
struct MF;
// 16x16x4 MFMA operand/result placement (wave64), both types:
//   A(i,k): lane = i + 16 k      B(k,j): lane = j + 16 k
// Result element r of lane l (measured on gfx950):
//   f64: D(i,j) with i = l / 16 + 4 r, j = l % 16
//   f32: D(i,j) with i = 4 (l / 16) + r, j = l % 16
template <>
struct MF<double> {
  typedef d4 acc;
  __device__ static __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct MF<float> {
  typedef f4 acc;
  __device__ static __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int row(int lane, int r) { return ((lane >> 4) << 2) + r; }
};
__device__ __forceinline__ int mf_col(int lane) { return lane & 15; }


// Epilogue of X = Phi D for one 128-point x 128-function tile: contributions
// rho += sum_n Phi(p,n) X(p,n), grad rho += 2 sum_n dPhi(p,n) X(p,n), summed
// over the tile's columns in a fixed order and written to part[tn][c][p].
template <typename T, typename ACC, int BM, int BN, bool GRAD>
__device__ __forceinline__ void density_epilogue(const ACC (&acc)[BM / 32][BN / 32],
                                                 const T *__restrict__ phi, long Mc, int NP,
                                                 long m0, int n0, int tn,
                                                 double *__restrict__ part,
                                                 double (&red)[2][4][BM]) {
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long cs = (long)NP * Mc;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = wm * WM + i * 16 + MF<T>::row(lane, r);
      const long p = m0 + row;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const long n = n0 + wn * WN + j * 16 + mf_col(lane);
        const double xv = (double)acc[i][j][r];
        const long q = n * Mc + p;
        s0 += (double)phi[q] * xv;
        if (GRAD) {
          s1 += 2.0 * (double)phi[cs + q] * xv;
          s2 += 2.0 * (double)phi[2 * cs + q] * xv;
          s3 += 2.0 * (double)phi[3 * cs + q] * xv;
        }
      }
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        if (GRAD) {
          s1 += __shfl_xor(s1, off);
          s2 += __shfl_xor(s2, off);
          s3 += __shfl_xor(s3, off);
        }
      }
      if ((lane & 15) == 0) {
        red[wn][0][row] = s0;
        if (GRAD) {
          red[wn][1][row] = s1;
          red[wn][2][row] = s2;
          red[wn][3][row] = s3;
        }
      }
    }
  __syncthreads();
  constexpr int NC = GRAD ? 4 : 1;
  for (int idx = tid; idx < NC * BM; idx += kThreads) {
    const int c = idx / BM, row = idx % BM;
    part[((long)tn * 4 + c) * Mc + m0 + row] = red[0][c][row] + red[1][c][row];
  }
}

// Epilogue of H = W^T Phi: the tile is added into the split's FP64 slot
// (stored on the first chunk).
template <typename T, typename ACC, int BM, int BN>
__device__ __forceinline__ void potential_epilogue(const ACC (&acc)[BM / 32][BN / 32], double *H,
                                                   int NP, int i0, int j0, int first) {
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long row = i0 + wm * WM + i * 16 + MF<T>::row(lane, r);
        const long col = j0 + wn * WN + j * 16 + mf_col(lane);
        const double v = (double)acc[i][j][r];
        const long q = row * NP + col;
        H[q] = first ? v : H[q] + v;
      }
}

}  // namespace pgxc

// Device helpers shared by the contraction kernels (FP64/FP32 MFMA and
// split-BF16): MFMA traits and result placement, XCD-aware tile order, and
// the two GEMM epilogues (density contraction, potential accumulation).
#pragma once
#include "pgxc_internal.h"

namespace pgxc {

// ------------------------------------------------------------------ helpers
// Global-memory access through pointers the compiler cannot place (they are
// read from descriptors): global_load / global_store instead of flat ones,
// which would also count against the LDS wait counter.
template <typename T>
__device__ __forceinline__ T gld(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ void gst(T *p, T v) {
  *(__attribute__((address_space(1))) T *)p = v;
}

// Workgroups are dispatched round-robin over the 8 XCDs; give every XCD a
// contiguous range of tile ids so that tiles sharing an operand panel share
// that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int per = nblk >> 3;
  if (bid >= (per << 3)) return bid;
  return (bid & 7) * per + (bid >> 3);
}

// A position of a queued list (thread 0 only) from the queues after the
// workgroup's own (blockIdx.x % kQueues), from `first` on; total when all are
// drained.
__device__ __forceinline__ long walk_claim_queue(const StageList &S, int first) {
  const int q = (int)(blockIdx.x % kQueues);
  for (int t = first; t < kQueues; ++t) {
    const int qq = (q + t) % kQueues;
    const long c = S.qbeg[qq] + (long)atomicAdd(S.ctr + qq, 1);
    if (c < S.qbeg[qq + 1]) return c;
  }
  return S.total;
}

// Walk of a stage list's positions.  Static: blockIdx.x, + gridDim.x, ...
// With queues: the next position of the workgroup's own queue; once that is
// drained, of the next queue that is not; total when all are.  Every thread
// of the workgroup calls it.
__device__ __forceinline__ long walk_next(const StageList &S, long prev, long *slot) {
  if (!S.ctr) return prev < 0 ? (long)blockIdx.x : prev + (long)gridDim.x;
  if (threadIdx.x == 0) *slot = walk_claim_queue(S, 0);
  __syncthreads();
  const long p = *slot;
  __syncthreads();
  return p;
}

// The same walk with the next position claimed during the current item:
// walk_ahead (every thread, before the item's last phase) issues thread 0's
// claim on its own queue without waiting for it, and walk_take (every thread,
// after the item) checks it (falling back to the other queues) and hands it
// to the workgroup, so the claim's round trip overlaps the last phase.  The
// slot is written again only after the next item's barriers.
__device__ __forceinline__ int walk_ahead(const StageList &S) {
  int raw = 0;  // thread 0: the claim's counter value, still in flight here
  if (S.ctr && threadIdx.x == 0) raw = atomicAdd(S.ctr + blockIdx.x % kQueues, 1);
  return raw;
}
__device__ __forceinline__ long walk_take(const StageList &S, long cur, int raw, long *slot) {
  if (!S.ctr) return cur + (long)gridDim.x;
  // the claim's value is first read here: an empty asm that takes it keeps
  // the compiler from converting it (and so waiting for it) where it arrives
  asm volatile("" : "+v"(raw));
  if (threadIdx.x == 0) {
    const int q = (int)(blockIdx.x % kQueues);
    const long p = S.qbeg[q] + (long)raw;
    *slot = p < S.qbeg[q + 1] ? p : walk_claim_queue(S, 1);
  }
  __syncthreads();
  return *slot;
}

// item at a position: the schedule's order, else (static walk) the
// XCD-blocked order, else (queues) the natural order
__device__ __forceinline__ long walk_item(const StageList &S, long pos) {
  if (S.order) return (long)S.order[pos];
  return S.ctr ? pos : (long)xcd_remap((int)pos, (int)S.total);
}

// per-item time log (wall clock ticks of the workgroup)
__device__ __forceinline__ void item_log(const StageList &S, long it, uint64_t t0) {
  if (!S.tlog) return;
  __syncthreads();
  if (threadIdx.x == 0) S.tlog[it] = (uint32_t)(wall_clock64() - t0);
}

// x = hi + lo with hi = bf16(x), lo = bf16(x - hi)
__device__ __forceinline__ void split_bf16(float x, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
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
  static constexpr bool kQuadRows = false;
  __device__ static __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct MF<float> {
  typedef f4 acc;
  __device__ static __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static constexpr bool kQuadRows = true;  // a lane's four results are consecutive rows
  __device__ static __forceinline__ int row(int lane, int r) { return ((lane >> 4) << 2) + r; }
};
__device__ __forceinline__ int mf_col(int lane) { return lane & 15; }


// ---------------------------------------------------------------- screening
// Order-preserving compaction over the workgroup: threads with act append
// value to list (after `base` entries); returns the new length.  Every thread
// of the workgroup calls it.
__device__ __forceinline__ int block_compact(bool act, int value, uint16_t *list, int base,
                                             int *wtot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(act);
  const int before = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wtot[wave] = __popcll(m);
  __syncthreads();
  int off = base;
  for (int w = 0; w < wave; ++w) off += wtot[w];
  if (act) list[off + before] = (uint16_t)value;
  const int total = base + wtot[0] + wtot[1] + wtot[2] + wtot[3];
  __syncthreads();
  return total;
}

// Copy n list entries into LDS (every thread calls it; ends with a barrier).
__device__ __forceinline__ void lds_list(const uint16_t *__restrict__ src, int n, uint16_t *dst) {
  for (int i = threadIdx.x; i < n; i += (int)blockDim.x) dst[i] = gld(src + i);
  __syncthreads();
}

// Function (row of Phi / W / D) of entry e of a gathered run list, or -1
// past the list's n runs.
// The list read is unconditional (clamped into the list), so the lookup
// compiles to a select rather than a branch.
__device__ __forceinline__ long run_fn(const uint16_t *runs, int n, int e) {
  const int r = e >> 3;
  const long v = (long)runs[min(r, max(n - 1, 0))] * kRun + (e & 7);
  return r < n ? v : -1;
}

// Potential item: the groups of its split on which both of its chunks can be
// nonzero (flags [G][nch]), ascending, into gl.  Returns the count.
__device__ __forceinline__ int potential_groups(const uint8_t *__restrict__ flags, int G, int nch,
                                                int ci, int cj, uint16_t *gl, int *wtot) {
  int cnt = 0;
  for (int b = 0; b < G; b += kThreads) {
    const int g = b + (int)threadIdx.x;
    const bool act = g < G && gld(flags + (long)g * nch + ci) != 0 && gld(flags + (long)g * nch + cj) != 0;
    cnt = block_compact(act, g, gl, cnt, wtot);
  }
  return cnt;
}

// Sum of M per-lane values (M = 4 or 16) over the 16 lanes of each lane
// group (the 16 columns of an MFMA result block) by recursive halving: at the
// exchange with lane l ^ o a lane keeps half of the values it still holds
// and sends the other half, so M = 16 values take 8 + 4 + 2 + 1 shuffles
// instead of 4 x 16.  Returns the total of value `index` in this lane
// (index = l & 15 for M = 16; for M = 4 the groups of 4 lanes l & ~3 share
// index (l & 12) / 4).  The addition order is fixed.
template <int O, int m>
struct GroupStep {  // one exchange with lane ^ O while m values are held
  template <int M>
  __device__ static __forceinline__ void run(double (&v)[M], int lane, int &index) {
    if constexpr (m > 1) {
      const bool hi = (lane & O) != 0;
#pragma unroll
      for (int k = 0; k < m / 2; ++k) {
        const double send = hi ? v[k] : v[k + m / 2];
        const double keep = hi ? v[k + m / 2] : v[k];
        v[k] = keep + __shfl_xor(send, O);
      }
      if (hi) index += m / 2;
    } else {
      v[0] += __shfl_xor(v[0], O);
    }
    if constexpr (O > 1) GroupStep<O / 2, (m > 1 ? m / 2 : 1)>::run(v, lane, index);
  }
};

template <int M>
__device__ __forceinline__ double group_reduce(double (&v)[M], int lane, int &index) {
  static_assert(M == 4 || M == 16, "4 or 16 values");
  index = 0;
  GroupStep<8, M>::run(v, lane, index);
  return v[0];
}

// Epilogue of X = Phi D for one point tile x gathered column chunk:
// contributions rho += sum_n Phi(p,n) X(p,n), grad rho += 2 sum_n dPhi(p,n)
// X(p,n), summed over the chunk's columns in a fixed order (j, then the
// 16-lane group reduction) and written to part[ch][c][p].  Column n of the
// chunk is function cfn[n] for n < ncol (absent columns add nothing).  With
// the FP32 result placement a lane holds four consecutive points of one
// column, read as one 16-byte load per component; the FP64 placement reads
// point by point.  A lane's columns are looked up once; the loads of IB row
// blocks (IB divides the wave's row blocks) are issued together and
// unconditionally (an absent column reads function 0 and adds nothing), so
// their latencies overlap instead of following one another.  Waves as in
// gemm_tile_bf16x3 (NT threads, WGN wave columns); every two wave columns
// make one partial slot: slot ch + s sums wave columns 2s and 2s + 1, and
// the first nslots slots are written.
template <typename T, typename ACC, int BM, int BN, bool GRAD, int IB = 1, int NT = kThreads,
          int WGN = 2, bool PW = false>
__device__ __forceinline__ void density_epilogue(
    const ACC (&acc)[BM / (NT / 64 / WGN) / 16][BN / WGN / 16], const T *__restrict__ phi,
    long Mc, int NP, long m0, const uint16_t *cfn, int ncol, int ch, double *__restrict__ part,
    double (&red)[WGN][4][BM], int nslots = 1) {
  constexpr int WGM = NT / 64 / WGN, WM = BM / WGM, WN = BN / WGN, TM = WM / 16, TN = WN / 16;
  constexpr int NC = GRAD ? 4 : 1, SLOTS = WGN / 2;
  constexpr bool QUAD = MF<T>::kQuadRows;
  static_assert(TM % IB == 0 && WGN % 2 == 0, "whole rounds of row blocks, wave column pairs");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const long cs = (long)NP * Mc;
  long col[TN];
  bool live[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int e = wn * WN + j * 16 + mf_col(lane);
    live[j] = e < ncol;
    col[j] = live[j] ? (long)cfn[min(e, ncol - 1)] * Mc : 0;
  }
  if constexpr (PW) {
    // one component plane at a time (a quarter of the registers of the loads;
    // each sum is formed in the same order)
    static_assert(!QUAD, "point-by-point placement");
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      double s[NC][4];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        T v[TN][4];
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const long q0 = col[j] + m0 + wm * WM + i * 16;
#pragma unroll
          for (int r = 0; r < 4; ++r) v[j][r] = gld(phi + c * cs + q0 + MF<T>::row(lane, r));
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) s[c][r] = 0.0;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          if (!live[j]) continue;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double xv = (double)acc[i][j][r];
            if (c == 0) s[0][r] += (double)v[j][r] * xv;
            else s[c][r] += 2.0 * (double)v[j][r] * xv;
          }
        }
      }
      double w[NC * 4];
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) w[c * 4 + r] = s[c][r];
      int idx;
      const double tot = group_reduce<NC * 4>(w, lane, idx);
      if (NC == 4 || (lane & 3) == 0) {
        const int c = idx / 4, r = idx % 4;
        red[wn][c][wm * WM + i * 16 + MF<T>::row(lane, r)] = tot;
      }
    }
  } else {
#pragma unroll
  for (int i0 = 0; i0 < TM; i0 += IB) {
    T v[IB][TN][NC][4];
#pragma unroll
    for (int ii = 0; ii < IB; ++ii)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const long q0 = col[j] + m0 + wm * WM + (i0 + ii) * 16;
        if constexpr (QUAD) {
          const long q = q0 + MF<T>::row(lane, 0);
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const f4 t = gld((const f4 *)(phi + c * cs + q));
            v[ii][j][c][0] = t.x; v[ii][j][c][1] = t.y; v[ii][j][c][2] = t.z; v[ii][j][c][3] = t.w;
          }
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < NC; ++c) v[ii][j][c][r] = gld(phi + c * cs + q0 + MF<T>::row(lane, r));
        }
      }
#pragma unroll
    for (int ii = 0; ii < IB; ++ii) {
      const int i = i0 + ii;
      double s[NC][4];
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) s[c][r] = 0.0;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        if (!live[j]) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double xv = (double)acc[i][j][r];
          s[0][r] += (double)v[ii][j][0][r] * xv;
#pragma unroll
          for (int c = 1; c < NC; ++c) s[c][r] += 2.0 * (double)v[ii][j][c][r] * xv;
        }
      }
      double w[NC * 4];
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) w[c * 4 + r] = s[c][r];
      int idx;
      const double tot = group_reduce<NC * 4>(w, lane, idx);
      if (NC == 4 || (lane & 3) == 0) {
        const int c = idx / 4, r = idx % 4;
        red[wn][c][wm * WM + i * 16 + MF<T>::row(lane, r)] = tot;
      }
    }
  }
  }
  __syncthreads();
  for (int idx = tid; idx < SLOTS * NC * BM; idx += NT) {
    const int sl = idx / (NC * BM), c = (idx / BM) % NC, row = idx % BM;
    if (sl < nslots)
      gst(part + ((long)(ch + sl) * 4 + c) * Mc + m0 + row,
          red[2 * sl][c][row] + red[2 * sl + 1][c][row]);
  }
}

// Epilogue of H = W^T Phi for one gathered row chunk x column chunk: the
// tile is added into the split's FP64 slot at the functions of its rows
// (rfn[m], m < nr) and columns (cfn[n], n < nc) (slots are zeroed at the
// start of a build).
template <typename T, typename ACC, int BM, int BN>
__device__ __forceinline__ void potential_epilogue(const ACC (&acc)[BM / 32][BN / 32], double *H,
                                                   int NP, const uint16_t *rfn, int nr,
                                                   const uint16_t *cfn, int nc) {
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
#pragma unroll
  for (int j = 0; j < TN; ++j) {
    const int n = wn * WN + j * 16 + mf_col(lane);
    if (n >= nc) continue;
    const long col = cfn[n];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = wm * WM + i * 16 + MF<T>::row(lane, r);
        if (m >= nr) continue;
        const long row = rfn[m];
        const long q = row * NP + col;
        gst(H + q, gld(H + q) + (double)acc[i][j][r]);
      }
  }
}

}  // namespace pgxc

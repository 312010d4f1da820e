// Detector-space rotations of NoiseSmoothedSVD (minkasi/mapmaking/noise.py:
// 650 svd, :667 v.d at setup, :704 v.d and :709 vT.d in apply_noise).
//
//  Hot path (apply_noise, every A.x): C = A.B with A = v or vT (ndet x ndet)
//  and B a TOD (ndet x n, FP32), on MFMA from split operands, accumulated in
//  FP32.  The default, mode 2, splits both into fp16 pairs (rot_f16_kernel
//  below).  Modes 0 and 1 use v_mfma_f32_32x32x16_bf16: A and B are cut into bf16 pieces
//  (h = RNE bf16 of x, m = bf16 of x-h, l = bf16 of x-h-m; for FP32 x the
//  three pieces are exact).  Mode 0 keeps the three products of order <= 1,
//  Ah.Bh + Ah.Bm + Am.Bh (about 2^-18 relative per product); mode 1 keeps
//  the six of order <= 2 (adds Ah.Bl + Al.Bh + Am.Bm), about FP32 accuracy,
//  which matters when a large common mode is rotated into modes the noise
//  filter later amplifies.  Workgroup tile 128x128, four waves of 64x64
//  (2x2 MFMA blocks), K staged through LDS 32 at a time; B is split while
//  it is staged.  Every TOD of a batch is one
//  segment of one launch; the block->tile map keeps the M-blocks that share
//  a column block of B on one XCD so B is re-read from that XCD's L2.
//
//  Setup (once per TOD): the Gram matrix d.d^T and the rotation v.d in
//  FP64 (plain tiled FMA kernels), feeding the host eigendecomposition that
//  stands in for np.linalg.svd's left singular vectors.
#include "common.h"

#include <algorithm>
#include <vector>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 32, KPAD = 8;  // LDS row = 40 bf16 = 80 B

struct RotSeg {
  const __bf16* a_hi;  // M_pad x lda (zero padded: M_pad = ceil(M/128)*128, lda % 32 == 0)
  const __bf16* a_mid;
  const __bf16* a_lo;
  int64_t lda;
  int64_t b_off, b_ld;  // element offsets into the exec-time B base
  int64_t c_off, c_ld;  // element offsets into the exec-time C base
  int32_t M, K, N, mblocks, nblocks, pad_;
  int64_t base;         // first global tile of this segment
};

__device__ inline int find_rseg(const RotSeg* segs, int nseg, int64_t g) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (segs[mid].base <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ inline void split3_bf16(float x, __bf16* h, __bf16* m, __bf16* l) {
  __bf16 hh = (__bf16)x;  // round to nearest even
  float r = x - (float)hh;  // exact
  __bf16 mm = (__bf16)r;
  *h = hh;
  *m = mm;
  *l = (__bf16)(r - (float)mm);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Buffer resource over [p, p + bytes) (raw buffer, gfx9 data format word).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// NP = number of products: 3 (mode 0) or 6 (mode 1).
//
// Staging of one K step (32 rows of B, 32 columns of A): thread t owns B
// column n0 + (t % 128) (clamped to N-1: columns past N only feed output
// columns that are never stored) and the two 8-row strips t/128 and
// t/128 + 2, and A row t/2, half t%2.  Its byte offsets are fixed per tile;
// the row and K-step parts are uniform and travel as the buffer instructions'
// scalar offsets.  The next step's loads are issued before the MFMAs of the
// current one.  The plan guarantees that every segment's B, C and A planes
// span less than 4 GiB, so 32-bit offsets hold.
template <int NP>
__global__ void __launch_bounds__(256) rot_bf16_kernel(const RotSeg* __restrict__ segs, int nseg, int64_t ntiles,
                                                       const float* __restrict__ Bbase, float* __restrict__ Cbase) {
  constexpr int NPL = NP == 6 ? 3 : 2;  // bf16 planes per operand
  __shared__ __attribute__((aligned(16))) __bf16 As[NPL][BM][BK + KPAD];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[NPL][BN][BK + KPAD];

  // XCD-aware remap: hardware dispatches consecutive blocks round-robin over
  // 8 XCDs; give each XCD a contiguous range of logical tiles.
  const int64_t G = gridDim.x;
  const int64_t per_xcd = (G + 7) / 8;
  const int64_t logical = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (logical >= ntiles) return;
  const int s = find_rseg(segs, nseg, logical);
  const RotSeg sg = segs[s];
  const int64_t local = logical - sg.base;
  const int mb = (int)(local % sg.mblocks), nb = (int)(local / sg.mblocks);
  const int m0 = mb * BM, n0 = nb * BN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  const int K = sg.K;

  const uint32_t ld = (uint32_t)sg.b_ld;
  const __amdgpu_buffer_rsrc_t brs = buffer_rsrc(Bbase + sg.b_off, (uint32_t)K * ld * 4u);
  const uint32_t abytes = (uint32_t)(sg.mblocks * BM) * (uint32_t)sg.lda * 2u;
  __amdgpu_buffer_rsrc_t ars[NPL];
  ars[0] = buffer_rsrc(sg.a_hi, abytes);
  ars[1] = buffer_rsrc(sg.a_mid, abytes);
  if (NPL == 3) ars[NPL - 1] = buffer_rsrc(sg.a_lo, abytes);

  const int nn = tid & (BN - 1), kg0 = tid >> 7;
  const int gn = min(n0 + nn, sg.N - 1);
  const uint32_t bvo = ((uint32_t)(kg0 * 8) * ld + (uint32_t)gn) * 4u;
  const int arow = tid >> 1, ahalf = tid & 1;
  const uint32_t avo = ((uint32_t)(m0 + arow) * (uint32_t)sg.lda + (uint32_t)(ahalf * 16)) * 2u;

  float xv[2][8];
  u32x4 av[NPL][2];
  auto load_stage = [&](int k0) {
#pragma unroll
    for (int pl = 0; pl < NPL; pl++) {
      av[pl][0] = __builtin_amdgcn_raw_buffer_load_b128(ars[pl], avo, (uint32_t)k0 * 2u, 0);
      av[pl][1] = __builtin_amdgcn_raw_buffer_load_b128(ars[pl], avo + 16u, (uint32_t)k0 * 2u, 0);
    }
    if (k0 + BK <= K) {
#pragma unroll
      for (int v = 0; v < 2; v++)
#pragma unroll
        for (int kk = 0; kk < 8; kk++)
          xv[v][kk] = __builtin_bit_cast(
              float, __builtin_amdgcn_raw_buffer_load_b32(brs, bvo, (uint32_t)(k0 + 16 * v + kk) * ld * 4u, 0));
    } else {  // last, partial step: rows past K are zero
#pragma unroll
      for (int v = 0; v < 2; v++)
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
          const int k = k0 + kg0 * 8 + 16 * v + kk;
          const uint32_t off = ((uint32_t)min(k, K - 1) * ld + (uint32_t)gn) * 4u;
          const float x = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(brs, off, 0u, 0));
          xv[v][kk] = k < K ? x : 0.f;
        }
    }
  };
  auto store_stage = [&]() {
#pragma unroll
    for (int pl = 0; pl < NPL; pl++) {
      *reinterpret_cast<u32x4*>(&As[pl][arow][ahalf * 16]) = av[pl][0];
      *reinterpret_cast<u32x4*>(&As[pl][arow][ahalf * 16 + 8]) = av[pl][1];
    }
#pragma unroll
    for (int v = 0; v < 2; v++) {
      bf16x8 ph, pm, plo;
#pragma unroll
      for (int kk = 0; kk < 8; kk++) {
        __bf16 h, m, l;
        split3_bf16(xv[v][kk], &h, &m, &l);
        ph[kk] = h;
        pm[kk] = m;
        plo[kk] = l;
      }
      const int kg = kg0 + 2 * v;
      *reinterpret_cast<bf16x8*>(&Bs[0][nn][kg * 8]) = ph;
      *reinterpret_cast<bf16x8*>(&Bs[1][nn][kg * 8]) = pm;
      if (NPL == 3) *reinterpret_cast<bf16x8*>(&Bs[NPL - 1][nn][kg * 8]) = plo;
    }
  };

  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

  // A wave's 32-row sub-blocks that lie wholly past the segment's last row
  // (the padding of M to the 128-row tile) skip their MFMAs; the condition
  // is uniform over the wave and nothing of those rows is stored.
  const bool live0 = m0 + wm * 64 < sg.M, live1 = m0 + wm * 64 + 32 < sg.M;

  load_stage(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    store_stage();
    __syncthreads();
    if (k0 + BK < K) load_stage(k0 + BK);  // in flight during the MFMAs below
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const int kq = ks + 8 * (lane >> 5);
      bf16x8 a[NPL][2], b[NPL][2];
#pragma unroll
      for (int pl = 0; pl < NPL; pl++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
          int r = wm * 64 + i * 32 + (lane & 31);
          a[pl][i] = *reinterpret_cast<const bf16x8*>(&As[pl][r][kq]);
          int c = wn * 64 + i * 32 + (lane & 31);
          b[pl][i] = *reinterpret_cast<const bf16x8*>(&Bs[pl][c][kq]);
        }
      // smallest products first
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          if (!(i == 0 ? live0 : live1)) continue;
          if (NP == 6) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[2][i], b[0][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1][i], b[1][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[NPL - 1][j], acc[i][j], 0, 0, 0);
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1][i], b[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  // epilogue: D[row][col], row = 8*(r/4) + 4*(lane/32) + r%4, col = lane%32
  // (Plain global stores: raw buffer stores of accumulator elements were
  // compiled to store element 0 of each block for every r.)
  float* C = Cbase + sg.c_off;
  const uint32_t cld = (uint32_t)sg.c_ld;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int col = n0 + wn * 64 + j * 32 + (lane & 31);
      if (col >= sg.N) continue;
      const uint32_t cvo = (uint32_t)(4 * (lane >> 5)) * cld + (uint32_t)col;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int row_u = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + (r & 3);  // uniform part of the row
        if (row_u + 4 * (lane >> 5) < sg.M) C[(uint32_t)row_u * cld + cvo] = acc[i][j][r];
      }
    }
}

// Mode 2: FP16 pairs.  x = h + l with h = RNE fp16 of x and l = fp16 of
// x - h represents an FP32 value to about 2^-22 relative, so the three
// products Al.Bh + Ah.Bl + Ah.Bh (v_mfma_f32_32x32x16_f16) give about FP32
// accuracy with half the MFMAs and two thirds of the LDS traffic of mode 1.
// FP16's range is narrow, so both operands are scaled by powers of two
// before the split: A (the rotation matrix) by a per-segment exponent fixed
// when it is split (its largest entry below 2^14), and each column of B by a
// running exponent.  The first K step sets a column's exponent from that
// step's largest magnitude (scaled into [2^13, 2^14)); a later step that
// would reach 2^15 lowers it (again below 2^14) and the column's
// accumulators are multiplied by the same power of two before that step's
// products, which is exact.  The epilogue undoes both scalings exactly
// (ldexp), so the result depends on them only through rounding, column by
// column as FP32 arithmetic would.  An infinity or NaN in a column reaches
// that column's outputs only.
//
// Staging and tiling are those of rot_bf16_kernel except that lanes c and
// c+32 of wave w stage column 32w+c (rows 0-7 and 16-23, 8-15 and 24-31 of
// a step), so a column's step maximum is one lane exchange.  A per-step
// flag (double-buffered by step parity, so that clearing the next step's
// flag never races with reading this one's) tells the MFMA waves whether
// any column's exponent changed in this step.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
constexpr int kF16Top = 14;  // scaled magnitudes stay below 2^14, rescaled at 2^15 (fp16 max ~2^16)

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) rot_f16_kernel(const RotSeg* __restrict__ segs, int nseg, int64_t ntiles,
                                                       const float* __restrict__ Bbase, float* __restrict__ Cbase) {
  __shared__ __attribute__((aligned(16))) _Float16 As[2][BM][BK + KPAD];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[2][BN][BK + KPAD];
  __shared__ int cdel[BN];
  __shared__ int anyd[2];
  __shared__ int cexp[BN];

  const int64_t G = gridDim.x;
  const int64_t per_xcd = (G + 7) / 8;
  const int64_t logical = (int64_t)(blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (logical >= ntiles) return;
  const int s = find_rseg(segs, nseg, logical);
  const RotSeg sg = segs[s];
  const int64_t local = logical - sg.base;
  const int mb = (int)(local % sg.mblocks), nb = (int)(local / sg.mblocks);
  const int m0 = mb * BM, n0 = nb * BN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  const int K = sg.K;

  const uint32_t ld = (uint32_t)sg.b_ld;
  const __amdgpu_buffer_rsrc_t brs = buffer_rsrc(Bbase + sg.b_off, (uint32_t)K * ld * 4u);
  const uint32_t abytes = (uint32_t)(sg.mblocks * BM) * (uint32_t)sg.lda * 2u;
  const __amdgpu_buffer_rsrc_t ars0 = buffer_rsrc(sg.a_hi, abytes);
  const __amdgpu_buffer_rsrc_t ars1 = buffer_rsrc(sg.a_mid, abytes);

  const int nn = (tid >> 6) * 32 + (lane & 31), kg0 = lane >> 5;
  const int gn = min(n0 + nn, sg.N - 1);
  const uint32_t bvo = ((uint32_t)(kg0 * 8) * ld + (uint32_t)gn) * 4u;
  const int arow = tid >> 1, ahalf = tid & 1;
  const uint32_t avo = ((uint32_t)(m0 + arow) * (uint32_t)sg.lda + (uint32_t)(ahalf * 16)) * 2u;

  float xv[2][8];
  u32x4 av[2][2];
  auto load_stage = [&](int k0) {
    av[0][0] = __builtin_amdgcn_raw_buffer_load_b128(ars0, avo, (uint32_t)k0 * 2u, 0);
    av[0][1] = __builtin_amdgcn_raw_buffer_load_b128(ars0, avo + 16u, (uint32_t)k0 * 2u, 0);
    av[1][0] = __builtin_amdgcn_raw_buffer_load_b128(ars1, avo, (uint32_t)k0 * 2u, 0);
    av[1][1] = __builtin_amdgcn_raw_buffer_load_b128(ars1, avo + 16u, (uint32_t)k0 * 2u, 0);
    if (k0 + BK <= K) {
#pragma unroll
      for (int v = 0; v < 2; v++)
#pragma unroll
        for (int kk = 0; kk < 8; kk++)
          xv[v][kk] = __builtin_bit_cast(
              float, __builtin_amdgcn_raw_buffer_load_b32(brs, bvo, (uint32_t)(k0 + 16 * v + kk) * ld * 4u, 0));
    } else {
#pragma unroll
      for (int v = 0; v < 2; v++)
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
          const int k = k0 + kg0 * 8 + 16 * v + kk;
          const uint32_t off = ((uint32_t)min(k, K - 1) * ld + (uint32_t)gn) * 4u;
          const float x = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(brs, off, 0u, 0));
          xv[v][kk] = k < K ? x : 0.f;
        }
    }
  };
  // Column exponent: set from the first step (scaled maximum in [2^13, 2^14)),
  // lowered when a later step would reach 2^15, again to below 2^14.
  int sx = 0;
  auto store_stage = [&](bool first, int k0s) {
    *reinterpret_cast<u32x4*>(&As[0][arow][ahalf * 16]) = av[0][0];
    *reinterpret_cast<u32x4*>(&As[0][arow][ahalf * 16 + 8]) = av[0][1];
    *reinterpret_cast<u32x4*>(&As[1][arow][ahalf * 16]) = av[1][0];
    *reinterpret_cast<u32x4*>(&As[1][arow][ahalf * 16 + 8]) = av[1][1];
    float m = 0.f;
#pragma unroll
    for (int v = 0; v < 2; v++)
#pragma unroll
      for (int kk = 0; kk < 8; kk++) m = fmaxf(m, fabsf(xv[v][kk]));
    m = fmaxf(m, __shfl_xor(m, 32));
    const int E = max((int)((__builtin_bit_cast(uint32_t, m) >> 23) & 0xff), 1);
    int del = 0;
    if (first) {
      sx = kF16Top + 126 - E;
    } else if (kF16Top + 1 + 126 - E < sx) {
      del = kF16Top + 126 - E - sx;
      sx += del;
    }
    if (kg0 == 0) {
      cdel[nn] = del;
      cexp[nn] = sx;
    }
    if (tid == 0) anyd[(k0s + 1) & 1] = 0;
    if (del != 0) anyd[k0s & 1] = 1;
#pragma unroll
    for (int v = 0; v < 2; v++) {
      f16x8 ph, pl;
#pragma unroll
      for (int kk = 0; kk < 8; kk++) {
        const float xs = ldexpf(xv[v][kk], sx);
        const _Float16 h = (_Float16)xs;
        ph[kk] = h;
        pl[kk] = (_Float16)(xs - (float)h);
      }
      const int kg = kg0 + 2 * v;
      *reinterpret_cast<f16x8*>(&Bs[0][nn][kg * 8]) = ph;
      *reinterpret_cast<f16x8*>(&Bs[1][nn][kg * 8]) = pl;
    }
  };

  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;
  const bool live0 = m0 + wm * 64 < sg.M, live1 = m0 + wm * 64 + 32 < sg.M;

  load_stage(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    store_stage(k0 == 0, k0 / BK);
    __syncthreads();
    if (k0 + BK < K) load_stage(k0 + BK);
    if (k0 > 0 && anyd[(k0 / BK) & 1]) {  // some column's exponent was lowered in this step
      const int d0 = cdel[wn * 64 + (lane & 31)], d1 = cdel[wn * 64 + 32 + (lane & 31)];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
          acc[i][0][r] = ldexpf(acc[i][0][r], d0);
          acc[i][1][r] = ldexpf(acc[i][1][r], d1);
        }
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const int kq = ks + 8 * (lane >> 5);
      f16x8 a[2][2], b[2][2];
#pragma unroll
      for (int pl = 0; pl < 2; pl++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
          const int r = wm * 64 + i * 32 + (lane & 31);
          a[pl][i] = *reinterpret_cast<const f16x8*>(&As[pl][r][kq]);
          const int c = wn * 64 + i * 32 + (lane & 31);
          b[pl][i] = *reinterpret_cast<const f16x8*>(&Bs[pl][c][kq]);
        }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          if (!(i == 0 ? live0 : live1)) continue;
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[1][i], b[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0][i], b[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  float* C = Cbase + sg.c_off;
  const uint32_t cld = (uint32_t)sg.c_ld;
  const int aexp = sg.pad_;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int cl = wn * 64 + j * 32 + (lane & 31);
      const int col = n0 + cl;
      if (col >= sg.N) continue;
      const int shift = -(aexp + cexp[cl]);
      const uint32_t cvo = (uint32_t)(4 * (lane >> 5)) * cld + (uint32_t)col;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int row_u = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + (r & 3);
        if (row_u + 4 * (lane >> 5) < sg.M) C[(uint32_t)row_u * cld + cvo] = ldexpf(acc[i][j][r], shift);
      }
    }
}

// hi/lo fp16 planes of 2^aexp times an FP64 matrix, padded row-major.
__global__ void split_matrix_f16_kernel(const double* __restrict__ a, int64_t M, int64_t K, int64_t lda_src,
                                        int aexp, _Float16* __restrict__ hi, _Float16* __restrict__ lo,
                                        int64_t M_pad, int64_t lda) {
  int64_t n = M_pad * lda;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = i / lda, c = i - r * lda;
    const double x = (r < M && c < K) ? ldexp(a[r * lda_src + c], aexp) : 0.0;
    const _Float16 h = (_Float16)(float)x;
    hi[i] = h;
    lo[i] = (_Float16)(float)(x - (double)(float)h);
  }
}

// hi/mid/lo split of an FP64 matrix into padded row-major bf16 planes.
__global__ void split_matrix_kernel(const double* __restrict__ a, int64_t M, int64_t K, int64_t lda_src,
                                    __bf16* __restrict__ hi, __bf16* __restrict__ mid, __bf16* __restrict__ lo,
                                    int64_t M_pad, int64_t lda) {
  int64_t n = M_pad * lda;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = i / lda, c = i - r * lda;
    double x = (r < M && c < K) ? a[r * lda_src + c] : 0.0;
    __bf16 h = (__bf16)(float)x;
    double r1 = x - (double)(float)h;
    __bf16 m = (__bf16)(float)r1;
    double r2 = r1 - (double)(float)m;
    hi[i] = h;
    mid[i] = m;
    lo[i] = (__bf16)(float)r2;
  }
}

// ---------------------------------------------------------------------------
// FP64 setup GEMMs

constexpr int TB = 64, TK = 16;

// C[M x N] = A[M x K] . B[K x N]; A FP64 row-major (lda), B of type TB_ (ldb),
// C FP64 (ldc).  64x64 tile per workgroup, 4x4 per thread.
template <typename TBT>
__global__ void __launch_bounds__(256) gemm_nn_f64(int M, int N, int K, const double* __restrict__ A, int64_t lda,
                                                   const TBT* __restrict__ B, int64_t ldb, double* __restrict__ C,
                                                   int64_t ldc) {
  __shared__ double As[TK][TB + 1];
  __shared__ double Bs[TK][TB + 1];
  const int bm = blockIdx.y * TB, bn = blockIdx.x * TB;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += TK) {
    for (int idx = threadIdx.x; idx < TK * TB; idx += 256) {
      int kk = idx / TB, mm = idx % TB;
      int gm = bm + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < K) ? A[(int64_t)gm * lda + gk] : 0.0;
      int gn = bn + mm;
      Bs[kk][mm] = (gk < K && gn < N) ? (double)B[(int64_t)gk * ldb + gn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int gm = bm + ty + 16 * i, gn = bn + tx + 16 * j;
      if (gm < M && gn < N) C[(int64_t)gm * ldc + gn] = acc[i][j];
    }
}

// Partial cross products over chunks of the long axis:
// P[chunk][M x N] = A[:, chunk] . B[:, chunk]^T (A: M rows, B: N rows, both
// row-major with n columns; the Gram matrix is A = B).
template <typename TA, typename TB2>
__global__ void __launch_bounds__(256) cross_partial_f64(int M, int N, int64_t n, const TA* __restrict__ A,
                                                         int64_t lda, const TB2* __restrict__ B, int64_t ldb,
                                                         int64_t chunk, double* __restrict__ P) {
  __shared__ double As[TK][TB + 1];
  __shared__ double Bs[TK][TB + 1];
  const int bm = blockIdx.y * TB, bn = blockIdx.x * TB;
  const int64_t c0 = (int64_t)blockIdx.z * chunk, c1 = min(n, c0 + chunk);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4] = {};
  for (int64_t k0 = c0; k0 < c1; k0 += TK) {
    for (int idx = threadIdx.x; idx < TK * TB; idx += 256) {
      int mm = idx / TK, kk = idx % TK;  // consecutive threads walk along n
      int64_t gk = k0 + kk;
      int gm = bm + mm, gn = bn + mm;
      As[kk][mm] = (gm < M && gk < c1) ? (double)A[(int64_t)gm * lda + gk] : 0.0;
      Bs[kk][mm] = (gn < N && gk < c1) ? (double)B[(int64_t)gn * ldb + gk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  double* out = P + (int64_t)blockIdx.z * M * N;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int gm = bm + ty + 16 * i, gn = bn + tx + 16 * j;
      if (gm < M && gn < N) out[(int64_t)gm * N + gn] = acc[i][j];
    }
}

__global__ void sum_partials_f64(const double* __restrict__ P, int64_t nchunk, int64_t mm, double* __restrict__ G) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mm; i += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int64_t c = 0; c < nchunk; c++) s += P[c * mm + i];
    G[i] = s;
  }
}

}  // namespace

// ---------------------------------------------------------------------------

struct mk_rot_plan {
  int nseg = 0;
  int mode = 0;
  int64_t ntiles = 0;
  RotSeg* d_segs = nullptr;
};

extern "C" int64_t mk_split_bf16_padded_dims(int64_t M, int64_t K, int64_t* M_pad, int64_t* lda) {
  if (M_pad) *M_pad = mk::cdiv(M, BM) * BM;
  if (lda) *lda = mk::cdiv(K, BK) * BK;
  return mk::cdiv(M, BM) * BM * mk::cdiv(K, BK) * BK;
}

extern "C" int mk_split_bf16_matrix(const double* a, int64_t M, int64_t K, int64_t lda_src, void* hi, void* mid,
                                    void* lo, void* stream) {
  MK_REQUIRE(a && hi && mid && lo && M > 0 && K > 0 && lda_src >= K, MK_ERR_ARG, "bad arguments");
  int64_t M_pad, lda;
  mk_split_bf16_padded_dims(M, K, &M_pad, &lda);
  split_matrix_kernel<<<mk::grid_for(M_pad * lda, 256), 256, 0, mk::as_stream(stream)>>>(
      a, M, K, lda_src, (__bf16*)hi, (__bf16*)mid, (__bf16*)lo, M_pad, lda);
  return mk::check_launch("split_matrix_kernel");
}

extern "C" int mk_split_f16_matrix(const double* a, int64_t M, int64_t K, int64_t lda_src, int aexp, void* hi,
                                   void* lo, void* stream) {
  MK_REQUIRE(a && hi && lo && M > 0 && K > 0 && lda_src >= K, MK_ERR_ARG, "bad arguments");
  MK_REQUIRE(aexp >= -512 && aexp <= 512, MK_ERR_ARG, "bad exponent %d", aexp);
  int64_t M_pad, lda;
  mk_split_bf16_padded_dims(M, K, &M_pad, &lda);
  split_matrix_f16_kernel<<<mk::grid_for(M_pad * lda, 256), 256, 0, mk::as_stream(stream)>>>(
      a, M, K, lda_src, aexp, (_Float16*)hi, (_Float16*)lo, M_pad, lda);
  return mk::check_launch("split_matrix_f16_kernel");
}

// desc per segment (11 int64): M, K, N, a_hi, a_mid, a_lo (device addresses), lda, b_off, b_ld, c_off, c_ld.
// Mode 2 reads a_hi, a_lo of mk_split_f16_matrix from fields 3 and 4 and the
// matrix's exponent from field 5.
extern "C" int mk_rot_plan_create(int nseg, const int64_t* desc, int mode, mk_rot_plan** out) {
  MK_REQUIRE(out, MK_ERR_ARG, "out is NULL");
  *out = nullptr;
  MK_REQUIRE(nseg >= 0 && (nseg == 0 || desc), MK_ERR_ARG, "bad segment table");
  MK_REQUIRE(mode == 0 || mode == 1 || mode == 2, MK_ERR_ARG,
             "mode must be 0 (3 bf16 products), 1 (6 bf16 products) or 2 (3 fp16 products)");
  std::vector<RotSeg> segs(nseg);
  int64_t base = 0;
  for (int i = 0; i < nseg; i++) {
    const int64_t* d = desc + 11 * i;
    RotSeg s{};
    s.M = (int32_t)d[0]; s.K = (int32_t)d[1]; s.N = (int32_t)d[2];
    s.a_hi = (const __bf16*)(uintptr_t)d[3];
    s.a_mid = (const __bf16*)(uintptr_t)d[4];
    if (mode == 2) {
      MK_REQUIRE(d[5] >= -512 && d[5] <= 512, MK_ERR_ARG, "segment %d: bad A exponent %lld", i, (long long)d[5]);
      s.a_lo = s.a_mid;
      s.pad_ = (int32_t)d[5];
    } else {
      s.a_lo = (const __bf16*)(uintptr_t)d[5];
    }
    s.lda = d[6];
    s.b_off = d[7]; s.b_ld = d[8]; s.c_off = d[9]; s.c_ld = d[10];
    MK_REQUIRE(d[0] > 0 && d[1] > 0 && d[2] >= 0 && d[0] < (1 << 20) && d[1] < (1 << 20) && d[2] < (1 << 30),
               MK_ERR_ARG, "segment %d: bad sizes", i);
    MK_REQUIRE(s.a_hi && s.a_mid && s.a_lo, MK_ERR_ARG, "segment %d: null A", i);
    MK_REQUIRE(s.lda % BK == 0 && s.lda >= s.K, MK_ERR_ARG, "segment %d: lda must be a multiple of %d >= K", i, BK);
    MK_REQUIRE(s.b_ld >= s.N && s.c_ld >= s.N && s.b_off >= 0 && s.c_off >= 0, MK_ERR_ARG,
               "segment %d: row strides must cover N and offsets be >= 0", i);
    MK_REQUIRE((int64_t)s.K * s.b_ld * 4 <= UINT32_MAX && (int64_t)s.M * s.c_ld * 4 <= UINT32_MAX &&
                   mk::cdiv(s.M, BM) * BM * s.lda * 2 <= UINT32_MAX,
               MK_ERR_ARG, "segment %d: B, C and A planes must each span less than 4 GiB", i);
    MK_REQUIRE(((uintptr_t)s.a_hi & 15) == 0 && ((uintptr_t)s.a_mid & 15) == 0 && ((uintptr_t)s.a_lo & 15) == 0,
               MK_ERR_ARG, "A must be 16-B aligned");
    s.mblocks = (int32_t)mk::cdiv(s.M, BM);
    s.nblocks = (int32_t)mk::cdiv(s.N, BN);
    s.base = base;
    base += (int64_t)s.mblocks * s.nblocks;
    segs[i] = s;
  }
  MK_REQUIRE(base < INT32_MAX, MK_ERR_ARG, "too many tiles");
  mk_rot_plan* p = new mk_rot_plan();
  p->nseg = nseg;
  p->mode = mode;
  p->ntiles = base;
  if (hipMalloc((void**)&p->d_segs, std::max(1, nseg) * sizeof(RotSeg)) != hipSuccess) {
    delete p;
    mk::set_error("rot plan: alloc failed");
    return MK_ERR_ALLOC;
  }
  if (nseg && hipMemcpy(p->d_segs, segs.data(), nseg * sizeof(RotSeg), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(p->d_segs);
    delete p;
    mk::set_error("rot plan: upload failed");
    return MK_ERR_HIP;
  }
  *out = p;
  return MK_OK;
}

extern "C" int mk_rot_plan_exec(const mk_rot_plan* p, const float* b_base, float* c_base, void* stream) {
  MK_REQUIRE(p && b_base && c_base, MK_ERR_ARG, "null argument");
  if (p->ntiles == 0) return MK_OK;
  int64_t grid = mk::cdiv(p->ntiles, 8) * 8;
  if (p->mode == 2) {
    rot_f16_kernel<<<(unsigned)grid, 256, 0, mk::as_stream(stream)>>>(p->d_segs, p->nseg, p->ntiles, b_base, c_base);
    return mk::check_launch("rot_f16_kernel");
  }
  if (p->mode == 1)
    rot_bf16_kernel<6><<<(unsigned)grid, 256, 0, mk::as_stream(stream)>>>(p->d_segs, p->nseg, p->ntiles, b_base,
                                                                          c_base);
  else
    rot_bf16_kernel<3><<<(unsigned)grid, 256, 0, mk::as_stream(stream)>>>(p->d_segs, p->nseg, p->ntiles, b_base,
                                                                          c_base);
  return mk::check_launch("rot_bf16_kernel");
}

extern "C" int mk_rot_plan_destroy(mk_rot_plan* p) {
  if (p) {
    if (p->d_segs) (void)hipFree(p->d_segs);
    delete p;
  }
  return MK_OK;
}

extern "C" int mk_gemm_nn_f64(int64_t M, int64_t N, int64_t K, const double* A, int64_t lda, const void* B,
                              int b_is_f32, int64_t ldb, double* C, int64_t ldc, void* stream) {
  MK_REQUIRE(M > 0 && N >= 0 && K > 0 && A && B && C, MK_ERR_ARG, "bad arguments");
  if (N == 0) return MK_OK;
  MK_REQUIRE(M < 65536 * TB, MK_ERR_ARG, "M too large");
  dim3 grid((unsigned)mk::cdiv(N, TB), (unsigned)mk::cdiv(M, TB));
  hipStream_t st = mk::as_stream(stream);
  if (b_is_f32)
    gemm_nn_f64<float><<<grid, 256, 0, st>>>((int)M, (int)N, (int)K, A, lda, (const float*)B, ldb, C, ldc);
  else
    gemm_nn_f64<double><<<grid, 256, 0, st>>>((int)M, (int)N, (int)K, A, lda, (const double*)B, ldb, C, ldc);
  return mk::check_launch("gemm_nn_f64");
}

extern "C" int64_t mk_gram_workspace_bytes(int64_t M, int64_t n) {
  int64_t chunk = 8192;
  int64_t nchunk = std::max<int64_t>(1, mk::cdiv(n, chunk));
  return nchunk * M * M * 8 + 256;
}

extern "C" int mk_gram_f64(int64_t M, int64_t n, const void* D, int d_is_f32, int64_t ldd, double* G, void* work,
                           int64_t work_bytes, void* stream) {
  MK_REQUIRE(M > 0 && n > 0 && D && G && work, MK_ERR_ARG, "bad arguments");
  MK_REQUIRE(work_bytes >= mk_gram_workspace_bytes(M, n), MK_ERR_ARG, "workspace too small");
  int64_t chunk = 8192;
  int64_t nchunk = std::max<int64_t>(1, mk::cdiv(n, chunk));
  MK_REQUIRE(nchunk < 65536, MK_ERR_ARG, "n too large");
  dim3 grid((unsigned)mk::cdiv(M, TB), (unsigned)mk::cdiv(M, TB), (unsigned)nchunk);
  hipStream_t st = mk::as_stream(stream);
  if (d_is_f32)
    cross_partial_f64<float, float><<<grid, 256, 0, st>>>((int)M, (int)M, n, (const float*)D, ldd, (const float*)D,
                                                          ldd, chunk, (double*)work);
  else
    cross_partial_f64<double, double><<<grid, 256, 0, st>>>((int)M, (int)M, n, (const double*)D, ldd,
                                                            (const double*)D, ldd, chunk, (double*)work);
  int r = mk::check_launch("cross_partial_f64");
  if (r) return r;
  sum_partials_f64<<<mk::grid_for(M * M, 256), 256, 0, st>>>((const double*)work, nchunk, M * M, G);
  return mk::check_launch("sum_partials_f64");
}

extern "C" int64_t mk_cross_workspace_bytes(int64_t M, int64_t N, int64_t n) {
  int64_t nchunk = std::max<int64_t>(1, mk::cdiv(n, 8192));
  return nchunk * M * N * 8 + 256;
}

extern "C" int mk_cross_f64(int64_t M, int64_t N, int64_t n, const double* A, int64_t lda, const double* B,
                            int64_t ldb, double* C, void* work, int64_t work_bytes, void* stream) {
  MK_REQUIRE(M > 0 && N > 0 && n > 0 && A && B && C && work, MK_ERR_ARG, "bad arguments");
  MK_REQUIRE(M < 65536 * TB && N < 65536 * TB && lda >= n && ldb >= n, MK_ERR_ARG, "bad shapes");
  MK_REQUIRE(work_bytes >= mk_cross_workspace_bytes(M, N, n), MK_ERR_ARG, "workspace too small");
  const int64_t chunk = 8192;
  const int64_t nchunk = std::max<int64_t>(1, mk::cdiv(n, chunk));
  MK_REQUIRE(nchunk < 65536, MK_ERR_ARG, "n too large");
  dim3 grid((unsigned)mk::cdiv(N, TB), (unsigned)mk::cdiv(M, TB), (unsigned)nchunk);
  hipStream_t st = mk::as_stream(stream);
  cross_partial_f64<double, double><<<grid, 256, 0, st>>>((int)M, (int)N, n, A, lda, B, ldb, chunk, (double*)work);
  int r = mk::check_launch("cross_partial_f64");
  if (r) return r;
  sum_partials_f64<<<mk::grid_for(M * N, 256), 256, 0, st>>>((const double*)work, nchunk, M * N, C);
  return mk::check_launch("sum_partials_f64");
}

// ---------------------------------------------------------------------------
// Deflation of dominant detector modes ahead of the FP32 rotation.
//
// The rows of v are eigenvectors of the detector covariance.  When a few
// modes carry most of the power (a strong common mode), the small rows of
// y = v.d come out of heavy cancellation, and FP32 accumulation leaves them
// with a relative error of about 1e-8 times the ratio of the largest row
// norm to theirs.  With U the leading k eigenvectors (the first k rows of v,
// transposed), c = U^T d (summed in FP64) and d_rest = d - U c,
//   v.d = v.d_rest + [c; 0]   (v U = [I; 0]),
// so the rotation sees d_rest, which holds no dominant mode, and c is added
// back into the first k rows of its output.

namespace {

constexpr int kMaxDefl = 8;
constexpr int kDeflCols = 256;  // columns (threads) per workgroup
constexpr int kDeflU = 8;       // rows whose loads are in flight together

struct DeflSeg {
  int64_t off;       // first element of the segment (rows of n, row stride n)
  int64_t u_off;     // U of the segment: ndet x k, row-major
  int64_t c_off;     // c of the segment: k x n
  int64_t blk_base;  // first workgroup of the segment
  int32_t ndet, n, k, pad;
};

__device__ inline int find_dseg(const DeflSeg* segs, int nseg, int64_t b) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].blk_base <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One workgroup per 256 sample columns of one segment (so the segment and U
// are uniform per workgroup), one thread per column: c = U^T x over the
// column (FP64), then out = x - U c, kDeflU rows' loads in flight at a time.
// out may equal in (each thread reads its own column before writing it).
__global__ void __launch_bounds__(kDeflCols) defl_kernel(const DeflSeg* __restrict__ segs, int nseg, const float* in,
                                                         float* out, const double* __restrict__ U,
                                                         double* __restrict__ C) {
  const DeflSeg s = segs[find_dseg(segs, nseg, blockIdx.x)];
  const int64_t t = (int64_t)(blockIdx.x - s.blk_base) * kDeflCols + threadIdx.x;
  if (t >= s.n) return;
  const float* x = in + s.off + t;
  float* y = out + s.off + t;
  const double* u = U + s.u_off;
  const int ndet = s.ndet, k = s.k;
  const int64_t n = s.n;
  double acc[kMaxDefl];
#pragma unroll
  for (int j = 0; j < kMaxDefl; j++) acc[j] = 0.0;
  int i = 0;
  for (; i + kDeflU <= ndet; i += kDeflU) {
    float xv[kDeflU];
#pragma unroll
    for (int q = 0; q < kDeflU; q++) xv[q] = x[(int64_t)(i + q) * n];
#pragma unroll
    for (int q = 0; q < kDeflU; q++)
#pragma unroll
      for (int j = 0; j < kMaxDefl; j++)
        if (j < k) acc[j] += u[(i + q) * k + j] * (double)xv[q];
  }
  for (; i < ndet; i++) {
    const double xv = (double)x[(int64_t)i * n];
#pragma unroll
    for (int j = 0; j < kMaxDefl; j++)
      if (j < k) acc[j] += u[i * k + j] * xv;
  }
#pragma unroll
  for (int j = 0; j < kMaxDefl; j++)
    if (j < k) C[s.c_off + (int64_t)j * n + t] = acc[j];
  i = 0;
  for (; i + kDeflU <= ndet; i += kDeflU) {
    float xv[kDeflU];
#pragma unroll
    for (int q = 0; q < kDeflU; q++) xv[q] = x[(int64_t)(i + q) * n];
#pragma unroll
    for (int q = 0; q < kDeflU; q++) {
      double r = (double)xv[q];
#pragma unroll
      for (int j = 0; j < kMaxDefl; j++)
        if (j < k) r -= u[(i + q) * k + j] * acc[j];
      y[(int64_t)(i + q) * n] = (float)r;
    }
  }
  for (; i < ndet; i++) {
    double r = (double)x[(int64_t)i * n];
#pragma unroll
    for (int j = 0; j < kMaxDefl; j++)
      if (j < k) r -= u[i * k + j] * acc[j];
    y[(int64_t)i * n] = (float)r;
  }
}

// y[j][t] += c[j][t] for the first k rows of every segment.
__global__ void __launch_bounds__(kDeflCols) defl_addback_kernel(const DeflSeg* __restrict__ segs, int nseg,
                                                                 float* __restrict__ Y, const double* __restrict__ C) {
  const DeflSeg s = segs[find_dseg(segs, nseg, blockIdx.x)];
  const int64_t t = (int64_t)(blockIdx.x - s.blk_base) * kDeflCols + threadIdx.x;
  if (t >= s.n) return;
  for (int j = 0; j < s.k; j++) {
    float* y = Y + s.off + (int64_t)j * s.n + t;
    *y = (float)((double)*y + C[s.c_off + (int64_t)j * s.n + t]);
  }
}

}  // namespace

struct mk_defl_plan {
  int nseg = 0;
  int64_t ncols = 0, nblk = 0;
  DeflSeg* d_segs = nullptr;
  double* u = nullptr;
  double* c = nullptr;
};

extern "C" int mk_defl_plan_destroy(mk_defl_plan* p) {
  if (p) {
    if (p->d_segs) (void)hipFree(p->d_segs);
    if (p->u) (void)hipFree(p->u);
    if (p->c) (void)hipFree(p->c);
    delete p;
  }
  return MK_OK;
}

extern "C" int mk_defl_plan_create(int nseg, const int64_t* desc, const double* u_host, mk_defl_plan** out) {
  MK_REQUIRE(out, MK_ERR_ARG, "out is NULL");
  *out = nullptr;
  MK_REQUIRE(nseg >= 0 && (nseg == 0 || (desc && u_host)), MK_ERR_ARG, "bad segment table");
  std::vector<DeflSeg> segs(nseg);
  int64_t ncols = 0, nu = 0, nc = 0, nblk = 0;
  for (int i = 0; i < nseg; i++) {
    const int64_t* d = desc + 4 * i;
    MK_REQUIRE(d[0] > 0 && d[0] < (1 << 20) && d[1] > 0 && d[1] < (1 << 30) && d[2] >= 0 && d[3] >= 1 &&
                   d[3] <= kMaxDefl && d[3] <= d[0],
               MK_ERR_ARG, "segment %d: need 1 <= k <= min(ndet, %d), sizes > 0, offset >= 0", i, kMaxDefl);
    DeflSeg& s = segs[i];
    s.ndet = (int32_t)d[0]; s.n = (int32_t)d[1]; s.off = d[2]; s.k = (int32_t)d[3]; s.pad = 0;
    s.u_off = nu; s.c_off = nc; s.blk_base = nblk;
    nu += (int64_t)s.ndet * s.k;
    nc += (int64_t)s.k * s.n;
    ncols += s.n;
    nblk += mk::cdiv(s.n, kDeflCols);
  }
  MK_REQUIRE(nblk < INT32_MAX, MK_ERR_ARG, "deflation plan: too many columns");
  mk_defl_plan* p = new mk_defl_plan();
  p->nseg = nseg;
  p->ncols = ncols;
  p->nblk = nblk;
  if (hipMalloc((void**)&p->d_segs, std::max<size_t>(1, nseg) * sizeof(DeflSeg)) != hipSuccess ||
      hipMalloc((void**)&p->u, std::max<int64_t>(1, nu) * 8) != hipSuccess ||
      hipMalloc((void**)&p->c, std::max<int64_t>(1, nc) * 8) != hipSuccess) {
    mk_defl_plan_destroy(p);
    mk::set_error("deflation plan: device allocation failed");
    return MK_ERR_ALLOC;
  }
  if ((nseg && hipMemcpy(p->d_segs, segs.data(), nseg * sizeof(DeflSeg), hipMemcpyHostToDevice) != hipSuccess) ||
      (nu && hipMemcpy(p->u, u_host, nu * 8, hipMemcpyHostToDevice) != hipSuccess)) {
    mk_defl_plan_destroy(p);
    mk::set_error("deflation plan: upload failed");
    return MK_ERR_HIP;
  }
  *out = p;
  return MK_OK;
}

extern "C" int mk_defl_plan_exec(const mk_defl_plan* p, const float* in, float* out, void* stream) {
  MK_REQUIRE(p, MK_ERR_ARG, "null plan");
  if (p->ncols == 0) return MK_OK;
  MK_REQUIRE(in && out, MK_ERR_ARG, "null pointer");
  defl_kernel<<<(unsigned)p->nblk, kDeflCols, 0, mk::as_stream(stream)>>>(p->d_segs, p->nseg, in, out, p->u, p->c);
  return mk::check_launch("defl_kernel");
}

extern "C" int mk_defl_plan_addback(const mk_defl_plan* p, float* y, void* stream) {
  MK_REQUIRE(p, MK_ERR_ARG, "null plan");
  if (p->ncols == 0) return MK_OK;
  MK_REQUIRE(y, MK_ERR_ARG, "null pointer");
  defl_addback_kernel<<<(unsigned)p->nblk, kDeflCols, 0, mk::as_stream(stream)>>>(p->d_segs, p->nseg, y, p->c);
  return mk::check_launch("defl_addback_kernel");
}

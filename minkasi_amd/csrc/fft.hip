// Batched four-step FFT engine and the two transforms built on it.
//
// NoiseSmoothedSVD.apply_noise (minkasi/mapmaking/noise.py:704-711) applies,
// per rotated detector mode y (length n), y'' = C W C y where C is the
// unnormalised DCT-I (FFTW REDFT00, minkasi/lib/mkfftw.c:174-212;
// scipy.fft.dct type 1, minkasi/tools/py_fft.py:239) and W = mywt.  With
// m = n-1 and N = 2m, C W C y = (even extension of y) circularly convolved
// (period N) with h = C W (because C C = N I).  The engine evaluates that
// circular convolution exactly as a linear one through a real FFT of
// L >= 3m+1 points: the input is the even periodic extension z' on
// [-m, 2m], the kernel is h placed symmetrically (so its spectrum B is real
// and even), and the outputs needed are the window [m, 2m].  This works for
// any n (no Bluestein, no smooth-length requirement) and fuses the forward
// transform, the weight multiply and the inverse transform.  L is the
// smallest 4096 P1 >= 3m+1 with P1 5-smooth (<= 256) on the FP32 hot path,
// else the next power of two.
//
// Stand-alone DCT-I (fft_r2r, minkasi/tools/fft.py:290-344) uses Bluestein's
// chirp-z algorithm (Bluestein 1970; Rabiner, Schafer & Rader 1969) on the
// same engine: Y_k = Re(chi_k sum_j c_j x_j chi_j conj(chi_{k-j})) with
// chi_j = exp(-i pi j^2 / N), c_j = 1 at j = 0, m and 2 otherwise, the
// convolution evaluated with a complex FFT of P = 2^q >= 2m+1 points.
//
// Four-step layout (P = P1*P2 complex points, j = P2*j1 + j2, k = k1 + P1*k2):
//   K1: column FFTs of length P1 (one workgroup = cw adjacent columns, so
//       global reads are contiguous runs), twiddle w_P^(j2*k1), store T[k1][j2].
//   K2: row FFTs of length P2 (rows k1 and P1-k1 together, which is where the
//       real-FFT unpacking partner C_(P-k) lives), pointwise spectral work,
//       inverse row FFTs, conjugate twiddle, store back.
//   K3: inverse column FFTs of length P1, 1/P scale, write the wanted outputs.
// Every LDS sub-FFT is a Stockham pass sequence (radix 8, then 4 or 2; radix
// 3 and 5 for mixed column lengths) with a stage-ordered twiddle table.
// Each launch covers every row of every segment (TOD) in the batch through
// a work list.
#include "common.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace {

template <typename T>
struct cx {
  T re, im;
};
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.re * s, a.im * s}; }

__device__ __forceinline__ void sincospi_t(float x, float* s, float* c) { sincospif(x, s, c); }
__device__ __forceinline__ void sincospi_t(double x, double* s, double* c) { sincospi(x, s, c); }

// exp(-i*pi*num/den) for integers 0 <= num < 2*den.  expmi2 requires den a
// power of two and num < 2^24, so -num * (1/den) is exact in T.
template <typename T>
__device__ __forceinline__ cx<T> expmi(int64_t num, int64_t den) {
  T s, c;
  sincospi_t((T)(-(double)num / (double)den), &s, &c);
  return {c, s};
}
template <typename T>
__device__ __forceinline__ cx<T> expmi2(int64_t num, T inv_den) {
  T s, c;
  sincospi_t(-(T)num * inv_den, &s, &c);
  return {c, s};
}

constexpr int kThreads = 512;                  // threads per FFT workgroup
constexpr int kMaxWG = 4096;                   // complex points per workgroup in an LDS FFT (8 per thread)
constexpr int kMaxLds = 144 * 1024;            // dynamic LDS budget per workgroup (gfx950 has 160 KB)

// LDS rows are padded: element e of a row lives at e + e/16 (measured best of
// e/8, e/16, e/32 and none on gfx950) and rows are
// row_stride(len) apart (odd), so the strided Stockham writes of the early
// radix-8 stages, the row-parallel writes of K1/K3 and the unit-stride reads
// all map to distinct banks.
#ifndef MK_FFT_PAD_SHIFT
#define MK_FFT_PAD_SHIFT 4
#endif
constexpr __host__ __device__ __forceinline__ int pad8(int e) { return e + (e >> MK_FFT_PAD_SHIFT); }
// Odd, so that rows read in parallel fall in different banks (for a power of
// two len this is len + len/16 + 1).
constexpr __host__ __device__ __forceinline__ int row_stride(int len) { return (len + (len >> MK_FFT_PAD_SHIFT)) | 1; }

// In-register DFTs (natural order in and out).
template <typename T, bool INV>
__device__ __forceinline__ void dft2(cx<T>* v) {
  cx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <typename T, bool INV>
__device__ __forceinline__ void dft4(cx<T>& a0, cx<T>& a1, cx<T>& a2, cx<T>& a3) {
  cx<T> b0 = cadd(a0, a2), b1 = csub(a0, a2), b2 = cadd(a1, a3), d = csub(a1, a3);
  cx<T> b3 = INV ? cx<T>{-d.im, d.re} : cx<T>{d.im, -d.re};  // (+/-i) * d
  a0 = cadd(b0, b2);
  a2 = csub(b0, b2);
  a1 = cadd(b1, b3);
  a3 = csub(b1, b3);
}
template <typename T, bool INV>
__device__ __forceinline__ void dft4v(cx<T>* v) { dft4<T, INV>(v[0], v[1], v[2], v[3]); }
template <typename T, bool INV>
__device__ __forceinline__ void dft8(cx<T>* v) {
  const T c = T(0.70710678118654752440);
  cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<T, INV>(e0, e1, e2, e3);
  dft4<T, INV>(o0, o1, o2, o3);
  if (!INV) {
    o1 = {c * (o1.re + o1.im), c * (o1.im - o1.re)};   // * exp(-i pi/4)
    o2 = {o2.im, -o2.re};                               // * -i
    o3 = {c * (o3.im - o3.re), -c * (o3.re + o3.im)};  // * exp(-3i pi/4)
  } else {
    o1 = {c * (o1.re - o1.im), c * (o1.re + o1.im)};   // * exp(i pi/4)
    o2 = {-o2.im, o2.re};                               // * i
    o3 = {-c * (o3.re + o3.im), c * (o3.re - o3.im)};  // * exp(3i pi/4)
  }
  v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
}

// One radix-R Stockham stage over nrows padded rows of length 2^l2len.
// Every thread loads all of its butterflies' inputs, the workgroup
// synchronises, then the outputs are written in place.
// tw: this stage's twiddles, tw[(r-1)*Ns + k] = exp(-2 pi i k r / (R Ns)),
// so the lanes of a wave (consecutive k) read consecutive words.
template <typename T, bool INV, int LR>
__device__ __forceinline__ void fft_stage(cx<T>* buf, const cx<T>* tw, int l2len, int nrows, int rs, int Ns,
                                          int lns) {
  constexpr int R = 1 << LR;
  constexpr int NB = kMaxWG / R / kThreads;
  const int lq = l2len - LR, q = 1 << lq;
  const int nbf = nrows << lq;
  cx<T> v[NB][R];
#pragma unroll
  for (int u = 0; u < NB; u++) {
    int bf = threadIdx.x + u * kThreads;
    if (bf < nbf) {
      int row = bf >> lq, j = bf & (q - 1), k = j & (Ns - 1);
      const cx<T>* base = buf + row * rs;
#pragma unroll
      for (int r = 0; r < R; r++) v[u][r] = base[pad8(j + r * q)];
      if (lns > 0) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          cx<T> w = tw[(r - 1) * Ns + k];
          if (INV) w = cconj(w);
          v[u][r] = cmul(v[u][r], w);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NB; u++) {
    int bf = threadIdx.x + u * kThreads;
    if (bf < nbf) {
      int row = bf >> lq, j = bf & (q - 1), k = j & (Ns - 1);
      if (LR == 3) dft8<T, INV>(v[u]);
      else if (LR == 2) dft4v<T, INV>(v[u]);
      else dft2<T, INV>(v[u]);
      cx<T>* base = buf + row * rs;
      int idx = ((j >> lns) << (lns + LR)) + k;
#pragma unroll
      for (int r = 0; r < R; r++) base[pad8(idx + r * Ns)] = v[u][r];
    }
  }
  __syncthreads();
}

// Radix of the stage that follows `rem` remaining levels: 8 while three or
// more levels remain, then 4 or 2.
__host__ __device__ __forceinline__ int stage_lr(int rem) { return rem >= 3 ? 3 : rem; }

// In-place FFT of nrows padded rows (row stride rs) of length 2^l2len,
// nrows * 2^l2len <= kMaxWG; tw = stage-ordered table (fill_stage_table).
template <typename T, bool INV>
__device__ void lds_fft(cx<T>* buf, const cx<T>* tw, int l2len, int nrows, int rs) {
  int Ns = 1, lns = 0, rem = l2len, off = 0;
  while (rem > 0) {
    const int lr = stage_lr(rem);
    if (lr == 3) fft_stage<T, INV, 3>(buf, tw + off, l2len, nrows, rs, Ns, lns);
    else if (lr == 2) fft_stage<T, INV, 2>(buf, tw + off, l2len, nrows, rs, Ns, lns);
    else fft_stage<T, INV, 1>(buf, tw + off, l2len, nrows, rs, Ns, lns);
    off += ((1 << lr) - 1) * Ns;
    Ns <<= lr; lns += lr; rem -= lr;
  }
}

// Twiddles of every stage of a length-2^l2len FFT, stage after stage:
// entry (r-1)*Ns + k of a stage is exp(-2 pi i k r / (R Ns)); 2^l2len - 1
// entries in all.
template <typename T>
__device__ void fill_stage_table(cx<T>* tab, int l2len) {
  const T inv = T(1) / (T)(1 << l2len);
  int Ns = 1, rem = l2len, off = 0;
  while (rem > 0) {
    const int lr = stage_lr(rem), R = 1 << lr;
    const int cnt = (R - 1) * Ns;
    const int mul = 1 << (l2len - lr - (l2len - rem));  // len / (R Ns)
    for (int e = threadIdx.x; e < cnt; e += kThreads) {
      int r = e / Ns + 1, k = e - (r - 1) * Ns;
      tab[off + e] = expmi2<T>(2 * (int64_t)k * r * mul, inv);
    }
    off += cnt;
    Ns <<= lr; rem -= lr;
  }
}

// exp(-2 pi i a / P) for 0 <= a < P, P = 2^lP <= 2^23.
template <typename T>
__device__ __forceinline__ cx<T> twiddle_P(int64_t a, T invP) { return expmi2<T>(2 * a, invP); }

// ---- Mixed radix (2, 3, 5) transform lengths ---------------------------------
// Radix of the next stage of an N-point Stockham FFT when rem = N / Ns points
// per sub-transform remain to be combined: the power-of-two part first (8
// while 8 divides rem, then one 4 or 2; for N = 2^q this is the stage_lr
// sequence), then 3s, then 5s.  Stage tables are laid out in this order.
constexpr __host__ __device__ int next_radix(int rem) {
  return (rem & 1) == 0 ? ((rem & 7) == 0 ? 8 : ((rem & 3) == 0 ? 4 : 2)) : (rem % 3 == 0 ? 3 : 5);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(cx<T>* v) {
  const T h = T(0.86602540378443864676);  // sin(2 pi / 3)
  cx<T> a = cadd(v[1], v[2]), d = csub(v[1], v[2]);
  cx<T> t = {v[0].re - T(0.5) * a.re, v[0].im - T(0.5) * a.im};
  cx<T> u = INV ? cx<T>{-h * d.im, h * d.re} : cx<T>{h * d.im, -h * d.re};  // (+/-i) sin(2pi/3) d
  v[0] = cadd(v[0], a);
  v[1] = cadd(t, u);
  v[2] = csub(t, u);
}
template <typename T, bool INV>
__device__ __forceinline__ void dft5(cx<T>* v) {
  const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);  // cos(2pi/5), cos(4pi/5)
  const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);   // sin(2pi/5), sin(4pi/5)
  cx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]), a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  cx<T> t1 = {v[0].re + c1 * a1.re + c2 * a2.re, v[0].im + c1 * a1.im + c2 * a2.im};
  cx<T> t2 = {v[0].re + c2 * a1.re + c1 * a2.re, v[0].im + c2 * a1.im + c1 * a2.im};
  cx<T> u1 = {s1 * b1.re + s2 * b2.re, s1 * b1.im + s2 * b2.im};
  cx<T> u2 = {s2 * b1.re - s1 * b2.re, s2 * b1.im - s1 * b2.im};
  // forward: y1 = t1 - i u1, y4 = t1 + i u1, y2 = t2 - i u2, y3 = t2 + i u2
  cx<T> iu1 = INV ? cx<T>{-u1.im, u1.re} : cx<T>{u1.im, -u1.re};
  cx<T> iu2 = INV ? cx<T>{-u2.im, u2.re} : cx<T>{u2.im, -u2.re};
  v[0] = cadd(v[0], cadd(a1, a2));
  v[1] = cadd(t1, iu1);
  v[4] = csub(t1, iu1);
  v[2] = cadd(t2, iu2);
  v[3] = csub(t2, iu2);
}
template <typename T, bool INV, int R>
__device__ __forceinline__ void dft_r(cx<T>* v) {
  if constexpr (R == 8) dft8<T, INV>(v);
  else if constexpr (R == 4) dft4v<T, INV>(v);
  else if constexpr (R == 2) dft2<T, INV>(v);
  else if constexpr (R == 3) dft3<T, INV>(v);
  else dft5<T, INV>(v);
}

// One radix-R stage of an N-point mixed-radix Stockham FFT over nrows padded
// rows (the runtime counterpart of fft_stage): butterfly j of a row reads
// j + r N/R, twiddles by tw[(r-1) Ns + k] (k = j mod Ns) and writes
// (j / Ns) Ns R + k + r Ns.
template <typename T, bool INV, int R>
__device__ __forceinline__ void fft_stage_mx(cx<T>* buf, const cx<T>* tw, int N, int nrows, int rs, int Ns) {
  constexpr int NB = (kMaxWG / R + kThreads - 1) / kThreads;
  const int q = N / R, nbf = nrows * q;
  cx<T> v[NB][R];
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int bf = threadIdx.x + u * kThreads;
    if (bf < nbf) {
      const int row = bf / q, j = bf - row * q, k = j % Ns;
      const cx<T>* base = buf + row * rs;
#pragma unroll
      for (int r = 0; r < R; r++) v[u][r] = base[pad8(j + r * q)];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          cx<T> w = tw[(r - 1) * Ns + k];
          if (INV) w = cconj(w);
          v[u][r] = cmul(v[u][r], w);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int bf = threadIdx.x + u * kThreads;
    if (bf < nbf) {
      const int row = bf / q, j = bf - row * q, k = j % Ns;
      dft_r<T, INV, R>(v[u]);
      cx<T>* base = buf + row * rs;
      const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) base[pad8(idx + r * Ns)] = v[u][r];
    }
  }
  __syncthreads();
}

// In-place N-point FFT (N 5-smooth) of nrows padded rows, nrows * N <= kMaxWG;
// tw = fill_stage_table_mx table.
template <typename T, bool INV>
__device__ void lds_fft_mx(cx<T>* buf, const cx<T>* tw, int N, int nrows, int rs) {
  int Ns = 1, off = 0;
  while (Ns < N) {
    const int R = next_radix(N / Ns);
    switch (R) {
      case 8: fft_stage_mx<T, INV, 8>(buf, tw + off, N, nrows, rs, Ns); break;
      case 4: fft_stage_mx<T, INV, 4>(buf, tw + off, N, nrows, rs, Ns); break;
      case 2: fft_stage_mx<T, INV, 2>(buf, tw + off, N, nrows, rs, Ns); break;
      case 3: fft_stage_mx<T, INV, 3>(buf, tw + off, N, nrows, rs, Ns); break;
      default: fft_stage_mx<T, INV, 5>(buf, tw + off, N, nrows, rs, Ns); break;
    }
    off += (R - 1) * Ns;
    Ns *= R;
  }
}

// Stage table of an N-point mixed-radix FFT (next_radix order): entry
// (r-1) Ns + k of a stage is exp(-2 pi i k r / (R Ns)); N - 1 entries.
template <typename T>
__device__ void fill_stage_table_mx(cx<T>* tab, int N) {
  int Ns = 1, off = 0;
  while (Ns < N) {
    const int R = next_radix(N / Ns), cnt = (R - 1) * Ns;
    for (int e = threadIdx.x; e < cnt; e += blockDim.x) {
      const int r = e / Ns + 1, k = e - (r - 1) * Ns;
      tab[off + e] = expmi<T>(2 * (int64_t)k * r, (int64_t)R * Ns);
    }
    off += cnt;
    Ns *= R;
  }
}

// exp(-2 pi i a / P) for 0 <= a < P and any P (FP64 argument reduction).
template <typename T>
__device__ __forceinline__ cx<T> twiddle_any(int64_t a, int64_t P) { return expmi<T>(2 * a, P); }

// Segment: `rows` rows sharing one length n and one plan.
struct Seg {
  int64_t in_off, in_stride;    // input rows (elements of the input type)
  int64_t out_off, out_stride;  // output rows
  int64_t t_off, t_stride;      // T buffer rows (complex)
  int64_t s_off, s_stride;      // spectrum rows (filter: P+1 reals; Bluestein: P complex, stride 0)
  void* s_ptr;                  // device address of this segment's spectrum rows (set at upload)
  void* tw_ptr;                 // hot K2: fine/coarse twiddle tables of this P (set at create)
  int32_t rows, n, l1, l2, cw, flags;  // l1 = log2(P1), or -1 when P1 is not a power of two
  int64_t base1, base2, base3;  // first global item of K1, K2, K3
  int32_t per1, per2;           // items per row for K1/K3, K2
  int32_t r2, p1;               // k1 values per K2 item; column length P1 (2^a 3^b 5^c)
};

enum Mode { kFilter = 0, kFSpec = 1, kBlue = 2, kBSpec = 3 };

__device__ inline int find_seg(const Seg* segs, int nseg, int64_t g, int which) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    int64_t b = which == 1 ? segs[mid].base1 : (which == 2 ? segs[mid].base2 : segs[mid].base3);
    if (b <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---- K1: column FFTs -------------------------------------------------------
// LDS: cw padded rows of length P1 (one per matrix column j2), then the P1
// twiddle table and the P2 table exp(-2 pi i j/P) for the four-step twiddle.
template <typename T, typename TIn, int MODE>
__global__ void __launch_bounds__(kThreads, sizeof(T) == 4 ? 4 : 1) k1_kernel(const Seg* __restrict__ segs, int nseg,
                                                      const TIn* __restrict__ in, cx<T>* __restrict__ tbuf) {
  extern __shared__ unsigned char smem_raw[];
  const int s = find_seg(segs, nseg, blockIdx.x, 1);
  const Seg sg = segs[s];
  const int64_t local = (int64_t)blockIdx.x - sg.base1;
  const int row = (int)(local / sg.per1), jb = (int)(local % sg.per1);
  const int l1 = sg.l1, l2 = sg.l2, P1 = sg.p1, P2 = 1 << l2, cw = sg.cw;
  // Mixed-radix column lengths reach this kernel only in the spectrum pass
  // (the filter runs them on the hot path), so only that mode carries the
  // mixed-radix code.
  const bool pow2 = MODE != kFSpec || l1 >= 0;
  const int lcw = 31 - __clz(cw);
  const int rs = row_stride(P1);
  const int64_t P = (int64_t)P1 * P2;
  const int64_t m = sg.n - 1, N = 2 * m;
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem_raw);  // [cw][rs]
  cx<T>* tw = buf + cw * rs;                          // [P1] stage table
  const TIn* x = in + sg.in_off + (int64_t)row * sg.in_stride;
  const int j2_0 = jb * cw;
  if (pow2) fill_stage_table<T>(tw, l1);
  else if constexpr (MODE == kFSpec) fill_stage_table_mx<T>(tw, P1);
  constexpr int UNR = 8;  // loads in flight per thread
  if (MODE == kFilter || MODE == kFSpec) {
    // real input packed two per complex point: a_(2j), a_(2j+1)
    const int64_t L = 2 * P;
    const int total = P1 * cw * 2;
    for (int base = threadIdx.x; base < total; base += kThreads * UNR) {
      T val[UNR];
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        int idx = base + u * kThreads;
        val[u] = T(0);
        if (idx < total) {
          int j1 = idx >> (lcw + 1), w = idx & (2 * cw - 1);
          int64_t i = 2 * ((int64_t)P2 * j1 + j2_0) + w;  // a-index
          if (MODE == kFilter) {
            if (i <= 3 * m) {
              int64_t t = i - m;
              int64_t src = t < 0 ? -t : (t <= m ? t : 2 * m - t);
              val[u] = (T)x[src];
            }
          } else {
            // symmetric embedding of h (rows of DCT-I(W)): b_i = h_i (i < m),
            // h_m/2 at i = m and L-m, h_(L-i) for L-m < i < L
            if (i < m) val[u] = (T)x[i];
            else if (i == m || i == L - m) val[u] = (T)x[m] * T(0.5);
            else if (i > L - m) val[u] = (T)x[L - i];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        int idx = base + u * kThreads;
        if (idx < total) {
          int j1 = idx >> (lcw + 1), w = idx & (2 * cw - 1);
          T* dst = reinterpret_cast<T*>(buf + (w >> 1) * rs + pad8(j1));
          dst[w & 1] = val[u];
        }
      }
    }
  } else {
    for (int idx = threadIdx.x; idx < P1 * cw; idx += kThreads) {
      int j1 = idx >> lcw, c = idx & (cw - 1);
      int64_t j = (int64_t)P2 * j1 + j2_0 + c;
      cx<T> val{T(0), T(0)};
      if (MODE == kBlue) {
        if (j <= m) {
          T wgt = (j == 0 || j == m) ? T(1) : T(2);
          cx<T> ch = expmi<T>((j * j) % (2 * N), N);
          val = cscale(ch, wgt * (T)x[j]);
        }
      } else {  // kBSpec: conj(chirp) at lags -m..m, placed mod P
        int64_t t = j <= m ? j : (j >= P - m ? P - j : -1);
        if (t >= 0) val = cconj(expmi<T>((t * t) % (2 * N), N));
      }
      buf[c * rs + pad8(j1)] = val;
    }
  }
  __syncthreads();
  if (pow2) lds_fft<T, false>(buf, tw, l1, cw, rs);
  else if constexpr (MODE == kFSpec) lds_fft_mx<T, false>(buf, tw, P1, cw, rs);
  cx<T>* trow = tbuf + sg.t_off + (int64_t)row * sg.t_stride;
  const T invP = T(1) / (T)P;
  for (int idx = threadIdx.x; idx < P1 * cw; idx += kThreads) {
    int k1 = idx >> lcw, c = idx & (cw - 1);
    int64_t j2 = j2_0 + c;
    int64_t a = (j2 * k1) % P;
    cx<T> v = cmul(buf[c * rs + pad8(k1)], pow2 ? twiddle_P<T>(a, invP) : twiddle_any<T>(a, P));
    trow[(int64_t)k1 * P2 + j2] = v;
  }
}

// ---- K2: row FFTs + spectral work + inverse row FFTs ----------------------
// One workgroup takes r2 consecutive k1 values of one row (TOD detector
// mode).  Real modes also load each partner row P1-k1, where the real-FFT
// unpacking partner C_(P-k) lives, so every element's partner is in LDS and
// one thread handles the pair (k, P-k) end to end: unpack, weight, repack,
// in place (w_L^(P-k) = -conj(w_L^k), so one sincospi per pair).
// TS: storage type of real spectra (the FP32 filter keeps FP32 spectra that
// were computed in FP64).
template <typename T, typename TS, int MODE>
__global__ void __launch_bounds__(kThreads, sizeof(T) == 4 ? 4 : 1) k2_kernel(const Seg* __restrict__ segs, int nseg,
                                                      cx<T>* __restrict__ tbuf) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ int slot_k1[16], slot_partner[16];
  __shared__ int nslot_sh;
  const int s = find_seg(segs, nseg, blockIdx.x, 2);
  const Seg sg = segs[s];
  const int64_t local = (int64_t)blockIdx.x - sg.base2;
  const int row = (int)(local / sg.per2), grp = (int)(local % sg.per2);
  const int l2 = sg.l2, P1 = sg.p1, P2 = 1 << l2;
  const bool pow2 = sg.l1 >= 0;
  const int rs = row_stride(P2);
  const int64_t P = (int64_t)P1 * P2;
  const bool real_mode = (MODE == kFilter || MODE == kFSpec);
  if (threadIdx.x == 0) {
    int ns = 0;
    const int kmax = real_mode ? P1 / 2 : P1 - 1;
    for (int q = 0; q < sg.r2; q++) {
      int k1 = grp * sg.r2 + q;
      if (k1 > kmax) break;
      int a = ns++;
      slot_k1[a] = k1;
      slot_partner[a] = a;
      int k1b = k1 == 0 ? 0 : P1 - k1;
      if (real_mode && k1b != k1) {
        int b = ns++;
        slot_k1[b] = k1b;
        slot_partner[a] = b;
        slot_partner[b] = a;
      }
    }
    nslot_sh = ns;
  }
  cx<T>* X = reinterpret_cast<cx<T>*>(smem_raw);             // [nslot][rs]
  cx<T>* tw = X + (real_mode ? 2 * sg.r2 : sg.r2) * rs;      // [P2] stage table
  // filter: this workgroup's spectrum rows, fetched with the T rows so their
  // latency overlaps the forward FFT ([nslot][P2] + B_P)
  TS* S = reinterpret_cast<TS*>(tw + P2);
  fill_stage_table<T>(tw, l2);
  __syncthreads();
  const int nslot = nslot_sh;
  cx<T>* trow = tbuf + sg.t_off + (int64_t)row * sg.t_stride;
  {
    constexpr int UNR = 8;
    const int total = nslot * P2;
    const TS* sbase = reinterpret_cast<const TS*>(sg.s_ptr) + (int64_t)row * sg.s_stride;
    for (int base = threadIdx.x; base < total; base += kThreads * UNR) {
      cx<T> v[UNR];
      TS b[UNR];
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        int idx = base + u * kThreads;
        if (idx < total) {
          int64_t off = (int64_t)slot_k1[idx >> l2] * P2 + (idx & (P2 - 1));
          v[u] = trow[off];
          if (MODE == kFilter) b[u] = sbase[off];
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        int idx = base + u * kThreads;
        if (idx < total) {
          X[(idx >> l2) * rs + pad8(idx & (P2 - 1))] = v[u];
          if (MODE == kFilter) S[idx] = b[u];
        }
      }
    }
    if (MODE == kFilter && threadIdx.x == 0) S[total] = sbase[P];
  }
  __syncthreads();
  lds_fft<T, false>(X, tw, l2, nslot, rs);
  auto at = [&](int idx) -> cx<T>& { return X[(idx >> l2) * rs + pad8(idx & (P2 - 1))]; };
  if (MODE == kBSpec) {
    cx<T>* srow = reinterpret_cast<cx<T>*>(sg.s_ptr);
    for (int idx = threadIdx.x; idx < nslot * P2; idx += kThreads)
      srow[(int64_t)slot_k1[idx >> l2] * P2 + (idx & (P2 - 1))] = at(idx);
    return;
  }
  if (MODE == kBlue) {
    const cx<T>* srow = reinterpret_cast<const cx<T>*>(sg.s_ptr);
    for (int idx = threadIdx.x; idx < nslot * P2; idx += kThreads)
      at(idx) = cmul(at(idx), srow[(int64_t)slot_k1[idx >> l2] * P2 + (idx & (P2 - 1))]);
    __syncthreads();
  } else {
    TS* srow = reinterpret_cast<TS*>(sg.s_ptr) + (int64_t)row * sg.s_stride;
    const T invP = T(1) / (T)P;
    for (int idx = threadIdx.x; idx < nslot * P2; idx += kThreads) {
      int sl = idx >> l2, k2 = idx & (P2 - 1);
      int kk1 = slot_k1[sl];
      int pidx = (kk1 == 0) ? (sl * P2 + ((P2 - k2) & (P2 - 1))) : (slot_partner[sl] * P2 + (P2 - 1 - k2));
      if (pidx < idx) continue;  // the pair is handled by its lower index
      int64_t k = kk1 + (int64_t)P1 * k2;
      cx<T> C = at(idx), Cq = at(pidx);
      // E_k = (C_k + conj C_(P-k))/2, O_k = (C_k - conj C_(P-k))/(2i); the
      // partner's E, O are conj(E_k), conj(O_k).
      cx<T> E = cscale(cadd(C, cconj(Cq)), T(0.5));
      cx<T> Dd = csub(C, cconj(Cq));
      cx<T> O = {Dd.im * T(0.5), -Dd.re * T(0.5)};
      cx<T> w = pow2 ? expmi2<T>(k, invP) : expmi<T>(k, P);  // w_L^k; w_L^(P-k) = -conj(w)
      cx<T> A = cadd(E, cmul(w, O));
      cx<T> Aq = csub(cconj(E), cmul(cconj(w), cconj(O)));
      int pk1 = slot_k1[pidx >> l2], pk2 = pidx & (P2 - 1);
      if (MODE == kFSpec) {
        srow[(int64_t)kk1 * P2 + k2] = (TS)A.re;
        if (pidx != idx) srow[(int64_t)pk1 * P2 + pk2] = (TS)Aq.re;  // self pair k = 0: Aq is A_P
        if (k == 0) srow[P] = (TS)(E.re - O.re);                       // A_P
        continue;
      }
      T b = (T)S[idx];
      T bq = (T)S[pidx];
      cx<T> Y = cscale(A, b), Yq = cscale(Aq, bq);
      if (k == 0) {
        // self pair: the inverse partner of Y_0 is Y_P (real)
        T yP = (E.re - O.re) * (T)S[nslot * P2];
        cx<T> Ep = {T(0.5) * (Y.re + yP), T(0.5) * Y.im};
        cx<T> Op = {T(0.5) * (Y.re - yP), T(0.5) * Y.im};
        at(idx) = {Ep.re - Op.im, Ep.im + Op.re};
        continue;
      }
      // D_k = E'_k + i O'_k, E'_k = (Y_k + conj Y_(P-k))/2,
      // O'_k = (Y_k - conj Y_(P-k)) conj(w_L^k)/2; same for P-k with w -> -conj(w)
      cx<T> Ep = cscale(cadd(Y, cconj(Yq)), T(0.5));
      cx<T> Op = cmul(cscale(csub(Y, cconj(Yq)), T(0.5)), cconj(w));
      cx<T> Eq = cscale(cadd(Yq, cconj(Y)), T(0.5));
      cx<T> Oq = cmul(cscale(csub(Yq, cconj(Y)), T(0.5)), cx<T>{-w.re, -w.im});
      at(idx) = {Ep.re - Op.im, Ep.im + Op.re};
      if (pidx != idx) at(pidx) = {Eq.re - Oq.im, Eq.im + Oq.re};
    }
    if (MODE == kFSpec) return;
    __syncthreads();
  }
  lds_fft<T, true>(X, tw, l2, nslot, rs);
  const T invP2 = T(1) / (T)P;
  for (int idx = threadIdx.x; idx < nslot * P2; idx += kThreads) {
    int sl = idx >> l2, j2 = idx & (P2 - 1);
    int kk1 = slot_k1[sl];
    int64_t a = ((int64_t)j2 * kk1) % P;
    trow[(int64_t)kk1 * P2 + j2] = cmul(at(idx), cconj(pow2 ? twiddle_P<T>(a, invP2) : twiddle_any<T>(a, P)));
  }
}

// ---- K2, FP32 filter hot path: rows of 2^11 points --------------------------
// Every workgroup holds exactly two rows: k1 and its partner P1-k1, or, for
// the first workgroup of a mode, the self-paired rows k1 = 0 and (even P1)
// P1/2; with odd P1 the second slot of that workgroup is idle.
// Stage and four-step twiddles come from tables built once per plan in FP64
// (no transcendental in the loop); the stage shapes are compile-time, so the
// LDS addresses fold into instruction offsets.  Each thread owns four
// (k, P-k) pairs of the spectral step (five for thread 0 of workgroup 0).
constexpr int kHotL2 = 11;
constexpr int kHotP2 = 1 << kHotL2;
constexpr int kHotRS = row_stride(kHotP2);
constexpr int kHotMaxL1 = 8;                  // P1 <= 256
constexpr int kFineBits = 10;                 // fine table: 1024 entries of w_2P
// X rows + fine/coarse tables; the stage twiddles are read from the global
// table (16 KB, L1-resident), which keeps three workgroups per CU.
constexpr int kHotLds = (2 * kHotRS + (1 << kFineBits) + (4 << kHotMaxL1)) * 8;

// Compile-time Stockham stage over NROWS padded rows (stride RS) of 2^L2LEN
// points; NROWS * 2^L2LEN = kMaxWG.
template <bool INV, int LR, int L2LEN, int LNS, int NROWS, int RS>
__device__ __forceinline__ void ct_stage(cx<float>* buf, const cx<float>* tw) {
  constexpr int R = 1 << LR;
  constexpr int NB = (NROWS << L2LEN) / R / kThreads;
  static_assert(NB >= 1 && NB * R * kThreads == (NROWS << L2LEN), "stage does not tile the workgroup");
  constexpr int LQ = L2LEN - LR, Q = 1 << LQ, NS = 1 << LNS;
  cx<float> v[NB][R];
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int bf = threadIdx.x + u * kThreads;
    const int row = bf >> LQ, j = bf & (Q - 1), k = j & (NS - 1);
    const cx<float>* base = buf + row * RS;
#pragma unroll
    for (int r = 0; r < R; r++) v[u][r] = base[pad8(j + r * Q)];
    if (LNS > 0) {
#pragma unroll
      for (int r = 1; r < R; r++) {
        cx<float> w = tw[(r - 1) * NS + k];
        if (INV) w = cconj(w);
        v[u][r] = cmul(v[u][r], w);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int bf = threadIdx.x + u * kThreads;
    const int row = bf >> LQ, j = bf & (Q - 1), k = j & (NS - 1);
    if (LR == 3) dft8<float, INV>(v[u]);
    else if (LR == 2) dft4v<float, INV>(v[u]);
    else dft2<float, INV>(v[u]);
    cx<float>* base = buf + row * RS;
    const int idx = ((j >> LNS) << (LNS + LR)) + k;
#pragma unroll
    for (int r = 0; r < R; r++) base[pad8(idx + r * NS)] = v[u][r];
  }
  __syncthreads();
}

// FFT of NROWS rows of 2^L2LEN points, stages of radix 8 then 4 or 2; st is
// the stage-ordered table (fill_stage_table order).
template <bool INV, int L2LEN, int NROWS, int RS, int LNS = 0, int OFF = 0>
__device__ __forceinline__ void ct_fft(cx<float>* buf, const cx<float>* st) {
  if constexpr (LNS < L2LEN) {
    constexpr int LR = (L2LEN - LNS) >= 3 ? 3 : (L2LEN - LNS);
    ct_stage<INV, LR, L2LEN, LNS, NROWS, RS>(buf, st + OFF);
    ct_fft<INV, L2LEN, NROWS, RS, LNS + LR, OFF + ((1 << LR) - 1) * (1 << LNS)>(buf, st);
  }
}

__device__ __forceinline__ cx<float> tw2P(const cx<float>* F, const cx<float>* C, int k) {
  return cmul(F[k & ((1 << kFineBits) - 1)], C[k >> kFineBits]);  // w_2P^k
}

__global__ void __launch_bounds__(kThreads, 4) k2_hot_kernel(const Seg* __restrict__ segs, int nseg,
                                                              cx<float>* __restrict__ tbuf,
                                                              const cx<float>* __restrict__ stab_g) {
  extern __shared__ unsigned char smem_raw[];
  const int s = find_seg(segs, nseg, blockIdx.x, 2);
  const Seg sg = segs[s];
  const int local = (int)((int64_t)blockIdx.x - sg.base2);
  const int row = local / sg.per2, g = local - row * sg.per2;
  const int P1 = sg.p1;
  const int P = P1 << kHotL2;
  const bool single = g == 0 && (P1 & 1);  // odd P1: column 0 is the only self-paired one
  const int kA = g == 0 ? 0 : g, kB = g == 0 ? (single ? 0 : P1 / 2) : P1 - g;
  cx<float>* X = reinterpret_cast<cx<float>*>(smem_raw);  // [2][kHotRS]
  cx<float>* F = X + 2 * kHotRS;                           // [1024] w_2P^l
  cx<float>* Ct = F + (1 << kFineBits);                    // [4 P1] w_2P^(1024 h)
  const cx<float>* st = stab_g;                            // stage table (global)
  const cx<float>* twF = reinterpret_cast<const cx<float>*>(sg.tw_ptr);
  const int nC = 4 * P1;
  for (int e = threadIdx.x; e < (1 << kFineBits); e += kThreads) F[e] = twF[e];
  for (int e = threadIdx.x; e < nC; e += kThreads) Ct[e] = twF[(1 << kFineBits) + e];
  cx<float>* trow = tbuf + sg.t_off + (int64_t)row * sg.t_stride;
  const float* srow = reinterpret_cast<const float*>(sg.s_ptr) + (int64_t)row * sg.s_stride;
  {
    cx<float> va[4], vb[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int k2 = threadIdx.x + u * kThreads;
      va[u] = trow[kA * kHotP2 + k2];
      vb[u] = trow[kB * kHotP2 + k2];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int k2 = threadIdx.x + u * kThreads;
      X[pad8(k2)] = va[u];
      X[kHotRS + pad8(k2)] = vb[u];
    }
  }
  // this thread's pairs: (slot, k2) <-> (slot_q, k2q)
  constexpr int NPAIR = 5;
  int p_sl[NPAIR], p_k2[NPAIR], p_slq[NPAIR], p_k2q[NPAIR];
  float bb[NPAIR], bq[NPAIR];
  float bP = 0.f;
#pragma unroll
  for (int u = 0; u < NPAIR; u++) {
    const int pl = threadIdx.x + u * kThreads;
    int sl = -1, k2 = 0, slq = 0, k2q = 0;
    if (g != 0) {
      if (pl < kHotP2) { sl = 0; k2 = pl; slq = 1; k2q = kHotP2 - 1 - pl; }
    } else if (pl <= kHotP2 / 2) {
      sl = 0; k2 = pl; slq = 0; k2q = (kHotP2 - pl) & (kHotP2 - 1);
    } else if (!single && pl < kHotP2 + 1) {
      sl = 1; k2 = pl - (kHotP2 / 2 + 1); slq = 1; k2q = kHotP2 - 1 - k2;
    }
    p_sl[u] = sl; p_k2[u] = k2; p_slq[u] = slq; p_k2q[u] = k2q;
    bb[u] = bq[u] = 0.f;
    if (sl >= 0) {
      bb[u] = srow[(sl ? kB : kA) * kHotP2 + k2];
      bq[u] = srow[(slq ? kB : kA) * kHotP2 + k2q];
    }
  }
  if (g == 0 && threadIdx.x == 0) bP = srow[P];
  __syncthreads();
  // twiddles factor along k2 = tid + 512 u:  w_2P^(k1 + P1 k2) =
  // w_2P^k1 * w_4096^tid * w_8^u, and w_P^(k1 j2) = w_P^(k1 tid) * (w_P^(512 k1))^u
  cx<float> w4096t;
  sincospif(-(float)threadIdx.x * (1.f / 2048.f), &w4096t.im, &w4096t.re);
  const cx<float> wkA = tw2P(F, Ct, kA), wkB = tw2P(F, Ct, kB);
  ct_fft<false, kHotL2, 2, kHotRS>(X, st);
#pragma unroll
  for (int u = 0; u < NPAIR; u++) {
    const int sl = p_sl[u];
    if (sl < 0) continue;
    const int k2 = p_k2[u], slq = p_slq[u], k2q = p_k2q[u];
    const int kk1 = sl ? kB : kA;
    const int idx = sl * kHotRS + pad8(k2), pidx = slq * kHotRS + pad8(k2q);
    const int k = kk1 + P1 * k2;
    cx<float> C = X[idx], Cq = X[pidx];
    cx<float> E = cscale(cadd(C, cconj(Cq)), 0.5f);
    cx<float> Dd = csub(C, cconj(Cq));
    cx<float> O = {Dd.im * 0.5f, -Dd.re * 0.5f};
    cx<float> w;
    if (k2 == (int)threadIdx.x + u * kThreads) {  // every pair but those of the row k1 = P1/2 of workgroup 0
      constexpr float c8 = 0.70710678118654752440f;
      const cx<float> w8[4] = {{1.f, 0.f}, {c8, -c8}, {0.f, -1.f}, {-c8, -c8}};
      w = cmul(cmul(sl ? wkB : wkA, w4096t), w8[u & 3]);
    } else {
      w = tw2P(F, Ct, k);
    }
    cx<float> A = cadd(E, cmul(w, O));
    cx<float> Aq = csub(cconj(E), cmul(cconj(w), cconj(O)));
    cx<float> Y = cscale(A, bb[u]), Yq = cscale(Aq, bq[u]);
    if (k == 0) {
      const float yP = (E.re - O.re) * bP;
      cx<float> Ep = {0.5f * (Y.re + yP), 0.5f * Y.im};
      cx<float> Op = {0.5f * (Y.re - yP), 0.5f * Y.im};
      X[idx] = {Ep.re - Op.im, Ep.im + Op.re};
      continue;
    }
    cx<float> Ep = cscale(cadd(Y, cconj(Yq)), 0.5f);
    cx<float> Op = cmul(cscale(csub(Y, cconj(Yq)), 0.5f), cconj(w));
    cx<float> Eq = cscale(cadd(Yq, cconj(Y)), 0.5f);
    cx<float> Oq = cmul(cscale(csub(Yq, cconj(Y)), 0.5f), cx<float>{-w.re, -w.im});
    X[idx] = {Ep.re - Op.im, Ep.im + Op.re};
    if (pidx != idx) X[pidx] = {Eq.re - Oq.im, Eq.im + Oq.re};
  }
  __syncthreads();
  ct_fft<true, kHotL2, 2, kHotRS>(X, st);
  // kA, kB < 256 and tid < 512, so the products fit in 32 bits
  const uint32_t uP = (uint32_t)P;
  cx<float> tA = tw2P(F, Ct, 2 * (int)(((uint32_t)kA * threadIdx.x) % uP));
  cx<float> tB = tw2P(F, Ct, 2 * (int)(((uint32_t)kB * threadIdx.x) % uP));
  const cx<float> sA = tw2P(F, Ct, 2 * (int)(((uint32_t)kA * kThreads) % uP));
  const cx<float> sB = tw2P(F, Ct, 2 * (int)(((uint32_t)kB * kThreads) % uP));
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int j2 = threadIdx.x + u * kThreads;
    trow[kA * kHotP2 + j2] = cmul(X[pad8(j2)], cconj(tA));
    if (!single) trow[kB * kHotP2 + j2] = cmul(X[kHotRS + pad8(j2)], cconj(tB));
    tA = cmul(tA, sA);
    tB = cmul(tB, sB);
  }
}

// ---- K1 / K3, FP32 filter hot path: columns of P1 = 2^a 3^b 5^c (rows of 2^11)
// One launch per column length; CW adjacent columns per workgroup (the
// largest power of two with CW P1 <= 4096, but at least 32), so global reads
// and writes are runs of CW points.  The LDS tile is column-major: point j
// of column c at j CW + c.  A stage's butterflies are numbered column-fastest,
// so the lanes of a half-wave touch 32 consecutive columns of one row (no
// bank conflicts, one shared twiddle), and every index is a shift or mask of
// the butterfly number.
constexpr int kHotMaxPts = 8192;  // points per K1/K3 workgroup (16 per thread)
constexpr int hot_cw(int p1) {
  int c = 1;
  while (2 * c * p1 <= kMaxWG && 2 * c <= kHotP2) c *= 2;
  return c < 32 ? 32 : c;
}
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }
// Column tile of the general kernels (the spectrum pass runs on the plan's
// geometry): the largest power of two with cw P1 <= 4096.
constexpr int generic_cw(int p1) {
  int c = 1;
  while (2 * c * p1 <= kMaxWG && 2 * c <= kHotP2) c *= 2;
  return c;
}

// One radix-R stage of N-point column FFTs over the CW columns of a
// column-major tile; the tail of the last butterfly round is masked when
// CW N / R is not a multiple of the workgroup size.
template <bool INV, int R, int N, int NS, int CW>
__device__ __forceinline__ void ct_stage_cm(cx<float>* buf, const cx<float>* tw) {
  constexpr int Q = N / R, TOT = CW * Q, NB = (TOT + kThreads - 1) / kThreads, LCW = ilog2c(CW);
  constexpr bool FULL = NB * kThreads == TOT;
  cx<float> v[NB][R];
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int bf = threadIdx.x + u * kThreads;
    if (FULL || bf < TOT) {
      const int c = bf & (CW - 1), j = bf >> LCW, k = j % NS;
#pragma unroll
      for (int r = 0; r < R; r++) v[u][r] = buf[(j + r * Q) * CW + c];
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
          cx<float> w = tw[(r - 1) * NS + k];
          if (INV) w = cconj(w);
          v[u][r] = cmul(v[u][r], w);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int bf = threadIdx.x + u * kThreads;
    if (FULL || bf < TOT) {
      const int c = bf & (CW - 1), j = bf >> LCW, k = j % NS;
      dft_r<float, INV, R>(v[u]);
      const int idx = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) buf[(idx + r * NS) * CW + c] = v[u][r];
    }
  }
  __syncthreads();
}

template <bool INV, int N, int CW, int NS = 1, int OFF = 0>
__device__ __forceinline__ void ct_fft_cm(cx<float>* buf, const cx<float>* st) {
  if constexpr (NS < N) {
    constexpr int R = next_radix(N / NS);
    ct_stage_cm<INV, R, N, NS, CW>(buf, st + OFF);
    ct_fft_cm<INV, N, CW, NS * R, OFF + (R - 1) * NS>(buf, st);
  }
}

// Blocks are taken in an XCD-aware order: the hardware deals consecutive
// block ids round-robin over the 8 XCDs, and the remap gives each XCD a
// contiguous range of (row, column block) tiles, so the pieces of one row
// run together on one XCD.  A row's input is then read from HBM once and
// its mirrored images (each sample appears up to three times in the even
// extension) come from that XCD's L2: K1's fetch drops from 28.4 to 7.8 GB
// per apply on the 64-TOD benchmark.
template <int P1>
__global__ void __launch_bounds__(kThreads, 2) k1_hot_kernel(const Seg* __restrict__ segs, int nseg,
                                                              const float* __restrict__ in,
                                                              cx<float>* __restrict__ tbuf,
                                                              const cx<float>* __restrict__ stab_g,
                                                              int64_t nblk) {
  constexpr int CW = hot_cw(P1), LCW = ilog2c(CW), NPTS = CW * P1;
  constexpr int P = P1 * kHotP2;
  static_assert(NPTS <= kHotMaxPts && kHotP2 % CW == 0, "K1 tile");
  extern __shared__ unsigned char smem_raw[];
  const int64_t per = ((int64_t)gridDim.x + 7) / 8;
  const int64_t bid = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (bid >= nblk) return;
  const int s = find_seg(segs, nseg, bid, 1);
  const Seg sg = segs[s];
  const int local = (int)(bid - sg.base1);
  const int row = local / sg.per1, jb = local - row * sg.per1;
  const int m = sg.n - 1;
  cx<float>* buf = reinterpret_cast<cx<float>*>(smem_raw);  // [P1][CW]
  cx<float>* st = buf + NPTS;                                // [P1] stage table
  cx<float>* F = st + P1;                                    // [1024] w_2P^l
  cx<float>* Ct = F + (1 << kFineBits);                      // [4 P1] w_2P^(1024 h)
  const cx<float>* twF = reinterpret_cast<const cx<float>*>(sg.tw_ptr);
  for (int e = threadIdx.x; e < P1 - 1; e += kThreads) st[e] = stab_g[e];
  for (int e = threadIdx.x; e < (1 << kFineBits); e += kThreads) F[e] = twF[e];
  for (int e = threadIdx.x; e < 4 * P1; e += kThreads) Ct[e] = twF[(1 << kFineBits) + e];
  const float* x = in + sg.in_off + (int64_t)row * sg.in_stride;
  const int j2_0 = jb * CW;
  // real input a_i, i = 2 (P2 j1 + j2) + w, from the even extension on
  // [-m, 2m]; reals 2 (j1 CW + c) + {0, 1} of the tile are consecutive.
  // A thread's w is the same for every u and j1 grows by kThreads / 2CW, so
  // i is affine in u.  The extension's sample is x[m - |m - |i - m||], and
  // past 3m that index is negative: as a 32-bit buffer offset it lies past
  // the row's bytes, and the buffer load returns the padding's zero.
  // (Wide tiles of short columns, 2CW > kThreads, compute i per u.)
  constexpr int NR = 2 * NPTS, NV = (NR + kThreads - 1) / kThreads;
  const __amdgpu_buffer_rsrc_t xrs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(x), (short)0, (int)((m + 1) * 4), 0x00020000);
  float val[NV];
  if constexpr (kThreads % (2 * CW) == 0) {
    constexpr int DI = 2 * kHotP2 * (kThreads / (2 * CW));  // i step per u
    const int t0 = 2 * (kHotP2 * ((int)threadIdx.x >> (LCW + 1)) + j2_0) + ((int)threadIdx.x & (2 * CW - 1)) - m;
#pragma unroll
    for (int u = 0; u < NV; u++) {
      const int t = t0 + u * DI;
      const int xi = m - abs(m - abs(t));
      const float v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(xrs, (uint32_t)xi * 4u, 0u, 0));
      val[u] = (NV * kThreads == NR || (int)threadIdx.x + u * kThreads < NR) ? v : 0.f;
    }
  } else {
#pragma unroll
    for (int u = 0; u < NV; u++) {
      const int idx = threadIdx.x + u * kThreads;
      const int t = 2 * (kHotP2 * (idx >> (LCW + 1)) + j2_0) + (idx & (2 * CW - 1)) - m;
      const int xi = m - abs(m - abs(t));
      const float v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(xrs, (uint32_t)xi * 4u, 0u, 0));
      val[u] = (NV * kThreads == NR || idx < NR) ? v : 0.f;
    }
  }
#pragma unroll
  for (int u = 0; u < NV; u++) {
    const int idx = threadIdx.x + u * kThreads;
    if (NV * kThreads == NR || idx < NR) reinterpret_cast<float*>(buf)[idx] = val[u];
  }
  __syncthreads();
  ct_fft_cm<false, P1, CW>(buf, st);
  cx<float>* trow = tbuf + sg.t_off + (int64_t)row * sg.t_stride;
  constexpr int NO = (NPTS + kThreads - 1) / kThreads;
  // The thread's column c is the same for every u (kThreads is a multiple of
  // CW) and k1 grows by kThreads / CW per step, so the twiddles w^(2 (j2_0+c) k1)
  // form a geometric sequence: one complex multiply per step, re-seeded from
  // the exact tables every 4 steps so FP32 rounding does not build up.
  if constexpr (kThreads % CW == 0) {
    const int c = threadIdx.x & (CW - 1), j2 = j2_0 + c;
    const cx<float> step = tw2P(F, Ct, 2 * ((j2 * (kThreads / CW)) % P));
    cx<float> tw{1.f, 0.f};
#pragma unroll
    for (int u = 0; u < NO; u++) {
      const int idx = threadIdx.x + u * kThreads;
      if (NO * kThreads != NPTS && idx >= NPTS) break;
      const int k1 = idx >> LCW;
      if ((u & 3) == 0) tw = tw2P(F, Ct, 2 * ((j2 * k1) % P));  // < 2^11 * 2^8
      trow[k1 * kHotP2 + j2] = cmul(buf[idx], tw);
      tw = cmul(tw, step);
    }
  } else {  // CW > kThreads (short columns): the column changes with u
#pragma unroll
    for (int u = 0; u < NO; u++) {
      const int idx = threadIdx.x + u * kThreads;
      if (NO * kThreads != NPTS && idx >= NPTS) break;
      const int k1 = idx >> LCW, c = idx & (CW - 1);
      const int a = ((j2_0 + c) * k1) % P;  // < 2^11 * 2^8
      trow[k1 * kHotP2 + j2_0 + c] = cmul(buf[idx], tw2P(F, Ct, 2 * a));
    }
  }
}

template <int P1>
__global__ void __launch_bounds__(kThreads, 2) k3_hot_kernel(const Seg* __restrict__ segs, int nseg,
                                                              const cx<float>* __restrict__ tbuf,
                                                              float* __restrict__ out,
                                                              const cx<float>* __restrict__ stab_g) {
  constexpr int CW = hot_cw(P1), LCW = ilog2c(CW), NPTS = CW * P1;
  constexpr int P = P1 * kHotP2;
  static_assert(NPTS <= kHotMaxPts && kHotP2 % CW == 0, "K3 tile");
  extern __shared__ unsigned char smem_raw[];
  const int s = find_seg(segs, nseg, blockIdx.x, 3);
  const Seg sg = segs[s];
  const int local = (int)((int64_t)blockIdx.x - sg.base3);
  const int row = local / sg.per1, jb = local - row * sg.per1;
  const int m = sg.n - 1;
  cx<float>* buf = reinterpret_cast<cx<float>*>(smem_raw);  // [P1][CW]
  cx<float>* st = buf + NPTS;                                // [P1] stage table
  for (int e = threadIdx.x; e < P1 - 1; e += kThreads) st[e] = stab_g[e];
  const cx<float>* trow = tbuf + sg.t_off + (int64_t)row * sg.t_stride;
  const int j2_0 = jb * CW;
  {
    constexpr int NV = (NPTS + kThreads - 1) / kThreads;
    cx<float> v[NV];
#pragma unroll
    for (int u = 0; u < NV; u++) {
      const int idx = threadIdx.x + u * kThreads;
      if (NV * kThreads == NPTS || idx < NPTS) v[u] = trow[(idx >> LCW) * kHotP2 + j2_0 + (idx & (CW - 1))];
    }
#pragma unroll
    for (int u = 0; u < NV; u++) {
      const int idx = threadIdx.x + u * kThreads;
      if (NV * kThreads == NPTS || idx < NPTS) buf[idx] = v[u];
    }
  }
  __syncthreads();
  ct_fft_cm<true, P1, CW>(buf, st);
  const float scale = 1.f / (float)P;
  const bool halve = sg.flags & 1;
  float* y = out + sg.out_off + (int64_t)row * sg.out_stride;
  constexpr int NR = 2 * NPTS, NO = (NR + kThreads - 1) / kThreads;
#pragma unroll
  for (int u = 0; u < NO; u++) {
    const int idx = threadIdx.x + u * kThreads;
    if (NO * kThreads != NR && idx >= NR) break;
    const int j1 = idx >> (LCW + 1), w = idx & (2 * CW - 1);
    const int k = 2 * (kHotP2 * j1 + j2_0) + w - m;
    if (k >= 0 && k <= m) {
      float v = reinterpret_cast<const float*>(buf)[idx] * scale;
      if (halve && (k == 0 || k == m)) v *= 0.5f;
      y[k] = v;
    }
  }
}

// Column lengths of the hot path: every 5-smooth P1 in [2, 256].
constexpr int kHotP1[] = {2,   3,   4,   5,   6,   8,   9,   10,  12,  15,  16,  18,  20,  24,  25,  27,  30,
                          32,  36,  40,  45,  48,  50,  54,  60,  64,  72,  75,  80,  81,  90,  96,  100, 108,
                          120, 125, 128, 135, 144, 150, 160, 162, 180, 192, 200, 216, 225, 240, 243, 250, 256};
constexpr int kNumHotP1 = sizeof(kHotP1) / sizeof(kHotP1[0]);
int hot_p1_index(int p1) {
  for (int i = 0; i < kNumHotP1; i++)
    if (kHotP1[i] == p1) return i;
  return -1;
}

// LDS of the hot K1 (tile, stage table, fine/coarse tables) and K3 (tile,
// stage table).
size_t hot13_lds(int p1, int which) {
  return (size_t)(hot_cw(p1) * p1 + p1 + (which == 1 ? (1 << kFineBits) + 4 * p1 : 0)) * sizeof(cx<float>);
}

// Stage table of an N-point mixed-radix FFT (fill_stage_table_mx order),
// computed in FP64.
__global__ void mixed_stage_table_kernel(cx<float>* out, int N) {
  int Ns = 1, off = 0;
  while (Ns < N) {
    const int R = next_radix(N / Ns), cnt = (R - 1) * Ns;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += gridDim.x * blockDim.x) {
      const int r = e / Ns + 1, k = e - (r - 1) * Ns;
      double sn, cs;
      sincospi(-2.0 * (double)(k * r) / (double)(R * Ns), &sn, &cs);
      out[off + e] = {(float)cs, (float)sn};
    }
    off += cnt;
    Ns *= R;
  }
}

// Twiddle tables of the hot K2 path, computed in FP64: the stage table of a
// 2^11-point FFT (fill_stage_table order) and, per P, w_2P^l (l < 1024) then
// w_2P^(1024 h) (h < 2P / 1024).
__global__ void hot_tables_kernel(cx<float>* out, int64_t P, int stage_l2) {
  if (stage_l2 > 0) {  // stage table of a 2^stage_l2-point FFT
    int Ns = 1, rem = stage_l2, off = 0;
    while (rem > 0) {
      const int lr = stage_lr(rem), R = 1 << lr, cnt = (R - 1) * Ns, mul = 1 << (rem - lr);
      for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += gridDim.x * blockDim.x) {
        int r = e / Ns + 1, k = e - (r - 1) * Ns;
        double sn, cs;
        sincospi(-2.0 * (double)((int64_t)k * r * mul) / (double)(1 << stage_l2), &sn, &cs);
        out[off + e] = {(float)cs, (float)sn};
      }
      off += cnt;
      Ns <<= lr; rem -= lr;
    }
    return;
  }
  const int64_t nF = 1 << kFineBits, nC = 2 * P / nF;
  for (int64_t e = blockIdx.x * blockDim.x + threadIdx.x; e < nF + nC; e += gridDim.x * blockDim.x) {
    int64_t t = e < nF ? e : (e - nF) * nF;  // exponent of w_2P
    double sn, cs;
    sincospi(-(double)t / (double)P, &sn, &cs);
    out[e] = {(float)cs, (float)sn};
  }
}

// ---- K3: inverse column FFTs + output -------------------------------------
template <typename T, typename TOut, int MODE>
__global__ void __launch_bounds__(kThreads, sizeof(T) == 4 ? 4 : 1) k3_kernel(const Seg* __restrict__ segs, int nseg,
                                                      const cx<T>* __restrict__ tbuf, TOut* __restrict__ out) {
  extern __shared__ unsigned char smem_raw[];
  const int s = find_seg(segs, nseg, blockIdx.x, 3);
  const Seg sg = segs[s];
  const int64_t local = (int64_t)blockIdx.x - sg.base3;
  const int row = (int)(local / sg.per1), jb = (int)(local % sg.per1);
  // power-of-two column lengths only (the plan routes mixed lengths to the
  // hot path)
  const int l1 = sg.l1, l2 = sg.l2, P1 = 1 << l1, P2 = 1 << l2, cw = sg.cw;
  const int lcw = 31 - __clz(cw);
  const int rs = row_stride(P1);
  const int64_t P = (int64_t)P1 * P2;
  const int64_t m = sg.n - 1, N = 2 * m;
  cx<T>* buf = reinterpret_cast<cx<T>*>(smem_raw);  // [cw][rs]
  cx<T>* tw = buf + cw * rs;                          // [P1] stage table
  const cx<T>* trow = tbuf + sg.t_off + (int64_t)row * sg.t_stride;
  const int j2_0 = jb * cw;
  fill_stage_table<T>(tw, l1);
  {
    constexpr int UNR = 8;
    const int total = P1 * cw;
    for (int base = threadIdx.x; base < total; base += kThreads * UNR) {
      cx<T> v[UNR];
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        int idx = base + u * kThreads;
        if (idx < total) {
          int k1 = idx >> lcw, c = idx & (cw - 1);
          v[u] = trow[(int64_t)k1 * P2 + j2_0 + c];
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        int idx = base + u * kThreads;
        if (idx < total) {
          int k1 = idx >> lcw, c = idx & (cw - 1);
          buf[c * rs + pad8(k1)] = v[u];
        }
      }
    }
  }
  __syncthreads();
  lds_fft<T, true>(buf, tw, l1, cw, rs);
  const T scale = T(1) / (T)P;
  TOut* y = out + sg.out_off + (int64_t)row * sg.out_stride;
  if (MODE == kFilter) {
    const bool halve = sg.flags & 1;
    for (int idx = threadIdx.x; idx < P1 * cw * 2; idx += kThreads) {
      int j1 = idx >> (lcw + 1), w = idx & (2 * cw - 1);
      int64_t i = 2 * ((int64_t)P2 * j1 + j2_0) + w;
      int64_t k = i - m;
      if (k >= 0 && k <= m) {
        const T* src = reinterpret_cast<const T*>(buf + (w >> 1) * rs + pad8(j1));
        T v = src[w & 1] * scale;
        if (halve && (k == 0 || k == m)) v *= T(0.5);
        y[k] = (TOut)v;
      }
    }
  } else {  // kBlue
    for (int idx = threadIdx.x; idx < P1 * cw; idx += kThreads) {
      int j1 = idx >> lcw, c = idx & (cw - 1);
      int64_t k = (int64_t)P2 * j1 + j2_0 + c;
      if (k <= m) {
        cx<T> w = cscale(buf[c * rs + pad8(j1)], scale);
        cx<T> ch = expmi<T>((k * k) % (2 * N), N);
        y[k] = (TOut)(ch.re * w.re - ch.im * w.im);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host-side planning

int log2_ceil(int64_t x) {
  int q = 0;
  while ((int64_t(1) << q) < x) q++;
  return q;
}

struct Geometry {
  int l1, l2, cw;  // l1 = log2(p1), or -1 when p1 is not a power of two
  int p1;
  int64_t P;
};

// P = 2^q complex points split as P1 * P2 with the row length P2 as long as
// K2 holds (2^11 points, a real-mode row and its partner fill one
// workgroup), so the columns are short and K1/K3 can take many adjacent
// columns at once: their global accesses are runs of cw complex points.
// MK_FFT_L2MAX (power-of-two exponent, 2..11) lowers the row-length cap.
//
// mixed: the FP32 hot path takes any 5-smooth P1 in [2, 256] with P2 = 2^11,
// so P = P1 2^11 is the smallest such multiple of 2^11 >= P_min (instead of
// the next power of two): about 1.4x fewer points on average.
Geometry hot_geometry(int p1) {
  Geometry g;
  g.p1 = p1;
  g.l1 = (p1 & (p1 - 1)) == 0 ? ilog2c(p1) : -1;
  g.l2 = kHotL2;
  g.P = (int64_t)p1 * kHotP2;
  g.cw = generic_cw(p1);
  return g;
}

Geometry geometry(int64_t P_min, int elem_bytes, bool mixed = false) {
  if (mixed && P_min > kHotP2) {
    const int64_t need = (P_min + kHotP2 - 1) / kHotP2;
    for (int i = 0; i < kNumHotP1; i++)
      if (kHotP1[i] >= need) return hot_geometry(kHotP1[i]);
  }
  int q = std::max(2, log2_ceil(P_min));
  static int l2max = [] {
    const char* e = getenv("MK_FFT_L2MAX");
    int v = e ? atoi(e) : 11;
    return std::min(11, std::max(2, v));
  }();
  Geometry g;
  g.l2 = std::min(l2max, q - 1);
  g.l1 = q - g.l2;
  g.p1 = 1 << g.l1;
  g.P = int64_t(1) << q;
  int P1 = g.p1, P2 = 1 << g.l2;
  int cw = std::min(std::max(1, kMaxWG / P1), P2);
  // keep the K1/K3 tile plus the twiddle table within the LDS budget
  while (cw > 1 && ((int64_t)cw * row_stride(P1) + P1) * 2 * elem_bytes > kMaxLds) cw >>= 1;
  g.cw = cw;
  return g;
}

template <typename K>
void allow_big_lds(K kernel) {
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            kMaxLds);
}

template <int I = 0>
void allow_lds_hot13() {
  if constexpr (I < kNumHotP1) {
    allow_big_lds(k1_hot_kernel<kHotP1[I]>);
    allow_big_lds(k3_hot_kernel<kHotP1[I]>);
    allow_lds_hot13<I + 1>();
  }
}

void allow_lds_all() {
  static bool done = false;
  if (done) return;
  done = true;
  allow_big_lds(k1_kernel<float, float, kFilter>);
  allow_big_lds(k1_kernel<double, double, kFilter>);
  allow_big_lds(k1_kernel<float, double, kFSpec>);
  allow_big_lds(k1_kernel<double, double, kFSpec>);
  allow_big_lds(k1_kernel<float, float, kBlue>);
  allow_big_lds(k1_kernel<double, double, kBlue>);
  allow_big_lds(k1_kernel<float, float, kBSpec>);
  allow_big_lds(k1_kernel<double, double, kBSpec>);
  allow_big_lds(k2_kernel<float, float, kFilter>);
  allow_big_lds(k2_kernel<double, double, kFilter>);
  allow_big_lds(k2_kernel<float, float, kFSpec>);
  allow_big_lds(k2_kernel<double, double, kFSpec>);
  allow_big_lds(k2_kernel<double, float, kFSpec>);
  allow_big_lds(k2_kernel<float, float, kBlue>);
  allow_big_lds(k2_kernel<double, double, kBlue>);
  allow_big_lds(k2_kernel<float, float, kBSpec>);
  allow_big_lds(k2_kernel<double, double, kBSpec>);
  allow_big_lds(k3_kernel<float, float, kFilter>);
  allow_big_lds(k3_kernel<double, double, kFilter>);
  allow_big_lds(k3_kernel<float, float, kBlue>);
  allow_big_lds(k3_kernel<double, double, kBlue>);
  allow_big_lds(k2_hot_kernel);
  allow_lds_hot13();
  (void)hipGetLastError();  // an unsupported attribute must not poison later launch checks
}

size_t lds_k13(const Geometry& g, int eb) {
  return (size_t)(g.cw * row_stride(g.p1) + g.p1) * 2 * eb;
}
// k1 values per K2 workgroup: slots (2 per k1 for real modes) * P2 <= kMaxWG
// LDS of K2: slot rows, the stage table and (real modes) the slots' spectrum
// values, one real each plus B_P.
size_t k2_bytes(int r, int P2, int eb, bool real) {
  int ns = real ? 2 * r : r;
  return (size_t)(ns * row_stride(P2) + P2) * 2 * eb + (real ? (size_t)(ns * P2 + 1) * eb : 0) + 64;
}
// k1 values per K2 workgroup: slots (2 per k1 for real modes) * P2 <= kMaxWG
int k2_r2(const Geometry& g, int eb, bool real) {
  int P2 = 1 << g.l2, P1 = g.p1;
  int r = 8;
  while (r > 1 && ((real ? 2 : 1) * r * P2 > kMaxWG || k2_bytes(r, P2, eb, real) > (size_t)kMaxLds)) r >>= 1;
  return std::min(r, real ? P1 / 2 + 1 : P1);
}
size_t lds_k2(const Geometry& g, int eb, bool real = true) {
  return k2_bytes(k2_r2(g, eb, real), 1 << g.l2, eb, real);
}

}  // namespace

// ---------------------------------------------------------------------------
// Filter plan: per segment (TOD) rows x n, input/output row offsets.

struct mk_filt_plan {
  int nseg = 0;
  int elem_bytes = 4;  // 4: FP32 spectra and T buffer, 8: FP64
  std::vector<Seg> h_segs;
  Seg* d_segs = nullptr;
  int64_t n1 = 0, n2 = 0;  // total K1/K3 items, K2 items
  int64_t t_elems = 0, s_elems = 0;
  void* tbuf = nullptr;    // complex
  bool tbuf_owned = true;  // false: the caller's buffer (mk_filt_plan_set_work)
  void* spec = nullptr;    // real, P+1 per row
  size_t lds1 = 0, lds2 = 0;
  bool have_spec = false;
  std::vector<char> borrowed;
  // FP32 K2 split: segments on the hot path (rows of 2^11, P1 <= 256) and the
  // rest, each list with its own K2 numbering; twiddle tables for the hot path
  Seg* d_hot = nullptr;
  Seg* d_cold = nullptr;
  Seg* d_hot13 = nullptr;   // hot segments grouped by column length, K1/K3 numbering per group
  int nhot = 0, ncold = 0;
  int64_t n2hot = 0, n2cold = 0, n1cold = 0;
  int g_off[kNumHotP1] = {0}, g_cnt[kNumHotP1] = {0};  // per column length (kHotP1 index)
  int64_t g_n1[kNumHotP1] = {0};
  int64_t g_stab[kNumHotP1] = {0};  // offset of that length's stage table in tw_tab
  // stage tables of 2^1..2^11 points (2^l - 2 offsets), then per-P fine/coarse
  // tables, then the mixed-radix stage tables of the hot column lengths
  void* tw_tab = nullptr;
};

namespace {
// The work buffer (rows * P complex) is allocated on first use and can be
// released between uses; spectra stay resident.
int ensure_tbuf(const mk_filt_plan* cp) {
  mk_filt_plan* p = const_cast<mk_filt_plan*>(cp);
  if (p->tbuf) return MK_OK;
  p->tbuf_owned = true;
  if (hipMalloc(&p->tbuf, std::max<int64_t>(1, p->t_elems) * 2 * p->elem_bytes) != hipSuccess) {
    p->tbuf = nullptr;
    mk::set_error("filter plan: work buffer allocation (%lld B) failed", (long long)(p->t_elems * 2 * p->elem_bytes));
    return MK_ERR_ALLOC;
  }
  return MK_OK;
}

bool hot_shape(const Seg& s) { return s.l2 == kHotL2 && hot_p1_index(s.p1) >= 0; }
bool is_hot(const mk_filt_plan* p, const Seg& s) {
  return p->elem_bytes == 4 && hot_shape(s) && s.tw_ptr != nullptr;
}

int upload(mk_filt_plan* p) {
  if (p->nseg && hipMemcpy(p->d_segs, p->h_segs.data(), p->nseg * sizeof(Seg), hipMemcpyHostToDevice) != hipSuccess) {
    mk::set_error("filter plan: descriptor upload failed");
    return MK_ERR_HIP;
  }
  if (p->elem_bytes != 4) return MK_OK;
  std::vector<Seg> hot, cold, hot13;
  int64_t b_hot = 0, b_cold2 = 0, b_cold1 = 0;
  for (const Seg& s0 : p->h_segs) {
    Seg s = s0;
    if (is_hot(p, s)) {
      s.per2 = (s.p1 + 1) / 2;  // pairs (k1, P1-k1), k1 = 1..; column 0 (+ P1/2) in item 0
      s.base2 = b_hot;
      b_hot += (int64_t)s.rows * s.per2;
      hot.push_back(s);
    } else {
      if (s.l1 < 0) {
        mk::set_error("filter plan: column length %d is only supported on the FP32 hot path", s.p1);
        return MK_ERR_ARG;
      }
      s.base2 = b_cold2;
      b_cold2 += (int64_t)s.rows * s.per2;
      s.base1 = s.base3 = b_cold1;
      b_cold1 += (int64_t)s.rows * s.per1;
      cold.push_back(s);
    }
  }
  for (int l = 0; l < kNumHotP1; l++) {
    p->g_off[l] = (int)hot13.size();
    int64_t b = 0;
    for (const Seg& s0 : hot)
      if (s0.p1 == kHotP1[l]) {
        Seg s = s0;
        s.per1 = kHotP2 / hot_cw(s.p1);  // the hot K1/K3 tile is wider than the general one
        s.base1 = s.base3 = b;
        b += (int64_t)s.rows * s.per1;
        hot13.push_back(s);
      }
    p->g_cnt[l] = (int)hot13.size() - p->g_off[l];
    p->g_n1[l] = b;
  }
  p->nhot = (int)hot.size();
  p->ncold = (int)cold.size();
  p->n2hot = b_hot;
  p->n2cold = b_cold2;
  p->n1cold = b_cold1;
  if ((p->nhot && hipMemcpy(p->d_hot, hot.data(), hot.size() * sizeof(Seg), hipMemcpyHostToDevice) != hipSuccess) ||
      (p->nhot &&
       hipMemcpy(p->d_hot13, hot13.data(), hot13.size() * sizeof(Seg), hipMemcpyHostToDevice) != hipSuccess) ||
      (p->ncold &&
       hipMemcpy(p->d_cold, cold.data(), cold.size() * sizeof(Seg), hipMemcpyHostToDevice) != hipSuccess)) {
    mk::set_error("filter plan: descriptor upload failed");
    return MK_ERR_HIP;
  }
  return MK_OK;
}

const cx<float>* stage_tab(const mk_filt_plan* p, int l) { return (const cx<float>*)p->tw_tab + ((1 << l) - 2); }

// Builds the hot-path twiddle tables of an FP32 plan and points each hot
// segment at its P's tables.
int build_hot_tables(mk_filt_plan* p) {
  std::vector<int64_t> Ps;
  std::vector<int> p1s;
  for (const Seg& s : p->h_segs)
    if (hot_shape(s)) {
      const int64_t P = (int64_t)s.p1 * kHotP2;
      if (std::find(Ps.begin(), Ps.end(), P) == Ps.end()) Ps.push_back(P);
      if (std::find(p1s.begin(), p1s.end(), s.p1) == p1s.end()) p1s.push_back(s.p1);
    }
  if (Ps.empty()) return MK_OK;
  std::vector<int64_t> offs;
  int64_t tot = 2 * kHotP2;  // stage tables of 2^1 .. 2^11 points first
  for (int64_t P : Ps) {
    offs.push_back(tot);
    tot += (1 << kFineBits) + 2 * P / (1 << kFineBits);
  }
  for (int v : p1s) {
    p->g_stab[hot_p1_index(v)] = tot;
    tot += v;  // v - 1 entries, rounded up
  }
  if (hipMalloc(&p->tw_tab, tot * sizeof(cx<float>)) != hipSuccess) {
    p->tw_tab = nullptr;
    mk::set_error("filter plan: twiddle table allocation failed");
    return MK_ERR_ALLOC;
  }
  cx<float>* base = (cx<float>*)p->tw_tab;
  for (int l = 1; l <= kHotL2; l++) hot_tables_kernel<<<8, 256>>>(base + ((1 << l) - 2), 0, l);
  for (size_t i = 0; i < Ps.size(); i++) hot_tables_kernel<<<64, 256>>>(base + offs[i], Ps[i], 0);
  for (int v : p1s) mixed_stage_table_kernel<<<1, 256>>>(base + p->g_stab[hot_p1_index(v)], v);
  int r = mk::check_launch("hot_tables_kernel");
  if (r) return r;
  MK_HIP(hipDeviceSynchronize());
  for (Seg& s : p->h_segs)
    if (hot_shape(s)) {
      const int64_t P = (int64_t)s.p1 * kHotP2;
      size_t i = std::find(Ps.begin(), Ps.end(), P) - Ps.begin();
      s.tw_ptr = base + offs[i];
    }
  return MK_OK;
}

void free_filt(mk_filt_plan* p) {
  if (!p) return;
  if (p->d_segs) (void)hipFree(p->d_segs);
  if (p->d_hot) (void)hipFree(p->d_hot);
  if (p->d_hot13) (void)hipFree(p->d_hot13);
  if (p->d_cold) (void)hipFree(p->d_cold);
  if (p->tw_tab) (void)hipFree(p->tw_tab);
  if (p->tbuf && p->tbuf_owned) (void)hipFree(p->tbuf);
  if (p->spec) (void)hipFree(p->spec);
  delete p;
}

// desc per segment: {rows, n, in_off, in_stride, out_off, out_stride}
// mixed: FP32 filter plans whose rows may take the hot path's 5-smooth
// column lengths.
// p1 (may be null): per segment a hot column length chosen by the caller
// (0 = the default, the smallest one that fits).
int build_segs(int nseg, const int64_t* desc, int eb, bool filter, std::vector<Seg>& segs, int64_t* n1,
               int64_t* n2, int64_t* t_elems, int64_t* s_elems, size_t* lds1, size_t* lds2, bool mixed = false,
               const int32_t* p1 = nullptr) {
  segs.clear();
  int64_t b1 = 0, b2 = 0, toff = 0, soff = 0;
  size_t l1max = 0, l2max = 0;
  for (int i = 0; i < nseg; i++) {
    const int64_t* d = desc + 6 * i;
    int64_t rows = d[0], n = d[1];
    MK_REQUIRE(rows >= 0 && n >= 2 && n < (int64_t(1) << 30), MK_ERR_ARG,
               "segment %d: need rows >= 0 and 2 <= n < 2^30 (rows=%lld n=%lld)", i, (long long)rows,
               (long long)n);
    int64_t m = n - 1;
    const int64_t P_min = filter ? (3 * m + 2) / 2 : 2 * m + 1;
    Geometry g = geometry(P_min, eb, mixed);
    if (p1 && p1[i] > 0) {
      MK_REQUIRE(mixed && hot_p1_index(p1[i]) >= 0, MK_ERR_ARG,
                 "segment %d: column length %d is not one of the FP32 hot lengths", i, (int)p1[i]);
      MK_REQUIRE((int64_t)p1[i] * kHotP2 >= P_min && P_min > kHotP2, MK_ERR_ARG,
                 "segment %d: column length %d is too short for n=%lld", i, (int)p1[i], (long long)n);
      g = hot_geometry(p1[i]);
    }
    MK_REQUIRE(g.P <= (int64_t(1) << 24), MK_ERR_ARG, "segment %d: n=%lld too long", i, (long long)n);
    Seg s{};
    s.in_off = d[2]; s.in_stride = d[3]; s.out_off = d[4]; s.out_stride = d[5];
    s.rows = (int32_t)rows; s.n = (int32_t)n; s.l1 = g.l1; s.l2 = g.l2; s.cw = g.cw; s.flags = 0; s.p1 = g.p1;
    s.t_off = toff; s.t_stride = g.P;
    s.s_off = soff; s.s_stride = filter ? g.P + 1 : 0;
    s.per1 = (int32_t)((int64_t(1) << g.l2) / g.cw);
    s.r2 = k2_r2(g, 8, filter);  // same grouping for FP32 and FP64 passes
    s.per2 = (int32_t)mk::cdiv(filter ? g.p1 / 2 + 1 : g.p1, s.r2);
    s.base1 = b1; s.base2 = b2; s.base3 = b1;
    b1 += rows * s.per1;
    b2 += rows * s.per2;
    toff += rows * g.P;
    soff += filter ? rows * (g.P + 1) : g.P;
    l1max = std::max(l1max, lds_k13(g, eb));
    l2max = std::max(l2max, lds_k2(g, eb, filter));
    MK_REQUIRE(l1max <= (size_t)kMaxLds && l2max <= (size_t)kMaxLds, MK_ERR_ARG,
               "segment %d: n=%lld needs more LDS than the %d-byte budget at %d-byte precision", i,
               (long long)n, kMaxLds, eb);
    segs.push_back(s);
  }
  *n1 = b1; *n2 = b2; *t_elems = toff; *s_elems = soff; *lds1 = l1max; *lds2 = l2max;
  return MK_OK;
}

// Spectrum pass in arithmetic type T into spectra of type TS; tb holds
// rows*P complex<T>; lds1/lds2 sized for T.
template <typename T, typename TS>
void launch_fspec(mk_filt_plan* p, const void* wdct, void* tb, size_t lds1, size_t lds2, hipStream_t st) {
  k1_kernel<T, double, kFSpec><<<(unsigned)p->n1, kThreads, lds1, st>>>(p->d_segs, p->nseg, (const double*)wdct,
                                                                        (cx<T>*)tb);
  k2_kernel<T, TS, kFSpec><<<(unsigned)p->n2, kThreads, lds2, st>>>(p->d_segs, p->nseg, (cx<T>*)tb);
}

template <int I>
void launch_hot13(const mk_filt_plan* p, int which, const float* in, float* out, hipStream_t st) {
  constexpr int P1 = kHotP1[I];
  const Seg* segs = p->d_hot13 + p->g_off[I];
  const unsigned grid = (unsigned)p->g_n1[I];
  const cx<float>* stab = (const cx<float>*)p->tw_tab + p->g_stab[I];
  if (which == 1)  // XCD-aware block order: the grid is rounded up to a multiple of 8 XCDs
    k1_hot_kernel<P1><<<(unsigned)(mk::cdiv(grid, 8) * 8), kThreads, hot13_lds(P1, 1), st>>>(
        segs, p->g_cnt[I], in, (cx<float>*)p->tbuf, stab, (int64_t)grid);
  else
    k3_hot_kernel<P1><<<grid, kThreads, hot13_lds(P1, 3), st>>>(segs, p->g_cnt[I], (const cx<float>*)p->tbuf, out,
                                                             stab);
}

template <int I = 0>
void launch_hot13_idx(const mk_filt_plan* p, int idx, int which, const float* in, float* out, hipStream_t st) {
  if constexpr (I < kNumHotP1) {
    if (idx == I) launch_hot13<I>(p, which, in, out, st);
    else launch_hot13_idx<I + 1>(p, idx, which, in, out, st);
  }
}

void launch_hot13_all(const mk_filt_plan* p, int which, const float* in, float* out, hipStream_t st) {
  for (int i = 0; i < kNumHotP1; i++)
    if (p->g_cnt[i] && p->g_n1[i]) launch_hot13_idx(p, i, which, in, out, st);
}

template <typename T, typename TIO>
void launch_filter(const mk_filt_plan* p, const TIO* in, TIO* out, hipStream_t st) {
  if constexpr (sizeof(T) == 4) {
    if (p->nhot + p->ncold == p->nseg) {
      launch_hot13_all(p, 1, in, out, st);
      if (p->n1cold)
        k1_kernel<T, TIO, kFilter><<<(unsigned)p->n1cold, kThreads, p->lds1, st>>>(p->d_cold, p->ncold, in,
                                                                                  (cx<T>*)p->tbuf);
      if (p->n2hot)
        k2_hot_kernel<<<(unsigned)p->n2hot, kThreads, kHotLds, st>>>(p->d_hot, p->nhot, (cx<float>*)p->tbuf,
                                                                    stage_tab(p, kHotL2));
      if (p->n2cold)
        k2_kernel<T, T, kFilter><<<(unsigned)p->n2cold, kThreads, p->lds2, st>>>(p->d_cold, p->ncold,
                                                                                (cx<T>*)p->tbuf);
      launch_hot13_all(p, 3, in, out, st);
      if (p->n1cold)
        k3_kernel<T, TIO, kFilter><<<(unsigned)p->n1cold, kThreads, p->lds1, st>>>(p->d_cold, p->ncold,
                                                                                  (const cx<T>*)p->tbuf, out);
      return;
    }
  }
  k1_kernel<T, TIO, kFilter><<<(unsigned)p->n1, kThreads, p->lds1, st>>>(p->d_segs, p->nseg, in, (cx<T>*)p->tbuf);
  k2_kernel<T, T, kFilter><<<(unsigned)p->n2, kThreads, p->lds2, st>>>(p->d_segs, p->nseg, (cx<T>*)p->tbuf);
  k3_kernel<T, TIO, kFilter><<<(unsigned)p->n1, kThreads, p->lds1, st>>>(p->d_segs, p->nseg,
                                                                         (const cx<T>*)p->tbuf, out);
}

}  // namespace

extern "C" int mk_filt_plan_create(int nseg, const int64_t* desc, int elem_bytes, mk_filt_plan** out) {
  return mk_filt_plan_create_p1(nseg, desc, elem_bytes, nullptr, out);
}

extern "C" int mk_filt_hot_p1(int32_t* out, int cap) {
  if (out)
    for (int i = 0; i < kNumHotP1 && i < cap; i++) out[i] = kHotP1[i];
  return kNumHotP1;
}

extern "C" int mk_filt_plan_create_p1(int nseg, const int64_t* desc, int elem_bytes, const int32_t* p1,
                                      mk_filt_plan** out) {
  MK_REQUIRE(out, MK_ERR_ARG, "out is NULL");
  *out = nullptr;
  MK_REQUIRE(nseg >= 0 && (nseg == 0 || desc), MK_ERR_ARG, "bad segment table");
  MK_REQUIRE(elem_bytes == 4 || elem_bytes == 8, MK_ERR_ARG, "elem_bytes must be 4 or 8");
  allow_lds_all();
  mk_filt_plan* p = new mk_filt_plan();
  p->nseg = nseg;
  p->borrowed.assign(std::max(1, nseg), 0);
  p->elem_bytes = elem_bytes;
  const bool hot = elem_bytes == 4 && getenv("MK_FFT_NO_HOT") == nullptr;
  int st = build_segs(nseg, desc, elem_bytes, true, p->h_segs, &p->n1, &p->n2, &p->t_elems, &p->s_elems, &p->lds1,
                      &p->lds2, hot, p1);
  if (st) { free_filt(p); return st; }
  MK_REQUIRE(p->n1 < INT32_MAX && p->n2 < INT32_MAX, (free_filt(p), MK_ERR_ARG), "too many work items");
  if (hipMalloc((void**)&p->d_segs, std::max<size_t>(1, nseg) * sizeof(Seg)) != hipSuccess ||
      hipMalloc((void**)&p->d_hot, std::max<size_t>(1, nseg) * sizeof(Seg)) != hipSuccess ||
      hipMalloc((void**)&p->d_cold, std::max<size_t>(1, nseg) * sizeof(Seg)) != hipSuccess ||
      hipMalloc((void**)&p->d_hot13, std::max<size_t>(1, nseg) * sizeof(Seg)) != hipSuccess) {
    free_filt(p);
    mk::set_error("filter plan: device allocation failed");
    return MK_ERR_ALLOC;
  }
  // spectra are allocated by mk_filt_plan_set_weights: a plan that borrows
  // every segment's spectrum never holds any of its own
  for (auto& sg : p->h_segs) sg.s_ptr = nullptr;
  if (hot) {
    st = build_hot_tables(p);
    if (st) { free_filt(p); return st; }
  }
  if (upload(p)) {
    free_filt(p);
    return MK_ERR_HIP;
  }
  *out = p;
  return MK_OK;
}

extern "C" int mk_filt_plan_borrow_spectrum(mk_filt_plan* p, int seg, const mk_filt_plan* src, int src_seg) {
  MK_REQUIRE(p && src, MK_ERR_ARG, "null plan");
  MK_REQUIRE(seg >= 0 && seg < p->nseg && src_seg >= 0 && src_seg < src->nseg, MK_ERR_ARG, "segment out of range");
  MK_REQUIRE(p->elem_bytes == src->elem_bytes, MK_ERR_ARG, "precision mismatch");
  MK_REQUIRE(src->have_spec, MK_ERR_ARG, "source plan has no weights");
  Seg& d = p->h_segs[seg];
  const Seg& s = src->h_segs[src_seg];
  MK_REQUIRE(d.rows == s.rows && d.n == s.n && d.p1 == s.p1 && d.l2 == s.l2, MK_ERR_ARG,
             "segment geometry differs (rows %d/%d, n %d/%d)", d.rows, s.rows, d.n, s.n);
  d.s_ptr = s.s_ptr;
  p->borrowed[seg] = 1;
  bool all = true;
  for (int i = 0; i < p->nseg; i++) all = all && p->borrowed[i];
  if (all) p->have_spec = true;
  return upload(p) ? MK_ERR_HIP : MK_OK;
}

extern "C" int mk_filt_plan_destroy(mk_filt_plan* p) {
  free_filt(p);
  return MK_OK;
}

extern "C" int mk_filt_plan_info(const mk_filt_plan* p, int64_t* info) {
  MK_REQUIRE(p && info, MK_ERR_ARG, "null pointer");
  info[0] = p->nseg;
  info[1] = p->n1;
  info[2] = p->n2;
  info[3] = p->t_elems * 2 * p->elem_bytes;
  info[4] = p->s_elems * p->elem_bytes;
  info[5] = p->elem_bytes;
  info[6] = (int64_t)p->lds1;
  info[7] = (int64_t)p->lds2;
  return MK_OK;
}

// Per-segment FFT geometry: out[4*i..] = {P1, P2, cw, P} of segment i.
extern "C" int mk_filt_plan_geometry(const mk_filt_plan* p, int64_t* out) {
  MK_REQUIRE(p && out, MK_ERR_ARG, "null pointer");
  for (int i = 0; i < p->nseg; i++) {
    const Seg& s = p->h_segs[i];
    out[4 * i] = s.p1;
    out[4 * i + 1] = int64_t(1) << s.l2;
    out[4 * i + 2] = s.cw;
    out[4 * i + 3] = s.t_stride;
  }
  return MK_OK;
}

extern "C" int mk_filt_plan_set_weights(mk_filt_plan* p, const double* wdct, void* stream) {
  MK_REQUIRE(p && (p->nseg == 0 || wdct), MK_ERR_ARG, "null argument");
  for (int i = 0; i < p->nseg; i++)
    MK_REQUIRE(!p->borrowed[i], MK_ERR_ARG, "plan borrows spectra; set_weights would overwrite the source plan");
  if (p->n1 == 0) { p->have_spec = true; return MK_OK; }
  hipStream_t st = mk::as_stream(stream);
  if (!p->spec) {
    if (hipMalloc(&p->spec, std::max<int64_t>(1, p->s_elems) * p->elem_bytes) != hipSuccess) {
      p->spec = nullptr;
      mk::set_error("filter plan: spectrum allocation (%lld B) failed", (long long)(p->s_elems * p->elem_bytes));
      return MK_ERR_ALLOC;
    }
    for (auto& sg : p->h_segs) sg.s_ptr = (unsigned char*)p->spec + sg.s_off * p->elem_bytes;
    int r = upload(p);
    if (r) return r;
  }
  // The spectrum is always computed in FP64 (a transient FP64 work buffer)
  // and stored in the plan's precision, unless FP64 LDS tiles would not fit.
  size_t l1_64 = 0, l2_64 = 0;
  for (const Seg& sg : p->h_segs) {
    Geometry g{sg.l1, sg.l2, sg.cw, sg.p1, (int64_t)sg.p1 << sg.l2};
    l1_64 = std::max(l1_64, lds_k13(g, 8));
    l2_64 = std::max(l2_64, lds_k2(g, 8));
  }
  const bool fp64_pass = l1_64 <= (size_t)kMaxLds && l2_64 <= (size_t)kMaxLds;
  void* tb = nullptr;
  if (fp64_pass) {
    MK_HIP(hipMalloc(&tb, std::max<int64_t>(1, p->t_elems) * 16));
    if (p->elem_bytes == 4) launch_fspec<double, float>(p, wdct, tb, l1_64, l2_64, st);
    else launch_fspec<double, double>(p, wdct, tb, l1_64, l2_64, st);
  } else {
    MK_HIP(hipMalloc(&tb, std::max<int64_t>(1, p->t_elems) * 8));
    launch_fspec<float, float>(p, wdct, tb, p->lds1, p->lds2, st);
  }
  int r = mk::check_launch("filter spectrum");
  if (r == MK_OK) p->have_spec = true;
  MK_HIP(hipStreamSynchronize(st));
  MK_HIP(hipFree(tb));
  return r;
}

extern "C" int mk_filt_plan_exec_f32(const mk_filt_plan* p, const float* in, float* out, void* stream) {
  MK_REQUIRE(p && in && out, MK_ERR_ARG, "null argument");
  MK_REQUIRE(p->have_spec, MK_ERR_ARG, "filter weights not set");
  MK_REQUIRE(p->elem_bytes == 4, MK_ERR_ARG, "plan was created for FP64");
  if (p->n1 == 0) return MK_OK;
  int r = ensure_tbuf(p);
  if (r) return r;
  launch_filter<float, float>(p, in, out, mk::as_stream(stream));
  return mk::check_launch("filter exec f32");
}

extern "C" int mk_filt_plan_exec_f64(const mk_filt_plan* p, const double* in, double* out, void* stream) {
  MK_REQUIRE(p && in && out, MK_ERR_ARG, "null argument");
  MK_REQUIRE(p->have_spec, MK_ERR_ARG, "filter weights not set");
  MK_REQUIRE(p->elem_bytes == 8, MK_ERR_ARG, "plan was created for FP32");
  if (p->n1 == 0) return MK_OK;
  int r = ensure_tbuf(p);
  if (r) return r;
  launch_filter<double, double>(p, in, out, mk::as_stream(stream));
  return mk::check_launch("filter exec f64");
}

extern "C" int mk_filt_plan_release_work(mk_filt_plan* p, void* stream) {
  MK_REQUIRE(p, MK_ERR_ARG, "null plan");
  if (p->tbuf) {
    MK_HIP(hipStreamSynchronize(mk::as_stream(stream)));
    if (p->tbuf_owned) MK_HIP(hipFree(p->tbuf));
    p->tbuf = nullptr;
  }
  return MK_OK;
}

extern "C" int mk_filt_plan_work_bytes(const mk_filt_plan* p, int64_t* bytes) {
  MK_REQUIRE(p && bytes, MK_ERR_ARG, "null pointer");
  *bytes = p->t_elems * 2 * p->elem_bytes;
  return MK_OK;
}

extern "C" int mk_filt_plan_set_work(mk_filt_plan* p, void* work, int64_t bytes, void* stream) {
  MK_REQUIRE(p, MK_ERR_ARG, "null plan");
  const int64_t need = p->t_elems * 2 * p->elem_bytes;
  MK_REQUIRE(work == nullptr || (bytes >= need && ((uintptr_t)work & 15) == 0), MK_ERR_ARG,
             "work buffer of %lld B (need %lld B, 16-byte aligned)", (long long)bytes, (long long)need);
  int r = mk_filt_plan_release_work(p, stream);
  if (r) return r;
  if (work) {
    p->tbuf = work;
    p->tbuf_owned = false;
  }
  return MK_OK;
}

extern "C" int mk_filt_plan_set_halve(mk_filt_plan* p, int halve_edges) {
  MK_REQUIRE(p, MK_ERR_ARG, "null plan");
  for (auto& s : p->h_segs) s.flags = halve_edges ? 1 : 0;
  return upload(p);
}

// ---------------------------------------------------------------------------
// Stand-alone DCT-I (Bluestein), rows of one length n.

extern "C" int64_t mk_dct1_workspace_bytes(int64_t nrow, int64_t n, int elem_bytes) {
  if (n < 2 || nrow < 0 || (elem_bytes != 4 && elem_bytes != 8)) return -1;
  Geometry g = geometry(2 * (n - 1) + 1, elem_bytes);
  return (nrow * g.P + g.P) * 2 * elem_bytes + 4096;
}

namespace {
template <typename T, typename TIO>
int dct1_impl(const TIO* x, TIO* y, int64_t nrow, int64_t n, void* work, int64_t work_bytes, void* stream) {
  MK_REQUIRE(nrow >= 0 && n >= 2, MK_ERR_ARG, "DCT-I needs n >= 2 (got %lld)", (long long)n);
  if (nrow == 0) return MK_OK;
  MK_REQUIRE(x && y && work, MK_ERR_ARG, "null argument");
  int64_t need = mk_dct1_workspace_bytes(nrow, n, sizeof(T));
  MK_REQUIRE(work_bytes >= need, MK_ERR_ARG, "workspace too small (%lld < %lld)", (long long)work_bytes,
             (long long)need);
  MK_REQUIRE(((uintptr_t)work & 255) == 0, MK_ERR_ARG, "workspace must be 256-byte aligned");
  allow_lds_all();
  hipStream_t st = mk::as_stream(stream);
  int64_t desc[6] = {nrow, n, 0, n, 0, n};
  std::vector<Seg> segs;
  int64_t n1, n2, te, se;
  size_t l1, l2;
  int r = build_segs(1, desc, sizeof(T), false, segs, &n1, &n2, &te, &se, &l1, &l2);
  if (r) return r;
  MK_REQUIRE(n1 < INT32_MAX && n2 < INT32_MAX, MK_ERR_ARG, "too many rows");
  // workspace: [segment descriptor (aligned)] [V spectrum P] [T buffer nrow*P]
  unsigned char* w = (unsigned char*)work;
  Seg* d_seg = (Seg*)w;
  cx<T>* V = (cx<T>*)(w + 4096);
  cx<T>* tb = V + segs[0].t_stride;
  Seg s0 = segs[0];
  s0.s_off = 0;
  s0.s_ptr = V;
  MK_HIP(hipMemcpyAsync(d_seg, &s0, sizeof(Seg), hipMemcpyHostToDevice, st));
  MK_HIP(hipStreamSynchronize(st));
  // spectrum of the conjugate chirp: a one-row pass (rows = 1) over the same geometry
  Seg sv = s0;
  sv.rows = 1;
  Seg* d_sv = d_seg + 1;
  MK_HIP(hipMemcpyAsync(d_sv, &sv, sizeof(Seg), hipMemcpyHostToDevice, st));
  MK_HIP(hipStreamSynchronize(st));
  k1_kernel<T, TIO, kBSpec><<<(unsigned)s0.per1, kThreads, l1, st>>>(d_sv, 1, x, V);
  k2_kernel<T, T, kBSpec><<<(unsigned)s0.per2, kThreads, l2, st>>>(d_sv, 1, V);
  k1_kernel<T, TIO, kBlue><<<(unsigned)n1, kThreads, l1, st>>>(d_seg, 1, x, tb);
  k2_kernel<T, T, kBlue><<<(unsigned)n2, kThreads, l2, st>>>(d_seg, 1, tb);
  k3_kernel<T, TIO, kBlue><<<(unsigned)n1, kThreads, l1, st>>>(d_seg, 1, tb, y);
  return mk::check_launch("dct1");
}
}  // namespace

// Static workspace of mk_dct1_f64_noalloc: the Bluestein workspace (segment
// descriptors, chirp spectrum, nrow x P complex FP64 rows) for callers that
// cannot allocate device memory.  64 MiB per device that runs the library,
// held until the process ends.  Host code never uses this name as an
// address: it asks hipGetSymbolAddress for each device's copy once, with
// that device current, and passes that pointer on.
constexpr int64_t kScratchBytes = int64_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (looked up on its first use).
int scratch_ptr(void** out) {
  int dev = -1;
  MK_HIP(hipGetDevice(&dev));
  MK_REQUIRE(dev >= 0 && dev < kMaxDevices, MK_ERR_ARG, "device index %d outside 0..%d", dev, kMaxDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MK_REQUIRE(p != nullptr && ((uintptr_t)p & 255) == 0, MK_ERR_HIP, "scratch address lookup gave %p", p);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MK_OK;
}
}  // namespace

extern "C" int mk_dct1_f64_noalloc(const double* x, double* y, int64_t nrow, int64_t n, void* stream) {
  MK_REQUIRE(nrow >= 0 && n >= 2, MK_ERR_ARG, "DCT-I needs nrow >= 0 and n >= 2 (got %lld, %lld)", (long long)nrow,
             (long long)n);
  if (nrow == 0) return MK_OK;
  const int64_t need = mk_dct1_workspace_bytes(nrow, n, 8);
  MK_REQUIRE(need > 0 && need <= kScratchBytes, MK_ERR_ARG,
             "mk_dct1_f64_noalloc: %lld rows of %lld points need %lld B of workspace, more than its %lld B static "
             "scratch; use mk_dct1_f64 with a caller workspace (or fewer rows per call)",
             (long long)nrow, (long long)n, (long long)need, (long long)kScratchBytes);
  void* ws = nullptr;
  int r = scratch_ptr(&ws);
  if (r) return r;
  return dct1_impl<double, double>(x, y, nrow, n, ws, kScratchBytes, stream);
}

extern "C" int mk_dct1_f64(const double* x, double* y, int64_t nrow, int64_t n, void* work, int64_t work_bytes,
                           void* stream) {
  return dct1_impl<double, double>(x, y, nrow, n, work, work_bytes, stream);
}
extern "C" int mk_dct1_f32(const float* x, float* y, int64_t nrow, int64_t n, void* work, int64_t work_bytes,
                           void* stream) {
  return dct1_impl<float, float>(x, y, nrow, n, work, work_bytes, stream);
}

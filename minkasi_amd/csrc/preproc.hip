// TOD preprocessing of minkasi's canonical pipeline
// (examples/minkasi_mpi_example.py:42-52): fit_cm_plus_poly
// (minkasi/tods/processing.py:510-579), the common-mode median over
// detectors plus a least-squares fit of Legendre drifts and the Legendre-
// weighted common mode (_linfit_2mat, processing.py:28-39).
//
//  median_cols   : per sample (column) the median over the detectors (rows),
//                  numpy semantics (mean of the two middle values for an even
//                  count).  One wave per column: the column is loaded into
//                  registers (64 x V values, padded with +inf) and sorted
//                  with a bitonic network, lanes exchanging through
//                  __shfl_xor and register pairs swapping in place.
//  design        : X[t][k] = P_k(x_t) (k <= ord) and X[t][ord+1+i] =
//                  P_i(x_t) cm[t] (i < cm_ord), x_t = (t - (n-1)/2) / ((n-1)/2),
//                  Legendre P by the three-term recurrence, FP64; also X^T.
//  design_dots   : D[k][j] = sum_t X[t][k] y_j[t] for rows y_j, FP64, per-row
//                  chunk partials summed in a fixed order.
//  design_apply  : out[j][t] = base[j][t] - sum_k X[t][c0+k] coef[k][j] (or
//                  the plain product without base).
#include "common.h"

#include <cmath>

namespace {

constexpr int kWave = 64;

template <typename T, int V>
__global__ void __launch_bounds__(256) median_cols_kernel(const T* __restrict__ dat, int64_t nrow, int64_t ncol,
                                                          int64_t ld, T* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= ncol) return;  // whole wave
  constexpr int N = kWave * V;
  T x[V];
#pragma unroll
  for (int e = 0; e < V; e++) {
    const int64_t i = (int64_t)e * kWave + lane;
    x[e] = i < nrow ? dat[i * ld + c] : T(INFINITY);
  }
  // bitonic sort of index i = e * 64 + lane, ascending
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= kWave) {  // partner in the same lane: register e ^ (j / 64)
        const int jj = j / kWave;
#pragma unroll
        for (int e = 0; e < V; e++) {
          if (e & jj) continue;
          const int i = e * kWave + lane;
          const bool up = (i & k) == 0;
          const T a = x[e], b = x[e ^ jj];
          const T lo = a < b ? a : b, hi = a < b ? b : a;
          x[e] = up ? lo : hi;
          x[e ^ jj] = up ? hi : lo;
        }
      } else {  // partner in lane ^ j, same register
#pragma unroll
        for (int e = 0; e < V; e++) {
          const int i = e * kWave + lane;
          const T y = __shfl_xor(x[e], j, 64);
          const bool up = (i & k) == 0, lower = (i & j) == 0;
          const T lo = x[e] < y ? x[e] : y, hi = x[e] < y ? y : x[e];
          x[e] = (up == lower) ? lo : hi;
        }
      }
    }
  }
  auto at = [&](int64_t i) -> T {  // sorted element i (uniform index)
    const int e = (int)(i / kWave), l = (int)(i % kWave);
    T v = x[0];
#pragma unroll
    for (int q = 1; q < V; q++)
      if (q == e) v = x[q];
    return __shfl(v, l, 64);
  };
  T med;
  if (nrow & 1) {
    med = at((nrow - 1) / 2);
  } else {
    const T a = at(nrow / 2 - 1), b = at(nrow / 2);
    med = (a + b) / T(2);
  }
  if (lane == 0) out[c] = med;
}

template <typename T>
__global__ void __launch_bounds__(256) design_kernel(const T* __restrict__ cm, int64_t n, int ord, int cm_ord,
                                                     double* __restrict__ X, double* __restrict__ Xt) {
  const int ncol = ord + 1 + cm_ord;
  const double half = 0.5 * (double)(n - 1);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const double x = ((double)t - half) / half;
    const int kmax = ord > cm_ord - 1 ? ord : cm_ord - 1;
    double pm2 = 1.0, pm1 = x;
    const double c = (double)cm[t];
    for (int k = 0; k <= kmax; k++) {
      double p;
      if (k == 0) p = 1.0;
      else if (k == 1) p = x;
      else {
        p = (pm1 * x * (2 * k - 1) - pm2 * (k - 1)) / k;
        pm2 = pm1;
        pm1 = p;
      }
      if (k <= ord) {
        X[t * ncol + k] = p;
        Xt[k * n + t] = p;
      }
      if (k < cm_ord) {
        X[t * ncol + ord + 1 + k] = p * c;
        Xt[(int64_t)(ord + 1 + k) * n + t] = p * c;
      }
    }
  }
}

constexpr int kDotChunk = 8192;  // samples per partial
constexpr int kMaxCol = 8;

// part[(j * nchunk + chunk) * ncol + k] = sum over the chunk of X[t][k] y_j[t]
template <typename T>
__global__ void __launch_bounds__(256) design_dots_kernel(const double* __restrict__ X, int64_t n, int ncol,
                                                          const T* __restrict__ y, int64_t ld, int64_t nchunk,
                                                          double* __restrict__ part) {
  __shared__ double red[kMaxCol][256];
  const int64_t j = blockIdx.y, chunk = blockIdx.x;
  const int64_t t0 = chunk * kDotChunk, t1 = min(n, t0 + kDotChunk);
  double acc[kMaxCol];
#pragma unroll
  for (int k = 0; k < kMaxCol; k++) acc[k] = 0.0;
  const T* row = y + j * ld;
  for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const double v = (double)row[t];
#pragma unroll
    for (int k = 0; k < kMaxCol; k++)
      if (k < ncol) acc[k] += X[t * ncol + k] * v;
  }
#pragma unroll
  for (int k = 0; k < kMaxCol; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < kMaxCol; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if ((int)threadIdx.x < ncol) part[(j * nchunk + chunk) * ncol + threadIdx.x] = red[threadIdx.x][0];
}

// D[k][j] = sum over chunks (in order) of the partials
__global__ void design_dots_reduce(const double* __restrict__ part, int64_t nrow, int ncol, int64_t nchunk,
                                   double* __restrict__ D) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nrow * ncol;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / ncol, k = e % ncol;
    double s = 0.0;
    for (int64_t c = 0; c < nchunk; c++) s += part[(j * nchunk + c) * ncol + k];
    D[k * nrow + j] = s;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) design_apply_kernel(const T* __restrict__ base, int64_t nrow, int64_t n,
                                                           int64_t ld, const double* __restrict__ X, int ncol, int c0,
                                                           int nk, const double* __restrict__ coef, T* __restrict__ out,
                                                           int64_t ld_out) {
  const int64_t total = nrow * n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / n, t = e - j * n;
    double p = 0.0;
    for (int k = 0; k < nk; k++) p += X[t * ncol + c0 + k] * coef[k * nrow + j];
    out[j * ld_out + t] = base ? (T)((double)base[j * ld + t] - p) : (T)p;
  }
}

template <typename T>
int median_cols_impl(const T* dat, int64_t nrow, int64_t ncol, int64_t ld, T* out, void* stream) {
  MK_REQUIRE(nrow >= 1 && nrow <= 1024 && ncol >= 0 && ld >= ncol, MK_ERR_ARG,
             "median over 1..1024 rows (got %lld rows, %lld columns, ld %lld)", (long long)nrow, (long long)ncol,
             (long long)ld);
  if (ncol == 0) return MK_OK;
  MK_REQUIRE(dat && out, MK_ERR_ARG, "null pointer");
  const unsigned grid = (unsigned)mk::cdiv(ncol, 4);
  hipStream_t st = mk::as_stream(stream);
  if (nrow <= 64) median_cols_kernel<T, 1><<<grid, 256, 0, st>>>(dat, nrow, ncol, ld, out);
  else if (nrow <= 128) median_cols_kernel<T, 2><<<grid, 256, 0, st>>>(dat, nrow, ncol, ld, out);
  else if (nrow <= 256) median_cols_kernel<T, 4><<<grid, 256, 0, st>>>(dat, nrow, ncol, ld, out);
  else if (nrow <= 512) median_cols_kernel<T, 8><<<grid, 256, 0, st>>>(dat, nrow, ncol, ld, out);
  else median_cols_kernel<T, 16><<<grid, 256, 0, st>>>(dat, nrow, ncol, ld, out);
  return mk::check_launch("median_cols_kernel");
}

template <typename T>
int design_impl(const T* cm, int64_t n, int ord, int cm_ord, double* X, double* Xt, void* stream) {
  MK_REQUIRE(n >= 2 && ord >= 0 && cm_ord >= 0 && ord + 1 + cm_ord <= kMaxCol, MK_ERR_ARG,
             "design needs n >= 2 and ord + 1 + cm_ord <= %d", kMaxCol);
  MK_REQUIRE(cm && X && Xt, MK_ERR_ARG, "null pointer");
  design_kernel<T><<<mk::grid_for(n, 256), 256, 0, mk::as_stream(stream)>>>(cm, n, ord, cm_ord, X, Xt);
  return mk::check_launch("design_kernel");
}

template <typename T>
int dots_impl(const double* X, int64_t n, int ncol, const T* y, int64_t nrow, int64_t ld, double* D, void* work,
              int64_t work_bytes, void* stream) {
  MK_REQUIRE(n >= 1 && ncol >= 1 && ncol <= kMaxCol && nrow >= 1 && nrow < 65536 && ld >= n, MK_ERR_ARG,
             "bad design dot sizes");
  MK_REQUIRE(X && y && D && work, MK_ERR_ARG, "null pointer");
  const int64_t nchunk = mk::cdiv(n, kDotChunk);
  MK_REQUIRE(work_bytes >= nrow * nchunk * ncol * 8, MK_ERR_ARG, "workspace too small");
  hipStream_t st = mk::as_stream(stream);
  design_dots_kernel<T><<<dim3((unsigned)nchunk, (unsigned)nrow), 256, 0, st>>>(X, n, ncol, y, ld, nchunk,
                                                                               (double*)work);
  int r = mk::check_launch("design_dots_kernel");
  if (r) return r;
  design_dots_reduce<<<mk::grid_for(nrow * ncol, 256), 256, 0, st>>>((const double*)work, nrow, ncol, nchunk, D);
  return mk::check_launch("design_dots_reduce");
}

template <typename T>
int apply_impl(const T* base, int64_t nrow, int64_t n, int64_t ld, const double* X, int ncol, int c0, int nk,
               const double* coef, T* out, int64_t ld_out, void* stream) {
  MK_REQUIRE(nrow >= 0 && n >= 1 && ncol >= 1 && c0 >= 0 && nk >= 0 && c0 + nk <= ncol && ld >= n && ld_out >= n,
             MK_ERR_ARG, "bad design apply sizes");
  if (nrow == 0) return MK_OK;
  MK_REQUIRE(X && coef && out, MK_ERR_ARG, "null pointer");
  design_apply_kernel<T><<<mk::grid_for(nrow * n, 256, 256 * 64), 256, 0, mk::as_stream(stream)>>>(
      base, nrow, n, ld, X, ncol, c0, nk, coef, out, ld_out);
  return mk::check_launch("design_apply_kernel");
}

}  // namespace

extern "C" int mk_median_cols_f32(const float* dat, int64_t nrow, int64_t ncol, int64_t ld, float* out,
                                  void* stream) {
  return median_cols_impl<float>(dat, nrow, ncol, ld, out, stream);
}
extern "C" int mk_median_cols_f64(const double* dat, int64_t nrow, int64_t ncol, int64_t ld, double* out,
                                  void* stream) {
  return median_cols_impl<double>(dat, nrow, ncol, ld, out, stream);
}
extern "C" int mk_cm_poly_design_f32(const float* cm, int64_t n, int ord, int cm_ord, double* X, double* Xt,
                                     void* stream) {
  return design_impl<float>(cm, n, ord, cm_ord, X, Xt, stream);
}
extern "C" int mk_cm_poly_design_f64(const double* cm, int64_t n, int ord, int cm_ord, double* X, double* Xt,
                                     void* stream) {
  return design_impl<double>(cm, n, ord, cm_ord, X, Xt, stream);
}
extern "C" int64_t mk_design_dots_workspace_bytes(int64_t n, int ncol, int64_t nrow) {
  return nrow * mk::cdiv(n, kDotChunk) * ncol * 8 + 256;
}
extern "C" int mk_design_dots_f32(const double* X, int64_t n, int ncol, const float* y, int64_t nrow, int64_t ld,
                                  double* D, void* work, int64_t work_bytes, void* stream) {
  return dots_impl<float>(X, n, ncol, y, nrow, ld, D, work, work_bytes, stream);
}
extern "C" int mk_design_dots_f64(const double* X, int64_t n, int ncol, const double* y, int64_t nrow, int64_t ld,
                                  double* D, void* work, int64_t work_bytes, void* stream) {
  return dots_impl<double>(X, n, ncol, y, nrow, ld, D, work, work_bytes, stream);
}
extern "C" int mk_design_apply_f32(const float* base, int64_t nrow, int64_t n, int64_t ld, const double* X, int ncol,
                                   int c0, int nk, const double* coef, float* out, int64_t ld_out, void* stream) {
  return apply_impl<float>(base, nrow, n, ld, X, ncol, c0, nk, coef, out, ld_out, stream);
}
extern "C" int mk_design_apply_f64(const double* base, int64_t nrow, int64_t n, int64_t ld, const double* X,
                                   int ncol, int c0, int nk, const double* coef, double* out, int64_t ld_out,
                                   void* stream) {
  return apply_impl<double>(base, nrow, n, ld, X, ncol, c0, nk, coef, out, ld_out, stream);
}

// Polarised maps of minkasi's PolMap (minkasi/maps/polmap.py:47-724).  A map
// holds ncomp values per pixel, pixel-major (index ncomp * ipix + c), with
// the polarisation angle of each sample given as twogamma:
//
//  polmap2tod (minkasi.c:146-175)
//    IQU : dat (+)= m[3p] + m[3p+1] cos 2g + m[3p+2] sin 2g
//    QU  : dat (+)= m[2p] cos 2g + m[2p+1] sin 2g
//  tod2polmap (minkasi.c:177-257), accumulating
//    QU         : (d c, d s)                      IQU        : (d, d c, d s)
//    QU_PRECON  : (d c c, d s s, d s c)           IQU_PRECON : (d, d c, d s, d c c, d s c, d s s)
//  The PRECON sums are the per-pixel blocks of P^T P: 2x2 [[QQ, QU], [QU, UU]]
//  and 3x3 [[I, Q, U], [Q, QQ, QU], [U, QU, UU]].
//  invert (polmap.py:300-384): each block replaced by its pseudo-inverse
//    (singular values at most thresh * the largest dropped, np.linalg.pinv's
//    rcond rule).  For IQU_PRECON, blocks whose determinant is below
//    thresh * (the largest determinant of the map) are zeroed, and those with
//    a positive I entry keep 1/I there.
//  precon_mul (polmap.py:624-684): the block times the (Q, U) or (I, Q, U)
//    vector of each pixel.
//
// tod2polmap scatters with hardware FP atomics: the summation order is not
// fixed (the reference sums serially).  The block pseudo-inverse comes from
// a cyclic Jacobi eigendecomposition in FP64: for a symmetric block the
// singular values are the absolute eigenvalues, so V diag(1/lambda) V^T over
// the kept eigenvalues is the pseudo-inverse numpy's SVD gives.
#include "common.h"

#include <cmath>

namespace {

__host__ __device__ constexpr int ncomp_of(int mode) {
  return mode == MK_POL_QU ? 2 : mode == MK_POL_IQU ? 3 : mode == MK_POL_QU_PRECON ? 3 : 6;
}

__device__ inline void sincos_t(float a, float* s, float* c) { sincosf(a, s, c); }
__device__ inline void sincos_t(double a, double* s, double* c) { sincos(a, s, c); }

template <typename T, int NPOL>
__global__ void __launch_bounds__(256) polmap2tod_kernel(T* __restrict__ dat, const T* __restrict__ map,
                                                         const T* __restrict__ twogamma, int64_t n,
                                                         const int32_t* __restrict__ ipix, int64_t npix, int do_add,
                                                         int64_t* nbad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t p = ipix[i];
    const bool bad = (uint64_t)p >= (uint64_t)npix;
    mk::count_bad(bad, nbad);
    if (bad) continue;
    T s, c;
    sincos_t(twogamma[i], &s, &c);
    const T* m = map + NPOL * p;
    const T v = NPOL == 3 ? m[0] + m[1] * c + m[2] * s : m[0] * c + m[1] * s;
    dat[i] = do_add ? dat[i] + v : v;
  }
}

template <typename T, int MODE>
__global__ void __launch_bounds__(256) tod2polmap_kernel(T* __restrict__ map, const T* __restrict__ dat,
                                                         const T* __restrict__ twogamma, int64_t n,
                                                         const int32_t* __restrict__ ipix, int64_t npix,
                                                         int64_t* nbad) {
  constexpr int NC = ncomp_of(MODE);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t p = ipix[i];
    const bool bad = (uint64_t)p >= (uint64_t)npix;
    mk::count_bad(bad, nbad);
    if (bad) continue;
    T s, c;
    sincos_t(twogamma[i], &s, &c);
    const T d = dat[i];
    T* m = map + NC * p;
    if (MODE == MK_POL_QU) {
      atomicAdd(m + 0, d * c);
      atomicAdd(m + 1, d * s);
    } else if (MODE == MK_POL_IQU) {
      atomicAdd(m + 0, d);
      atomicAdd(m + 1, d * c);
      atomicAdd(m + 2, d * s);
    } else if (MODE == MK_POL_QU_PRECON) {
      atomicAdd(m + 0, d * c * c);
      atomicAdd(m + 1, d * s * s);
      atomicAdd(m + 2, d * s * c);
    } else {
      atomicAdd(m + 0, d);
      atomicAdd(m + 1, d * c);
      atomicAdd(m + 2, d * s);
      atomicAdd(m + 3, d * c * c);
      atomicAdd(m + 4, d * s * c);
      atomicAdd(m + 5, d * s * s);
    }
  }
}

// ---------------------------------------------------------------------------
// Per-pixel blocks.

// Cyclic Jacobi on a symmetric N x N block (N = 2 or 3): a is diagonalised
// in place (eigenvalues on the diagonal), v collects the rotations (columns
// are eigenvectors).
template <int N>
__device__ inline void sym_eig(double (&a)[3][3], double (&v)[3][3]) {
  for (int r = 0; r < 3; r++)
    for (int q = 0; q < 3; q++) v[r][q] = r == q ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 16; sweep++) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < N; p++) {
      diag += a[p][p] * a[p][p];
      for (int q = p + 1; q < N; q++) off += a[p][q] * a[p][q];
    }
    if (off <= 1e-36 * diag || off == 0.0) break;
    for (int p = 0; p < N - 1; p++) {
      for (int q = p + 1; q < N; q++) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        const double t = fabs(theta) > 1e150 ? 0.5 / theta
                                              : (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        a[p][p] -= t * apq;
        a[q][q] += t * apq;
        a[p][q] = a[q][p] = 0.0;
        for (int r = 0; r < N; r++) {
          if (r != p && r != q) {
            const double arp = a[r][p], arq = a[r][q];
            a[r][p] = a[p][r] = c * arp - s * arq;
            a[r][q] = a[q][r] = s * arp + c * arq;
          }
          const double vrp = v[r][p], vrq = v[r][q];
          v[r][p] = c * vrp - s * vrq;
          v[r][q] = s * vrp + c * vrq;
        }
      }
    }
  }
}

// Pseudo-inverse of a symmetric block with numpy's rcond rule.
template <int N>
__device__ inline void sym_pinv(double (&a)[3][3], double rcond) {
  double v[3][3];
  sym_eig<N>(a, v);
  double lam[3], smax = 0.0;
  for (int k = 0; k < N; k++) {
    lam[k] = a[k][k];
    smax = fmax(smax, fabs(lam[k]));
  }
  const double cut = rcond * smax;
  double inv[3];
  for (int k = 0; k < N; k++) inv[k] = fabs(lam[k]) > cut ? 1.0 / lam[k] : 0.0;
  for (int r = 0; r < N; r++)
    for (int q = 0; q < N; q++) {
      double acc = 0.0;
      for (int k = 0; k < N; k++) acc += v[r][k] * inv[k] * v[q][k];
      a[r][q] = acc;
    }
}

template <typename T>
__device__ inline double det_iqu(const T* m) {
  const double i = m[0], q = m[1], u = m[2], qq = m[3], qu = m[4], uu = m[5];
  return i * (qq * uu - qu * qu) - q * (q * uu - qu * u) + u * (q * qu - qq * u);
}

// Order-preserving map of a double onto an unsigned 64-bit key (larger
// double -> larger key), so the map-wide maximum is one integer atomicMax.
__device__ inline unsigned long long dkey(double d) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(d);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double dunkey(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

__global__ void key_init_kernel(unsigned long long* key) {
  if (blockIdx.x == 0 && threadIdx.x == 0) key[0] = 0ull;
}

template <typename T>
__global__ void __launch_bounds__(256) det_max_kernel(const T* __restrict__ map, int64_t npix,
                                                      unsigned long long* key) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double mx = -INFINITY;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += stride)
    mx = fmax(mx, det_iqu(map + 6 * p));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0 && mx > -INFINITY) atomicMax(key, dkey(mx));
}

template <typename T>
__global__ void __launch_bounds__(256) invert_iqu_kernel(T* __restrict__ map, int64_t npix, double thresh,
                                                         const unsigned long long* __restrict__ key) {
  const double lim = thresh * dunkey(key[0]);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += stride) {
    T* m = map + 6 * p;
    const double det = det_iqu(m);
    if (det < lim) {
      const double i0 = m[0];
      m[0] = i0 > 0 ? (T)(1.0 / i0) : T(0);
      for (int k = 1; k < 6; k++) m[k] = T(0);
      continue;
    }
    double a[3][3] = {{m[0], m[1], m[2]}, {m[1], m[3], m[4]}, {m[2], m[4], m[5]}};
    sym_pinv<3>(a, thresh);
    m[0] = (T)a[0][0];
    m[1] = (T)a[0][1];
    m[2] = (T)a[0][2];
    m[3] = (T)a[1][1];
    m[4] = (T)a[1][2];
    m[5] = (T)a[2][2];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) invert_qu_kernel(T* __restrict__ map, int64_t npix, double thresh) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += stride) {
    T* m = map + 3 * p;
    double a[3][3] = {{m[0], m[2], 0.0}, {m[2], m[1], 0.0}, {0.0, 0.0, 0.0}};
    sym_pinv<2>(a, thresh);
    m[0] = (T)a[0][0];
    m[1] = (T)a[1][1];
    m[2] = (T)a[0][1];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) precon_mul_kernel(T* __restrict__ out, const T* __restrict__ pre,
                                                         const T* __restrict__ x, int64_t npix, int iqu) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += stride) {
    if (iqu) {
      const T* a = pre + 6 * p;
      const T x0 = x[3 * p], x1 = x[3 * p + 1], x2 = x[3 * p + 2];
      out[3 * p] = a[0] * x0 + a[1] * x1 + a[2] * x2;
      out[3 * p + 1] = a[1] * x0 + a[3] * x1 + a[4] * x2;
      out[3 * p + 2] = a[2] * x0 + a[4] * x1 + a[5] * x2;
    } else {
      const T* a = pre + 3 * p;
      const T x0 = x[2 * p], x1 = x[2 * p + 1];
      out[2 * p] = a[0] * x0 + a[2] * x1;
      out[2 * p + 1] = a[2] * x0 + a[1] * x1;
    }
  }
}

template <typename T>
int polmap2tod_impl(T* dat, const T* map, const T* twogamma, int64_t n, const int32_t* ipix, int64_t npix,
                    int npol, int do_add, int64_t* nbad, void* stream) {
  MK_REQUIRE(n >= 0 && npix >= 0, MK_ERR_ARG, "negative size");
  MK_REQUIRE(npol == 2 || npol == 3, MK_ERR_ARG, "polmap2tod: npol must be 2 (QU) or 3 (IQU), got %d", npol);
  if (n == 0) return MK_OK;
  MK_REQUIRE(dat && map && twogamma && ipix, MK_ERR_ARG, "null pointer");
  const unsigned g = mk::grid_for(n, 256, 256 * 32);
  hipStream_t st = mk::as_stream(stream);
  if (npol == 3)
    polmap2tod_kernel<T, 3><<<g, 256, 0, st>>>(dat, map, twogamma, n, ipix, npix, do_add, nbad);
  else
    polmap2tod_kernel<T, 2><<<g, 256, 0, st>>>(dat, map, twogamma, n, ipix, npix, do_add, nbad);
  return mk::check_launch("polmap2tod_kernel");
}

template <typename T>
int tod2polmap_impl(T* map, const T* dat, const T* twogamma, int64_t n, const int32_t* ipix, int64_t npix,
                    int mode, int64_t* nbad, void* stream) {
  MK_REQUIRE(n >= 0 && npix >= 0, MK_ERR_ARG, "negative size");
  MK_REQUIRE(mode >= MK_POL_QU && mode <= MK_POL_IQU_PRECON, MK_ERR_ARG, "tod2polmap: unknown mode %d", mode);
  if (n == 0) return MK_OK;
  MK_REQUIRE(dat && map && twogamma && ipix, MK_ERR_ARG, "null pointer");
  const unsigned g = mk::grid_for(n, 256, 256 * 32);
  hipStream_t st = mk::as_stream(stream);
  switch (mode) {
    case MK_POL_QU: tod2polmap_kernel<T, MK_POL_QU><<<g, 256, 0, st>>>(map, dat, twogamma, n, ipix, npix, nbad); break;
    case MK_POL_IQU: tod2polmap_kernel<T, MK_POL_IQU><<<g, 256, 0, st>>>(map, dat, twogamma, n, ipix, npix, nbad); break;
    case MK_POL_QU_PRECON:
      tod2polmap_kernel<T, MK_POL_QU_PRECON><<<g, 256, 0, st>>>(map, dat, twogamma, n, ipix, npix, nbad);
      break;
    default:
      tod2polmap_kernel<T, MK_POL_IQU_PRECON><<<g, 256, 0, st>>>(map, dat, twogamma, n, ipix, npix, nbad);
  }
  return mk::check_launch("tod2polmap_kernel");
}

template <typename T>
int invert_impl(T* map, int64_t npix, int mode, double thresh, void* work, int64_t work_bytes, void* stream) {
  MK_REQUIRE(npix >= 0, MK_ERR_ARG, "negative size");
  MK_REQUIRE(mode == MK_POL_QU_PRECON || mode == MK_POL_IQU_PRECON, MK_ERR_ARG,
             "polmap_invert: mode must be QU_PRECON or IQU_PRECON, got %d", mode);
  if (npix == 0) return MK_OK;
  MK_REQUIRE(map, MK_ERR_ARG, "null pointer");
  const unsigned g = mk::grid_for(npix, 256, 256 * 16);
  hipStream_t st = mk::as_stream(stream);
  if (mode == MK_POL_QU_PRECON) {
    invert_qu_kernel<T><<<g, 256, 0, st>>>(map, npix, thresh);
    return mk::check_launch("invert_qu_kernel");
  }
  MK_REQUIRE(work && work_bytes >= 8, MK_ERR_ARG, "polmap_invert: IQU_PRECON needs an 8-byte device work buffer");
  unsigned long long* key = (unsigned long long*)work;
  key_init_kernel<<<1, 64, 0, st>>>(key);
  det_max_kernel<T><<<g, 256, 0, st>>>(map, npix, key);
  invert_iqu_kernel<T><<<g, 256, 0, st>>>(map, npix, thresh, key);
  return mk::check_launch("invert_iqu_kernel");
}

template <typename T>
int precon_mul_impl(T* out, const T* pre, const T* x, int64_t npix, int mode, void* stream) {
  MK_REQUIRE(npix >= 0, MK_ERR_ARG, "negative size");
  MK_REQUIRE(mode == MK_POL_QU_PRECON || mode == MK_POL_IQU_PRECON, MK_ERR_ARG,
             "polmap_precon_mul: mode must be QU_PRECON or IQU_PRECON, got %d", mode);
  if (npix == 0) return MK_OK;
  MK_REQUIRE(out && pre && x, MK_ERR_ARG, "null pointer");
  precon_mul_kernel<T><<<mk::grid_for(npix, 256, 256 * 16), 256, 0, mk::as_stream(stream)>>>(
      out, pre, x, npix, mode == MK_POL_IQU_PRECON);
  return mk::check_launch("precon_mul_kernel");
}

}  // namespace

extern "C" int mk_polmap2tod_f32(float* dat, const float* map, const float* twogamma, int64_t nsamp,
                                 const int32_t* ipix, int64_t npix, int npol, int do_add, int64_t* dev_nbad,
                                 void* stream) {
  return polmap2tod_impl<float>(dat, map, twogamma, nsamp, ipix, npix, npol, do_add, dev_nbad, stream);
}
extern "C" int mk_polmap2tod_f64(double* dat, const double* map, const double* twogamma, int64_t nsamp,
                                 const int32_t* ipix, int64_t npix, int npol, int do_add, int64_t* dev_nbad,
                                 void* stream) {
  return polmap2tod_impl<double>(dat, map, twogamma, nsamp, ipix, npix, npol, do_add, dev_nbad, stream);
}
extern "C" int mk_tod2polmap_atomic_f32(float* map, const float* dat, const float* twogamma, int64_t nsamp,
                                        const int32_t* ipix, int64_t npix, int mode, int64_t* dev_nbad,
                                        void* stream) {
  return tod2polmap_impl<float>(map, dat, twogamma, nsamp, ipix, npix, mode, dev_nbad, stream);
}
extern "C" int mk_tod2polmap_atomic_f64(double* map, const double* dat, const double* twogamma, int64_t nsamp,
                                        const int32_t* ipix, int64_t npix, int mode, int64_t* dev_nbad,
                                        void* stream) {
  return tod2polmap_impl<double>(map, dat, twogamma, nsamp, ipix, npix, mode, dev_nbad, stream);
}
extern "C" int mk_polmap_invert_f32(float* map, int64_t npix, int mode, double thresh, void* work,
                                    int64_t work_bytes, void* stream) {
  return invert_impl<float>(map, npix, mode, thresh, work, work_bytes, stream);
}
extern "C" int mk_polmap_invert_f64(double* map, int64_t npix, int mode, double thresh, void* work,
                                    int64_t work_bytes, void* stream) {
  return invert_impl<double>(map, npix, mode, thresh, work, work_bytes, stream);
}
extern "C" int mk_polmap_precon_mul_f32(float* out, const float* pre, const float* x, int64_t npix, int mode,
                                        void* stream) {
  return precon_mul_impl<float>(out, pre, x, npix, mode, stream);
}
extern "C" int mk_polmap_precon_mul_f64(double* out, const double* pre, const double* x, int64_t npix, int mode,
                                        void* stream) {
  return precon_mul_impl<double>(out, pre, x, npix, mode, stream);
}

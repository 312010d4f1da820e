// Map-space vector operations of PCG with FP64 scalars kept on the device.
//
// The reference runs these in numpy on FP64 maps (minkasi/mapmaking/pcg.py:
// 143-170 via Mapset/SkyMap dot, axpy, copy and __mul__).  Here maps are
// FP32, every inner product accumulates in FP64, alpha/beta are formed on the
// device so an iteration needs no host round trip, and each reduction is a
// fixed two-stage tree (kGroups partials, then one workgroup), so results are
// bitwise reproducible.  All kernels are HBM-bound streaming passes.
#include "common.h"

namespace {

constexpr int kGroups = 1024;  // stage-1 workgroups (fixed: the order of summation depends only on n)
constexpr int kBlock = 256;

enum Mode { kDot = 0, kAlpha, kXR, kXRZJ, kBeta, kInit };

struct Args {
  float* x; float* r; float* z; float* p;
  double* x64;  // when set, x accumulates here in FP64 and x holds its FP32 rounding
  const float* Ap; const float* w;
  const float* a; const float* b;
  const double* a64; const double* b64;
  const double* sc;
  int64_t n;
};

__device__ inline double block_sum(double v, double* sh) {
  // fixed-shape tree over the 256 threads
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double out = sh[0];
  __syncthreads();
  return out;
}

template <int M>
__global__ void __launch_bounds__(kBlock) stage1(Args A, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  double s0 = 0.0, s1 = 0.0;
  const int64_t stride = (int64_t)kGroups * kBlock;
  double alpha = 0.0;
  if (M == kXR || M == kXRZJ) alpha = A.sc[MK_SC_ALPHA];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < A.n; i += stride) {
    if (M == kDot) {
      s0 += (double)A.a[i] * (double)A.b[i];
    } else if (M == kAlpha) {
      s0 += (double)A.p[i] * (double)A.Ap[i];
    } else if (M == kXR || M == kXRZJ) {
      float pi = A.p[i], api = A.Ap[i];
      if (A.x64) {
        const double xi = A.x64[i] + alpha * (double)pi;
        A.x64[i] = xi;
        A.x[i] = (float)xi;
      } else {
        A.x[i] = (float)((double)A.x[i] + alpha * (double)pi);
      }
      float ri = (float)((double)A.r[i] - alpha * (double)api);
      A.r[i] = ri;
      s1 += (double)ri * (double)ri;
      if (M == kXRZJ) {
        float zi = A.w[i] * ri;
        A.z[i] = zi;
        s0 += (double)ri * (double)zi;
      }
    } else if (M == kBeta) {
      s0 += (double)A.r[i] * (double)A.z[i];
    } else if (M == kInit) {
      float ri = A.r[i], zi = A.z[i];
      A.p[i] = zi;
      s0 += (double)ri * (double)zi;
      s1 += (double)ri * (double)ri;
    }
  }
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kGroups + blockIdx.x] = s1;
  }
}

__global__ void __launch_bounds__(kBlock) stage1_f64(const double* __restrict__ a, const double* __restrict__ b,
                                                     int64_t n, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  double s0 = 0.0;
  const int64_t stride = (int64_t)kGroups * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) s0 += a[i] * b[i];
  s0 = block_sum(s0, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kGroups + blockIdx.x] = 0.0;
  }
}

template <int M>
__global__ void __launch_bounds__(kBlock) stage2(const double* __restrict__ part, double* __restrict__ sc) {
  __shared__ double sh[kBlock];
  double s0 = 0.0, s1 = 0.0;
  for (int k = 0; k < kGroups / kBlock; k++) {
    s0 += part[threadIdx.x * (kGroups / kBlock) + k];
    s1 += part[kGroups + threadIdx.x * (kGroups / kBlock) + k];
  }
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  if (threadIdx.x != 0) return;
  if (M == kDot) {
    sc[MK_SC_DOT] = s0;
  } else if (M == kAlpha) {
    sc[MK_SC_PAP] = s0;
    double zr = sc[MK_SC_ZR];
    bool ok = (s0 > 0.0) && isfinite(s0) && isfinite(zr);
    sc[MK_SC_ALPHA] = ok ? zr / s0 : 0.0;
    if (!ok) sc[MK_SC_FLAG] = 1.0;
  } else if (M == kXR) {
    sc[MK_SC_RR] = s1;
    if (!isfinite(s1)) sc[MK_SC_FLAG] = 2.0;
  } else if (M == kXRZJ || M == kBeta) {
    if (M == kXRZJ) {
      sc[MK_SC_RR] = s1;
      if (!isfinite(s1)) sc[MK_SC_FLAG] = 2.0;
    }
    double zr = sc[MK_SC_ZR];
    bool ok = zr != 0.0 && isfinite(s0);
    sc[MK_SC_BETA] = ok ? s0 / zr : 0.0;
    sc[MK_SC_ZR] = s0;
    if (!isfinite(s0)) sc[MK_SC_FLAG] = 3.0;
  } else if (M == kInit) {
    sc[MK_SC_ZR] = s0;
    sc[MK_SC_RR] = s1;
    sc[MK_SC_FLAG] = (isfinite(s0) && isfinite(s1)) ? 0.0 : 3.0;
  }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) axpy_kernel(T* __restrict__ y, const T* __restrict__ x, double a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = (T)((double)y[i] + a * (double)x[i]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) mul_kernel(T* __restrict__ o, const T* __restrict__ a,
                                                     const T* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) o[i] = a[i] * b[i];
}

// FP32 product with 16-byte loads and stores (all three arrays 16-byte
// aligned); elements past the last full group of 4 are done one by one.
__global__ void __launch_bounds__(kBlock) mul4_kernel(float* __restrict__ o, const float* __restrict__ a,
                                                      const float* __restrict__ b, int64_t n) {
  const int64_t n4 = n >> 2;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const float4* a4 = reinterpret_cast<const float4*>(a);
  const float4* b4 = reinterpret_cast<const float4*>(b);
  float4* o4 = reinterpret_cast<float4*>(o);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const float4 x = a4[i], y = b4[i];
    o4[i] = make_float4(x.x * y.x, x.y * y.y, x.z * y.z, x.w * y.w);
  }
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) o[i] = a[i] * b[i];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) addmul_kernel(T* __restrict__ y, const T* __restrict__ a,
                                                        const T* __restrict__ b, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const T ab = a[i] * b[i];
    y[i] = y[i] + ab;
  }
}

template <typename T>
int axpy_impl(T* y, const T* x, double a, int64_t n, void* stream) {
  MK_REQUIRE(n >= 0 && (n == 0 || (x && y)), MK_ERR_ARG, "bad arguments");
  if (n == 0) return MK_OK;
  axpy_kernel<T><<<mk::grid_for(n, kBlock), kBlock, 0, mk::as_stream(stream)>>>(y, x, a, n);
  return mk::check_launch("axpy_kernel");
}

template <typename T>
int mul_impl(T* out, const T* a, const T* b, int64_t n, void* stream) {
  MK_REQUIRE(n >= 0 && (n == 0 || (out && a && b)), MK_ERR_ARG, "bad arguments");
  if (n == 0) return MK_OK;
  if constexpr (sizeof(T) == 4) {
    if ((((uintptr_t)out | (uintptr_t)a | (uintptr_t)b) & 15) == 0) {
      mul4_kernel<<<mk::grid_for(mk::cdiv(n, 4), kBlock, 8192), kBlock, 0, mk::as_stream(stream)>>>(out, a, b, n);
      return mk::check_launch("mul4_kernel");
    }
  }
  mul_kernel<T><<<mk::grid_for(n, kBlock), kBlock, 0, mk::as_stream(stream)>>>(out, a, b, n);
  return mk::check_launch("mul_kernel");
}

template <typename T>
int addmul_impl(T* y, const T* a, const T* b, int64_t n, void* stream) {
  MK_REQUIRE(n >= 0 && (n == 0 || (y && a && b)), MK_ERR_ARG, "bad arguments");
  if (n == 0) return MK_OK;
  addmul_kernel<T><<<mk::grid_for(n, kBlock), kBlock, 0, mk::as_stream(stream)>>>(y, a, b, n);
  return mk::check_launch("addmul_kernel");
}

__global__ void __launch_bounds__(kBlock) p_kernel(float* __restrict__ p, const float* __restrict__ z,
                                                   const double* __restrict__ sc, int64_t n) {
  const double beta = sc[MK_SC_BETA];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = (float)((double)z[i] + beta * (double)p[i]);
}

template <int M>
int run(const Args& A, double* sc, void* work, void* stream_v) {
  MK_REQUIRE(A.n >= 0, MK_ERR_ARG, "n < 0");
  MK_REQUIRE(sc && work, MK_ERR_ARG, "null scalar block or workspace");
  hipStream_t s = mk::as_stream(stream_v);
  double* part = (double*)work;
  stage1<M><<<kGroups, kBlock, 0, s>>>(A, part);
  int st = mk::check_launch("pcg stage1");
  if (st) return st;
  stage2<M><<<1, kBlock, 0, s>>>(part, sc);
  return mk::check_launch("pcg stage2");
}

}  // namespace

extern "C" int64_t mk_reduce_workspace_bytes(void) { return 2 * kGroups * 8 + 256; }

extern "C" int mk_dot_f32(const float* a, const float* b, int64_t n, double* sc, void* work, void* stream) {
  MK_REQUIRE(n == 0 || (a && b), MK_ERR_ARG, "null input");
  Args A{};
  A.a = a; A.b = b; A.n = n;
  return run<kDot>(A, sc, work, stream);
}

extern "C" int mk_dot_f64(const double* a, const double* b, int64_t n, double* sc, void* work, void* stream_v) {
  MK_REQUIRE(n >= 0 && sc && work, MK_ERR_ARG, "bad arguments");
  MK_REQUIRE(n == 0 || (a && b), MK_ERR_ARG, "null input");
  hipStream_t s = mk::as_stream(stream_v);
  stage1_f64<<<kGroups, kBlock, 0, s>>>(a, b, n, (double*)work);
  int st = mk::check_launch("dot_f64 stage1");
  if (st) return st;
  stage2<kDot><<<1, kBlock, 0, s>>>((double*)work, sc);
  return mk::check_launch("dot_f64 stage2");
}

extern "C" int mk_axpy_f32(float* y, const float* x, double a, int64_t n, void* stream) {
  return axpy_impl<float>(y, x, a, n, stream);
}
extern "C" int mk_axpy_f64(double* y, const double* x, double a, int64_t n, void* stream) {
  return axpy_impl<double>(y, x, a, n, stream);
}

extern "C" int mk_mul_f32(float* out, const float* a, const float* b, int64_t n, void* stream) {
  return mul_impl<float>(out, a, b, n, stream);
}
extern "C" int mk_mul_f64(double* out, const double* a, const double* b, int64_t n, void* stream) {
  return mul_impl<double>(out, a, b, n, stream);
}

extern "C" int mk_addmul_f32(float* y, const float* a, const float* b, int64_t n, void* stream) {
  return addmul_impl<float>(y, a, b, n, stream);
}
extern "C" int mk_addmul_f64(double* y, const double* a, const double* b, int64_t n, void* stream) {
  return addmul_impl<double>(y, a, b, n, stream);
}

extern "C" int mk_pcg_alpha(const float* p, const float* Ap, int64_t n, double* sc, void* work, void* stream) {
  MK_REQUIRE(n == 0 || (p && Ap), MK_ERR_ARG, "null input");
  Args A{};
  A.p = const_cast<float*>(p); A.Ap = Ap; A.n = n;
  return run<kAlpha>(A, sc, work, stream);
}

extern "C" int mk_pcg_xr(float* x, float* r, const float* p, const float* Ap, int64_t n, double* sc, void* work,
                         void* stream) {
  MK_REQUIRE(n == 0 || (x && r && p && Ap), MK_ERR_ARG, "null input");
  Args A{};
  A.x = x; A.r = r; A.p = const_cast<float*>(p); A.Ap = Ap; A.sc = sc; A.n = n;
  return run<kXR>(A, sc, work, stream);
}

extern "C" int mk_pcg_xrz_jacobi(float* x, float* r, float* z, const float* p, const float* Ap, const float* w,
                                 int64_t n, double* sc, void* work, void* stream) {
  MK_REQUIRE(n == 0 || (x && r && z && p && Ap && w), MK_ERR_ARG, "null input");
  Args A{};
  A.x = x; A.r = r; A.z = z; A.p = const_cast<float*>(p); A.Ap = Ap; A.w = w; A.sc = sc; A.n = n;
  return run<kXRZJ>(A, sc, work, stream);
}

extern "C" int mk_pcg_beta(const float* r, const float* z, int64_t n, double* sc, void* work, void* stream) {
  MK_REQUIRE(n == 0 || (r && z), MK_ERR_ARG, "null input");
  Args A{};
  A.r = const_cast<float*>(r); A.z = const_cast<float*>(z); A.n = n;
  return run<kBeta>(A, sc, work, stream);
}

extern "C" int mk_pcg_p(float* p, const float* z, int64_t n, const double* sc, void* stream) {
  MK_REQUIRE(n >= 0 && (n == 0 || (p && z)) && sc, MK_ERR_ARG, "bad arguments");
  if (n == 0) return MK_OK;
  p_kernel<<<mk::grid_for(n, kBlock), kBlock, 0, mk::as_stream(stream)>>>(p, z, sc, n);
  return mk::check_launch("p_kernel");
}

extern "C" int mk_pcg_init(const float* r, const float* z, float* p, int64_t n, double* sc, void* work,
                           void* stream) {
  MK_REQUIRE(n == 0 || (r && z && p), MK_ERR_ARG, "null input");
  Args A{};
  A.r = const_cast<float*>(r); A.z = const_cast<float*>(z); A.p = p; A.n = n;
  return run<kInit>(A, sc, work, stream);
}

namespace {
// r = fl32(b - Ax) from FP64 b and Ax; rr = r.r (FP64) into the scalar block.
__global__ void __launch_bounds__(kBlock) set_r_stage1(float* __restrict__ r, const double* __restrict__ b,
                                                       const double* __restrict__ ax, int64_t n,
                                                       double* __restrict__ part) {
  __shared__ double sh[kBlock];
  double s1 = 0.0;
  const int64_t stride = (int64_t)kGroups * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const float ri = (float)(b[i] - ax[i]);
    r[i] = ri;
    s1 += (double)ri * (double)ri;
  }
  s1 = block_sum(s1, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = 0.0;
    part[kGroups + blockIdx.x] = s1;
  }
}
}  // namespace

extern "C" int mk_pcg_set_r(float* r, const double* b, const double* ax, int64_t n, double* sc, void* work,
                            void* stream_v) {
  MK_REQUIRE(n >= 0 && sc && work, MK_ERR_ARG, "bad arguments");
  MK_REQUIRE(n == 0 || (r && b && ax), MK_ERR_ARG, "null input");
  hipStream_t s = mk::as_stream(stream_v);
  set_r_stage1<<<kGroups, kBlock, 0, s>>>(r, b, ax, n, (double*)work);
  int st = mk::check_launch("set_r stage1");
  if (st) return st;
  stage2<kXR><<<1, kBlock, 0, s>>>((double*)work, sc);
  return mk::check_launch("set_r stage2");
}

extern "C" int mk_pcg_xr_x64(double* x64, float* x, float* r, const float* p, const float* Ap, int64_t n,
                             double* sc, void* work, void* stream) {
  MK_REQUIRE(n == 0 || (x64 && x && r && p && Ap), MK_ERR_ARG, "null input");
  Args A{};
  A.x64 = x64; A.x = x; A.r = r; A.p = const_cast<float*>(p); A.Ap = Ap; A.sc = sc; A.n = n;
  return run<kXR>(A, sc, work, stream);
}

extern "C" int mk_pcg_xrz_jacobi_x64(double* x64, float* x, float* r, float* z, const float* p, const float* Ap,
                                     const float* w, int64_t n, double* sc, void* work, void* stream) {
  MK_REQUIRE(n == 0 || (x64 && x && r && z && p && Ap && w), MK_ERR_ARG, "null input");
  Args A{};
  A.x64 = x64; A.x = x; A.r = r; A.z = z; A.p = const_cast<float*>(p); A.Ap = Ap; A.w = w; A.sc = sc; A.n = n;
  return run<kXRZJ>(A, sc, work, stream);
}

namespace {
__global__ void __launch_bounds__(kBlock) project_kernel(float* __restrict__ x, const float* __restrict__ u,
                                                         const double* __restrict__ sc, double inv_uu, int64_t n) {
  const double c = sc[MK_SC_DOT] * inv_uu;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x[i] = (float)((double)x[i] - c * (double)u[i]);
}
}  // namespace

extern "C" int mk_project_out_f32(float* x, const float* u, int64_t n, double inv_uu, double* sc, void* work,
                                  void* stream) {
  MK_REQUIRE(n >= 0 && (n == 0 || (x && u)) && sc && work, MK_ERR_ARG, "bad arguments");
  if (n == 0) return MK_OK;
  Args A{};
  A.a = x; A.b = u; A.n = n;
  int st = run<kDot>(A, sc, work, stream);
  if (st) return st;
  project_kernel<<<mk::grid_for(n, kBlock), kBlock, 0, mk::as_stream(stream)>>>(x, u, sc, inv_uu, n);
  return mk::check_launch("project_kernel");
}

// Spectrum smoothing for NoiseSmoothedSVD setup (minkasi/mapmaking/noise.py:
// 674-678 with smooth_many_vecs, minkasi/tools/smooth.py:49-78).
//
// smooth_many_vecs computes DCT-I(DCT-I(P) * DCT-I(g)) / (2(n-1)) with g the
// r2r-normalised Gaussian.  Because DCT-I o DCT-I = 2(n-1) I, that equals the
// circular (period 2(n-1)) convolution of the even extensions of P and g,
// S_k = sum_{t in (-m, m]} g[|t|] P_ext(k - t), m = n-1.  The Gaussian
// underflows far from its centre, so the sum is evaluated directly over the
// taps the host passes (|t| up to ~13 sigma, or the whole period when the
// row is shorter), in FP64, with the input window staged in LDS.  Optionally
// the input is squared first (P = DCT(v.d)^2) and the output inverted with
// bin 0 set to zero (mywt = 1/S, mywt[:, 0] = 0).
#include "common.h"

namespace {

constexpr int kChunk = 1024;

__device__ inline int64_t fold(int64_t i, int64_t m) {
  // even periodic extension with period N = 2m: index into [0, m]
  int64_t N = 2 * m;
  int64_t ii = i % N;
  if (ii < 0) ii += N;
  return ii > m ? N - ii : ii;
}

__global__ void __launch_bounds__(256) smooth_kernel(const double* __restrict__ x, int64_t n, int64_t chunks,
                                                     const double* __restrict__ g, int t_lo, int t_hi, int flags,
                                                     double* __restrict__ out, float* __restrict__ out32) {
  extern __shared__ double win[];  // [kChunk + t_hi - t_lo]
  const int64_t row = blockIdx.x / chunks;
  const int64_t k0 = (blockIdx.x % chunks) * kChunk;
  const int64_t m = n - 1;
  const double* xr = x + row * n;
  const int wlen = kChunk + t_hi - t_lo;
  // win[w] = P_ext(k0 - t_hi + w)
  for (int w = threadIdx.x; w < wlen; w += blockDim.x) {
    double v = xr[fold(k0 - t_hi + w, m)];
    win[w] = (flags & 1) ? v * v : v;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kChunk; c += blockDim.x) {
    int64_t k = k0 + c;
    if (k >= n) break;
    // S_k = sum_t g[|t|] P(k - t): P(k - t) = win[c + t_hi - t]
    double s = 0.0;
    for (int t = t_lo; t <= t_hi; t++) s = fma(g[t < 0 ? -t : t], win[c + t_hi - t], s);
    if (flags & 2) s = (k == 0) ? 0.0 : 1.0 / s;
    out[row * n + k] = s;
    if (out32) out32[row * n + k] = (float)s;
  }
}

}  // namespace

extern "C" int mk_smooth_rows_f64(const double* x, double* out, float* out32, int64_t nrow, int64_t n,
                                  const double* g, int32_t t_lo, int32_t t_hi, int32_t flags, void* stream) {
  MK_REQUIRE(nrow >= 0 && n >= 2, MK_ERR_ARG, "need n >= 2");
  if (nrow == 0) return MK_OK;
  MK_REQUIRE(x && out && g, MK_ERR_ARG, "null argument");
  MK_REQUIRE(t_lo <= 0 && t_hi >= 0 && t_hi - t_lo < 2 * (n - 1) + 1 && -t_lo <= n - 1 && t_hi <= n - 1,
             MK_ERR_ARG, "taps must satisfy -(n-1) <= t_lo <= 0 <= t_hi <= n-1");
  MK_REQUIRE(t_hi - t_lo <= 16384, MK_ERR_ARG, "too many taps");
  MK_REQUIRE(x != out, MK_ERR_ARG, "smoothing cannot run in place");
  int64_t chunks = mk::cdiv(n, kChunk);
  size_t lds = (size_t)(kChunk + t_hi - t_lo) * sizeof(double);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(smooth_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            144 * 1024);
  (void)hipGetLastError();
  smooth_kernel<<<(unsigned)(nrow * chunks), 256, lds, mk::as_stream(stream)>>>(x, n, chunks, g, t_lo, t_hi, flags,
                                                                               out, out32);
  return mk::check_launch("smooth_kernel");
}

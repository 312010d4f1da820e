// Learned map-space preconditioner: linear dilated 3x3 convolutions.
//
// The preconditioner is M = S (alpha I + K^T K) S with S = diag(sqrt(1/hits))
// (zero on unhit pixels) and K a chain of linear 3x3 convolutions (no
// activations), so M is symmetric positive semidefinite by construction and
// positive definite on hit pixels when alpha > 0.  Layer l has dilation
// dil_l: tap d = 3*di + dj reads offset (dil*(di-1), dil*(dj-1)).
//
//   conv3x3: y[b,o,p] (+)= sum_{c,d} w[o,c,d] x[b,c,p+off_d]   (zero padding)
//            The adjoint layer is the same kernel with the weights
//            transposed in (o,c) and flipped in d (mk_conv3x3_adjoint_weights).
//   wgrad:   g[o,c,d] = sum_{b,p} a[b,o,p] x[b,c,p+off_d]
//            the weight gradient of a forward layer (a = output gradient,
//            x = layer input) and of an adjoint layer (a = layer input,
//            x = output gradient).
//
// Tiling is polyphase: a workgroup owns the 16x16 lattice of pixels
// (i0 + dil*a, j0 + dil*c) inside a 16*dil square super-tile, so a dilated
// tap is a +-1 lattice step and the LDS tile is 18x18 per channel whatever
// the dilation.  Activations are planar FP32 [B][C][H][W]; arithmetic is FP32
// with the weights read through uniform (scalar) loads.  wgrad reduces fixed
// per-workgroup partials in a fixed order, so every result is deterministic.
#include "common.h"

namespace {

constexpr int TS = 16;            // lattice tile side
constexpr int HS = TS + 2;        // with the +-1 halo
constexpr int XP = HS + 1;        // LDS row pitch of an input tile
constexpr int XCH = HS * XP;      // LDS channel stride of an input tile (342)
constexpr int AP = TS + 1;        // LDS row pitch of an output-gradient tile
constexpr int ACH = TS * AP + 1;  // LDS channel stride of an output-gradient tile (273)
constexpr int MAXC = 16;
constexpr int kWgradGroups = 2048;

struct Tiling {
  int H, W, dil, ntj, per_img;  // per_img = ntiles_i * ntiles_j * dil^2
  __device__ void locate(int bid, int& b, int& i0, int& j0) const {
    b = bid / per_img;
    int rem = bid - b * per_img;
    int d2 = dil * dil;
    int T = rem / d2, ph = rem - T * d2;
    int ti = T / ntj, tj = T - ti * ntj;
    i0 = ti * TS * dil + ph / dil;
    j0 = tj * TS * dil + ph % dil;
  }
};

inline Tiling make_tiling(int H, int W, int dil) {
  Tiling t;
  t.H = H; t.W = W; t.dil = dil;
  int nti = (int)mk::cdiv(H, (int64_t)TS * dil);
  t.ntj = (int)mk::cdiv(W, (int64_t)TS * dil);
  t.per_img = nti * t.ntj * dil * dil;
  return t;
}

// Stages channels [0, C) of the lattice tile with halo into LDS (zero outside
// the image).
__device__ inline void stage_halo(float* xs, const float* __restrict__ xb, int C, const Tiling& t, int i0,
                                  int j0) {
  const int64_t plane = (int64_t)t.H * t.W;
  for (int k = threadIdx.x; k < C * HS * HS; k += 256) {
    int c = k / (HS * HS), rq = k - c * (HS * HS);
    int r = rq / HS, q = rq - r * HS;
    int gi = i0 + t.dil * (r - 1), gj = j0 + t.dil * (q - 1);
    float v = 0.f;
    if (gi >= 0 && gi < t.H && gj >= 0 && gj < t.W) v = xb[c * plane + (int64_t)gi * t.W + gj];
    xs[c * XCH + r * XP + q] = v;
  }
}

template <int CIN, int COUT>
__global__ void __launch_bounds__(256) conv3x3_kernel(const float* __restrict__ x, const float* __restrict__ w,
                                                      float* __restrict__ y, Tiling t, int accumulate) {
  __shared__ float xs[CIN * XCH];
  int b, i0, j0;
  t.locate(blockIdx.x, b, i0, j0);
  const int64_t plane = (int64_t)t.H * t.W;
  stage_halo(xs, x + (int64_t)b * CIN * plane, CIN, t, i0, j0);
  __syncthreads();
  const int li = threadIdx.x / TS, lj = threadIdx.x % TS;
  float acc[COUT];
#pragma unroll
  for (int o = 0; o < COUT; o++) acc[o] = 0.f;
#pragma unroll 1
  for (int c = 0; c < CIN; c++) {
    float v[9];
#pragma unroll
    for (int d = 0; d < 9; d++) v[d] = xs[c * XCH + (li + d / 3) * XP + lj + d % 3];
#pragma unroll
    for (int o = 0; o < COUT; o++) {
      const float* wo = w + (o * CIN + c) * 9;
#pragma unroll
      for (int d = 0; d < 9; d++) acc[o] = fmaf(wo[d], v[d], acc[o]);
    }
  }
  const int gi = i0 + t.dil * li, gj = j0 + t.dil * lj;
  if (gi < t.H && gj < t.W) {
    float* yb = y + (int64_t)b * COUT * plane + (int64_t)gi * t.W + gj;
#pragma unroll
    for (int o = 0; o < COUT; o++) {
      float v = acc[o];
      if (accumulate) v += yb[o * plane];
      yb[o * plane] = v;
    }
  }
}

// Thread k < CO*CI owns weights (o, c) = (k / CI, k % CI) and keeps their 9
// taps in registers over every tile the workgroup visits; the 3x3 input
// window slides along each lattice row.
template <int CO, int CI>
__global__ void __launch_bounds__(256) wgrad_kernel(const float* __restrict__ a, const float* __restrict__ x,
                                                    Tiling t, int total, float* __restrict__ part) {
  __shared__ float as[CO * ACH];
  __shared__ float xs[CI * XCH];
  const int64_t plane = (int64_t)t.H * t.W;
  const int k = threadIdx.x;
  const bool active = k < CO * CI;
  const int o = active ? k / CI : 0, c = active ? k % CI : 0;
  float g[9];
#pragma unroll
  for (int d = 0; d < 9; d++) g[d] = 0.f;
  for (int bid = blockIdx.x; bid < total; bid += gridDim.x) {
    int b, i0, j0;
    t.locate(bid, b, i0, j0);
    const float* ab = a + (int64_t)b * CO * plane;
    for (int e = threadIdx.x; e < CO * TS * TS; e += 256) {
      int oo = e / (TS * TS), rq = e - oo * (TS * TS);
      int r = rq / TS, q = rq - r * TS;
      int gi = i0 + t.dil * r, gj = j0 + t.dil * q;
      float v = 0.f;
      if (gi < t.H && gj < t.W) v = ab[oo * plane + (int64_t)gi * t.W + gj];
      as[oo * ACH + r * AP + q] = v;
    }
    stage_halo(xs, x + (int64_t)b * CI * plane, CI, t, i0, j0);
    __syncthreads();
    if (active) {
      const float* ao = as + o * ACH;
      const float* xc = xs + c * XCH;
#pragma unroll 1
      for (int r = 0; r < TS; r++) {
        float w0[3], w1[3], w2[3];  // window columns q-1+{0,1,2} (halo coordinates q, q+1, q+2)
#pragma unroll
        for (int di = 0; di < 3; di++) {
          w0[di] = xc[(r + di) * XP + 0];
          w1[di] = xc[(r + di) * XP + 1];
        }
#pragma unroll 4
        for (int q = 0; q < TS; q++) {
#pragma unroll
          for (int di = 0; di < 3; di++) w2[di] = xc[(r + di) * XP + q + 2];
          const float av = ao[r * AP + q];
#pragma unroll
          for (int di = 0; di < 3; di++) {
            g[3 * di + 0] = fmaf(av, w0[di], g[3 * di + 0]);
            g[3 * di + 1] = fmaf(av, w1[di], g[3 * di + 1]);
            g[3 * di + 2] = fmaf(av, w2[di], g[3 * di + 2]);
          }
#pragma unroll
          for (int di = 0; di < 3; di++) {
            w0[di] = w1[di];
            w1[di] = w2[di];
          }
        }
      }
    }
    __syncthreads();
  }
  if (active) {
    float* out = part + (int64_t)blockIdx.x * (CO * CI * 9) + (o * CI + c) * 9;
#pragma unroll
    for (int d = 0; d < 9; d++) out[d] = g[d];
  }
}

__global__ void __launch_bounds__(256) wgrad_reduce_kernel(const float* __restrict__ part, int nparts, int nw,
                                                           float* __restrict__ g, int accumulate) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  double s = 0.0;  // fixed order; FP64 sum of the FP32 partials
  for (int p = 0; p < nparts; p++) s += (double)part[(int64_t)p * nw + k];
  g[k] = accumulate ? (float)((double)g[k] + s) : (float)s;
}

__global__ void adjoint_weights_kernel(const float* __restrict__ w, int Co, int Ci, float* __restrict__ wt) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Co * Ci * 9) return;
  int o = k / (Ci * 9), c = (k / 9) % Ci, d = k % 9;
  wt[(c * Co + o) * 9 + (8 - d)] = w[k];
}

// z = s * (exp(log_alpha) * u + g)   (g may be NULL)
__global__ void __launch_bounds__(256) precon_finish_kernel(float* __restrict__ z, const float* __restrict__ s,
                                                            const float* __restrict__ log_alpha,
                                                            const float* __restrict__ u, const float* __restrict__ g,
                                                            int64_t n) {
  const float alpha = expf(log_alpha[0]);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float v = alpha * u[i];
    if (g) v += g[i];
    z[i] = s[i] * v;
  }
}

constexpr int kLossGroups = 1024;

__device__ inline double block_sum256(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  double out = sh[0];
  __syncthreads();
  return out;
}

// Loss L = sum m * (s*(alpha u + g) - e)^2; q = dL/d(alpha u + g) = 2 m s delta;
// partials of L and of dL/dalpha = sum q u.
__global__ void __launch_bounds__(256) precon_loss_stage1(const float* __restrict__ g, const float* __restrict__ u,
                                                          const float* __restrict__ s, const float* __restrict__ e,
                                                          const float* __restrict__ m,
                                                          const float* __restrict__ log_alpha, int64_t n,
                                                          float* __restrict__ q, double* __restrict__ part) {
  __shared__ double sh[256];
  const float alpha = expf(log_alpha[0]);
  double sl = 0.0, sa = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kLossGroups * 256) {
    float ui = u[i], si = s[i], mi = m[i];
    float del = si * fmaf(alpha, ui, g[i]) - e[i];
    float qi = 2.f * mi * si * del;
    q[i] = qi;
    sl += (double)mi * (double)del * (double)del;
    sa += (double)qi * (double)ui;
  }
  sl = block_sum256(sl, sh);
  sa = block_sum256(sa, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = sl;
    part[kLossGroups + blockIdx.x] = sa;
  }
}

__global__ void __launch_bounds__(256) precon_loss_stage2(const double* __restrict__ part, double* __restrict__ sc) {
  __shared__ double sh[256];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < kLossGroups; i += 256) {
    a += part[i];
    b += part[kLossGroups + i];
  }
  a = block_sum256(a, sh);
  b = block_sum256(b, sh);
  if (threadIdx.x == 0) {
    sc[0] = a;
    sc[1] = b;
  }
}

// Energy-norm loss of one preconditioned step: with z = M r and Az = A z,
// L = (z - e).(Az - r) = ||z - e||_A^2 (A e = r), q = dL/dt = 2 s (Az - r)
// for z = s t, and dL/dalpha = sum q u.
__global__ void __launch_bounds__(256) precon_aloss_stage1(const float* __restrict__ z, const float* __restrict__ az,
                                                           const float* __restrict__ r, const float* __restrict__ e,
                                                           const float* __restrict__ s, const float* __restrict__ u,
                                                           int64_t n, float* __restrict__ q,
                                                           double* __restrict__ part) {
  __shared__ double sh[256];
  double sl = 0.0, sa = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kLossGroups * 256) {
    double res = (double)az[i] - (double)r[i];
    float qi = (float)(2.0 * (double)s[i] * res);
    q[i] = qi;
    sl += ((double)z[i] - (double)e[i]) * res;
    sa += (double)qi * (double)u[i];
  }
  sl = block_sum256(sl, sh);
  sa = block_sum256(sa, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = sl;
    part[kLossGroups + blockIdx.x] = sa;
  }
}

// Box restriction by f (stride f): y[b,I,J] = f^-2 sum_{a,c<f} x[b, fI+a, fJ+c],
// zero beyond the fine grid; the coarse grid is ceil(H/f) x ceil(W/f).
__global__ void __launch_bounds__(256) restrict_kernel(const float* __restrict__ x, int H, int W, int f, int Hc,
                                                       int Wc, int64_t total, float* __restrict__ y) {
  const float inv = 1.f / (float)(f * f);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = k / ((int64_t)Hc * Wc);
    int rem = (int)(k - b * Hc * Wc);
    int I = rem / Wc, J = rem - I * Wc;
    const float* xb = x + b * (int64_t)H * W;
    float s = 0.f;
    for (int a = 0; a < f; a++) {
      int i = I * f + a;
      if (i >= H) break;
      for (int c = 0; c < f; c++) {
        int j = J * f + c;
        if (j < W) s += xb[(int64_t)i * W + j];
      }
    }
    y[k] = s * inv;
  }
}

// Adjoint of restrict_kernel: x[b,i,j] (+)= f^-2 y[b, i/f, j/f].
__global__ void __launch_bounds__(256) prolong_kernel(const float* __restrict__ y, int Hc, int Wc, int f, int H, int W,
                                                      int64_t total, float* __restrict__ x, int accumulate) {
  const float inv = 1.f / (float)(f * f);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = k / ((int64_t)H * W);
    int rem = (int)(k - b * H * W);
    int i = rem / W, j = rem - i * W;
    float v = y[b * (int64_t)Hc * Wc + (int64_t)(i / f) * Wc + j / f] * inv;
    x[k] = accumulate ? x[k] + v : v;
  }
}

// Full-weighting (tent) restriction by f: coarse point I is centred on fine
// position f I + (f-1)/2 and fine pixel i weighs max(0, 2f - |d|) / (2 f^2),
// d = 2i - (2 f I + f - 1), per axis (the weights of a coarse point sum to
// 1 in the interior).  Its adjoint is bilinear interpolation, so coarse
// corrections come back smooth instead of blocky.
__device__ __forceinline__ float tent_w(int i, int I, int f) {
  int d = 2 * i - (2 * f * I + f - 1);
  d = d < 0 ? -d : d;
  return d < 2 * f ? (float)(2 * f - d) : 0.f;
}

__global__ void __launch_bounds__(256) restrict_tent_kernel(const float* __restrict__ x, int H, int W, int f, int Hc,
                                                            int Wc, int64_t total, float* __restrict__ y) {
  const float norm = 1.f / (4.f * (float)f * (float)f * (float)f * (float)f);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = k / ((int64_t)Hc * Wc);
    int rem = (int)(k - b * Hc * Wc);
    int I = rem / Wc, J = rem - I * Wc;
    const float* xb = x + b * (int64_t)H * W;
    const int i0 = max(0, f * I - f / 2 - 1), i1 = min(H - 1, f * I + f + f / 2);
    const int j0 = max(0, f * J - f / 2 - 1), j1 = min(W - 1, f * J + f + f / 2);
    float s = 0.f;
    for (int i = i0; i <= i1; i++) {
      const float wi = tent_w(i, I, f);
      if (wi == 0.f) continue;
      float r = 0.f;
      for (int j = j0; j <= j1; j++) r = fmaf(tent_w(j, J, f), xb[(int64_t)i * W + j], r);
      s = fmaf(wi, r, s);
    }
    y[k] = s * norm;
  }
}

__global__ void __launch_bounds__(256) prolong_tent_kernel(const float* __restrict__ y, int Hc, int Wc, int f, int H,
                                                           int W, int64_t total, float* __restrict__ x,
                                                           int accumulate) {
  const float norm = 1.f / (4.f * (float)f * (float)f * (float)f * (float)f);
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (int64_t)gridDim.x * blockDim.x) {
    int64_t b = k / ((int64_t)H * W);
    int rem = (int)(k - b * H * W);
    int i = rem / W, j = rem - i * W;
    const float* yb = y + b * (int64_t)Hc * Wc;
    // coarse points whose support holds i: I with |2i - (2fI + f - 1)| < 2f
    const int Ilo = max(0, (2 * i - 3 * f + 1) / (2 * f) - 1), Ihi = min(Hc - 1, (2 * i + f + 1) / (2 * f) + 1);
    const int Jlo = max(0, (2 * j - 3 * f + 1) / (2 * f) - 1), Jhi = min(Wc - 1, (2 * j + f + 1) / (2 * f) + 1);
    float v = 0.f;
    for (int I = Ilo; I <= Ihi; I++) {
      const float wi = tent_w(i, I, f);
      if (wi == 0.f) continue;
      for (int J = Jlo; J <= Jhi; J++) v = fmaf(wi * tent_w(j, J, f), yb[(int64_t)I * Wc + J], v);
    }
    v *= norm;
    x[k] = accumulate ? x[k] + v : v;
  }
}

// Adam on a flat FP32 parameter vector.  Element n-1 is log(alpha); its
// gradient is alpha * dL/dalpha, taken from dalpha (FP64, on the device).
__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, int n,
                                                   const double* __restrict__ dalpha, float lr, float b1, float b2,
                                                   float eps, float bc1, float bc2, float gscale) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float gk;
  if (k == n - 1)
    gk = dalpha ? (float)((double)expf(p[k]) * dalpha[0]) : 0.f;
  else
    gk = g[k];
  gk *= gscale;
  float mk_ = b1 * m[k] + (1.f - b1) * gk;
  float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk_;
  v[k] = vk;
  p[k] -= lr * (mk_ / bc1) / (sqrtf(vk / bc2) + eps);
}

template <int CIN, int COUT>
int launch_conv(const float* x, const float* w, float* y, const Tiling& t, int B, int accumulate, hipStream_t st) {
  int64_t grid = (int64_t)B * t.per_img;
  MK_REQUIRE(grid > 0 && grid < INT32_MAX, MK_ERR_ARG, "conv3x3: %lld tiles out of range", (long long)grid);
  conv3x3_kernel<CIN, COUT><<<(unsigned)grid, 256, 0, st>>>(x, w, y, t, accumulate);
  return mk::check_launch("conv3x3_kernel");
}

template <int CO, int CI>
int launch_wgrad(const float* a, const float* x, const Tiling& t, int B, float* part, int groups, hipStream_t st) {
  int total = B * t.per_img;
  wgrad_kernel<CO, CI><<<groups, 256, 0, st>>>(a, x, t, total, part);
  return mk::check_launch("wgrad_kernel");
}

int wgrad_groups(int B, const Tiling& t) {
  int64_t total = (int64_t)B * t.per_img;
  return (int)(total < kWgradGroups ? total : kWgradGroups);
}

bool chan_ok(int c) { return c == 1 || c == MAXC; }

}  // namespace

extern "C" int mk_conv3x3_f32(const float* x, int32_t B, int32_t Cin, int32_t H, int32_t W, const float* w,
                              int32_t Cout, int32_t dil, float* y, int32_t accumulate, void* stream) {
  MK_REQUIRE(x && w && y, MK_ERR_ARG, "conv3x3: null argument");
  MK_REQUIRE(B > 0 && H > 0 && W > 0, MK_ERR_ARG, "conv3x3: bad shape B=%d H=%d W=%d", B, H, W);
  MK_REQUIRE(chan_ok(Cin) && chan_ok(Cout), MK_ERR_ARG, "conv3x3: channels must be 1 or %d (got %d -> %d)", MAXC,
             Cin, Cout);
  MK_REQUIRE(dil >= 1 && dil <= 4096, MK_ERR_ARG, "conv3x3: dilation %d out of range", dil);
  MK_REQUIRE((int64_t)B * Cin * H * W < ((int64_t)1 << 40), MK_ERR_ARG, "conv3x3: tensor too large");
  MK_REQUIRE(x != y, MK_ERR_ARG, "conv3x3 cannot run in place");
  Tiling t = make_tiling(H, W, dil);
  hipStream_t st = mk::as_stream(stream);
  if (Cin == 1 && Cout == MAXC) return launch_conv<1, MAXC>(x, w, y, t, B, accumulate, st);
  if (Cin == MAXC && Cout == MAXC) return launch_conv<MAXC, MAXC>(x, w, y, t, B, accumulate, st);
  if (Cin == MAXC && Cout == 1) return launch_conv<MAXC, 1>(x, w, y, t, B, accumulate, st);
  return launch_conv<1, 1>(x, w, y, t, B, accumulate, st);
}

extern "C" int mk_conv3x3_adjoint_weights(const float* w, int32_t Co, int32_t Ci, float* wt, void* stream) {
  MK_REQUIRE(w && wt && w != wt, MK_ERR_ARG, "adjoint_weights: bad pointers");
  MK_REQUIRE(Co >= 1 && Co <= MAXC && Ci >= 1 && Ci <= MAXC, MK_ERR_ARG, "adjoint_weights: bad channels");
  int nw = Co * Ci * 9;
  adjoint_weights_kernel<<<(unsigned)mk::cdiv(nw, 256), 256, 0, mk::as_stream(stream)>>>(w, Co, Ci, wt);
  return mk::check_launch("adjoint_weights_kernel");
}

extern "C" int64_t mk_conv3x3_wgrad_workspace_bytes(int32_t B, int32_t Co, int32_t Ci, int32_t H, int32_t W,
                                                    int32_t dil) {
  if (B <= 0 || H <= 0 || W <= 0 || dil < 1 || Co < 1 || Ci < 1) return -1;
  Tiling t = make_tiling(H, W, dil);
  return (int64_t)wgrad_groups(B, t) * Co * Ci * 9 * 4 + 256;
}

extern "C" int mk_conv3x3_wgrad_f32(const float* a, int32_t Co, const float* x, int32_t Ci, int32_t B, int32_t H,
                                    int32_t W, int32_t dil, float* g, int32_t accumulate, void* work,
                                    int64_t work_bytes, void* stream) {
  MK_REQUIRE(a && x && g && work, MK_ERR_ARG, "wgrad: null argument");
  MK_REQUIRE(chan_ok(Co) && chan_ok(Ci) && B > 0 && H > 0 && W > 0, MK_ERR_ARG, "wgrad: bad shape");
  MK_REQUIRE(dil >= 1 && dil <= 4096, MK_ERR_ARG, "wgrad: dilation %d out of range", dil);
  MK_REQUIRE((int64_t)B * MAXC * H * W < ((int64_t)1 << 40), MK_ERR_ARG, "wgrad: tensor too large");
  MK_REQUIRE(work_bytes >= mk_conv3x3_wgrad_workspace_bytes(B, Co, Ci, H, W, dil), MK_ERR_ARG,
             "wgrad: workspace too small");
  Tiling t = make_tiling(H, W, dil);
  MK_REQUIRE((int64_t)B * t.per_img < INT32_MAX, MK_ERR_ARG, "wgrad: too many tiles");
  int groups = wgrad_groups(B, t);
  hipStream_t st = mk::as_stream(stream);
  float* part = (float*)work;
  int r;
  if (Co == MAXC && Ci == MAXC) r = launch_wgrad<MAXC, MAXC>(a, x, t, B, part, groups, st);
  else if (Co == MAXC && Ci == 1) r = launch_wgrad<MAXC, 1>(a, x, t, B, part, groups, st);
  else if (Co == 1 && Ci == MAXC) r = launch_wgrad<1, MAXC>(a, x, t, B, part, groups, st);
  else r = launch_wgrad<1, 1>(a, x, t, B, part, groups, st);
  if (r) return r;
  int nw = Co * Ci * 9;
  wgrad_reduce_kernel<<<(unsigned)mk::cdiv(nw, 256), 256, 0, st>>>(part, groups, nw, g, accumulate);
  return mk::check_launch("wgrad_reduce_kernel");
}

extern "C" int mk_precon_finish_f32(float* z, const float* s, const float* log_alpha, const float* u, const float* g,
                                    int64_t n, void* stream) {
  MK_REQUIRE(z && s && log_alpha && u && n >= 0, MK_ERR_ARG, "precon_finish: null argument");
  if (n == 0) return MK_OK;
  precon_finish_kernel<<<mk::grid_for(n, 256), 256, 0, mk::as_stream(stream)>>>(z, s, log_alpha, u, g, n);
  return mk::check_launch("precon_finish_kernel");
}

extern "C" int64_t mk_precon_loss_workspace_bytes(void) { return (int64_t)2 * kLossGroups * sizeof(double); }

extern "C" int mk_precon_loss_f32(const float* g, const float* u, const float* s, const float* e, const float* m,
                                  const float* log_alpha, int64_t n, float* q, double* sc, void* work,
                                  int64_t work_bytes, void* stream) {
  MK_REQUIRE(g && u && s && e && m && log_alpha && q && sc && work, MK_ERR_ARG, "precon_loss: null argument");
  MK_REQUIRE(n > 0, MK_ERR_ARG, "precon_loss: empty input");
  MK_REQUIRE(work_bytes >= mk_precon_loss_workspace_bytes(), MK_ERR_ARG, "precon_loss: workspace too small");
  hipStream_t st = mk::as_stream(stream);
  precon_loss_stage1<<<kLossGroups, 256, 0, st>>>(g, u, s, e, m, log_alpha, n, q, (double*)work);
  int r = mk::check_launch("precon_loss_stage1");
  if (r) return r;
  precon_loss_stage2<<<1, 256, 0, st>>>((const double*)work, sc);
  return mk::check_launch("precon_loss_stage2");
}

extern "C" int mk_adam_f32(float* p, const float* g, float* m, float* v, int32_t n, const double* dalpha, float lr,
                           float beta1, float beta2, float eps, int32_t step, float gscale, void* stream) {
  MK_REQUIRE(p && g && m && v && n > 0 && step >= 1, MK_ERR_ARG, "adam: bad argument");
  float bc1 = 1.f - powf(beta1, (float)step), bc2 = 1.f - powf(beta2, (float)step);
  adam_kernel<<<(unsigned)mk::cdiv(n, 256), 256, 0, mk::as_stream(stream)>>>(p, g, m, v, n, dalpha, lr, beta1, beta2,
                                                                            eps, bc1, bc2, gscale);
  return mk::check_launch("adam_kernel");
}

extern "C" int mk_precon_aloss_f32(const float* z, const float* az, const float* r, const float* e, const float* s,
                                   const float* u, int64_t n, float* q, double* sc, void* work, int64_t work_bytes,
                                   void* stream) {
  MK_REQUIRE(z && az && r && e && s && u && q && sc && work, MK_ERR_ARG, "precon_aloss: null argument");
  MK_REQUIRE(n > 0, MK_ERR_ARG, "precon_aloss: empty input");
  MK_REQUIRE(work_bytes >= mk_precon_loss_workspace_bytes(), MK_ERR_ARG, "precon_aloss: workspace too small");
  hipStream_t st = mk::as_stream(stream);
  precon_aloss_stage1<<<kLossGroups, 256, 0, st>>>(z, az, r, e, s, u, n, q, (double*)work);
  int rc = mk::check_launch("precon_aloss_stage1");
  if (rc) return rc;
  precon_loss_stage2<<<1, 256, 0, st>>>((const double*)work, sc);
  return mk::check_launch("precon_loss_stage2");
}

extern "C" int mk_restrict_box_f32(const float* x, int32_t B, int32_t H, int32_t W, int32_t f, float* y, void* stream) {
  MK_REQUIRE(x && y && x != y, MK_ERR_ARG, "restrict: bad pointers");
  MK_REQUIRE(B > 0 && H > 0 && W > 0 && f >= 1 && f <= 64, MK_ERR_ARG, "restrict: bad shape");
  int Hc = (int)mk::cdiv(H, f), Wc = (int)mk::cdiv(W, f);
  int64_t total = (int64_t)B * Hc * Wc;
  restrict_kernel<<<mk::grid_for(total, 256), 256, 0, mk::as_stream(stream)>>>(x, H, W, f, Hc, Wc, total, y);
  return mk::check_launch("restrict_kernel");
}

extern "C" int mk_prolong_box_f32(const float* y, int32_t B, int32_t H, int32_t W, int32_t f, float* x,
                                  int32_t accumulate, void* stream) {
  MK_REQUIRE(x && y && x != y, MK_ERR_ARG, "prolong: bad pointers");
  MK_REQUIRE(B > 0 && H > 0 && W > 0 && f >= 1 && f <= 64, MK_ERR_ARG, "prolong: bad shape");
  int Hc = (int)mk::cdiv(H, f), Wc = (int)mk::cdiv(W, f);
  int64_t total = (int64_t)B * H * W;
  prolong_kernel<<<mk::grid_for(total, 256), 256, 0, mk::as_stream(stream)>>>(y, Hc, Wc, f, H, W, total, x,
                                                                               accumulate);
  return mk::check_launch("prolong_kernel");
}

extern "C" int mk_restrict_tent_f32(const float* x, int32_t B, int32_t H, int32_t W, int32_t f, float* y,
                                    void* stream) {
  MK_REQUIRE(x && y && x != y, MK_ERR_ARG, "restrict_tent: bad pointers");
  MK_REQUIRE(B > 0 && H > 0 && W > 0 && f >= 1 && f <= 64, MK_ERR_ARG, "restrict_tent: bad shape");
  int Hc = (int)mk::cdiv(H, f), Wc = (int)mk::cdiv(W, f);
  int64_t total = (int64_t)B * Hc * Wc;
  restrict_tent_kernel<<<mk::grid_for(total, 256), 256, 0, mk::as_stream(stream)>>>(x, H, W, f, Hc, Wc, total, y);
  return mk::check_launch("restrict_tent_kernel");
}

extern "C" int mk_prolong_tent_f32(const float* y, int32_t B, int32_t H, int32_t W, int32_t f, float* x,
                                   int32_t accumulate, void* stream) {
  MK_REQUIRE(x && y && x != y, MK_ERR_ARG, "prolong_tent: bad pointers");
  MK_REQUIRE(B > 0 && H > 0 && W > 0 && f >= 1 && f <= 64, MK_ERR_ARG, "prolong_tent: bad shape");
  int Hc = (int)mk::cdiv(H, f), Wc = (int)mk::cdiv(W, f);
  int64_t total = (int64_t)B * H * W;
  prolong_tent_kernel<<<mk::grid_for(total, 256), 256, 0, mk::as_stream(stream)>>>(y, Hc, Wc, f, H, W, total, x,
                                                                                    accumulate);
  return mk::check_launch("prolong_tent_kernel");
}

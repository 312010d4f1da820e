// Pointing -> pixel for a CAR WCS, in FP64 on the device.
//
// The reference pixelizes with astropy's wcs_world2pix (wcslib) on the host,
// one thread, once per TOD (minkasi/maps/skymap.py:457-490).  This kernel
// computes the same celestial->native rotation (Calabretta & Greisen 2002,
// A&A 395, 1077, eqs. 5 and 14), the CAR projection and the diagonal linear
// transform with wcslib's degree-based arithmetic, then applies the
// reference's -1 / round-half-even / x*ny+y / int32 steps.  Each sample reads
// two doubles and writes one int32 (20 B/sample).
#include "common.h"

namespace {

constexpr double kPi = 3.141592653589793238462643;
constexpr double kD2R = kPi / 180.0;
constexpr double kR2D = 180.0 / kPi;

struct CarWcs {
  double eul[6];   // alpha_p, 90-delta_p, phi_p, cos(90-delta_p), sin(90-delta_p), scale
  double cdelt[2];
  double crpix[2];
};

__device__ inline void sincosd(double a, double* s, double* c) {
  if (fmod(a, 90.0) == 0.0) {
    int i = abs((int)floor(a / 90.0 + 0.5)) % 4;
    if (i == 0) { *s = 0.0; *c = 1.0; }
    else if (i == 1) { *s = (a > 0.0) ? 1.0 : -1.0; *c = 0.0; }
    else if (i == 2) { *s = 0.0; *c = -1.0; }
    else { *s = (a > 0.0) ? -1.0 : 1.0; *c = 0.0; }
    return;
  }
  *s = sin(a * kD2R);
  *c = cos(a * kD2R);
}

__device__ inline double atan2d(double y, double x) {
  if (y == 0.0) return (x >= 0.0) ? 0.0 : 180.0;
  if (x == 0.0) return (y > 0.0) ? 90.0 : -90.0;
  return atan2(y, x) * kR2D;
}

__device__ inline double asind(double v) {
  if (v <= -1.0) { if (v + 1.0 > -1e-10) return -90.0; }
  else if (v == 0.0) return 0.0;
  else if (v >= 1.0) { if (v - 1.0 < 1e-10) return 90.0; }
  return asin(v) * kR2D;
}

__device__ inline double acosd(double v) {
  if (v >= 1.0) { if (v - 1.0 < 1e-10) return 0.0; }
  else if (v == 0.0) return 90.0;
  else if (v <= -1.0) { if (v + 1.0 > -1e-10) return 180.0; }
  return acos(v) * kR2D;
}

__device__ inline void car_one(double ra, double dec, const CarWcs& w, double* px, double* py) {
  const double* eul = w.eul;
  double lng = ra * 180 / kPi;
  double lat = dec * 180 / kPi;
  double dlng = lng - eul[0];
  double sinlat, coslat, sinlng, coslng;
  sincosd(lat, &sinlat, &coslat);
  sincosd(dlng, &sinlng, &coslng);
  double costhe3 = coslat * eul[3], costhe4 = coslat * eul[4];
  double sinthe3 = sinlat * eul[3], sinthe4 = sinlat * eul[4];
  double x = sinthe4 - costhe3 * coslng;
  if (fabs(x) < 1.0e-5) {
    double s2, c2;
    sincosd(lat + eul[1], &s2, &c2);
    x = -c2 + costhe3 * (1.0 - coslng);
  }
  double y = -coslat * sinlng;
  double dphi;
  if (x != 0.0 || y != 0.0) dphi = atan2d(y, x);
  else dphi = (eul[1] < 90.0) ? dlng - 180.0 : -dlng;
  double phi = fmod(eul[2] + dphi, 360.0);
  if (phi > 180.0) phi -= 360.0; else if (phi < -180.0) phi += 360.0;
  double theta;
  if (fmod(dlng, 180.0) == 0.0) {
    theta = lat + coslng * eul[1];
    if (theta > 90.0) theta = 180.0 - theta;
    if (theta < -90.0) theta = -180.0 - theta;
  } else {
    double z = sinthe3 + costhe4 * coslng;
    if (fabs(z) > 0.99) theta = copysign(acosd(sqrt(x * x + y * y)), z);
    else theta = asind(z);
  }
  double X = eul[5] * phi, Y = eul[5] * theta;
  *px = X / w.cdelt[0] + w.crpix[0];
  *py = Y / w.cdelt[1] + w.crpix[1];
}

__device__ inline int32_t to_ipix(double px, double py, int32_t nx, int32_t ny, bool* bad) {
  double xp = rint(px - 1.0);  // round half to even, as np.round
  double yp = rint(py - 1.0);
  *bad = !(xp >= 0.0 && xp < (double)nx && yp >= 0.0 && yp < (double)ny);
  return (int32_t)(xp * (double)ny + yp);
}

using mk::count_bad;

__global__ void __launch_bounds__(256) pix_kernel(const double* __restrict__ ra, const double* __restrict__ dec,
                                                  int64_t n, CarWcs w, int32_t nx, int32_t ny,
                                                  int32_t* __restrict__ ipix, int64_t* nbad) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    int64_t i = base + threadIdx.x;
    bool bad = false;
    if (i < n) {
      double px, py;
      car_one(ra[i], dec[i], w, &px, &py);
      ipix[i] = to_ipix(px, py, nx, ny, &bad);
    }
    count_bad(bad, nbad);
  }
}

__global__ void __launch_bounds__(256) raster_kernel(const double* __restrict__ ra_b, const double* __restrict__ dec_b,
                                                     const double* __restrict__ dra, const double* __restrict__ ddec,
                                                     int64_t ndet, int64_t ndata, CarWcs w, int32_t nx, int32_t ny,
                                                     int32_t* __restrict__ ipix, int64_t* nbad) {
  int64_t n = ndet * ndata;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    int64_t i = base + threadIdx.x;
    bool bad = false;
    if (i < n) {
      int64_t d = i / ndata, t = i - d * ndata;
      double px, py;
      car_one(ra_b[t] + dra[d], dec_b[t] + ddec[d], w, &px, &py);
      ipix[i] = to_ipix(px, py, nx, ny, &bad);
    }
    count_bad(bad, nbad);
  }
}

__global__ void __launch_bounds__(256) w2p_kernel(const double* __restrict__ ra, const double* __restrict__ dec,
                                                  int64_t n, CarWcs w, double* __restrict__ px,
                                                  double* __restrict__ py) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    car_one(ra[i], dec[i], w, px + i, py + i);
}

int load_wcs(const double* wcs10, CarWcs* w) {
  MK_REQUIRE(wcs10 != nullptr, MK_ERR_ARG, "wcs10 is NULL");
  for (int k = 0; k < 6; k++) w->eul[k] = wcs10[k];
  w->cdelt[0] = wcs10[6]; w->cdelt[1] = wcs10[7];
  w->crpix[0] = wcs10[8]; w->crpix[1] = wcs10[9];
  MK_REQUIRE(w->cdelt[0] != 0.0 && w->cdelt[1] != 0.0, MK_ERR_ARG, "cdelt must be nonzero");
  return MK_OK;
}

}  // namespace

extern "C" int mk_pix_from_radec_car(const double* ra, const double* dec, int64_t nsamp, const double* wcs10,
                                     int32_t nx, int32_t ny, int32_t* ipix, int64_t* dev_nbad, void* stream) {
  MK_REQUIRE(nsamp >= 0, MK_ERR_ARG, "nsamp < 0");
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(ra && dec && ipix, MK_ERR_ARG, "null pointer");
  MK_REQUIRE(nx > 0 && ny > 0 && (int64_t)nx * ny <= INT32_MAX, MK_ERR_ARG, "bad map size %d x %d", nx, ny);
  CarWcs w;
  int st = load_wcs(wcs10, &w);
  if (st) return st;
  pix_kernel<<<mk::grid_for(nsamp, 256, 256 * 8), 256, 0, mk::as_stream(stream)>>>(ra, dec, nsamp, w, nx, ny, ipix, dev_nbad);
  return mk::check_launch("pix_kernel");
}

extern "C" int mk_pix_from_raster_car(const double* ra_b, const double* dec_b, const double* dra, const double* ddec,
                                      int64_t ndet, int64_t ndata, const double* wcs10, int32_t nx, int32_t ny,
                                      int32_t* ipix, int64_t* dev_nbad, void* stream) {
  MK_REQUIRE(ndet >= 0 && ndata >= 0, MK_ERR_ARG, "negative size");
  if (ndet * ndata == 0) return MK_OK;
  MK_REQUIRE(ra_b && dec_b && dra && ddec && ipix, MK_ERR_ARG, "null pointer");
  MK_REQUIRE(nx > 0 && ny > 0 && (int64_t)nx * ny <= INT32_MAX, MK_ERR_ARG, "bad map size %d x %d", nx, ny);
  CarWcs w;
  int st = load_wcs(wcs10, &w);
  if (st) return st;
  raster_kernel<<<mk::grid_for(ndet * ndata, 256, 256 * 8), 256, 0, mk::as_stream(stream)>>>(
      ra_b, dec_b, dra, ddec, ndet, ndata, w, nx, ny, ipix, dev_nbad);
  return mk::check_launch("raster_kernel");
}

extern "C" int mk_world2pix_car(const double* ra, const double* dec, int64_t nsamp, const double* wcs10, double* px,
                                double* py, void* stream) {
  MK_REQUIRE(nsamp >= 0, MK_ERR_ARG, "nsamp < 0");
  if (nsamp == 0) return MK_OK;
  MK_REQUIRE(ra && dec && px && py, MK_ERR_ARG, "null pointer");
  CarWcs w;
  int st = load_wcs(wcs10, &w);
  if (st) return st;
  w2p_kernel<<<mk::grid_for(nsamp, 256, 256 * 8), 256, 0, mk::as_stream(stream)>>>(ra, dec, nsamp, w, px, py);
  return mk::check_launch("w2p_kernel");
}

/*
 * oracle.c -- FP64 CPU restatement of the minkasi mapmaking projections and
 * of the CAR pixelization, used ONLY as test infrastructure: the parity
 * tests, the smoke check and the CPU-baseline leg of bench.py call it as the
 * checker.  Nothing in the product (minkasi_amd/) links or calls this file.
 *
 * Pins (see tests/test_oracle.py):
 *   or_map2tod*       -- closed form m.ravel()[ipix]; cross-checked against the
 *                        reference's own minkasi.c built into oracle/_ref.
 *   or_tod2map*       -- np.bincount(ipix, weights); cross-checked against the
 *                        reference's tod2map_simple built into oracle/_ref.
 *   or_car_*          -- astropy 4.3.1 / wcslib 7.6 fixtures in tests/golden.
 *   or_dct1_direct    -- brute-force DCT-I definition (REDFT00), pinned against
 *                        scipy.fft.dct(type=1) and DCT-I^2 = 2(n-1) I.
 *   or_tod2cuts,      -- closed forms vec = dat.ravel()[imap] and
 *   or_cuts2tod          dat.ravel()[imap] = vec; bit-equal to the reference's
 *                        tod2cuts / cuts2tod built into oracle/_ref.
 *   or_polmap2tod,    -- closed forms m[3p] + m[3p+1] cos + m[3p+2] sin and
 *   or_tod2polmap        np.bincount with the cos/sin weights; bit-equal to the
 *                        reference's map2tod_{iqu,qu}_omp and
 *                        tod2map_{qu,iqu}{,_precon}_simple built into oracle/_ref.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ---------------------------------------------------------------------- */
/* Projections.  Follows minkasi/lib/minkasi.c:116-144 (map2tod_simple,
 * map2tod_omp) and :50-55, :97-112 (tod2map_simple, tod2map_omp).  Sizes are
 * 64-bit here (the reference computes ndet*ndata in int).                  */

void or_map2tod(double *dat, const double *map, int64_t ndet, int64_t ndata,
                const int32_t *pix, int do_add, int nthread)
{
  int64_t nn = ndet * ndata;
  if (nthread < 1) nthread = 1;
  if (do_add) {
#pragma omp parallel for num_threads(nthread) schedule(static)
    for (int64_t i = 0; i < nn; i++) dat[i] += map[pix[i]];
  } else {
#pragma omp parallel for num_threads(nthread) schedule(static)
    for (int64_t i = 0; i < nn; i++) dat[i] = map[pix[i]];
  }
}

void or_tod2map_simple(double *map, const double *dat, int64_t ndet,
                       int64_t ndata, const int32_t *pix)
{
  int64_t nn = ndet * ndata;
  for (int64_t i = 0; i < nn; i++) map[pix[i]] += dat[i];
}

/* Per-thread private maps as in tod2map_omp (minkasi.c:97-112), but the
 * private maps are summed in thread order after the scatter, so the result
 * does not depend on which thread reaches the reduction first. */
void or_tod2map_omp(double *map, const double *dat, int64_t ndet,
                    int64_t ndata, const int32_t *pix, int64_t npix,
                    int nthread)
{
  int64_t nn = ndet * ndata;
  if (nthread < 1) nthread = 1;
  double *priv = (double *)calloc((size_t)nthread * (size_t)npix, sizeof(double));
  if (!priv) return;
#pragma omp parallel num_threads(nthread)
  {
    int t = omp_get_thread_num();
    double *mine = priv + (size_t)t * (size_t)npix;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < nn; i++) mine[pix[i]] += dat[i];
#pragma omp barrier
#pragma omp for schedule(static)
    for (int64_t p = 0; p < npix; p++) {
      double s = map[p];
      for (int k = 0; k < nthread; k++) s += priv[(size_t)k * (size_t)npix + p];
      map[p] = s;
    }
  }
  free(priv);
}

/* ---------------------------------------------------------------------- */
/* CAR world->pixel, restated from Calabretta & Greisen 2002 (A&A 395, 1077)
 * with the degree-based arithmetic that wcslib uses, because the reference
 * pixelizes through astropy.wcs (minkasi/maps/skymap.py:457-490, WCS from
 * minkasi/maps/utils.py:13-67).  Every step is done in the same order as the
 * device kernel in minkasi_amd/csrc/pointing.hip so that both give the same
 * doubles up to the last-ulp differences of the platform's libm.           */

#define OR_PI 3.141592653589793238462643
static const double or_d2r = OR_PI / 180.0;
static const double or_r2d = 180.0 / OR_PI;

static void or_sincosd(double a, double *s, double *c)
{
  if (fmod(a, 90.0) == 0.0) {
    int i = abs((int)floor(a / 90.0 + 0.5)) % 4;
    switch (i) {
      case 0: *s = 0.0; *c = 1.0; break;
      case 1: *s = (a > 0.0) ? 1.0 : -1.0; *c = 0.0; break;
      case 2: *s = 0.0; *c = -1.0; break;
      default: *s = (a > 0.0) ? -1.0 : 1.0; *c = 0.0; break;
    }
    return;
  }
  *s = sin(a * or_d2r);
  *c = cos(a * or_d2r);
}

static double or_cosd(double a) { double s, c; or_sincosd(a, &s, &c); return c; }
static double or_sind(double a) { double s, c; or_sincosd(a, &s, &c); return s; }

static double or_atan2d(double y, double x)
{
  if (y == 0.0) {
    if (x >= 0.0) return 0.0;
    return 180.0;
  } else if (x == 0.0) {
    return (y > 0.0) ? 90.0 : -90.0;
  }
  return atan2(y, x) * or_r2d;
}

static double or_asind(double v)
{
  if (v <= -1.0) { if (v + 1.0 > -1e-10) return -90.0; }
  else if (v == 0.0) return 0.0;
  else if (v >= 1.0) { if (v - 1.0 < 1e-10) return 90.0; }
  return asin(v) * or_r2d;
}

static double or_acosd(double v)
{
  if (v >= 1.0) { if (v - 1.0 < 1e-10) return 0.0; }
  else if (v == 0.0) return 90.0;
  else if (v <= -1.0) { if (v + 1.0 > -1e-10) return 180.0; }
  return acos(v) * or_r2d;
}

/* Celestial->native Euler angles for a CAR projection (phi0 = theta0 = 0),
 * default LONPOLE/LATPOLE.  out[0..4] = alpha_p, 90-delta_p, phi_p,
 * cos(90-delta_p), sin(90-delta_p); out[5] = CAR scale r0*d2r.
 * Returns 0 on success. */
int or_car_euler(double lng0, double lat0, double *out)
{
  const double phi0 = 0.0, theta0 = 0.0, latpole = 90.0, tol = 1.0e-10;
  double phip = ((lat0 < theta0) ? 180.0 : 0.0) + phi0;
  double slat0, clat0, sthe0, cthe0, sphip, cphip, u, v;
  or_sincosd(lat0, &slat0, &clat0);
  or_sincosd(theta0, &sthe0, &cthe0);
  if (phip == phi0) {
    sphip = 0.0; cphip = 1.0; u = theta0; v = 90.0 - lat0;
  } else {
    or_sincosd(phip - phi0, &sphip, &cphip);
    double x = cthe0 * cphip, y = sthe0, z = sqrt(x * x + y * y);
    if (z == 0.0) return 1;
    u = or_atan2d(y, x);
    v = or_acosd(slat0 / z);
  }
  double latp1 = u + v;
  if (latp1 > 180.0) latp1 -= 360.0; else if (latp1 < -180.0) latp1 += 360.0;
  double latp2 = u - v;
  if (latp2 > 180.0) latp2 -= 360.0; else if (latp2 < -180.0) latp2 += 360.0;
  double latp;
  if (fabs(latp1) < 90.0 + tol && fabs(latp2) < 90.0 + tol)
    latp = (fabs(latp1 - latpole) < fabs(latp2 - latpole)) ? latp1 : latp2;
  else if (fabs(latp1) < 90.0 + tol) latp = latp1;
  else if (fabs(latp2) < 90.0 + tol) latp = latp2;
  else return 2;
  if (latp > 90.0) latp = 90.0; else if (latp < -90.0) latp = -90.0;
  double z = or_cosd(latp) * clat0, lngp;
  if (fabs(z) < tol) return 3;
  {
    double x = (sthe0 - or_sind(latp) * slat0) / z;
    double y = sphip * cthe0 / clat0;
    if (x == 0.0 && y == 0.0) return 4;
    lngp = lng0 - or_atan2d(y, x);
  }
  if (lng0 >= 0.0) {
    if (lngp < 0.0) lngp += 360.0; else if (lngp > 360.0) lngp -= 360.0;
  } else {
    if (lngp > 0.0) lngp -= 360.0; else if (lngp < -360.0) lngp += 360.0;
  }
  out[0] = lngp;
  out[1] = 90.0 - latp;
  out[2] = phip;
  or_sincosd(out[1], &out[4], &out[3]);
  out[5] = or_r2d * or_d2r; /* CAR: x = w0*phi with w0 = r0*d2r, r0 = r2d */
  return 0;
}

/* One sample: (ra, dec) in radians -> 1-based pixel coordinates, following
 * the general branch of the celestial->native rotation, the CAR projection and
 * the diagonal linear transform px = x/cdelt + crpix.                      */
static void or_car_one(double ra, double dec, const double *eul,
                       const double *lin, double *px, double *py)
{
  double lng = ra * 180 / OR_PI;
  double lat = dec * 180 / OR_PI;
  double dlng = lng - eul[0];
  double sinlat, coslat, sinlng, coslng;
  or_sincosd(lat, &sinlat, &coslat);
  or_sincosd(dlng, &sinlng, &coslng);
  double costhe3 = coslat * eul[3], costhe4 = coslat * eul[4];
  double sinthe3 = sinlat * eul[3], sinthe4 = sinlat * eul[4];
  double x = sinthe4 - costhe3 * coslng;
  if (fabs(x) < 1.0e-5) x = -or_cosd(lat + eul[1]) + costhe3 * (1.0 - coslng);
  double y = -coslat * sinlng;
  double dphi;
  if (x != 0.0 || y != 0.0) dphi = or_atan2d(y, x);
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
    if (fabs(z) > 0.99) theta = copysign(or_acosd(sqrt(x * x + y * y)), z);
    else theta = or_asind(z);
  }
  double X = eul[5] * phi, Y = eul[5] * theta;
  *px = X / lin[0] + lin[2];
  *py = Y / lin[1] + lin[3];
}

/* lin = {cdelt_x, cdelt_y, crpix_x, crpix_y} (degrees, 1-based). */
void or_car_world2pix(const double *ra, const double *dec, int64_t n,
                      const double *eul, const double *lin, double *px,
                      double *py)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) or_car_one(ra[i], dec[i], eul, lin, px + i, py + i);
}

/* skymap.py:483-489: x = px-1, y = py-1, np.round (half to even),
 * ipix = int32(x*ny + y) computed in FP64. */
void or_pix_from_radec(const double *ra, const double *dec, int64_t n,
                       const double *eul, const double *lin, int32_t ny,
                       int32_t *ipix, int nthread)
{
  if (nthread < 1) nthread = 1;
#pragma omp parallel for num_threads(nthread) schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double px, py;
    or_car_one(ra[i], dec[i], eul, lin, &px, &py);
    double xp = nearbyint(px - 1.0);
    double yp = nearbyint(py - 1.0);
    ipix[i] = (int32_t)(xp * (double)ny + yp);
  }
}

/* ---------------------------------------------------------------------- */
/* DCT-I (FFTW REDFT00, unnormalised) by its definition, for tiny n:
 * Y_k = x_0 + (-1)^k x_{n-1} + 2 sum_{j=1}^{n-2} x_j cos(pi j k/(n-1)).
 * The reference calls it through minkasi/lib/mkfftw.c:174-212 or
 * scipy.fft.dct(type=1) (minkasi/tools/py_fft.py:239).                     */
void or_dct1_direct(const double *x, int64_t nrow, int64_t n, double *out)
{
  int64_t N = 2 * (n - 1);
  for (int64_t r = 0; r < nrow; r++) {
    const double *xr = x + r * n;
    double *yr = out + r * n;
    for (int64_t k = 0; k < n; k++) {
      double s = xr[0] + ((k & 1) ? -xr[n - 1] : xr[n - 1]);
      for (int64_t j = 1; j < n - 1; j++) {
        int64_t ph = (j * k) % N; /* exact phase index */
        s += 2.0 * xr[j] * cos(OR_PI * (double)(2 * ph) / (double)N);
      }
      yr[k] = s;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* Cut samples (CutsCompact, minkasi/tods/cuts.py:201-498).  Follows
 * minkasi.c:260-284 (tod2cuts, cuts2tod): serial loops in index order.     */
void or_tod2cuts(double *vec, const double *dat, const int64_t *imap,
                 int64_t ncut, int do_add)
{
  for (int64_t i = 0; i < ncut; i++)
    vec[i] = do_add ? vec[i] + dat[imap[i]] : dat[imap[i]];
}

void or_cuts2tod(double *dat, const double *vec, const int64_t *imap,
                 int64_t ncut, int do_add)
{
  for (int64_t i = 0; i < ncut; i++)
    dat[imap[i]] = do_add ? dat[imap[i]] + vec[i] : vec[i];
}

/* ---------------------------------------------------------------------- */
/* Polarised projections (PolMap, minkasi/maps/polmap.py:464-545).  Follows
 * minkasi.c:146-175 (map2tod_iqu_omp, map2tod_qu_omp) and :177-257
 * (tod2map_{qu,iqu}_simple, tod2map_{qu,iqu}_precon_simple).  mode: 1 QU,
 * 2 IQU, 3 QU_PRECON, 4 IQU_PRECON; ncomp values per pixel, pixel-major.   */
void or_polmap2tod(double *dat, const double *map, const double *twogamma,
                   int64_t nn, const int32_t *pix, int npol, int do_add)
{
  for (int64_t i = 0; i < nn; i++) {
    const double c = cos(twogamma[i]), s = sin(twogamma[i]);
    const double *m = map + (int64_t)npol * pix[i];
    const double v = npol == 3 ? m[0] + m[1] * c + m[2] * s : m[0] * c + m[1] * s;
    dat[i] = do_add ? dat[i] + v : v;
  }
}

void or_tod2polmap(double *map, const double *dat, const double *twogamma,
                   int64_t nn, const int32_t *pix, int mode)
{
  const int nc = mode == 1 ? 2 : mode == 2 ? 3 : mode == 3 ? 3 : 6;
  for (int64_t i = 0; i < nn; i++) {
    const double c = cos(twogamma[i]), s = sin(twogamma[i]), d = dat[i];
    double *m = map + (int64_t)nc * pix[i];
    switch (mode) {
      case 1: m[0] += d * c; m[1] += d * s; break;
      case 2: m[0] += d; m[1] += d * c; m[2] += d * s; break;
      case 3: m[0] += d * c * c; m[1] += d * s * s; m[2] += d * s * c; break;
      default:
        m[0] += d; m[1] += d * c; m[2] += d * s;
        m[3] += d * c * c; m[4] += d * s * c; m[5] += d * s * s;
    }
  }
}

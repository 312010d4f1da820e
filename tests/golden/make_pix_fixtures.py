"""Writes tests/golden/pix_car_fixtures.npz: pixel-index golden vectors made
with astropy.wcs (wcslib), the library the reference pixelizes through.

Run with an interpreter that has astropy (here /opt/conda/bin/python3.9,
astropy 4.3.1 / wcslib 7.6, with small numpy-compat shims):

    /opt/conda/bin/python3.9 tests/golden/make_pix_fixtures.py

The WCS is built exactly as minkasi/maps/utils.py:47-66 builds it for
proj='CAR' (crpix [1,1], crval (ra_max, dec_min) in degrees, cdelt
[-pix, pix] in degrees, cosdec forced to 1), map sizes follow
minkasi/maps/skymap.py:206-234 (corners -> wcs_world2pix -> round -> max+pad)
and indices follow skymap.py:476-489 (-1, np.round, x*ny+y, int32).  This
script does not import the reference; it restates those few lines around the
astropy calls.  Only inputs and astropy outputs are stored.
"""
import os
import sys

import numpy as np

# numpy >= 1.24 removed aliases that astropy 4.3 still touches
np.asscalar = lambda a: a.item()
np.alen = lambda a: len(a)
for _n, _t in [("float", float), ("int", int), ("bool", bool),
               ("object", object), ("str", str), ("complex", complex)]:
    setattr(np, _n, _t)

from astropy import wcs  # noqa: E402


def make_wcs(lims, pixsize):
    w = wcs.WCS(naxis=2)
    cosdec = 1.0
    w.wcs.crpix = [1.0, 1.0]
    w.wcs.crval = [lims[1] * 180 / np.pi, lims[2] * 180 / np.pi]
    w.wcs.cdelt = [-pixsize / cosdec * 180 / np.pi, pixsize * 180 / np.pi]
    w.wcs.ctype = ["RA---CAR", "DEC--CAR"]
    return w


def npix(w, lims, pad):
    corners = np.zeros([4, 2])
    corners[0, :] = [lims[0], lims[2]]
    corners[1, :] = [lims[0], lims[3]]
    corners[2, :] = [lims[1], lims[2]]
    corners[3, :] = [lims[1], lims[3]]
    pc = np.round(w.wcs_world2pix(corners * 180 / np.pi, 1))
    return int(pc[:, 0].max() + pad), int(pc[:, 1].max() + pad)


def raster(rng, ndet, n, ra0, dec0, half_width, fs=100.0, speed=30.0):
    """Constant-velocity raster in RA coordinate (radians)."""
    arcsec = np.pi / 180 / 3600
    t = np.arange(n) / fs
    leg = 2 * half_width / speed  # seconds per leg
    ph = (t / leg) % 2.0
    tri = np.where(ph < 1.0, ph, 2.0 - ph) * 2.0 - 1.0
    ra_b = ra0 + tri * half_width * arcsec / np.cos(dec0)
    legno = np.floor(t / leg)
    dec_b = dec0 + (((legno * 10.0) % 480.0) - 240.0) * arcsec
    r = 60.0 * np.sqrt(rng.uniform(0, 1, ndet))
    th = rng.uniform(0, 2 * np.pi, ndet)
    ddec = r * np.sin(th) * arcsec
    dra = r * np.cos(th) * arcsec / np.cos(dec0)
    ra = ra_b[None, :] + dra[:, None]
    dec = dec_b[None, :] + ddec[:, None]
    return ra, dec


def main():
    out = {}
    cases = [
        # name, ra0 deg, dec0 deg, pix arcsec, ndet, n, seed
        ("c0_2arcsec", 180.0, 60.0, 2.0, 16, 512, 1),
        ("c1_1arcsec", 180.0, 60.0, 1.0, 16, 512, 2),
        ("south", 45.0, -30.0, 1.0, 16, 512, 3),
        ("equator", 10.0, 0.2, 1.5, 16, 512, 4),
        ("highdec", 300.0, 80.0, 1.0, 16, 512, 5),
        ("ra_small", 0.4, 20.0, 2.0, 16, 512, 6),
    ]
    arcsec = np.pi / 180 / 3600
    for name, ra0d, dec0d, pixas, ndet, n, seed in cases:
        rng = np.random.default_rng(seed)
        ra0, dec0 = np.deg2rad(ra0d), np.deg2rad(dec0d)
        ra, dec = raster(rng, ndet, n, ra0, dec0, 360.0)
        lims = [ra.min(), ra.max(), dec.min(), dec.max()]
        pixsize = pixas * arcsec
        w = make_wcs(lims, pixsize)
        nx, ny = npix(w, lims, 2)
        nn = ndet * n
        coords = np.zeros([nn, 2])
        coords[:, 0] = np.reshape(ra * 180 / np.pi, nn)
        coords[:, 1] = np.reshape(dec * 180 / np.pi, nn)
        pix = np.asarray(w.wcs_world2pix(coords, 1))
        xpix = np.reshape(pix[:, 0], [ndet, n]) - 1
        ypix = np.reshape(pix[:, 1], [ndet, n]) - 1
        ipix = np.asarray(np.round(xpix) * ny + np.round(ypix), dtype="int32")
        w.wcs.set()
        out[name + "_ra"] = ra
        out[name + "_dec"] = dec
        out[name + "_lims"] = np.asarray(lims)
        out[name + "_pixsize"] = np.asarray(pixsize)
        out[name + "_nxny"] = np.asarray([nx, ny])
        out[name + "_px"] = pix[:, 0].reshape(ndet, n)
        out[name + "_py"] = pix[:, 1].reshape(ndet, n)
        out[name + "_ipix"] = ipix
        out[name + "_lonlatpole"] = np.asarray([w.wcs.lonpole, w.wcs.latpole])
    out["cases"] = np.asarray([c[0] for c in cases])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pix_car_fixtures.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, "astropy", sys.modules["astropy"].__version__)


if __name__ == "__main__":
    main()

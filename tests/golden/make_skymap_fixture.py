"""Writes the FITS file SkyMap.write must reproduce, with astropy (the
library the reference writes maps with, minkasi/maps/skymap.py:705-722):

    tests/golden/skymap_write.fits   PrimaryHDU(map.T, header=wcs.to_header())
    tests/golden/skymap_write.npz    lims, pixsize, the map (nx, ny), and the
                                     header cards astropy wrote

The WCS is built with the steps of minkasi/maps/utils.py:13-67 (CAR,
ref_equ=False: crpix [1, 1], crval (lims[1], lims[2]) in degrees, cdelt
[-pix, pix] in degrees, cosdec forced to 1); the reference is not imported.
Run with an interpreter that has astropy (here /opt/conda/bin/python3.9,
astropy 4.3.1, with small numpy-compat shims):

    /opt/conda/bin/python3.9 tests/golden/make_skymap_fixture.py
"""
import os

import numpy as np

np.asscalar = lambda a: a.item()
np.alen = lambda a: len(a)
for _n, _t in [("float", float), ("int", int), ("bool", bool), ("object", object), ("str", str),
               ("complex", complex)]:
    setattr(np, _n, _t)

from astropy import wcs  # noqa: E402
from astropy.io import fits  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
rng = np.random.default_rng([2024, 12])
arcsec = np.pi / 180 / 3600
lims = [np.deg2rad(179.95), np.deg2rad(180.05), np.deg2rad(59.97), np.deg2rad(60.03)]
pixsize = 2 * arcsec
w = wcs.WCS(naxis=2)
w.wcs.crpix = [1.0, 1.0]
w.wcs.crval = [lims[1] * 180 / np.pi, lims[2] * 180 / np.pi]
w.wcs.cdelt = [-pixsize / 1.0 * 180 / np.pi, pixsize * 180 / np.pi]
w.wcs.ctype = ["RA---CAR", "DEC--CAR"]
nx, ny = 7, 5
m = rng.standard_normal((nx, ny))
hdu = fits.PrimaryHDU(m.transpose().copy(), header=w.to_header())
fname = os.path.join(HERE, "skymap_write.fits")
hdu.writeto(fname, overwrite=True)
h = fits.getheader(fname)
keys = np.array([k for k in h.keys()])
vals = np.array([repr(h[k]) for k in h.keys()])
np.savez(os.path.join(HERE, "skymap_write.npz"), lims=np.array(lims), pixsize=pixsize, map=m, keys=keys,
         vals=vals)
print("wrote", fname, "cards:", list(zip(keys, vals)))

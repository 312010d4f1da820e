"""Host-side CarWCS (product) against the astropy fixtures and the oracle."""
import os

import numpy as np

from minkasi_amd.maps.wcs import CarWCS
from oracle import core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_carwcs_matches_oracle_and_fixtures():
    z = np.load(os.path.join(ROOT, "tests/golden/pix_car_fixtures.npz"))
    for c in z["cases"]:
        lims = z[c + "_lims"]
        pix = float(z[c + "_pixsize"])
        w = CarWCS.from_lims(lims, pix)
        wo = core.car_wcs(lims, pix)
        assert np.array_equal(w.crval, wo["crval"]) and np.array_equal(w.cdelt, wo["cdelt"])
        assert np.array_equal(w.euler, core.car_euler(wo)), c
        ra = z[c + "_ra"][:, :64].ravel()
        dec = z[c + "_dec"][:, :64].ravel()
        px, py = w.world2pix(ra, dec)
        qx, qy = core.car_world2pix(wo, ra, dec)
        assert np.abs(px - qx).max() < 1e-9 and np.abs(py - qy).max() < 1e-9
        assert (w.lonpole, round(w.latpole, 9)) == (z[c + "_lonlatpole"][0], round(z[c + "_lonlatpole"][1], 9))

"""Pins the CPU oracle against golden vectors, the reference's own C code
(oracle/_ref, when built here) and closed forms."""
import ctypes

import numpy as np
import pytest
import scipy.fft

from oracle import core
from tests.helpers import raster_pointing

FIX = "tests/golden/pix_car_fixtures.npz"


@pytest.fixture(scope="module")
def fixtures():
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return np.load(os.path.join(root, FIX))


def test_car_pixels_match_astropy_fixtures(fixtures):
    z = fixtures
    for c in z["cases"]:
        lims = z[c + "_lims"]
        w = core.car_wcs(lims, float(z[c + "_pixsize"]))
        nx, ny = core.car_npix(w, lims, 2)
        assert (nx, ny) == tuple(int(v) for v in z[c + "_nxny"]), c
        ip = core.pix_from_radec(w, ny, z[c + "_ra"], z[c + "_dec"])
        assert np.array_equal(ip, z[c + "_ipix"]), c
        px, py = core.car_world2pix(w, z[c + "_ra"], z[c + "_dec"])
        # raw coordinates agree with wcslib to ~1e-9 px (last-ulp differences
        # in the degree arithmetic); ipix equality above is the contract
        assert np.abs(px - z[c + "_px"]).max() < 1e-8, c
        assert np.abs(py - z[c + "_py"]).max() < 1e-8, c
        lonpole, latpole = z[c + "_lonlatpole"]
        eul = core.car_euler(w)
        assert eul[2] == lonpole
        assert abs((90.0 - eul[1]) - latpole) < 1e-12


def test_car_pixels_catch_planted_mistakes(fixtures):
    """A transposed axis or a truncating round must fail the fixtures."""
    z = fixtures
    c = z["cases"][0]
    lims = z[c + "_lims"]
    w = core.car_wcs(lims, float(z[c + "_pixsize"]))
    nx, ny = core.car_npix(w, lims, 2)
    px, py = core.car_world2pix(w, z[c + "_ra"], z[c + "_dec"])
    good = (np.round(px - 1) * ny + np.round(py - 1)).astype(np.int32)
    assert np.array_equal(good, z[c + "_ipix"])
    swapped = (np.round(py - 1) * nx + np.round(px - 1)).astype(np.int32)
    trunc = (np.floor(px - 1) * ny + np.floor(py - 1)).astype(np.int32)
    assert not np.array_equal(swapped, z[c + "_ipix"])
    assert not np.array_equal(trunc, z[c + "_ipix"])


def _ref_or_skip():
    L = core.ref_lib()
    if L is None:
        pytest.skip("reference minkasi.c not built into oracle/_ref (reference tree absent)")
    return L


def test_projections_match_closed_form_and_reference():
    rng = np.random.default_rng(5)
    ndet, n, npix = 8, 3001, 977
    ipix = rng.integers(0, npix, size=(ndet, n)).astype(np.int32)
    m = rng.standard_normal(npix)
    d = np.zeros((ndet, n))
    core.map2tod(d, m, ipix, do_add=False)
    assert np.array_equal(d, m[ipix])
    d2 = d.copy()
    core.map2tod(d2, m, ipix, do_add=True)
    assert np.array_equal(d2, 2 * m[ipix])
    tod = rng.standard_normal((ndet, n))
    mp = np.zeros(npix)
    core.tod2map(mp, tod, ipix)
    ref = np.bincount(ipix.ravel(), tod.ravel(), minlength=npix)
    assert np.allclose(mp, ref, rtol=1e-12, atol=1e-12)
    L = core.ref_lib()
    if L is not None:
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        dr = np.zeros((ndet, n))
        L.map2tod_simple(p(dr), p(m), ndet, n, p(ipix), 0)
        assert np.array_equal(dr, d)
        mr = np.zeros(npix)
        L.tod2map_simple(p(mr), p(tod), ndet, n, p(ipix))
        assert np.allclose(mr, mp, rtol=1e-12, atol=1e-12)
        ms = np.zeros(npix)
        core.tod2map_simple(ms, tod, ipix)
        assert np.array_equal(ms, mr)  # same serial order -> bitwise equal


def test_dct1_definition_matches_scipy_and_involution():
    rng = np.random.default_rng(1)
    for n in (2, 3, 5, 17, 64, 65):
        x = rng.standard_normal((3, n))
        y = core.dct1_direct(x)
        ys = scipy.fft.dct(x, type=1, axis=-1)
        assert np.allclose(y, ys, rtol=1e-11, atol=1e-11 * np.abs(ys).max()), n
        if n > 1:
            yy = core.dct1_direct(y)
            assert np.allclose(yy, 2 * (n - 1) * x, rtol=1e-10, atol=1e-10 * n), n
    # known answer: DCT-I of a delta at 0 is all ones; of ones is 2(n-1) delta
    n = 9
    e0 = np.zeros((1, n)); e0[0, 0] = 1
    assert np.allclose(core.dct1_direct(e0), np.ones((1, n)))
    ones = np.ones((1, n))
    expect = np.zeros((1, n)); expect[0, 0] = 2 * (n - 1)
    assert np.allclose(core.dct1_direct(ones), expect, atol=1e-12)


def test_oracle_pix_large_raster_in_range():
    rng = np.random.default_rng(0)
    ra_b, dec_b, dra, ddec = raster_pointing(rng, 16, 8192)
    ra = ra_b[None, :] + dra[:, None]
    dec = dec_b[None, :] + ddec[:, None]
    lims = [ra.min(), ra.max(), dec.min(), dec.max()]
    w = core.car_wcs(lims, 2 * np.pi / 180 / 3600)
    nx, ny = core.car_npix(w, lims, 2)
    ip = core.pix_from_radec(w, ny, ra, dec)
    assert ip.min() >= 0 and ip.max() < nx * ny

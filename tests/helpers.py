"""Shared test helpers (inputs and comparisons)."""
import numpy as np

ARCSEC = np.pi / 180 / 3600


def raster_pointing(rng, ndet, n, ra0_deg=180.0, dec0_deg=60.0, half_width=360.0, fs=100.0, speed=30.0,
                    radius=60.0):
    """Constant-velocity RA raster with a disc of detector offsets (radians)."""
    ra0, dec0 = np.deg2rad(ra0_deg), np.deg2rad(dec0_deg)
    t = np.arange(n) / fs
    leg = 2 * half_width / speed
    ph = (t / leg) % 2.0
    tri = np.where(ph < 1.0, ph, 2.0 - ph) * 2.0 - 1.0
    ra_b = ra0 + tri * half_width * ARCSEC / np.cos(dec0)
    legno = np.floor(t / leg)
    dec_b = dec0 + (((legno * 10.0) % 480.0) - 240.0) * ARCSEC
    r = radius * np.sqrt(rng.uniform(0, 1, ndet))
    th = rng.uniform(0, 2 * np.pi, ndet)
    ddec = r * np.sin(th) * ARCSEC
    dra = r * np.cos(th) * ARCSEC / np.cos(dec0)
    return ra_b, dec_b, dra, ddec


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def max_rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    mb = np.abs(b).max()
    return np.abs(a - b).max() / (mb if mb > 0 else 1.0)

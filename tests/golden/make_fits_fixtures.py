"""Writes small FITS TODs with astropy.io.fits (the library the reference
reads them with) and the values the reference's readers derive from them:

    tests/golden/mustang_tod.fits        MUSTANG layout, calinfo keywords, masked samples
    tests/golden/mustang_tod_nocal.fits  no calinfo, no ELEV, extra unsigned / logical / string columns
    tests/golden/cbass_tod.fits          C-BASS layout with the nm20 cut table in HDU 3
    tests/golden/fits_expected.npz       expected arrays (astropy's reading + the readers' steps)

Run with an interpreter that has astropy (here /opt/conda/bin/python3.9,
astropy 4.3.1, with small numpy-compat shims):

    /opt/conda/bin/python3.9 tests/golden/make_fits_fixtures.py

The expected values restate the steps of minkasi/tods/io.py:160-284
(read_tod_from_fits) and :40-157 (read_tod_from_fits_cbass, without
qpoint) on the columns astropy returns; the reference is not imported.
"""
import json
import os

import numpy as np

np.asscalar = lambda a: a.item()
np.alen = lambda a: len(a)
for _n, _t in [("float", float), ("int", int), ("bool", bool), ("object", object), ("str", str),
               ("complex", complex)]:
    setattr(np, _n, _t)

from astropy.io import fits  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
rng = np.random.default_rng([2024, 11])


def mustang(ndet, ndata, cal, elev, extra, masked):
    pixid = np.repeat(rng.permutation(np.arange(100, 100 + 3 * ndet, 3))[:ndet], ndata).astype(np.int32)
    t = np.tile(np.arange(ndata) * 0.01 + 5.0e4, ndet) + rng.normal(0, 1e-6, ndet * ndata)
    dx = (np.deg2rad(359.99) + rng.normal(0, 2e-4, ndet * ndata)) % (2 * np.pi)   # straddles RA = 0
    dy = np.deg2rad(-5.0) + rng.normal(0, 2e-4, ndet * ndata)
    fnu = rng.normal(0, 1, ndet * ndata).astype(np.float32)
    ufnu = np.abs(rng.normal(0, 1, ndet * ndata)).astype(np.float32)
    if masked:
        ufnu[rng.choice(ndet * ndata, 7, replace=False)] = 1e6
    cols = [fits.Column(name="PIXID", format="J", array=pixid),
            fits.Column(name="DX", format="D", array=dx),
            fits.Column(name="DY", format="D", array=dy),
            fits.Column(name="TIME", format="D", array=t),
            fits.Column(name="FNU", format="E", array=fnu),
            fits.Column(name="UFNU", format="E", array=ufnu)]
    if elev:
        cols.append(fits.Column(name="ELEV", format="E", array=(45 + rng.normal(0, 1, ndet * ndata)).astype(np.float32)))
    if extra:
        cols.append(fits.Column(name="COUNTS", format="I", bzero=32768, array=rng.integers(0, 65535, ndet * ndata).astype(np.uint16)))
        cols.append(fits.Column(name="GOOD", format="L", array=rng.random(ndet * ndata) > 0.3))
        cols.append(fits.Column(name="TAG", format="6A", array=np.array(["ab", "cdef", "x y"] * (ndet * ndata // 3) + ["z"] * (ndet * ndata % 3))))
        cols.append(fits.Column(name="VEC", format="3E", array=rng.normal(0, 1, (ndet * ndata, 3)).astype(np.float32)))
        cols.append(fits.Column(name="SCALED", format="J", array=rng.integers(-400, 400, ndet * ndata).astype(np.int32)))
    hdu = fits.BinTableHDU.from_columns(cols)
    if cal:
        for i, k in enumerate(("scan", "bunit", "azimuth", "elevatio", "beameff", "apereff", "antgain", "gainunc",
                               "bmaj", "bmin", "bpa", "parang", "beamvol", "beamvunc")):
            hdu.header[k] = "K'S" if k == "bunit" else (17 if k == "scan" else 0.5 + 1.25 * i)
    return fits.HDUList([fits.PrimaryHDU(), hdu])


def expect_mustang(fname, branch, out, tag):
    with fits.open(fname) as hdul:
        raw = hdul[1].data
        pixid = np.array(raw["PIXID"])
        ndet = len(np.unique(pixid))
        ndata = len(pixid) // ndet
        out[tag + "pixid"] = pixid.reshape(ndet, ndata)[:, 0]
        dx = np.array(raw["DX"], dtype=np.float64)
        if branch is not None:
            bb = branch * np.pi / 180
            dx[dx > bb] -= 2 * np.pi
        out[tag + "dx"] = dx.reshape(ndet, ndata)
        out[tag + "dy"] = np.array(raw["DY"], dtype=np.float64).reshape(ndet, ndata)
        if "ELEV" in raw.names:
            out[tag + "elev"] = (np.array(raw["ELEV"]) * np.pi / 180).astype(np.float64).reshape(ndet, ndata)
        out[tag + "dt"] = np.median(np.diff(np.array(raw["TIME"]).reshape(ndet, ndata)[0]))
        out[tag + "dat_calib"] = np.array(raw["FNU"], dtype=np.float64).reshape(ndet, ndata)
        ufnu = np.array(raw["UFNU"])
        if (ufnu > 9e5).any():
            m = (ufnu < 9e5).reshape(ndet, ndata)
            out[tag + "mask"] = m
            out[tag + "mask_sum"] = m.sum(axis=0)
        for c in ("COUNTS", "GOOD", "TAG", "VEC", "SCALED"):
            if c in raw.names:
                v = np.array(raw[c])
                out[tag + "col_" + c] = v.astype(str) if v.dtype.kind in "SU" else v
        cal = {}
        try:
            for k in ("scan", "bunit", "azimuth", "elevatio", "beameff", "apereff", "antgain", "gainunc", "bmaj",
                      "bmin", "bpa", "parang", "beamvol", "beamvunc"):
                cal[k] = hdul[1].header[k]
        except KeyError:
            pass
        out[tag + "calinfo"] = np.array(json.dumps(cal))


def cbass(n):
    mjd = 56000.0 + np.arange(n) * (0.01 / 86400) + rng.normal(0, 1e-12, n)
    cols = [fits.Column(name=k, format="D", array=v) for k, v in (
        ("RA", rng.uniform(0, 6, n)), ("DEC", rng.uniform(-1, 1, n)), ("MJD", mjd),
        ("I1", rng.normal(0, 1, n)), ("I2", rng.normal(0, 1, n)), ("Q1", rng.normal(0, 1, n)),
        ("Q2", rng.normal(0, 1, n)), ("U1", rng.normal(0, 1, n)), ("U2", rng.normal(0, 1, n)),
        ("AZ", rng.uniform(0, 360, n)), ("EL", rng.uniform(20, 80, n)))]
    cols.append(fits.Column(name="FLAG", format="J", array=(rng.random(n) > 0.8).astype(np.int32)))
    start = np.array([3, 20, 41], dtype=np.int32)
    end = np.array([15, 33, 57], dtype=np.int32)
    cut = fits.BinTableHDU.from_columns([fits.Column(name="START", format="J", array=start),
                                         fits.Column(name="END", format="J", array=end)])
    empty = fits.BinTableHDU.from_columns([fits.Column(name="X", format="E", array=np.zeros(1, np.float32))])
    return fits.HDUList([fits.PrimaryHDU(), fits.BinTableHDU.from_columns(cols), empty, cut])


def expect_cbass(fname, out):
    with fits.open(fname) as hdul:
        raw = hdul[1].data
        ra, dec, mjd = np.array(raw["RA"]), np.array(raw["DEC"]), np.array(raw["MJD"])
        tvec = (mjd - 2455977.5 + 2400000.5) * 86400 + 1329696000
        n = len(ra)
        out["cb_dx"], out["cb_dy"] = ra.reshape(1, n), dec.reshape(1, n)
        out["cb_dt"], out["cb_ctime"] = np.median(np.diff(tvec)), tvec
        out["cb_I"] = (0.5 * (raw["I1"] + raw["I2"])).reshape(1, n)
        Q, U = 0.5 * (raw["Q1"] + raw["Q2"]), 0.5 * (raw["U1"] + raw["U2"])
        out["cb_pol_v34"] = np.vstack([-U, Q])
        out["cb_pol"] = np.vstack([Q, U])
        out["cb_az"], out["cb_el"] = raw["AZ"] * np.pi / 180, raw["EL"] * np.pi / 180
        flag = np.array(raw["FLAG"])
        out["cb_mask"] = (1 - flag).astype(np.int8).reshape(1, n)
        keep = np.zeros(n, dtype=np.int8)
        for a, b in zip(hdul[3].data["START"], hdul[3].data["END"]):
            keep[a:b] = 1
        out["cb_mask_nm20"] = (1 - flag).astype(np.int8).reshape(1, n) * keep


if __name__ == "__main__":
    out = {}
    f1 = os.path.join(HERE, "mustang_tod.fits")
    mustang(5, 48, cal=True, elev=True, extra=False, masked=True).writeto(f1, overwrite=True)
    expect_mustang(f1, None, out, "m1_")
    expect_mustang(f1, 180.0, out, "m1b_")
    f2 = os.path.join(HERE, "mustang_tod_nocal.fits")
    mustang(3, 30, cal=False, elev=False, extra=True, masked=False).writeto(f2, overwrite=True)
    with fits.open(f2, mode="update") as hdul:  # SCALED: stored integers, read back as 0.25 * raw + 10
        k = 1 + hdul[1].columns.names.index("SCALED")
        hdul[1].header["TSCAL%d" % k] = 0.25
        hdul[1].header["TZERO%d" % k] = 10.0
    expect_mustang(f2, None, out, "m2_")
    f3 = os.path.join(HERE, "cbass_tod.fits")
    cbass(64).writeto(f3, overwrite=True)
    expect_cbass(f3, out)
    np.savez_compressed(os.path.join(HERE, "fits_expected.npz"), **out)
    print("wrote", sorted(out))

"""TOD readers (minkasi/tods/io.py) against FITS files written by astropy
and the values astropy reads back (tests/golden/make_fits_fixtures.py), and
the host-side helpers (gapfill_lin, octave structs, blacklists)."""
import json
import os

import numpy as np
import pytest

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def exp():
    return np.load(os.path.join(G, "fits_expected.npz"))


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if a.dtype.kind in "SU" or b.dtype.kind in "SU":
        assert list(a.astype(str).ravel()) == list(b.astype(str).ravel())
    else:
        assert np.array_equal(a, b), (a.ravel()[:5], b.ravel()[:5])


@pytest.mark.parametrize("branch,tag", [(None, "m1_"), (180.0, "m1b_")])
def test_mustang_reader_matches_astropy(exp, branch, tag):
    from minkasi_amd.tods.io import read_tod_from_fits
    dat = read_tod_from_fits(os.path.join(G, "mustang_tod.fits"), branch=branch)
    for k in ("pixid", "dx", "dy", "elev", "dat_calib", "mask", "mask_sum"):
        _same(dat[k], exp[tag + k])
    assert dat["dx"].dtype == np.float64 and dat["dat_calib"].dtype == np.float64
    assert dat["dx"].flags.c_contiguous and dat["dat_calib"].flags.c_contiguous
    assert dat["dt"] == float(exp[tag + "dt"])
    cal = json.loads(str(exp[tag + "calinfo"]))
    assert dat["calinfo"]["calinfo"] is True
    assert {k: v for k, v in dat["calinfo"].items() if k != "calinfo"} == cal
    if branch is not None:  # every RA above the branch moved down by 2 pi
        assert dat["dx"].max() <= np.deg2rad(branch) and dat["dx"].min() < 0


def test_mustang_without_calinfo_elev_or_cuts(exp, capsys):
    from minkasi_amd.tods.io import read_tod_from_fits
    dat = read_tod_from_fits(os.path.join(G, "mustang_tod_nocal.fits"))
    for k in ("pixid", "dx", "dy", "dat_calib"):
        _same(dat[k], exp["m2_" + k])
    assert "elev" not in dat and "mask" not in dat
    # the reference leaves calinfo['calinfo'] True even when the keywords are missing
    assert dat["calinfo"] == {"calinfo": True}
    assert "calinfo information not found" in capsys.readouterr().out


def test_bintable_column_types_match_astropy(exp):
    """Unsigned (TZERO 32768), logical, character, vector and TSCAL/TZERO-
    scaled columns read as astropy reads them."""
    from minkasi_amd.tods.io import fits_open
    h = fits_open(os.path.join(G, "mustang_tod_nocal.fits"))
    assert len(h) == 2 and h[0].data is None
    t = h[1].data
    for c in ("COUNTS", "GOOD", "TAG", "VEC", "SCALED"):
        _same(t[c], exp["m2_col_" + c])
    assert t["COUNTS"].dtype == np.uint16 and t["GOOD"].dtype == bool and t["SCALED"].dtype == np.float64
    assert t["pixid"].shape == (90,)  # case-insensitive names
    with pytest.raises(KeyError):
        t["NOPE"]
    with pytest.raises(KeyError):
        h[1].header["BEAMVOL"]


@pytest.mark.parametrize("dopol,v34", [(False, True), (True, True), (True, False)])
def test_cbass_reader_matches_astropy(exp, dopol, v34):
    from minkasi_amd.tods.io import read_tod_from_fits_cbass
    dat = read_tod_from_fits_cbass(os.path.join(G, "cbass_tod.fits"), dopol=dopol, v34=v34)
    assert dat["dt"] == float(exp["cb_dt"])
    _same(dat["ctime"], exp["cb_ctime"])
    if dopol:
        _same(dat["dx"], np.vstack([exp["cb_dx"]] * 2))
        _same(dat["dat_calib"], exp["cb_pol_v34"] if v34 else exp["cb_pol"])
        _same(dat["az"], exp["cb_az"])
        _same(dat["el"], exp["cb_el"])
        _same(dat["mask"], np.vstack([exp["cb_mask"]] * 2))
    else:
        _same(dat["dx"], exp["cb_dx"])
        _same(dat["dy"], exp["cb_dy"])
        _same(dat["dat_calib"], exp["cb_I"])
        _same(dat["mask"], exp["cb_mask"])
    assert "ra" not in dat and "twogamma_saved" not in dat  # qpoint is not available


def test_cbass_nm20_cuts(exp):
    from minkasi_amd.tods.io import read_tod_from_fits_cbass
    dat = read_tod_from_fits_cbass(os.path.join(G, "cbass_tod.fits"), nm20=True)
    _same(dat["mask"], exp["cb_mask_nm20"])
    assert list(dat["nm20_start"]) == [3, 20, 41] and list(dat["nm20_stop"]) == [15, 33, 57]


def test_fits_open_rejects_non_fits(tmp_path):
    from minkasi_amd.tods.io import fits_open
    p = tmp_path / "x.fits"
    p.write_bytes(b"NOTFITS = T" + b" " * 2869)
    with pytest.raises(ValueError):
        fits_open(str(p))
    q = tmp_path / "y.fits"
    q.write_bytes(b"SIMPLE  =                    T" + b" " * 2850)  # no END card
    with pytest.raises(ValueError):
        fits_open(str(q))


def test_fits_header_values():
    from minkasi_amd.tods.io import _parse_value
    assert _parse_value("                   T / flag") is True
    assert _parse_value("                  42 / n") == 42
    assert _parse_value("             1.5D-03") == 1.5e-3
    assert _parse_value("'O''HARA '           / name") == "O'HARA"
    assert _parse_value("") is None


def test_gapfill_lin_interior_runs_are_linear():
    """Interior NaN runs become the straight line between their neighbours;
    cuts name those neighbours (tods/processing.py:286-354)."""
    from minkasi_amd.tods.processing import gapfill_lin
    x = np.arange(20, dtype=float) ** 2
    d = np.vstack([x, x + 1, x * 0 + 3])
    d[0, 4:7] = np.nan
    d[0, 10] = np.nan
    d[1, 15:18] = np.nan
    out, cuts = gapfill_lin(d.copy())
    assert cuts == {0: [[3, 7], [9, 11]], 1: [[14, 18]]}
    ref = np.vstack([x, x + 1, x * 0 + 3])
    for (r, lo, hi) in ((0, 3, 7), (0, 9, 11), (1, 14, 18)):
        t = np.arange(lo + 1, hi)
        line = ref[r, lo] + (ref[r, hi] - ref[r, lo]) * (t - lo) / (hi - lo)
        assert np.allclose(out[r, lo + 1:hi], line, rtol=0, atol=1e-12)
    mask = ~np.isnan(d)
    assert np.array_equal(out[mask], ref[mask])


def test_gapfill_lin_edge_cases_follow_the_reference():
    from minkasi_amd.tods.processing import gapfill_lin
    d = np.ones((1, 8))
    d[0, :2] = np.nan  # a run at the start: left edge wraps to the last sample, nothing is filled
    out, cuts = gapfill_lin(d.copy())
    assert cuts == {0: [[7, 2]]} and np.isnan(out[0, :2]).all()
    e = np.ones((1, 8))
    e[0, 6:] = np.nan  # a run at the end has no right neighbour
    with pytest.raises(IndexError):
        gapfill_lin(e)
    with pytest.raises(ValueError):
        gapfill_lin(np.full((1, 5), np.nan))


def _write_octave(path, fields):
    with open(path, "wb") as f:
        np.array([len(fields)], dtype=np.int32).tofile(f)
        for key, arr, code in fields:
            f.write((key + "\n").encode())
            dims = np.array(arr.shape[::-1], dtype=np.int32)  # stored reversed
            np.array([len(dims)], dtype=np.int32).tofile(f)
            dims.tofile(f)
            np.array([code], dtype=np.int32).tofile(f)
            arr.tofile(f)


def test_octave_struct_round_trip(tmp_path):
    from minkasi_amd.tods.io import read_octave_struct, todvec_from_files_octave
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 7))
    b = rng.integers(-5, 5, size=(2, 4)).astype(np.int32)
    c = rng.normal(size=(5,)).astype(np.float32)
    p = str(tmp_path / "tod.oct")
    _write_octave(p, [("dat_calib", a, 8), ("pixid", b, -4), ("dt", c, 4)])
    dat = read_octave_struct(p)
    assert np.array_equal(dat["dat_calib"], a) and np.array_equal(dat["pixid"], b)
    assert np.array_equal(dat["dt"], c) and dat["dt"].dtype == np.float32
    tv = todvec_from_files_octave([p, p])
    assert tv.ntod == 2 and np.array_equal(tv.tods[1].info["dat_calib"], a)


def test_cut_blacklist(capsys):
    from minkasi_amd.tods.io import cut_blacklist
    names = ["/d/b.fits", "/d/a.fits", "/e/c.fits"]
    assert cut_blacklist(names, ["x/a.fits", "zz.fits"]) == ["/d/b.fits", "/e/c.fits"]
    assert "deleted  1  bad files." in capsys.readouterr().out


def test_skymap_write_matches_astropy_fixture(tmp_path):
    """SkyMap.write's file against the one astropy writes for the same CAR
    WCS and map (fits.PrimaryHDU(map.T, header=wcs.to_header()),
    minkasi/maps/skymap.py:705-722; tests/golden/make_skymap_fixture.py):
    the same cards in the same order with the same values (astropy prints
    14 significant digits), and the same data, bit for bit."""
    from minkasi_amd.maps.skymap import write_map_fits
    from minkasi_amd.maps.wcs import get_wcs
    from minkasi_amd.tods.io import fits_open
    fx = np.load(os.path.join(G, "skymap_write.npz"))
    w = get_wcs(list(fx["lims"]), float(fx["pixsize"]))
    fname = str(tmp_path / "map.fits")
    write_map_fits(fname, fx["map"], w)
    ours, ref = fits_open(fname)[0], fits_open(os.path.join(G, "skymap_write.fits"))[0]
    keys = [c[:8].strip() for c in ref.header.cards]
    assert keys == list(fx["keys"])
    assert [c[:8].strip() for c in ours.header.cards] == keys
    for k in keys:
        a, b = ours.header[k], ref.header[k]
        if isinstance(b, float):
            assert abs(a - b) <= 1e-12 * max(abs(b), 1e-300), (k, a, b)
        else:
            assert a == b and type(a) == type(b), (k, a, b)
    assert ours.data.dtype == ref.data.dtype and np.array_equal(ours.data, ref.data)
    assert np.array_equal(ours.data.T, fx["map"])
    with open(fname, "rb") as f:
        assert len(f.read()) % 2880 == 0

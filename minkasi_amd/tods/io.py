"""Loading TODs from disk (minkasi/tods/io.py).

The reference reads FITS through astropy.io.fits.  astropy is not part of
this package's dependencies, so FITS files are read by a small reader of
the standard's primary array / IMAGE / BINTABLE layout written here
(`fits_open`): 2880-byte blocks of 80-character header cards, big-endian
data, binary-table rows of fixed-width fields (FITS Standard 4.0, sections
4, 5 and 7.3).  Table columns come back as numpy arrays with TSCALn/TZEROn
applied as astropy applies them (unsigned integers for the 2^(b-1) zero
offsets, float64 otherwise).  Heap (variable-length array) columns are not
supported.

The readers return the reference's info dicts (numpy, FP64 pointing and
data as the reference casts them) for Tod(info); moving the data to the
device is Tod's job.
"""
import re

import numpy as np

from .core import Tod, TodVec

_BLOCK = 2880

# BINTABLE TFORM type codes -> (numpy big-endian type, bytes per element)
_TFORM = {"L": ("i1", 1), "B": ("u1", 1), "I": (">i2", 2), "J": (">i4", 4), "K": (">i8", 8), "A": ("S1", 1),
          "E": (">f4", 4), "D": (">f8", 8), "C": (">c8", 8), "M": (">c16", 16)}
_BITPIX = {8: "u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def _parse_value(s):
    """Value field of a header card (columns 11-80): string, logical,
    integer or real; the comment after '/' is dropped."""
    s = s.strip()
    if s.startswith("'"):
        out, i = [], 1
        while i < len(s):
            if s[i] == "'":
                if i + 1 < len(s) and s[i + 1] == "'":  # '' is a quote inside the string
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(s[i])
            i += 1
        return "".join(out).rstrip()
    s = s.split("/", 1)[0].strip()
    if s == "T":
        return True
    if s == "F":
        return False
    if s == "":
        return None
    try:
        return int(s)
    except ValueError:
        pass
    try:
        return float(s.replace("D", "E").replace("d", "e"))
    except ValueError:
        return s


class FitsHeader(dict):
    """Header keywords -> values (the last card wins); keys upper case.
    Indexing a missing keyword raises KeyError, as astropy's headers do."""

    def __init__(self):
        super().__init__()
        self.cards = []

    def __getitem__(self, key):
        return dict.__getitem__(self, key.upper())

    def __contains__(self, key):
        return dict.__contains__(self, key.upper())

    def get(self, key, default=None):
        return dict.get(self, key.upper(), default)


def _read_header(buf, pos):
    """Parses the header starting at byte pos; returns (header, offset of
    the data that follows it)."""
    hdr = FitsHeader()
    while True:
        if pos + _BLOCK > len(buf):
            raise ValueError("FITS header runs past the end of the file (no END card)")
        block = bytes(buf[pos:pos + _BLOCK]).decode("ascii", errors="replace")
        pos += _BLOCK
        for c in range(0, _BLOCK, 80):
            card = block[c:c + 80]
            key = card[:8].strip().upper()
            if key == "END":
                return hdr, pos
            hdr.cards.append(card)
            if card[8:10] == "= ":
                dict.__setitem__(hdr, key, _parse_value(card[10:]))


class FitsTable:
    """Rows of a BINTABLE HDU; table[name] is one column (case-insensitive
    name) as a numpy array of shape (rows,) or (rows, repeat)."""

    def __init__(self, rec, names, scale, zero, kinds):
        self._rec = rec
        self.names = names
        self._lower = {n.lower(): n for n in names}
        self._scale, self._zero, self._kinds = scale, zero, kinds

    def __len__(self):
        return len(self._rec)

    def __getitem__(self, name):
        key = self._lower.get(str(name).lower())
        if key is None:
            raise KeyError("no column %r in the table (columns: %s)" % (name, ", ".join(self.names)))
        col = np.asarray(self._rec[key])
        kind = self._kinds.get(key)
        if kind == "L":  # logical: 'T' true, 'F' or 0 false
            return col == ord("T")
        if kind == "A":  # character: trailing blanks dropped
            return np.char.rstrip(col.astype(str))
        sc, z = self._scale.get(key, 1.0), self._zero.get(key, 0.0)
        if col.dtype.kind in "iu" and sc == 1.0 and z != 0.0:
            bits = 8 * col.dtype.itemsize
            if col.dtype.kind == "i" and z == float(1 << (bits - 1)):   # unsigned integers in signed storage
                return (col.astype(col.dtype.newbyteorder("=")).view("u%d" % col.dtype.itemsize)
                        ^ np.array(1 << (bits - 1), dtype="u%d" % col.dtype.itemsize))
            if col.dtype.kind == "u" and bits == 8 and z == -128.0:     # signed bytes in unsigned storage
                return (col ^ np.uint8(0x80)).view(np.int8)
        if sc != 1.0 or z != 0.0:
            return col.astype(np.float64) * sc + z
        return col.astype(col.dtype.newbyteorder("="))


class FitsHDU:
    def __init__(self, header, data):
        self.header = header
        self.data = data


def _table_dtype(hdr):
    nf = int(hdr.get("TFIELDS", 0))
    names, formats, offsets, scale, zero, kinds = [], [], [], {}, {}, {}
    off = 0
    for i in range(1, nf + 1):
        form = str(hdr["TFORM%d" % i]).strip().upper()
        m = re.match(r"^(\d*)([LXBIJKAEDCMPQ])", form)
        if not m:
            raise ValueError("unsupported TFORM%d %r" % (i, form))
        rep = int(m.group(1)) if m.group(1) else 1
        code = m.group(2)
        name = str(hdr.get("TTYPE%d" % i, "col%d" % i)).strip() or "col%d" % i
        if code in "PQ":
            raise ValueError("column %s: variable-length (heap) arrays are not supported" % name)
        if code == "X":
            nbytes = (rep + 7) // 8
            dt = np.dtype(("u1", (nbytes,))) if nbytes > 1 else np.dtype("u1")
        elif code == "A":
            dt, nbytes = np.dtype("S%d" % max(rep, 1)), rep
        else:
            t, size = _TFORM[code]
            dt = np.dtype((t, (rep,))) if rep > 1 else np.dtype(t)
            nbytes = rep * size
        if nbytes == 0:
            continue
        names.append(name)
        formats.append(dt)
        offsets.append(off)
        off += nbytes
        if code in "LA":
            kinds[name] = code
        if "TSCAL%d" % i in hdr:
            scale[name] = float(hdr["TSCAL%d" % i])
        if "TZERO%d" % i in hdr:
            zero[name] = float(hdr["TZERO%d" % i])
    width = int(hdr["NAXIS1"])
    if off > width:
        raise ValueError("BINTABLE fields need %d bytes per row, NAXIS1 is %d" % (off, width))
    return (np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": width}), names, scale, zero,
            kinds)


def fits_open(fname, memmap=True):
    """All HDUs of a FITS file: a list of FitsHDU(header, data), data being
    a numpy array (primary / IMAGE, BSCALE/BZERO applied), a FitsTable
    (BINTABLE) or None."""
    if memmap:
        buf = np.memmap(fname, dtype=np.uint8, mode="r")
    else:
        with open(fname, "rb") as f:
            buf = np.frombuffer(f.read(), dtype=np.uint8)
    hdus, pos = [], 0
    while pos + _BLOCK <= len(buf):
        if not bytes(buf[pos:pos + 8]).strip(b" \0"):
            break  # trailing padding
        hdr, pos = _read_header(buf, pos)
        if not hdus and hdr.get("SIMPLE") is None:
            raise ValueError("%s: not a FITS file (no SIMPLE card)" % fname)
        naxis = int(hdr.get("NAXIS", 0))
        dims = [int(hdr["NAXIS%d" % k]) for k in range(1, naxis + 1)]
        bitpix = int(hdr["BITPIX"])
        nelem = int(np.prod(dims)) if naxis else 0
        size = abs(bitpix) // 8 * int(hdr.get("GCOUNT", 1)) * (int(hdr.get("PCOUNT", 0)) + nelem)
        if pos + size > len(buf):
            raise ValueError("%s: HDU %d data runs past the end of the file" % (fname, len(hdus)))
        xt = str(hdr.get("XTENSION", "")).strip().upper()
        data = None
        if xt == "BINTABLE":
            dt, names, scale, zero, kinds = _table_dtype(hdr)
            rows = dims[1] if naxis > 1 else 0
            rec = np.frombuffer(buf, dtype=dt, count=rows, offset=pos) if rows else np.zeros(0, dt)
            data = FitsTable(rec, names, scale, zero, kinds)
        elif nelem and xt in ("", "IMAGE"):
            arr = np.frombuffer(buf, dtype=_BITPIX[bitpix], count=nelem, offset=pos).reshape(dims[::-1])
            bs, bz = float(hdr.get("BSCALE", 1.0)), float(hdr.get("BZERO", 0.0))
            data = arr.astype(np.float64) * bs + bz if (bs != 1.0 or bz != 0.0) else arr.astype(arr.dtype.newbyteorder("="))
        hdus.append(FitsHDU(hdr, data))
        pos += -(-size // _BLOCK) * _BLOCK
    return hdus


def _lims_line(fname, ndet, ndata, npixid, dx, dy):
    ff = 180 / np.pi
    print("ndata and ndet are ", ndet, ndata, npixid, " on ", fname, "with lims ",
          dx.min() * ff, dx.max() * ff, dy.min() * ff, dy.max() * ff)


_CALINFO_KEYS = ("scan", "bunit", "azimuth", "elevatio", "beameff", "apereff", "antgain", "gainunc", "bmaj", "bmin",
                 "bpa", "parang", "beamvol", "beamvunc")


def read_tod_from_fits(fname, hdu=1, branch=None):
    """A MUSTANG TOD (tods/io.py:160-284): columns PIXID, DX, DY, [ELEV],
    TIME, FNU, UFNU of the BINTABLE in HDU `hdu`, detector-major rows.
    Pointing and data are cast to FP64; RA above `branch` degrees wraps by
    -2 pi; samples with UFNU > 9e5 are masked."""
    hdul = fits_open(fname)
    raw = hdul[hdu].data
    if raw is None or not getattr(raw, "names", None):
        raise ValueError("TOD seems to be empty")
    calinfo = {"calinfo": False}
    try:  # per-scan calibration keywords, when present (io.py:188-214)
        calinfo["calinfo"] = True
        for kwd in _CALINFO_KEYS:
            calinfo[kwd] = hdul[hdu].header[kwd]
    except KeyError:
        print("WARNING - calinfo information not found in fits file header - to track JytoK etc you may need to "
              "reprocess the fits files using mustangidl > revision 932")
    dat = {}
    pixid = np.array(raw["PIXID"])
    dets = np.unique(pixid)
    ndet = len(dets)
    ndata = int(len(pixid) / len(dets))
    dat["pixid"] = np.reshape(pixid, [ndet, ndata])[:, 0]
    dx = np.array(raw["DX"])
    if branch is not None:
        bb = branch * np.pi / 180.0
        dx[dx > bb] = dx[dx > bb] - 2 * np.pi
    dat["dx"] = np.ascontiguousarray(np.reshape(dx, [ndet, ndata]), dtype=np.float64)
    dy = np.array(raw["DY"])
    dat["dy"] = np.ascontiguousarray(np.reshape(dy, [ndet, ndata]), dtype=np.float64)
    if "ELEV" in [n.upper() for n in raw.names]:
        elev = np.array(raw["ELEV"]) * np.pi / 180
        dat["elev"] = np.ascontiguousarray(np.reshape(elev, [ndet, ndata]), dtype=np.float64)
    tt = np.reshape(np.array(raw["TIME"]), [ndet, ndata])[0, :]
    dat["dt"] = np.median(np.diff(tt))
    dat["dat_calib"] = np.ascontiguousarray(np.reshape(np.array(raw["FNU"]), [ndet, ndata]), dtype=np.float64)
    ufnu = np.array(raw["UFNU"])
    if np.sum(np.array(ufnu > 9e5)) > 0:
        dat["mask"] = np.reshape(ufnu < 9e5, dat["dat_calib"].shape)
        dat["mask_sum"] = np.sum(dat["mask"], axis=0)
    dat["fname"] = fname
    dat["calinfo"] = calinfo
    _lims_line(fname, ndet, ndata, len(pixid), dx, dy)
    return dat


def read_tod_from_fits_cbass(fname, dopol=False, lat=37.2314, lon=-118.2941, v34=True, nm20=False):
    """A C-BASS TOD (tods/io.py:40-157): one detector of I = (I1 + I2)/2, or
    with dopol the -U, Q pair (Q, U with v34=False), masks from FLAG, and
    the nm20 cut ranges of HDU 3.  The reference re-derives RA/Dec and the
    polarisation angle with qpoint when it is installed; qpoint is not
    available here, so those keys (ra, dec, twogamma_saved) are not made,
    as in the reference without qpoint."""
    hdul = fits_open(fname)
    raw = hdul[1].data
    ra, dec = raw["RA"], raw["DEC"]
    I = 0.5 * (raw["I1"] + raw["I2"])
    mjd = raw["MJD"]
    tvec = (mjd - 2455977.5 + 2400000.5) * 86400 + 1329696000
    dat = {"dx": np.reshape(np.asarray(ra, dtype="float64"), [1, len(ra)]),
           "dy": np.reshape(np.asarray(dec, dtype="float64"), [1, len(dec)]),
           "dt": np.median(np.diff(tvec)), "ctime": tvec}
    if dopol:
        dat["dx"] = np.vstack([dat["dx"], dat["dx"]])
        dat["dy"] = np.vstack([dat["dy"], dat["dy"]])
        Q = 0.5 * (raw["Q1"] + raw["Q2"])
        U = 0.5 * (raw["U1"] + raw["U2"])
        dat["dat_calib"] = np.zeros([2, len(Q)])
        if v34:
            dat["dat_calib"][0, :] = -U
            dat["dat_calib"][1, :] = Q
        else:
            dat["dat_calib"][0, :] = Q
            dat["dat_calib"][1, :] = U
        dat["az"] = raw["AZ"] * np.pi / 180
        dat["el"] = raw["EL"] * np.pi / 180
        dat["mask"] = np.zeros([2, len(Q)], dtype="int8")
        dat["mask"][0, :] = 1 - raw["FLAG"]
        dat["mask"][1, :] = 1 - raw["FLAG"]
    else:
        dat["dat_calib"] = np.zeros([1, len(I)])
        dat["dat_calib"][:] = I
        dat["mask"] = np.zeros([1, len(I)], dtype="int8")
        dat["mask"][:] = 1 - raw["FLAG"]
    dat["pixid"] = [0]
    dat["fname"] = fname
    if nm20:
        try:  # extra cuts in HDU 3 (io.py:139-155): samples outside every [START, END) are masked
            cut = hdul[3].data
            dat["nm20_start"] = cut["START"]
            dat["nm20_stop"] = cut["END"]
            keep = 0 * dat["mask"]
            for a, b in zip(dat["nm20_start"], dat["nm20_stop"]):
                keep[:, a:b] = 1
            dat["mask"] = dat["mask"] * keep
        except (IndexError, KeyError, AttributeError, TypeError):
            print("missing nm20 for ", fname)
    return dat


def read_tod_from_fits_NIKA2(fname):
    """A NIKA2 TOD from an IDL save file (tods/io.py:287-387), read with
    scipy.io.readsav as the reference reads it: detectors that are NaN
    throughout are dropped, RA/Dec go from degrees to radians (RA wrapped to
    [0, 360) first), and NaN gaps are filled linearly (gapfill_lin)."""
    from scipy.io import readsav
    from .processing import gapfill_lin
    m = re.match(r"(.*)stru\_scan([0-9]+)s([0-9]+)\_2mm\_tot\.xdr", fname)
    if m is None:
        raise ValueError("%s does not look like a NIKA2 stru_scan<N>s<M>_2mm_tot.xdr file" % fname)
    calinfo = {"calinfo": True, "scan": m.groups()[-1]}
    raw = readsav(fname)
    data = raw.get("scan", None)
    if data is None:
        raise ValueError("TOD seems to be empty")
    dat = {}
    dat_calib = data["SIGNAL_PLW"][0]
    pixid = np.arange(dat_calib.shape[0])
    good = np.argwhere(~np.all(np.isnan(dat_calib), axis=-1))
    if len(good) == 0:
        print("No good detectors in tod {}".format(fname))
        return None
    good = good[:, 0]
    ndet, ndata = len(good), dat_calib.shape[1]
    dat["pixid"] = pixid[good]
    subscans = data["IND_SUBSCANS"][0]
    dat["subscans"] = subscans
    dat["turnarounds"] = np.where(np.diff(subscans))[0]
    dx = np.array(data["XRA_BOLOS_PLW"][0])[good, :]
    if dx[0, 0] < 0:
        dx += 360
    dx *= np.pi / 180.0
    dat["dx"] = np.ascontiguousarray(np.reshape(dx, [ndet, ndata]), dtype=np.float64)
    dy = np.array(data["YDEC_BOLOS_PLW"][0])[good, :]
    dy *= np.pi / 180.0
    dat["dy"] = np.ascontiguousarray(np.reshape(dy, [ndet, ndata]), dtype=np.float64)
    if "ELEV" in data.dtype.names:
        print("WARNING: El currently not correctly implemented")
        dat["elev"] = np.ones([ndet, ndata], dtype="float64") * data["ELEV"][0]
    dat["dt"] = np.median(np.diff(data["TIME_S"][0]))
    filled, cuts = gapfill_lin(np.array(dat_calib[good, :], dtype=np.float64))
    dat["dat_calib"] = filled
    dat["cuts"] = cuts
    dat["fname"] = fname
    dat["calinfo"] = calinfo
    _lims_line(fname, ndet, ndata, len(pixid), dx, dy)
    return dat


def _get_type(nbyte):
    """Element type code of an octave struct field (io.py:25-37)."""
    types = {8: np.dtype("float64"), 4: np.dtype("float32"), -4: np.dtype("int32"), -8: np.dtype("int64"),
             1: np.dtype("S1")}
    if nbyte not in types:
        print("Unsupported nbyte " + repr(nbyte) + " in get_type")
        return None
    return types[nbyte]


def read_octave_struct(fname):
    """A TOD saved as an octave struct (io.py:390-422): int32 key count,
    then per key its name line, int32 ndim, int32 dims (reversed on read),
    int32 element code and the elements."""
    dat = {}
    with open(fname, "rb") as f:
        nkey = int(np.fromfile(f, "int32", 1)[0])
        for _ in range(nkey):
            key = f.readline().strip().decode()
            ndim = int(np.fromfile(f, "int32", 1)[0])
            dims = np.flipud(np.fromfile(f, "int32", ndim))
            dtype = _get_type(int(np.fromfile(f, "int32", 1)[0]))
            if dtype is None:
                raise ValueError("%s: field %s has an unsupported element code" % (fname, key))
            tmp = np.fromfile(f, dtype, int(np.prod(dims)))
            dat[key] = np.reshape(tmp, dims)
    return dat


def todvec_from_files_octave(fnames):
    """A TodVec of octave-struct TODs (io.py:425-440)."""
    todvec = TodVec()
    for fname in fnames:
        todvec.add_tod(Tod(read_octave_struct(fname)))
    return todvec


def cut_blacklist(tod_names, blacklist):
    """tod_names without the blacklisted file names, sorted (io.py:443-470)."""
    tods = {nm.split("/")[-1]: nm for nm in tod_names}
    ncut = 0
    for nm in blacklist:
        tt = nm.split("/")[-1]
        if tt in tods:
            ncut += 1
            del tods[tt]
    print("deleted ", ncut, " bad files.")
    return sorted(tods.values())

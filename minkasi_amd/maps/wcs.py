"""CAR world coordinate system without astropy.

The reference builds an astropy WCS (minkasi/maps/utils.py:13-67: CAR,
crpix [1, 1], crval (ra_max, dec_min) in degrees, cdelt [-pix, pix] with
cosdec forced to 1) and pixelizes through wcslib.  ``CarWCS`` holds the same
numbers, derives the celestial->native Euler angles of a CAR projection with
default LONPOLE/LATPOLE (Calabretta & Greisen 2002, A&A 395, 1077, sec. 2.4),
and evaluates world->pixel on the host for the handful of points the map
sizing needs (minkasi/maps/skymap.py:206-234).  Per-sample pixelization runs
on the GPU with the same arithmetic (csrc/pointing.hip).
"""
import math

import numpy as np

_D2R = math.pi / 180.0
_R2D = 180.0 / math.pi


def _sincosd(a):
    if math.fmod(a, 90.0) == 0.0:
        i = abs(int(math.floor(a / 90.0 + 0.5))) % 4
        return [(0.0, 1.0), ((1.0 if a > 0 else -1.0), 0.0), (0.0, -1.0), ((-1.0 if a > 0 else 1.0), 0.0)][i]
    return math.sin(a * _D2R), math.cos(a * _D2R)


def _atan2d(y, x):
    if y == 0.0:
        return 0.0 if x >= 0.0 else 180.0
    if x == 0.0:
        return 90.0 if y > 0.0 else -90.0
    return math.atan2(y, x) * _R2D


def _asind(v):
    if v <= -1.0:
        if v + 1.0 > -1e-10:
            return -90.0
    elif v == 0.0:
        return 0.0
    elif v >= 1.0:
        if v - 1.0 < 1e-10:
            return 90.0
    return math.asin(v) * _R2D


def _acosd(v):
    if v >= 1.0:
        if v - 1.0 < 1e-10:
            return 0.0
    elif v == 0.0:
        return 90.0
    elif v <= -1.0:
        if v + 1.0 > -1e-10:
            return 180.0
    return math.acos(v) * _R2D


class CarWCS:
    """crval/cdelt in degrees, crpix 1-based, as get_wcs sets them."""

    def __init__(self, crval, cdelt, crpix=(1.0, 1.0)):
        self.crval = np.asarray(crval, dtype=np.float64).copy()
        self.cdelt = np.asarray(cdelt, dtype=np.float64).copy()
        self.crpix = np.asarray(crpix, dtype=np.float64).copy()
        self.ctype = ["RA---CAR", "DEC--CAR"]
        self.euler = self._euler()
        self.lonpole = self.euler[2]
        self.latpole = 90.0 - self.euler[1]

    @classmethod
    def from_lims(cls, lims, pixsize, cosdec=None):
        """get_wcs(lims, pixsize, 'CAR', cosdec, ref_equ=False)."""
        cosdec = 1.0  # CAR forces cosdec = 1 (utils.py:53)
        crval = [lims[1] * 180 / np.pi, lims[2] * 180 / np.pi]
        cdelt = [-pixsize / cosdec * 180 / np.pi, pixsize * 180 / np.pi]
        return cls(crval, cdelt, [1.0, 1.0])

    def _euler(self):
        lng0, lat0 = float(self.crval[0]), float(self.crval[1])
        phi0 = theta0 = 0.0
        latpole, tol = 90.0, 1e-10
        phip = (180.0 if lat0 < theta0 else 0.0) + phi0
        slat0, clat0 = _sincosd(lat0)
        sthe0, cthe0 = _sincosd(theta0)
        if phip == phi0:
            sphip, cphip, u, v = 0.0, 1.0, theta0, 90.0 - lat0
        else:
            sphip, cphip = _sincosd(phip - phi0)
            x, y = cthe0 * cphip, sthe0
            z = math.sqrt(x * x + y * y)
            if z == 0.0:
                raise ValueError("invalid CAR WCS")
            u = _atan2d(y, x)
            v = _acosd(slat0 / z)
        latp1 = u + v
        latp1 = latp1 - 360.0 if latp1 > 180.0 else (latp1 + 360.0 if latp1 < -180.0 else latp1)
        latp2 = u - v
        latp2 = latp2 - 360.0 if latp2 > 180.0 else (latp2 + 360.0 if latp2 < -180.0 else latp2)
        ok1, ok2 = abs(latp1) < 90.0 + tol, abs(latp2) < 90.0 + tol
        if ok1 and ok2:
            latp = latp1 if abs(latp1 - latpole) < abs(latp2 - latpole) else latp2
        elif ok1:
            latp = latp1
        elif ok2:
            latp = latp2
        else:
            raise ValueError("invalid CAR WCS: no native pole")
        latp = min(90.0, max(-90.0, latp))
        z = _sincosd(latp)[1] * clat0
        if abs(z) < tol:
            raise ValueError("invalid CAR WCS: degenerate pole")
        x = (sthe0 - _sincosd(latp)[0] * slat0) / z
        y = sphip * cthe0 / clat0
        lngp = lng0 - _atan2d(y, x)
        if lng0 >= 0.0:
            if lngp < 0.0:
                lngp += 360.0
            elif lngp > 360.0:
                lngp -= 360.0
        else:
            if lngp > 0.0:
                lngp -= 360.0
            elif lngp < -360.0:
                lngp += 360.0
        e1 = 90.0 - latp
        s1, c1 = _sincosd(e1)
        return np.array([lngp, e1, phip, c1, s1, _R2D * _D2R])

    def params10(self):
        """The 10 doubles mk_pix_from_radec_car takes."""
        return np.ascontiguousarray(np.concatenate([self.euler, self.cdelt, self.crpix]), dtype=np.float64)

    def world2pix(self, ra, dec):
        """1-based pixel coordinates for radians (host, few points)."""
        ra = np.atleast_1d(np.asarray(ra, dtype=np.float64))
        dec = np.atleast_1d(np.asarray(dec, dtype=np.float64))
        eul = self.euler
        px = np.empty(ra.shape)
        py = np.empty(ra.shape)
        for i in range(ra.size):
            lng = float(ra.flat[i]) * 180 / math.pi
            lat = float(dec.flat[i]) * 180 / math.pi
            dlng = lng - eul[0]
            sinlat, coslat = _sincosd(lat)
            sinlng, coslng = _sincosd(dlng)
            costhe3, costhe4 = coslat * eul[3], coslat * eul[4]
            sinthe3, sinthe4 = sinlat * eul[3], sinlat * eul[4]
            x = sinthe4 - costhe3 * coslng
            if abs(x) < 1e-5:
                x = -_sincosd(lat + eul[1])[1] + costhe3 * (1.0 - coslng)
            y = -coslat * sinlng
            if x != 0.0 or y != 0.0:
                dphi = _atan2d(y, x)
            else:
                dphi = dlng - 180.0 if eul[1] < 90.0 else -dlng
            phi = math.fmod(eul[2] + dphi, 360.0)
            if phi > 180.0:
                phi -= 360.0
            elif phi < -180.0:
                phi += 360.0
            if math.fmod(dlng, 180.0) == 0.0:
                theta = lat + coslng * eul[1]
                if theta > 90.0:
                    theta = 180.0 - theta
                if theta < -90.0:
                    theta = -180.0 - theta
            else:
                zz = sinthe3 + costhe4 * coslng
                if abs(zz) > 0.99:
                    theta = math.copysign(_acosd(math.sqrt(x * x + y * y)), zz)
                else:
                    theta = _asind(zz)
            px.flat[i] = eul[5] * phi / self.cdelt[0] + self.crpix[0]
            py.flat[i] = eul[5] * theta / self.cdelt[1] + self.crpix[1]
        return px, py

    def to_header(self):
        """FITS header cards for the map (SkyMap.write): the cards, in the
        order, that astropy's WCS.to_header() gives for this CAR WCS (the
        reference writes its maps with it, skymap.py:705-722): the axis
        count, reference pixel, increments and units, projection types,
        reference value, the native pole (LONPOLE/LATPOLE), MJDREF 0 and
        RADESYS ICRS (astropy's defaults when no equinox or date is set)."""
        return [
            ("WCSAXES", 2),
            ("CRPIX1", float(self.crpix[0])), ("CRPIX2", float(self.crpix[1])),
            ("CDELT1", float(self.cdelt[0])), ("CDELT2", float(self.cdelt[1])),
            ("CUNIT1", "deg"), ("CUNIT2", "deg"),
            ("CTYPE1", self.ctype[0]), ("CTYPE2", self.ctype[1]),
            ("CRVAL1", float(self.crval[0])), ("CRVAL2", float(self.crval[1])),
            ("LONPOLE", float(self.lonpole)), ("LATPOLE", float(self.latpole)),
            ("MJDREF", 0.0), ("RADESYS", "ICRS"),
        ]

    def __deepcopy__(self, memo):
        return CarWCS(self.crval, self.cdelt, self.crpix)


def get_wcs(lims, pixsize, proj="CAR", cosdec=None, ref_equ=False):
    """minkasi/maps/utils.py:13-67 for proj='CAR', ref_equ=False."""
    if proj != "CAR":
        raise ValueError(f"Unknown projection type, {proj}, in get_wcs.")
    if ref_equ:
        raise NotImplementedError("ref_equ=True is not supported")
    return CarWCS.from_lims(lims, pixsize, cosdec)

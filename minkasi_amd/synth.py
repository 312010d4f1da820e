"""Seeded synthetic TODs for tests and benchmarks.

Scan: constant-velocity raster in RA coordinate, +-half_width arcsec of RA
at `speed` arcsec/s of RA, sampled at fs; Dec steps `dec_step` arcsec per leg and
reflects at +-dec_span/2.  Detector offsets are uniform in a disc of
`radius` arcsec on the sky (dRA = dx / cos dec).  Signal: one Gaussian
source (FWHM 10 arcsec, amplitude 1) evaluated with the formula of the
reference's fill_gauss_src (minkasi/lib/minkasi.c:302-319: dx scaled by
cos(dec0), cutoff at exponent -20), plus white noise sigma and a common-mode
random walk with step sigma*2*sin(pi*fknee/fs), times per-detector gains.
`corr_frac`>0 adds a rank-5 random-walk term of amplitude corr_amp*sigma to
that fraction of detectors.  Per-TOD random numbers come from
numpy.random.default_rng([config_id, tod_index]); the per-sample white
noise comes from the same generator on the host (`backend="numpy"`) or from
a torch generator seeded from it on the GPU (`backend="torch"`, for sizes the
host cannot hold).
"""
import numpy as np

ARCSEC = np.pi / 180 / 3600


class ScanSpec:
    def __init__(self, ra0_deg=180.0, dec0_deg=60.0, half_width=360.0, speed=30.0, fs=100.0, dec_step=10.0,
                 dec_span=480.0, radius=60.0, center_jitter=0.0, src_fwhm=10.0, src_amp=1.0, sigma=0.1,
                 fknee=1.0, gain_sigma=0.0, corr_frac=0.0, corr_rank=5, corr_amp=10.0):
        self.__dict__.update(locals())
        del self.__dict__["self"]


def make_tod_info(config_id, tod_index, ndet, n, spec=None, dtype=np.float64, backend="numpy", device=None,
                  want_radec=True):
    """Returns a Tod-style info dict: dat_calib (ndet x n), dx, dy (radians,
    FP64), dt, ctime, plus the raster factors (ra_b, dec_b, dra, ddec)."""
    spec = spec or ScanSpec()
    rng = np.random.default_rng([int(config_id), int(tod_index)])
    ra0 = np.deg2rad(spec.ra0_deg)
    dec0 = np.deg2rad(spec.dec0_deg)
    if spec.center_jitter:
        ra0 = ra0 + rng.uniform(-1, 1) * spec.center_jitter * ARCSEC / np.cos(dec0)
        dec0c = dec0 + rng.uniform(-1, 1) * spec.center_jitter * ARCSEC
    else:
        dec0c = dec0
    t = np.arange(n) / spec.fs
    leg = 2 * spec.half_width / spec.speed
    ph = (t / leg) % 2.0
    tri = np.where(ph < 1.0, ph, 2.0 - ph) * 2.0 - 1.0
    ra_b = ra0 + tri * spec.half_width * ARCSEC  # sweep in RA coordinate
    legno = np.floor(t / leg)
    span = spec.dec_span
    u = (legno * spec.dec_step) % (2 * span)
    tri_dec = np.where(u < span, u, 2 * span - u) - span / 2
    dec_b = dec0c + tri_dec * ARCSEC
    r = spec.radius * np.sqrt(rng.uniform(0, 1, ndet))
    th = rng.uniform(0, 2 * np.pi, ndet)
    ddec = r * np.sin(th) * ARCSEC
    dra = r * np.cos(th) * ARCSEC / np.cos(dec0)
    gains = 1.0 + spec.gain_sigma * rng.standard_normal(ndet) if spec.gain_sigma else np.ones(ndet)
    step = spec.sigma * 2 * np.sin(np.pi * spec.fknee / spec.fs)
    cm = np.cumsum(rng.standard_normal(n)) * step
    modes = mix = None
    if spec.corr_frac > 0:
        modes = np.cumsum(rng.standard_normal((spec.corr_rank, n)), axis=1) * step * spec.corr_amp
        ncorr = int(round(spec.corr_frac * ndet))
        which = rng.choice(ndet, ncorr, replace=False)
        mix = np.zeros((ndet, spec.corr_rank))
        mix[which] = rng.normal(0, np.sqrt(1.0 / spec.corr_rank), (ncorr, spec.corr_rank))
    seed = int(rng.integers(0, 2 ** 62))
    sig = spec.src_fwhm / np.sqrt(8 * np.log(2)) * ARCSEC
    info = dict(dt=1.0 / spec.fs, ctime=t, ra_b=ra_b, dec_b=dec_b, dra=dra, ddec=ddec, gains=gains)
    if backend == "numpy":
        dx = ra_b[None, :] + dra[:, None]
        dy = dec_b[None, :] + ddec[:, None]
        sigm = _gauss_src(dx, dy, np.deg2rad(spec.ra0_deg), np.deg2rad(spec.dec0_deg), sig, spec.src_amp, np)
        dat = (sigm + cm[None, :]) * gains[:, None]
        if modes is not None:
            dat = dat + mix @ modes
        white = np.random.default_rng(seed).standard_normal((ndet, n)) * spec.sigma
        info["dat_calib"] = (dat + white).astype(dtype)
        if want_radec:
            info["dx"], info["dy"] = dx, dy
        return info
    import torch
    dev = device or torch.device("cuda", torch.cuda.current_device())
    tdt = torch.float32 if dtype in (np.float32, "float32", torch.float32) else torch.float64
    rb = torch.from_numpy(ra_b).to(dev)
    db = torch.from_numpy(dec_b).to(dev)
    dxr = torch.from_numpy(dra).to(dev)
    dyr = torch.from_numpy(ddec).to(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    dat = torch.empty((ndet, n), dtype=tdt, device=dev)
    rows = max(1, (1 << 26) // max(n, 1))
    gd = torch.from_numpy(gains).to(dev)
    cmd = torch.from_numpy(cm).to(dev)
    md = torch.from_numpy(modes).to(dev) if modes is not None else None
    mixd = torch.from_numpy(mix).to(dev) if mix is not None else None
    for r0 in range(0, ndet, rows):
        r1 = min(ndet, r0 + rows)
        dx = rb[None, :] + dxr[r0:r1, None]
        dy = db[None, :] + dyr[r0:r1, None]
        blk = _gauss_src(dx, dy, np.deg2rad(spec.ra0_deg), np.deg2rad(spec.dec0_deg), sig, spec.src_amp, torch)
        blk = (blk + cmd[None, :]) * gd[r0:r1, None]
        if md is not None:
            blk = blk + mixd[r0:r1] @ md
        blk = blk + torch.randn((r1 - r0, n), generator=g, device=dev, dtype=torch.float64) * spec.sigma
        dat[r0:r1] = blk.to(tdt)
    info["dat_calib"] = dat
    if want_radec:
        info["dx"] = rb[None, :] + dxr[:, None]
        info["dy"] = db[None, :] + dyr[:, None]
    return info


def _gauss_src(dx, dy, x0, y0, sig, amp, xp):
    fac = -0.5 / sig / sig
    cosdec = np.cos(y0)
    delx = (x0 - dx) * cosdec
    dely = y0 - dy
    arg = (delx * delx + dely * dely) * fac
    val = amp * xp.exp(arg)
    return xp.where(arg > -20, val, 0.0 * val)


CONFIGS = {
    # id: (ntod, ndet range, n range, map pixels, pixsize arcsec, fwhm bins, iterations, dtype, spec kwargs)
    0: dict(ntod=1, ndet=(64, 64), n=(65536, 65536), npix=256, pix=2.0, fwhm=10.0, iters=20, dtype="float64",
            spec=dict()),
    1: dict(ntod=32, ndet=(215, 215), n=(40000, 160000), npix=1024, pix=1.0, fwhm=50.0, iters=50, dtype="float32",
            spec=dict(center_jitter=120.0, gain_sigma=0.05)),
    2: dict(ntod=256, ndet=(64, 512), n=(10000, 200000), npix=2048, pix=1.0, fwhm=50.0, iters=50, dtype="float32",
            spec=dict(center_jitter=120.0, gain_sigma=0.05)),
    3: dict(ntod=8, ndet=(1024, 1024), n=(262144, 262144), npix=1024, pix=1.0, fwhm=50.0, iters=100,
            dtype="float32", spec=dict(center_jitter=120.0, gain_sigma=0.05, corr_frac=0.2)),
    4: dict(ntod=64, ndet=(64, 512), n=(10000, 200000), npix=2048, pix=1.0, fwhm=50.0, iters=1000,
            dtype="float32", spec=dict(center_jitter=120.0, gain_sigma=0.05)),
}


def tod_shapes(config_id, ntod=None, seed=10 ** 6):
    """(ndet, n) per TOD, drawn from default_rng([config_id, seed])."""
    c = CONFIGS[config_id]
    ntod = c["ntod"] if ntod is None else ntod
    rng = np.random.default_rng([int(config_id), int(seed)])
    out = []
    for _ in range(ntod):
        ndet = int(rng.integers(c["ndet"][0], c["ndet"][1] + 1))
        n = int(rng.integers(c["n"][0], c["n"][1] + 1))
        out.append((ndet, n))
    return out

"""NoiseSmoothedSVD on the GPU (minkasi/mapmaking/noise.py:584-772).

Setup (once per TOD, FP64 on the device):
  u      -- eigenvectors of the Gram matrix d.d^T (the left singular vectors
            np.linalg.svd returns, noise.py:650), ordered by decreasing
            eigenvalue; the ndet x ndet eigendecomposition runs on the host.
  mywt   -- 1 / smooth(DCT-I(v.d)^2), bin 0 zeroed (noise.py:667-678).
  C.mywt -- DCT-I of each mywt row, the kernel whose spectrum the filter
            plan keeps (see csrc/fft.hip).
apply_noise (every A.x):
  FP32 data:  split-BF16 MFMA rotation v.d -> batched FFT filter C mywt C
              with the edge halving folded in -> split-BF16 rotation vT.(.)
  FP64 data:  the same steps with FP64 GEMMs and an FP64 filter.
prewhiten=True folds the per-detector noise vector into the rotation
matrices (v.diag(1/w) and diag(1/w).vT), which is algebraically the
reference's divide-before / divide-after.
"""
import numpy as np
import torch

from .. import device as dev
from ..lib import mkgpu
from ..lib.plans import DeflPlan, FiltPlan, RotPlan, SplitMatrix
from ..tools import fft as gfft
from ..tools.smooth import smooth_rows

try:
    from typing import Protocol, runtime_checkable
except ImportError:  # pragma: no cover
    from typing_extensions import Protocol, runtime_checkable


@runtime_checkable
class NoiseModelType(Protocol):
    """noise.py:88-95."""

    def __init__(self, dat):
        pass

    def apply_noise(self, dat):
        return None


@runtime_checkable
class WithDetWeights(NoiseModelType, Protocol):
    """noise.py:97-100."""

    def get_det_weights(self):
        return None


# A mode is deflated ahead of the FP32 rotation when its rotated row is more
# than DEFLATE_RATIO times the median row norm: the FP32 accumulation error of
# the small rows grows as about 1e-8 times the largest such ratio left in the
# data, so 500 keeps apply_noise within about 5e-6 of FP64 while TODs whose
# modes stay below it skip the extra pass (12 bytes per sample).
DEFLATE_RATIO = 500.0
MAX_DEFLATE = 8


def deflation_rank(row_norms):
    """Number of leading rows of v.d (modes in decreasing power) whose norm
    exceeds DEFLATE_RATIO times the median row norm (at most MAX_DEFLATE)."""
    row_norms = np.asarray(row_norms, dtype=np.float64)
    med = float(np.median(row_norms)) if row_norms.size else 0.0
    if not med > 0:
        return 0
    k = 0
    while k < min(MAX_DEFLATE, row_norms.size - 1) and row_norms[k] > DEFLATE_RATIO * med:
        k += 1
    return k


def _gram(d):
    ndet, n = d.shape
    G = torch.empty((ndet, ndet), dtype=torch.float64, device=d.device)
    nb = int(mkgpu.lib.mk_gram_workspace_bytes(ndet, n))
    work = torch.empty(nb // 8 + 32, dtype=torch.float64, device=d.device)
    mkgpu.check("mk_gram_f64", ndet, n, d.data_ptr(), int(d.dtype == torch.float32), n, G.data_ptr(),
                work.data_ptr(), work.numel() * 8, dev.stream_ptr())
    return G


def _gemm64(A, B):
    """FP64 device GEMM A (M x K, FP64) . B (K x N, FP32 or FP64)."""
    M, K = A.shape
    N = B.shape[1]
    C = torch.empty((M, N), dtype=torch.float64, device=B.device)
    mkgpu.check("mk_gemm_nn_f64", M, N, K, A.data_ptr(), K, B.data_ptr(), int(B.dtype == torch.float32), N,
                C.data_ptr(), N, dev.stream_ptr())
    return C


class NoiseSmoothedSVD:
    """GPU noise model with the reference's constructor and methods.

    Attributes (reference names): v, vT (numpy FP64, ndet x ndet), mywt
    (numpy FP32, ndet x n, in host memory), noisevec (numpy or None); and
    mode_wt, the FP64 row sums of mywt.
    """

    def __init__(self, dat_use, fwhm=50.0, prewhiten=False, fit_powlaw=False, u_in=None, keep_fp64=None):
        if fit_powlaw:
            raise NotImplementedError("fit_powlaw relies on minkasi.fitting, outside the mapmaking inner loop")
        d = dat_use if isinstance(dat_use, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(dat_use))
        d = d.to(dev.device())
        if d.dtype not in (torch.float32, torch.float64):
            d = d.to(torch.float64)
        d = d.contiguous()
        if d.dim() != 2 or d.shape[1] < 2:
            raise ValueError("dat_use must be ndet x ndata with ndata >= 2")
        ndet, n = d.shape
        self.ndet, self.n = ndet, n
        self.fwhm = float(fwhm)
        self.dtype = d.dtype
        if prewhiten:
            # np.median semantics: mean of the two middle values for an even count
            srt = torch.sort(torch.abs(torch.diff(d.to(torch.float64), dim=1)), dim=1).values
            cnt = srt.shape[1]
            nv = srt[:, cnt // 2] if cnt % 2 else 0.5 * (srt[:, cnt // 2 - 1] + srt[:, cnt // 2])
            self.noisevec = nv.cpu().numpy().copy()
            d = (d.to(torch.float64) / nv[:, None]).to(d.dtype).contiguous()
        else:
            self.noisevec = None
        if u_in is None:
            G = _gram(d).cpu().numpy()
            evals, evecs = np.linalg.eigh(0.5 * (G + G.T))
            u = evecs[:, ::-1]
        else:
            u = np.asarray(u_in, dtype=np.float64)
            assert u.shape[0] == u.shape[1]
        self.v = np.zeros([ndet, ndet])
        self.v[:] = u.transpose()
        self.vT = self.v.T.copy() if u_in is None else np.linalg.inv(self.v)
        # spectra, FP64 on the device
        v_dev = torch.from_numpy(np.ascontiguousarray(self.v)).to(d.device)
        y = _gemm64(v_dev, d)                                # dat_rot (noise.py:667)
        # modes that dominate the rotated data are deflated ahead of the FP32
        # rotation (DeflPlan); the row norms of v.d are the modes' amplitudes
        self.defl_k = 0 if self.noisevec is not None else deflation_rank(
            torch.linalg.vector_norm(y, dim=1).cpu().numpy())
        Y = gfft._dct1_tensor(y)                              # dat_trans (noise.py:674)
        del y
        W = smooth_rows(Y, self.fwhm, square=True, invert=True)  # mywt in FP64 (noise.py:675-677)
        del Y
        # get_det_weights needs only the row sums; mywt itself is kept in host
        # memory (FP32), where the reference keeps its numpy array, so that it
        # does not hold 4 bytes of HBM per sample
        self.mode_wt = W.sum(dim=1).cpu().numpy()
        self.mywt = W.to(torch.float32).cpu().numpy()
        wdct = gfft._dct1_tensor(W)
        del W
        self._build_apply(wdct, keep_fp64 if keep_fp64 is not None else (self.dtype == torch.float64))

    # ------------------------------------------------------------------
    def _rot_mats(self):
        v, vT = self.v, self.vT
        if self.noisevec is not None:
            inv = 1.0 / self.noisevec
            v = v * inv[None, :]
            vT = vT * inv[:, None]
        return v, vT

    def _build_apply(self, wdct, keep_fp64):
        ndet, n = self.ndet, self.n
        seg = [(ndet, n, 0, n, 0, n)]
        self.filt32 = FiltPlan(seg, 4)
        self.filt32.set_weights(wdct)
        self.filt32.set_halve(True)
        self.filt64 = None
        if keep_fp64:
            self.filt64 = FiltPlan(seg, 8)
            self.filt64.set_weights(wdct)
            self.filt64.set_halve(True)
        v, vT = self._rot_mats()
        self.v_split = SplitMatrix(v)
        self.vT_split = SplitMatrix(vT)
        k = getattr(self, "defl_k", 0)
        self.defl = DeflPlan([(ndet, n, 0, self.v[:k].T)]) if k else None
        self.rot_v = RotPlan([(self.v_split, n, 0, n, 0, n)])
        self.rot_vT = RotPlan([(self.vT_split, n, 0, n, 0, n)])
        d = dev.device()
        self._v64 = torch.from_numpy(np.ascontiguousarray(v)).to(d)
        self._vT64 = torch.from_numpy(np.ascontiguousarray(vT)).to(d)

    # ------------------------------------------------------------------
    def apply_noise(self, dat):
        """noise.py:680-714.  Returns a new device tensor of dat's dtype."""
        x = dat if isinstance(dat, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(dat))
        x = x.to(dev.device()).contiguous()
        if tuple(x.shape) != (self.ndet, self.n):
            raise ValueError("data shape %s does not match noise model %s" % (tuple(x.shape), (self.ndet, self.n)))
        if x.dtype == torch.float32:
            tmp = torch.empty_like(x)
            tmp2 = torch.empty_like(x)
            return self.apply_noise_wscratch(x, tmp, tmp2)
        return self._apply64(x.to(torch.float64))

    def apply_noise_wscratch(self, dat, tmp, tmp2):
        """FP32 apply with caller buffers (noise.py:716-760); result in tmp2."""
        if self.defl is not None:
            self.defl.exec(dat, tmp2)
            self.rot_v.exec(tmp2, tmp)
            self.defl.addback(tmp)
        else:
            self.rot_v.exec(dat, tmp)
        self.filt32.exec(tmp, tmp)
        self.rot_vT.exec(tmp, tmp2)
        return tmp2

    def _apply64(self, x):
        if self.filt64 is None:
            raise RuntimeError("FP64 apply needs keep_fp64=True at construction (or FP64 input data)")
        y = _gemm64(self._v64, x)
        self.filt64.exec(y, y)
        return _gemm64(self._vT64, y)

    def get_det_weights(self):
        """noise.py:762-772."""
        tmp = np.dot(self.vT, np.dot(np.diag(self.mode_wt), self.v))
        return np.diag(tmp).copy() * 2.0

"""Timestream models as Mapset members: tsModel and tsMultiModel
(minkasi/mapmaking/timestream.py:842-1000).

A tsModel keeps one model per TOD (for example one CutsCompact each) in
``data``, keyed by the TOD's ``info['fname']``.  A TOD without a file name is
keyed by ``info['tag']``, the index TodVec.add_tod gives it.  Mapset calls
map2tod/tod2map per TOD and dot/axpy/clear/copy/__mul__ on the whole model,
as for a SkyMap.  The member models run their own device kernels.

One GPU holds every TOD, so dot needs no all-reduce (the reference calls
comm.allreduce under MPI, timestream.py:885-886) and mpi_reduce does nothing.
"""
import numpy as np


def tod_key(tod):
    """The name a tsModel files a TOD's model under."""
    info = tod.info
    if "fname" in info:
        return info["fname"]
    if "tag" in info:
        return info["tag"]
    raise KeyError("the TOD has neither info['fname'] nor info['tag'] to key its timestream model")


class tsModel:
    def __init__(self, todvec=None, modelclass=None, *args, **kwargs):
        self.data = {}
        if todvec is None:
            return
        for tod in todvec.tods:
            self.data[tod_key(tod)] = modelclass(tod, *args, **kwargs)

    def copy(self):
        new = tsModel()
        for nm in self.data.keys():
            new.data[nm] = self.data[nm].copy()
        return new

    def tod2map(self, tod, dat, do_add=True, do_omp=False):
        """timestream.py:857-861: a do_add=False call clears every model first."""
        if not do_add:
            self.clear()
        self.data[tod_key(tod)].tod2map(tod, dat, do_add, do_omp)

    def map2tod(self, tod, dat, do_add=True, do_omp=True):
        """timestream.py:863-867: do_add=False zeroes dat first."""
        if not do_add:
            dat[:] = 0.0
        self.data[tod_key(tod)].map2tod(tod, dat, do_add, do_omp)

    def apply_prior(self, x, Ax):
        for nm in self.data.keys():
            self.data[nm].apply_prior(x.data[nm], Ax.data[nm])

    def dot(self, tsmodels=None):
        tot = 0.0
        for nm in self.data.keys():
            if tsmodels is None:
                tot = tot + self.data[nm].dot(self.data[nm])
            else:
                if nm not in tsmodels.data:
                    raise KeyError("tsModel.dot: model %r missing from the other tsModel" % (nm,))
                tot = tot + self.data[nm].dot(tsmodels.data[nm])
        return tot

    def clear(self):
        for nm in self.data.keys():
            self.data[nm].clear()

    def axpy(self, tsmodel, a):
        for nm in self.data.keys():
            self.data[nm].axpy(tsmodel.data[nm], a)

    def __mul__(self, tsmodel):
        """Used as precon * r (timestream.py:898-906)."""
        tt = self.copy()
        for nm in self.data.keys():
            tt.data[nm] = self.data[nm] * tsmodel.data[nm]
        return tt

    def mpi_reduce(self):
        return None

    def get_caches(self):
        for nm in self.data.keys():
            f = getattr(self.data[nm], "get_caches", None)
            if f is not None:
                f()

    def clear_caches(self):
        for nm in self.data.keys():
            f = getattr(self.data[nm], "clear_caches", None)
            if f is not None:
                f()


class tsMultiModel(tsModel):
    """Models shared between groups of TODs (timestream.py:928-1000): TOD i
    uses the model of todtags[i], stored in tod.info[tag]."""

    def __init__(self, todvec=None, todtags=None, modelclass=None, tag="ts_multi_model", *args, **kwargs):
        self.data = {}
        self.tag = tag
        if todtags is not None:
            alltags = np.unique(np.hstack([np.asarray(todtags)]))
            if modelclass is not None:
                for mytag in alltags:
                    self.data[mytag] = modelclass(*args, **kwargs)
            if todvec is not None:
                for i, tod in enumerate(todvec.tods):
                    tod.info[tag] = todtags[i]

    def copy(self):
        new = tsMultiModel(tag=self.tag)
        for nm in self.data.keys():
            new.data[nm] = self.data[nm].copy()
        return new

    def tod2map(self, tod, dat, do_add=True, do_omp=False):
        self.data[tod.info[self.tag]].tod2map(tod, dat, do_add, do_omp)

    def map2tod(self, tod, dat, do_add=True, do_omp=False):
        self.data[tod.info[self.tag]].map2tod(tod, dat, do_add, do_omp)

    def __mul__(self, tsmodel):
        tt = self.copy()
        for nm in self.data.keys():
            tt.data[nm] = self.data[nm] * tsmodel.data[nm]
        return tt

"""Mapset (minkasi/maps/mapset.py:19-231): a list of maps acting as one PCG
vector.  Same semantics: add_map stores a copy, copy() deep-copies, dot sums
the members' dots, __mul__ multiplies member-wise into a copy."""


class _MapsetBase:
    def __init__(self):
        self.nmap = 0
        self.maps = []

    def add_map(self, map):
        self.maps.append(map.copy())
        self.nmap = self.nmap + 1

    def copy(self):
        new_mapset = self.__class__()
        for i in range(self.nmap):
            new_mapset.add_map(self.maps[i].copy())
        return new_mapset


class Mapset(_MapsetBase):
    def clear(self):
        for i in range(self.nmap):
            self.maps[i].clear()

    def _check(self, mapset):
        if self.nmap != mapset.nmap:
            raise ValueError(
                f"nmaps don't match, this Mapset has {self.nmap} but the provided Mapset has {mapset.nmap}")

    def dot(self, mapset):
        self._check(mapset)
        tot = 0.0
        for i in range(self.nmap):
            tot = tot + self.maps[i].dot(mapset.maps[i])
        return tot

    def axpy(self, mapset, a):
        self._check(mapset)
        for i in range(self.nmap):
            self.maps[i].axpy(mapset.maps[i], a)

    def __add__(self, mapset):
        mm = self.copy()
        mm.axpy(mapset, 1.0)
        return mm

    def __sub__(self, mapset):
        mm = self.copy()
        mm.axpy(mapset, -1.0)
        return mm

    def __mul__(self, mapset):
        self._check(mapset)
        mm = mapset.copy()
        for i in range(self.nmap):
            mm.maps[i] = self.maps[i] * mapset.maps[i]
        return mm

    def apply_prior(self, x, Ax):
        """mapset.py:248-275: each map's apply_prior, with the whole Mapset
        for maps whose prior is global (isglobal_prior) and map by map
        otherwise."""
        for i in range(self.nmap):
            m = self.maps[i]
            if m is None:
                continue
            if getattr(m, "isglobal_prior", False):
                m.apply_prior(x, Ax)
            else:
                m.apply_prior(x.maps[i], Ax.maps[i])

    def get_caches(self):
        for i in range(self.nmap):
            self.maps[i].get_caches()

    def clear_caches(self):
        for i in range(self.nmap):
            self.maps[i].clear_caches()

    def mpi_reduce(self):
        """Single GPU: nothing to reduce (mapset.py:224-231)."""
        return None

"""Neighbour list with the reference's interface and output format
(dmff/common/nblist.py:12-70): allocate(coords) returns (capacity, 3) pairs
[i, j, nbond] with i < j, padded with (N, N, 0) rows to a capacity fixed at
1.3x the first call's pair count.  The pair set is built on the GPU by a cell
list with an FP64 distance test.
"""
import numpy as np
import torch

from .engine import Engine


def _cov_to_exclusions(cov_map):
    if cov_map is None:
        return {}
    if isinstance(cov_map, dict):
        return cov_map
    cm = np.asarray(cov_map.detach().cpu() if isinstance(cov_map, torch.Tensor) else cov_map)
    ii, jj = np.nonzero(cm)
    return {(int(i), int(j)): int(cm[i, j]) for i, j in zip(ii, jj)}


class NeighborList:
    def __init__(self, box, rcut, cov_map, padding=True, device="cuda"):
        self.box = box
        self.rcut = float(rcut)
        self.capacity_multiplier = None
        self.padding = padding
        self.cov_map = cov_map
        self.device = device
        self._eng = None
        self._excl = _cov_to_exclusions(cov_map)

    def _engine(self, n):
        if self._eng is None or self._eng.n_atoms != n:
            e = Engine(n, 0, False, False, 0.0, (0, 0, 0), self.rcut, device=self.device)
            e.set_topology(None, None, self._excl)
            self._eng = e
        return self._eng

    def allocate(self, coords, box=None):
        pos = torch.as_tensor(np.asarray(coords.detach().cpu() if isinstance(coords, torch.Tensor) else coords),
                              dtype=torch.float64)
        b = self.box if box is None else box
        b = torch.as_tensor(np.asarray(b.detach().cpu() if isinstance(b, torch.Tensor) else b), dtype=torch.float64)
        n = pos.shape[0]
        eng = self._engine(n)
        eng.build_neighbors(pos, b)
        pairs = eng.get_pairs(0).to(torch.int64)
        self._positions = coords
        if self.capacity_multiplier is None:
            self.capacity_multiplier = int(pairs.shape[0] * 1.3)
        if not self.padding:
            self._pairs = pairs
            return self._pairs
        self.capacity_multiplier = max(self.capacity_multiplier, pairs.shape[0])
        pad = self.capacity_multiplier - pairs.shape[0]
        if pad > 0:
            padrows = torch.tensor([[n, n, 0]], dtype=torch.int64, device=pairs.device).repeat(pad, 1)
            pairs = torch.cat([pairs, padrows], dim=0)
        self._pairs = pairs
        return self._pairs

    def update(self, positions, box=None):
        self.allocate(positions, box)

    @property
    def pairs(self):
        return self._pairs

    @property
    def scaled_pairs(self):
        return self._pairs

    @property
    def positions(self):
        return self._positions


NeighborListFreud = NeighborList


class NoCutoffNeighborList:
    """dmff/common/nblist.py:73-120: every pair i < j, with the same padding and
    nbond column.  A host enumeration, as in the reference (O(N^2), for clusters
    and small molecules)."""

    def __init__(self, cov_map, padding=True):
        self.capacity_multiplier = None
        self.padding = padding
        self.cov_map = cov_map
        self._excl = _cov_to_exclusions(cov_map)

    def _candidates(self, coords):
        n = coords.shape[0]
        i, j = np.triu_indices(n, k=1)
        return np.stack([i, j], axis=1).astype(np.int64)

    def _finish(self, nlist, n):
        if self.capacity_multiplier is None:
            self.capacity_multiplier = int(nlist.shape[0] * 1.3)
        if self.padding:
            self.capacity_multiplier = max(self.capacity_multiplier, nlist.shape[0])
            pad = self.capacity_multiplier - nlist.shape[0]
            if pad > 0:
                nlist = np.vstack([nlist, np.full((pad, 2), n, dtype=np.int64)])
        nb = np.array([self._excl.get((int(a), int(b)), 0) for a, b in nlist], dtype=np.int64)
        self._pairs = torch.as_tensor(np.concatenate([nlist, nb[:, None]], axis=1))
        return self._pairs

    def allocate(self, coords, box=None):
        self._positions = coords
        c = np.asarray(coords.detach().cpu() if isinstance(coords, torch.Tensor) else coords, dtype=np.float64)
        return self._finish(self._candidates(c), c.shape[0])

    def update(self, positions, box=None):
        self.allocate(positions)

    @property
    def pairs(self):
        return self._pairs

    @property
    def scaled_pairs(self):
        return self._pairs

    @property
    def positions(self):
        return self._positions


class NoPeriodicNeighborList(NoCutoffNeighborList):
    """dmff/common/nblist.py:129-160: pairs i < j with a plain (not periodic)
    distance below rcut."""

    def __init__(self, rcut, cov_map, padding=True):
        super().__init__(cov_map, padding)
        self.rcut = float(rcut)

    def allocate(self, coords, box=None):
        self._positions = coords
        c = np.asarray(coords.detach().cpu() if isinstance(coords, torch.Tensor) else coords, dtype=np.float64)
        nl = self._candidates(c)
        d = np.linalg.norm(c[nl[:, 0]] - c[nl[:, 1]], axis=1)
        return self._finish(nl[d < self.rcut], c.shape[0])

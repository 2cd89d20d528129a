"""Row-slab decomposition of one periodic 2-D grid, held on one device.

This is how a grid too large for one device is split: each device keeps a slab
of consecutive rows plus h = 4R = 2 (order - 1) ghost rows above and below it
(``native.slab_halo_rows``), advances the slab by one RK4 step on the fused
kernel (``native.slab_step``, which writes only the slab's own rows), and then
exchanges h rows with each of its two neighbours.  One step depends on the
state within 4R rows, so one exchange per step keeps every slab equal, bit for
bit, to the same rows of the whole-grid solve on the fused engine.

Here every slab lives on the same device and the exchange is a device-to-device
copy of the same rows that a multi-device run would send to its neighbours, so
the decomposition (ghost depth, exchange pattern, kernel at the slab's real
shape) is exercised and timed on one GPU.

The reference has no counterpart: its solvers (nls/nls.f90:797-919) are serial.
"""
import numpy as np

from . import native


class SlabGrid(object):
    """An n x n periodic grid split into row slabs.

    desc      ProblemDesc of the whole grid (dim 2, periodic, RK4; desc.n = n)
    rows      number of slabs (equal split) or the list of slab heights,
              each at least h, summing to n
    P, psi    torch tensors of shape (n, n) on the device; nres likewise
              (dynamic reservoir only, else None)
    """

    def __init__(self, desc, rows, P, psi, nres=None):
        import torch
        n = int(desc.n)
        if desc.dim != 2 or desc.bc != native.BC["periodic"] or desc.scheme != native.SCHEME["rk4"]:
            raise ValueError("row slabs are 2-D, periodic and RK4")
        if isinstance(rows, int):
            k = rows
            rows = [n // k + (1 if i < n % k else 0) for i in range(k)]
        rows = [int(r) for r in rows]
        if sum(rows) != n:
            raise ValueError("slab heights %s do not sum to n=%d" % (rows, n))
        self.h = h = native.slab_halo_rows(desc.order)
        if min(rows) < h:
            raise ValueError("every slab needs at least h=%d rows (got %s)" % (h, rows))
        self.desc, self.n, self.rows = desc, n, rows
        self.dyn = bool(desc.dynamic)
        if self.dyn and nres is None:
            raise ValueError("the dynamic reservoir needs nres")
        self.starts = list(np.cumsum([0] + rows[:-1]))
        dev = psi.device
        self.P, self.cur, self.alt, self.ncur, self.nalt = [], [], [], [], []
        for r0, r in zip(self.starts, rows):
            idx = torch.tensor([(r0 - h + i) % n for i in range(r + 2 * h)], device=dev, dtype=torch.long)
            self.P.append(P.index_select(0, idx).contiguous())
            self.cur.append(psi.index_select(0, idx).contiguous())
            self.alt.append(torch.zeros_like(self.cur[-1]))
            if self.dyn:
                self.ncur.append(nres.index_select(0, idx).contiguous())
                self.nalt.append(torch.zeros_like(self.ncur[-1]))

    def ghost_bytes_per_step(self):
        """Bytes each slab receives per step (h rows from each neighbour)."""
        c = self.cur[0].element_size()
        r = self.ncur[0].element_size() if self.dyn else 0
        return 2 * self.h * self.n * (c + r)

    def exchange(self):
        """Ghost rows of every slab from its neighbours' own rows (periodic)."""
        h, k = self.h, len(self.rows)
        fields = [self.cur] + ([self.ncur] if self.dyn else [])
        for f in fields:
            for i in range(k):
                up, down = (i - 1) % k, (i + 1) % k
                ru = self.rows[up]
                f[i][0:h].copy_(f[up][ru:ru + h])                       # last h own rows of the slab above
                f[i][h + self.rows[i]:2 * h + self.rows[i]].copy_(f[down][h:2 * h])   # first h of the one below

    def step(self, steps=1):
        for _ in range(int(steps)):
            self.exchange()
            for i, r in enumerate(self.rows):
                native.slab_step(self.desc, r, self.P[i], self.cur[i], self.ncur[i] if self.dyn else None,
                                 self.alt[i], self.nalt[i] if self.dyn else None)
            self.cur, self.alt = self.alt, self.cur
            if self.dyn:
                self.ncur, self.nalt = self.nalt, self.ncur

    def gather(self):
        """(psi, nres) of the whole grid, from the slabs' own rows."""
        import torch
        h = self.h
        psi = torch.cat([c[h:h + r] for c, r in zip(self.cur, self.rows)], 0)
        nres = torch.cat([c[h:h + r] for c, r in zip(self.ncur, self.rows)], 0) if self.dyn else None
        return psi, nres

"""Ensembles: many independent basins advanced together in one model.

The member basins are laid end to end -- nodes, elements and open
boundaries renumbered with offsets -- into one union basin whose zeta
matrix is block diagonal, and the library is told the member ranges
(shyhip_set_members), so colouring and upwind tables are built per member.
Every kernel launch of a step then covers all members at once: one launch
sequence per step for the whole ensemble instead of one per member.

Each member keeps its own forcing: boundary levels (one boundary id range
per member, set together with shyhip_set_boundary_levels), wind stress
(per node), tracers.  The zeta solve runs on the union system; its
convergence test is on the union residual, so a member's zeta agrees with
its stand-alone run to the solver tolerance, not bit for bit.
"""
from __future__ import annotations

from dataclasses import dataclass

import math

import numpy as np

from .mesh import Basin


@dataclass
class Union:
    basin: Basin
    node_start: np.ndarray      # [nmem+1]
    elem_start: np.ndarray      # [nmem+1]
    bnd_start: np.ndarray       # [nmem+1] first global boundary id - 1 of each member


def union_basin(basins) -> Union:
    """Lay `basins` end to end.  All members must share the vertical grid."""
    b0 = basins[0]
    for b in basins[1:]:
        if b.nlv != b0.nlv or not np.array_equal(b.hlv, b0.hlv):
            raise ValueError("ensemble members must share the vertical grid (nlv, hlv)")
    ns = np.zeros(len(basins) + 1, np.int64)
    es = np.zeros(len(basins) + 1, np.int64)
    bs = np.zeros(len(basins) + 1, np.int64)
    for m, b in enumerate(basins):
        ns[m + 1] = ns[m] + b.nkn
        es[m + 1] = es[m] + b.nel
        bs[m + 1] = bs[m] + (len(b.ibtyp) - 1)
    ibnd = []
    for m, b in enumerate(basins):
        ib = np.asarray(b.ibnd, np.int64)
        ibnd.append(np.where(ib > 0, ib + bs[m], 0))
    ibtyp = np.concatenate([[0]] + [np.asarray(b.ibtyp[1:], np.int32) for b in basins]).astype(np.int32)
    u = Basin(
        nkn=int(ns[-1]), nel=int(es[-1]), nlv=b0.nlv,
        xgv=np.concatenate([b.xgv for b in basins]).astype(np.float32),
        ygv=np.concatenate([b.ygv for b in basins]).astype(np.float32),
        nen3v=np.concatenate([b.nen3v + ns[m] for m, b in enumerate(basins)]).astype(np.int32),
        hm3v=np.concatenate([b.hm3v for b in basins]).astype(np.float32),
        hev=np.concatenate([b.hev for b in basins]).astype(np.float32),
        hlv=b0.hlv, hldv=b0.hldv,
        ilhv=np.concatenate([b.ilhv for b in basins]).astype(np.int32),
        ilhkv=np.concatenate([b.ilhkv for b in basins]).astype(np.int32),
        ibnd=np.concatenate(ibnd).astype(np.int32),
        ibtyp=ibtyp,
        iextb=np.concatenate([b.iextb for b in basins]).astype(np.int32),
        bnd_nodes=np.concatenate([b.bnd_nodes + ns[m] for m, b in enumerate(basins)]).astype(np.int32),
        seed=b0.seed, name=f"ensemble_{len(basins)}")
    return Union(u, ns.astype(np.int32), es.astype(np.int32), bs.astype(np.int32))


class Ensemble:
    """Members stepped together on the GPU through one libshyhip model.

    members: list of dicts with keys basin, wind (taux, tauy), level (a
    function t -> boundary level of each of the member's boundaries, or a
    float function for one boundary), tracers (dict name -> (nkn, nlv)
    array, same names for every member)."""

    def __init__(self, members, precision="fp32", params=None, czv=2.5e-3, visv=0.0, difv=0.0, deferred=False):
        import shyhip
        self.members = members
        self.u = union_basin([m["basin"] for m in members])
        names = list(members[0].get("tracers", {}) or {})
        tracers = None
        if names:
            tracers = {n: np.concatenate([np.asarray(m["tracers"][n], np.float64) for m in members]) for n in names}
        self.model = shyhip.Model(self.u.basin, precision, params=params, czv=czv, visv=visv, difv=difv,
                                  tracers=tracers, members=(self.u.node_start, self.u.elem_start))
        # wind stress per node
        tx = np.zeros(self.u.basin.nkn, np.float32)
        ty = np.zeros(self.u.basin.nkn, np.float32)
        for m, mem in enumerate(members):
            w = mem.get("wind", (0.0, 0.0))
            tx[self.u.node_start[m]:self.u.node_start[m + 1]] = w[0]
            ty[self.u.node_start[m]:self.u.node_start[m + 1]] = w[1]
        self.model.push("tauxnv", tx)
        self.model.push("tauynv", ty)
        if deferred:    # steps queue behind each other; transport overlaps the next hydro
            self.model.set_deferred_checks(True)
        self.nbnd = int(self.u.bnd_start[-1])
        self.t = 0.0
        self.level_series = False

    def levels(self, t):
        lv = np.zeros(self.nbnd, np.float32)
        for m, mem in enumerate(self.members):
            f = mem.get("level")
            if f is None:
                continue
            v = np.atleast_1d(np.asarray(f(t), np.float32))
            b0, b1 = self.u.bnd_start[m], self.u.bnd_start[m + 1]
            lv[b0:b1] = v if len(v) == b1 - b0 else v[0]
        return lv

    def set_level_series(self, t_end):
        """Every member's boundary levels held on the device as boundary
        files: records at every time step from the current time to t_end,
        interpolated with four records (the default for a level boundary);
        the steps then upload no levels."""
        dt = self.model.params_dict["dt"]
        n = int(math.ceil((t_end - self.t) / dt))
        rec_t = self.t + dt * np.arange(-2, n + 4)
        lv = np.stack([self.levels(t) for t in rec_t])
        ibtyp = self.u.basin.ibtyp
        for b in range(self.nbnd):
            if ibtyp[b + 1] == 1:
                self.model.set_boundary_series(b + 1, rec_t, lv[:, b], intpol=4, dtime0=self.t)
        self.level_series = True

    def step(self, istot=1):
        """One step of every member; istot <= 0: each member's transport
        takes the substeps scalar_stability gives it alone."""
        dt = self.model.params_dict["dt"]
        if self.nbnd and not self.level_series:
            self.model.set_boundary_levels(self.levels(self.t + dt))
        self.model.step(None, istot)
        self.t = self.model.t

    def pull(self, name, m, dtype=np.float64):
        """Field `name` of member m, in the member's own numbering."""
        a = self.model.pull(name, dtype)
        b = self.members[m]["basin"]
        if a.shape[0] == self.u.basin.nkn:
            return a[self.u.node_start[m]:self.u.node_start[m + 1]]
        if a.shape[0] == self.u.basin.nel:
            return a[self.u.elem_start[m]:self.u.elem_start[m + 1]]
        raise ValueError(f"{name}: not a node or element field ({a.shape}, member nkn {b.nkn})")

    def member_istot(self):
        return self.model.member_istot(len(self.members))

    def sync(self):
        self.model.sync()

    def close(self):
        self.model.close()

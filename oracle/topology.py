"""Topology oracle: MRAG leaf hierarchy, refine/compress planning, data transfer.

TEST INFRASTRUCTURE ONLY.  A Python restatement, written independently of the
product's C++ host code, of:

  * Grid::_setup / getBlocksInfo         MRAGcore/MRAGrid.inl:93-111,191-219
  * Grid::_computeNeighborhood           MRAGcore/MRAGrid.inl:841-1149 (periodic
                                         images, adjacency MRAGridHelpers.h:63-72)
  * Refiner_SpaceExtension::createPlan   MRAGscience/MRAGRefiner_SpaceExtension.h:25-67
  * Refiner::createPlan (chain reaction) MRAGcore/MRAGRefiner.h:43-198
  * Compressor::createPlan               MRAGcore/MRAGCompressor.h:68-146
  * Grid::refine / Grid::compress        MRAGcore/MRAGrid.inl:248-486
  * BlockSplitter::_splitBlocks          MRAGcore/MRAGBlockSplitter.h:24-99
  * BlockCollapser::_collapseBlocks      MRAGcore/MRAGBlockCollapser.h:24-85
  * Science::AutomaticRefinement/Compression  MRAGscience/MRAGScienceCore.h:42-177

Pins (tests/test_oracle_planner.py, tests/test_oracle_math_pins.py): the
planners plan for plan against the reference's Refiner, Refiner_SpaceExtension
and Compressor compiled from source (oracle/ref_harness_plan.cpp); the split
filter by polynomial reproduction; the AutomaticRefinement/Compression loops
by scripted-detail known answers.

Block data are numpy arrays (B,B,B) per leaf keyed by blockID, indexed [z,y,x].
Block IDs are handed out by a monotonically increasing counter, as
BlockCollection::_createIDs (MRAGBlockCollection.inl:93-101).
"""
import itertools

import numpy as np

from . import oracle as O

HS4 = {-3: -1 / 16., -2: 0., -1: 9 / 16., 0: 1., 1: 9 / 16., 2: 0., 3: -1 / 16.}
HS2 = {-1: 0.5, 0: 1., 1: 0.5}


class Topology:
    def __init__(self, bpd, max_level, B=16, order=4, res_jump=1):
        l = 0
        while (1 << l) < bpd:
            l += 1
        assert (1 << l) == bpd, "bpd must be a power of two"
        self.B = B
        self.order = order
        self.max_level = max_level
        self.res_jump = res_jump
        self._next = 0
        self.leaves = {}  # (l,x,y,z) -> blockID
        for x, y, z in itertools.product(range(bpd), repeat=3):
            self.leaves[(l, x, y, z)] = self._new_id()

    def _new_id(self):
        i = self._next
        self._next += 1
        return i

    # -- queries -------------------------------------------------------
    def keys(self):
        return sorted(self.leaves)

    def finest(self):
        return max(k[0] for k in self.leaves)

    def _lut(self):
        sig = (id(self.leaves), len(self.leaves), self._next)
        if getattr(self, "_lut_sig", None) != sig:
            L = max(self.max_level, self.finest())
            nF = 1 << L
            lut = np.empty((nF, nF, nF), object)
            for k in self.leaves:
                d = L - k[0]
                x, y, z = (c << d for c in k[1:])
                s = 1 << d
                for zz in range(z, z + s):
                    for yy in range(y, y + s):
                        for xx in range(x, x + s):
                            lut[zz, yy, xx] = k
            self._lut_cache = (L, nF, lut)
            self._lut_sig = sig
        return self._lut_cache

    def neighbors(self, key):
        """Real leaves adjacent to `key` (boxes touching at a face, edge or
        corner) directly or through a periodic image, as a set of
        (neighbor_key, via_image) pairs.  A leaf may be adjacent to its own
        periodic image on coarse levels, as in the reference."""
        L, nF, lut = self._lut()
        d = L - key[0]
        bs = [c << d for c in key[1:]]
        be = [(c + 1) << d for c in key[1:]]
        out = set()
        for z in range(bs[2] - 1, be[2] + 1):
            for y in range(bs[1] - 1, be[1] + 1):
                for x in range(bs[0] - 1, be[0] + 1):
                    inside = bs[0] <= x < be[0] and bs[1] <= y < be[1] and bs[2] <= z < be[2]
                    if inside:
                        continue
                    img = not (0 <= x < nF and 0 <= y < nF and 0 <= z < nF)
                    k2 = lut[z % nF, y % nF, x % nF]
                    if k2 == key and not img:
                        continue
                    out.add((k2, img))
        return out

    def leaf_grid(self):
        """(keys, LeafGrid) with leaves in sorted key order."""
        keys = self.keys()
        lv = [k[0] for k in keys]
        ix = [k[1:] for k in keys]
        return keys, O.LeafGrid(lv, ix, self.B, max(self.max_level, max(lv)), self.order)

    # -- planning --------------------------------------------------------
    def plan_refine(self, flagged_ids):
        """Set of leaf keys that Grid::refine would split for `flagged_ids`."""
        flag = {k for k, i in self.leaves.items() if i in flagged_ids}
        # Refiner_SpaceExtension step 3: a leaf with a finer, non-image neighbour
        for k in self.leaves:
            for n, img in self.neighbors(k):
                if not img and n[0] > k[0]:
                    flag.add(k)
                    break
        # step 4: never refine leaves on the finest allowed level
        flag = {k for k in flag if k[0] != self.max_level}
        # Refiner::createPlan step 3: chain reaction on the level jump
        victims = set(flag)
        old = set(flag)
        while True:
            new = set()
            for v in old:
                nl = v[0] + 1
                for n, _ in self.neighbors(v):
                    if abs(n[0] - nl) > self.res_jump and n not in victims:
                        new.add(n)
                        victims.add(n)
            if not new:
                break
            old = new
        return victims

    def plan_compress(self, flagged_ids):
        """Parents (l,x,y,z) that Compressor::createPlan would collapse."""
        flag = {k for k, i in self.leaves.items() if i in flagged_ids}
        parents = set()
        for k in self.leaves:
            if k[0] > 0:
                parents.add((k[0] - 1, k[1] // 2, k[2] // 2, k[3] // 2))
        out = []
        for p in sorted(parents):
            ch = [(p[0] + 1, 2 * p[1] + c[0], 2 * p[2] + c[1], 2 * p[3] + c[2])
                  for c in itertools.product((0, 1), repeat=3)]
            if not all(c in self.leaves for c in ch):
                continue
            if not all(c in flag for c in ch):
                continue
            ok = True
            for c in ch:
                for n, _ in self.neighbors(c):
                    if abs(n[0] - p[0]) > self.res_jump:
                        ok = False
                        break
                if not ok:
                    break
            if ok:
                out.append(p)
        return out


class AdaptiveField:
    """Leaf data (phi, pw, pg) on a Topology with MRAG's split/collapse."""

    FIELDS = ("phi", "pw", "pg")

    def __init__(self, topo):
        self.t = topo
        self.data = {}  # blockID -> dict field -> (B,B,B) float64

    def set_from_volume(self, name, vol_fn):
        """vol_fn(level, origin_xyz, h) -> (B,B,B) array [z,y,x]."""
        B = self.t.B
        for k, i in self.t.leaves.items():
            l = k[0]
            h = 2.0 ** (-l) / B
            org = np.array(k[1:]) * 2.0 ** (-l)
            self.data.setdefault(i, {})[name] = np.asarray(vol_fn(l, org, h), np.float64)

    def stacked(self, name, keys):
        return np.stack([self.data[self.t.leaves[k]][name] for k in keys])

    def fwt_details(self, ids=None):
        """blockID -> detail (FP64 oracle FWT on the current topology)."""
        keys, g = self.t.leaf_grid()
        phi = self.stacked("phi", keys)
        pos = {self.t.leaves[k]: i for i, k in enumerate(keys)}
        which = None if ids is None else [pos[i] for i in ids]
        det = g.fwt_details(phi, which)
        if ids is None:
            return {self.t.leaves[k]: det[i] for i, k in enumerate(keys)}
        return {i: det[pos[i]] for i in ids}

    def _split(self, parents):
        """BlockSplitter::_splitBlocks as a gather: every child point d gets
        sum over parent lab points s, in the reference's loop order (s_z, s_y,
        s_x increasing), of lab(s) * ((wZ*wY)*wX) with w = Hs(2s - d).  Terms
        with weight Hs(+-2) = 0 add an exact zero and are skipped."""
        B = self.t.B
        keys, g = self.t.leaf_grid()
        kidx = {k: i for i, k in enumerate(keys)}
        if self.t.order == 4:
            taps = {0: [(0, 1.0)], 1: [(-1, -1 / 16.), (0, 9 / 16.), (1, 9 / 16.), (2, -1 / 16.)]}
        else:
            taps = {0: [(0, 1.0)], 1: [(0, 0.5), (1, 0.5)]}
        m = np.arange(B)
        newdata = {}
        stacked = {f: self.stacked(f, keys) for f in self.FIELDS}
        for p in sorted(parents):
            child = {}
            for f in self.FIELDS:
                lab = g.fill_lab(stacked[f], kidx[p], -1, 3)  # lab[s+1] for s in [-1, B+2)
                out = np.zeros((2 * B, 2 * B, 2 * B))
                for pz, py, px in itertools.product((0, 1), repeat=3):
                    acc = np.zeros((B, B, B))
                    for oz, wz in taps[pz]:
                        for oy, wy in taps[py]:
                            wzy = wz * wy
                            for ox, wx in taps[px]:
                                w = wzy * wx
                                acc = acc + lab[np.ix_(m + oz + 1, m + oy + 1, m + ox + 1)] * w
                    out[pz::2, py::2, px::2] = acc
                child[f] = out
            newdata[p] = child
        return newdata

    def refine(self, flagged_ids):
        parents = self.t.plan_refine(flagged_ids)
        if not parents:
            return 0
        B = self.t.B
        newdata = self._split(parents)
        for p in sorted(parents):
            pid = self.t.leaves.pop(p)
            del self.data[pid]
            for c in itertools.product((0, 1), repeat=3):
                ck = (p[0] + 1, 2 * p[1] + c[0], 2 * p[2] + c[1], 2 * p[3] + c[2])
                cid = self.t._new_id()
                self.t.leaves[ck] = cid
                self.data[cid] = {f: newdata[p][f][c[2] * B:(c[2] + 1) * B, c[1] * B:(c[1] + 1) * B,
                                                   c[0] * B:(c[0] + 1) * B].copy() for f in self.FIELDS}
        return len(parents)

    def compress(self, flagged_ids):
        plan = self.t.plan_compress(flagged_ids)
        B = self.t.B
        h = B // 2
        for p in plan:
            pid = self.t._new_id()
            nd = {f: np.zeros((B, B, B)) for f in self.FIELDS}
            for c in itertools.product((0, 1), repeat=3):
                ck = (p[0] + 1, 2 * p[1] + c[0], 2 * p[2] + c[1], 2 * p[3] + c[2])
                cid = self.t.leaves.pop(ck)
                for f in self.FIELDS:
                    nd[f][c[2] * h:(c[2] + 1) * h, c[1] * h:(c[1] + 1) * h, c[0] * h:(c[0] + 1) * h] = \
                        self.data[cid][f][0::2, 0::2, 0::2]
                del self.data[cid]
            self.t.leaves[p] = pid
            self.data[pid] = nd
        return len(plan)  # Grid::compress reports the number of collapses

    def automatic_refinement(self, tol, max_level, max_loops=1):
        """Science::AutomaticRefinement<0,0> (MRAGScienceCore.h:42-108)."""
        nice = set()
        loops = 0
        total = 0
        while True:
            ids = [i for k, i in self.t.leaves.items() if not (max_level >= 0 and max_level <= k[0])
                   and i not in nice]
            if not ids:
                break
            det = self.fwt_details(ids)
            flagged = {i for i in ids if det[i] >= tol}
            nice |= {i for i in ids if det[i] < tol}
            if not flagged:
                break
            total += self.refine(flagged)
            loops += 1
            if max_loops > 0 and loops >= max_loops:
                break
        return total

    def automatic_compression(self, tol, max_loops=-1):
        """Science::AutomaticCompression<0,0> (MRAGScienceCore.h:117-177)."""
        nice = set()
        loops = 0
        total = 0
        while True:
            ids = [i for i in self.t.leaves.values() if i not in nice]
            det = self.fwt_details(ids)
            flagged = {i for i in ids if det[i] < tol}
            nice |= {i for i in ids if det[i] >= tol}
            if not flagged:
                break
            n = self.compress(flagged)
            if n == 0:
                break
            total += n
            loops += 1
            if max_loops > 0 and loops >= max_loops:
                break
        return total


def run_adaptive(field, Dw, Dg, rho, dt, n_steps, adapt_every, tol_refine, tol_compress, max_level, threads=8,
                 log=None):
    """Glioma_HG_UQ::run (Glioma/Glioma_HG_UQ.cpp:544-575) in FP64 on the
    oracle's adaptive field, with the adaptation after every full chunk of
    `adapt_every` steps (AutomaticRefinement with one loop, then
    AutomaticCompression without a loop limit) and the final refinement.
    Returns the leaf count after every adaptation and after the final one."""
    history = []
    done = 0
    while done < n_steps:
        m = min(adapt_every, n_steps - done)
        keys, g = field.t.leaf_grid()
        st = {f: field.stacked(f, keys) for f in field.FIELDS}
        phi = g.fk_steps(st["phi"], st["pw"], st["pg"], Dw, Dg, rho, dt, m, threads)
        for i, k in enumerate(keys):
            field.data[field.t.leaves[k]]["phi"] = phi[i]
        done += m
        if m == adapt_every:
            field.automatic_refinement(tol_refine, max_level, 1)
            field.automatic_compression(tol_compress, -1)
            history.append(len(field.t.leaves))
            if log:
                log("oracle: step %d leaves %d" % (done, len(field.t.leaves)))
    field.automatic_refinement(tol_refine, max_level, 1)
    history.append(len(field.t.leaves))
    return history


def run_adaptive_days(field, Dw, Dg, rho, dt, T_days, dumpfreq, tol_refine, tol_compress, max_level, threads=8):
    """Glioma_HG_UQ::run with the reference's own time schedule
    (Glioma/Glioma_HG_UQ.cpp:541-575): `while (t <= tend)` one step, t += dt,
    and whenever t >= whenToWrite the adaptation (refine with one loop,
    compress without a loop limit) and whenToWrite += dumpfreq; the final
    refinement.  Returns (steps, leaf count after every adaptation)."""
    t, when, steps = 0.0, float(dumpfreq), 0
    history = []
    pending = 0

    def flush(m):
        keys, g = field.t.leaf_grid()
        st = {f: field.stacked(f, keys) for f in field.FIELDS}
        phi = g.fk_steps(st["phi"], st["pw"], st["pg"], Dw, Dg, rho, dt, m, threads)
        for i, k in enumerate(keys):
            field.data[field.t.leaves[k]]["phi"] = phi[i]

    while t <= T_days:
        pending += 1
        t += dt
        steps += 1
        if t >= when:
            flush(pending)
            pending = 0
            field.automatic_refinement(tol_refine, max_level, 1)
            field.automatic_compression(tol_compress, -1)
            history.append(len(field.t.leaves))
            when += dumpfreq
    if pending:
        flush(pending)
    field.automatic_refinement(tol_refine, max_level, 1)
    history.append(len(field.t.leaves))
    return steps, history

"""Shared test helpers: toy hierarchies and lab builders (CPU only)."""
import itertools

import numpy as np


def balanced_leaves(base, refine_points, max_level):
    """Uniform grid at `base`, then refine the leaves containing each point
    (x,y,z in [0,1)) down to max_level, keeping neighbouring levels within one
    (periodic); returns sorted list of (level, x, y, z)."""
    leaves = {(base,) + t for t in itertools.product(range(1 << base), repeat=3)}

    def containing(p, lv):
        for l in range(lv + 1):
            k = (l,) + tuple(int(np.floor(c * (1 << l))) for c in p)
            if k in leaves:
                return k
        return None

    def split(k):
        leaves.remove(k)
        l, x, y, z = k
        for c in itertools.product((0, 1), repeat=3):
            leaves.add((l + 1, 2 * x + c[0], 2 * y + c[1], 2 * z + c[2]))

    def balance():
        changed = True
        while changed:
            changed = False
            for k in sorted(leaves):
                l, x, y, z = k
                n = 1 << l
                for d in itertools.product((-1, 0, 1), repeat=3):
                    if d == (0, 0, 0):
                        continue
                    # neighbour region centre at level l, check coarser leaves there
                    q = [((c + dd) % n + 0.5) / n for c, dd in zip((x, y, z), d)]
                    kk = containing(q, l)
                    if kk is not None and kk[0] < l - 1:
                        split(kk)
                        changed = True
                        break
                if changed:
                    break

    for p in refine_points:
        while True:
            k = containing(p, max_level)
            if k[0] >= max_level:
                break
            split(k)
            balance()
    return sorted(leaves)


def leaf_arrays(leaves):
    lv = np.array([k[0] for k in leaves], np.int32)
    ix = np.array([k[1:] for k in leaves], np.int32)
    return lv, ix


def periodic_lab(vol, bx, by, bz, B=16, halo=1):
    """Lab of a block of a uniform periodic volume vol[z,y,x] with `halo` cells."""
    N = vol.shape[0]
    r = np.arange(-halo, B + halo)
    iz = (bz * B + r) % N
    iy = (by * B + r) % N
    ix = (bx * B + r) % N
    return vol[np.ix_(iz, iy, ix)]

"""Views of a state's particles in the kernels' internal (half-cell) sorted
order. clsph_download_sorted_aos returns the reference's sorted buffer,
ordered stably by cell id; the internal order groups each cell by octant.
The particle at internal position k sits at reference position slot[k],
where the reference order is the stable sort of the canonical order by cell
id, i.e. by (cell id, canonical index)."""
import numpy as np


def slots(st):
    keys, perm = st.sorted()
    cell = keys >> 3  # single scene: key_base is 0
    order = np.lexsort((perm, cell))
    slot = np.empty(len(keys), np.int64)
    slot[order] = np.arange(len(keys))
    return slot, perm


def internal_sorted(st):
    """(particles in internal order, perm: internal -> canonical index)."""
    slot, perm = slots(st)
    return st.download_sorted()[slot], perm

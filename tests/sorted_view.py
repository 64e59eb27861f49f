"""Views of a state's particles in the kernels' internal (half-cell) sorted
order. clsph_download_sorted_aos and clsph_read_sorted return the reference's
sorted buffer, ordered stably by cell id; the kernels order each cell by
octant as well. clsph_read_order maps internal position k to reference
position slot[k] and canonical index perm[k]."""


def internal_sorted(st):
    """(particles in internal order, perm: internal -> canonical index)."""
    slot, perm = st.order()
    return st.download_sorted()[slot.astype("int64")], perm.astype("int64")

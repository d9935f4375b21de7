"""Slab partition helper (reference src/utilites.py:18-24).

Only ``get_index`` carries over: it is the slab decomposition a multi-GPU
build would use; on one GPU it returns (0, N).  Rank r of n gets
N//n points, plus one for the first N%n ranks.
"""


def get_index(N, n, rank):
    base, extra = divmod(N, n)
    start = rank * base + min(rank, extra)
    return (start, start + base + (1 if rank < extra else 0))

"""Inputs for the scan regression test: volumes whose list of non-empty
tiles has a length at which two scan workgroups' slot indices reach its last
entry (see tests/test_scan_part.py).  The tile decomposition is the
extraction's: a tile is 512 consecutive in-plane cube ids of one plane z < Z-1;
the scan splits the list over ceil(tiles / 1024) workgroups."""
import numpy as np

TILE = 512
SCAN_ITEMS = 1024


def tile_counts(nv):
    """Vertex count of every tile of one volume, from the per-cube counts
    (Z, Y, X) of the oracle's classify."""
    Z, Y, X = nv.shape
    if Z < 2:
        return np.zeros(0, np.int64)
    R = -(-X * Y // TILE)
    planes = nv[:Z - 1].reshape(Z - 1, X * Y).astype(np.int64)
    planes = np.pad(planes, ((0, 0), (0, R * TILE - X * Y)))
    return planes.reshape(Z - 1, R, TILE).sum(axis=2).reshape(-1)


def index_writers(nz, tiles):
    """Workgroups whose slot indices (in range or not) reach entry nz - 1."""
    nblocks = -(-tiles // SCAN_ITEMS)
    L = -(-nz // nblocks)
    c = -(-L // 256)
    n = 0
    for blk in range(nblocks):
        lb = min(blk * L, nz)
        if lb <= nz - 1 < lb + 256 * c:
            n += 1
    return n

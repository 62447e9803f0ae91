"""Tile-shape autotuner for the XC build.

Times every candidate layout -- function tile (64 / 96 / 128) x contraction
variant (K step 16 with two LDS stages, 32 with one, 32 with two) -- on a
sample chunk of the real grid and basis, with the real density, and returns
the fastest.  Tiles that pad the basis by more than 25 % over the best
padding are skipped.  Results are cached per (nbas, sample, precisions) in the
process; tune() is cheap next to one full build of a large system.
"""
import numpy as np

from .xc import BF16X3, FP32, FP64, PBE, XCBuilder

TILES = (64, 96, 128)
VARIANTS = (0, 1, 2)
_CACHE = {}


def _padded(n, t):
    return -(-n // t) * t


def candidate_tiles(nbas, slack=1.25):
    best = min(_padded(nbas, t) for t in TILES)
    return [t for t in TILES if _padded(nbas, t) <= slack * best]


def _time(basis, grid, C, occ, tile, var64, var32, precision, reps):
    with XCBuilder(basis, grid, tile=tile, variants=(var64, var32)) as xc:
        xc.set_orbitals(C, occ)
        xc.build(PBE, precision, fetch=False)  # warm-up (allocation, plan)
        ts = []
        for _ in range(reps):
            xc.build(PBE, precision, fetch=False)
            ts.append(xc.timing()["total"])
    return min(ts)


def tune(basis, grid, C, occ, sample_points=65536, reps=2, mixed=FP32, log=None):
    """Returns dict(tile, variant_fp64, variant_fp32, table) for this basis.
    The tile is chosen for the mixed-precision build (the headline path);
    the FP64 and FP32 variants are then chosen for that tile."""
    n = basis.nbas
    npts = min(sample_points, grid.shape[0])
    # a contiguous sample from the middle of the grid: inner and outer shells
    lo = (grid.shape[0] - npts) // 2
    sample = np.asfortranarray(grid[lo:lo + npts])
    key = (n, npts, mixed, float(np.sum(sample[:, 3])))
    if key in _CACHE:
        return _CACHE[key]
    table = {}
    mixed_prec = FP32 if mixed == FP32 else BF16X3
    best_tile, best_t = None, None
    for tile in candidate_tiles(n):
        # the split-BF16 kernels have a single variant; FP32 sweeps them
        vs = VARIANTS if mixed_prec == FP32 else (1,)
        for v in vs:
            t = _time(basis, sample, C, occ, tile, 0, v, mixed_prec, reps)
            table[("mixed", tile, v)] = t
            if log:
                log("autotune mixed tile=%d variant=%d: %.3f ms" % (tile, v, t))
            if best_t is None or t < best_t:
                best_tile, best_t, best32 = tile, t, v
    best64, b64 = 0, None
    for v in VARIANTS:
        t = _time(basis, sample, C, occ, best_tile, v, best32, FP64, reps)
        table[("fp64", best_tile, v)] = t
        if log:
            log("autotune fp64 tile=%d variant=%d: %.3f ms" % (best_tile, v, t))
        if b64 is None or t < b64:
            best64, b64 = v, t
    if mixed_prec == BF16X3:
        best32 = 1
        for v in VARIANTS:
            t = _time(basis, sample, C, occ, best_tile, best64, v, FP32, reps)
            table[("fp32", best_tile, v)] = t
            if t < table.get(("fp32", best_tile, best32), np.inf):
                best32 = v
    res = dict(tile=best_tile, variant_fp64=best64, variant_fp32=best32,
               table={"%s/tile%d/var%d" % k: v for k, v in table.items()})
    _CACHE[key] = res
    return res

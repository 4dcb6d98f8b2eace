"""Measured choice of the tile kernels' tuning knobs.

``autotune`` times the planned grid + degrid round trip of a caller's own
data (device-resident coordinates, visibilities and kernel bank) for a small
set of (tile, chunk, group_kib) settings and keeps the fastest on the
context.  Results do not depend on the knobs (gridding is bit-identical for
every setting), so the choice is purely one of speed.
"""

import itertools
import time

import torch

from . import _lib
from . import synthesis as G

TILES = (0, 48, 32)          # 0: the largest tile the support allows
CHUNKS = (512, 1024, 2048)   # visibilities per work item
GROUPS = (0,)                # plane groups: automatic (one per XCD for large banks)


def autotune(bank, N, uvw, lam, vis, mode=_lib.PLANE_SINGLE, vplane=None, wstep=1.0, wbin0=0,
             repeats=3, tiles=TILES, chunks=CHUNKS, groups=GROUPS, verbose=False):
    """Time every (tile, chunk, group_kib) setting on this workload and set
    the fastest.  Returns ((tile, chunk, group_kib), {setting: ms})."""
    guv = torch.zeros((N, N), dtype=torch.complex64, device=uvw.device)
    out = torch.empty(uvw.shape[0], dtype=torch.complex64, device=uvw.device)
    plan = G.GridPlan(bank, N)
    times = {}
    for tile, chunk, gk in itertools.product(tiles, chunks, groups):
        if tile and tile + bank.max_support - 1 > 95:
            continue
        _lib.set_tuning(tile, chunk, gk)
        best = float('inf')
        for r in range(repeats + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plan.bin(uvw, lam, mode, vplane, wstep, wbin0)
            guv.zero_()
            plan.grid(guv, vis)
            plan.degrid(guv, out)
            torch.cuda.synchronize()
            if r:  # first run warms the setting's buffers
                best = min(best, 1e3 * (time.perf_counter() - t0))
        times[(tile, chunk, gk)] = best
        if verbose:
            print('tile %2d chunk %5d group_kib %d: %.3f ms' % (tile, chunk, gk, best))
    choice = min(times, key=times.get)
    _lib.set_tuning(*choice)
    return choice, times

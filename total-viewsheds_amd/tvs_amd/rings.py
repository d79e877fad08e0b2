"""Ring records: compact form, the fixed-slot layout, and single-viewshed
reconstruction.

A band's ring data (kernel.rs:141-243) is its PoV followed by alternating
close/open DEM ids; ring i opens at ids[2i] (the PoV for i = 0) and closes at
ids[2i+1].  `Engine.compact_rings(sector)` returns, per band k (forward bands
0..N-1, then backward), the count slot (1 + number of events) and all
events back to back.
"""
from __future__ import annotations

import numpy as np


def band_povs(geometry) -> np.ndarray:
    """PoV DEM id of every band k (kernel.rs:117-131)."""
    N = geometry.computable_points_count
    t = np.arange(N, dtype=np.int64)
    pov = (t // geometry.tvs_width + geometry.max_los_as_points) * geometry.dem_width + \
        (t % geometry.tvs_width + geometry.max_los_as_points)
    return np.concatenate([pov, pov]).astype(np.uint32)


def offsets(counts: np.ndarray) -> np.ndarray:
    """Start of each band's events in the packed list."""
    ev = counts.astype(np.int64) - 1
    return np.concatenate([[0], np.cumsum(ev)[:-1]]) if ev.size else ev


def to_records(geometry, counts: np.ndarray, events: np.ndarray, R: int) -> np.ndarray:
    """The fixed-slot layout (compute.rs:46-47, kernel.rs:141-243): u32[2N*R]
    with [count, PoV, events...] per band, zero elsewhere; events past R
    slots are dropped, as tvs_read_rings does."""
    bands = counts.size
    rec = np.zeros((bands, R), np.uint32)
    if R == 0:
        return rec.reshape(-1)
    rec[:, 0] = counts
    if R > 1:
        rec[:, 1] = band_povs(geometry)
    off = offsets(counts)
    for k in range(bands):
        n = int(counts[k]) - 1
        keep = min(n, R - 2)
        if keep > 0:
            rec[k, 2:2 + keep] = events[off[k]:off[k] + keep]
    return rec.reshape(-1)


def viewshed_rings(geometry, pov_id: int, sectors) -> list[tuple[int, int]]:
    """Every (opening, closing) DEM-id pair of the observer at `pov_id` over
    the given sectors' compact rings (a list of (counts, events))."""
    povs = band_povs(geometry)
    pairs = []
    for counts, events in sectors:
        off = offsets(counts)
        for k in np.nonzero(povs == pov_id)[0]:
            ids = [pov_id] + [int(x) for x in events[off[k]:off[k] + int(counts[k]) - 1]]
            pairs.extend((ids[2 * i], ids[2 * i + 1]) for i in range(len(ids) // 2))
    return pairs


def observer_rings(pov_id: int, counts: np.ndarray, events: np.ndarray) -> list[tuple[int, int]]:
    """Every (opening, closing) pair of one observer's bands as
    `Engine.observer_walk` returns them (counts and packed events)."""
    off = offsets(counts)
    pairs = []
    for k in range(counts.size):
        ids = [pov_id] + [int(x) for x in events[off[k]:off[k] + int(counts[k]) - 1]]
        pairs.extend((ids[2 * i], ids[2 * i + 1]) for i in range(len(ids) // 2))
    return pairs


def ascii_viewshed(geometry, pov_id: int, sectors=None, pairs=None) -> list[str]:
    """One observer's viewshed as text rows: '+' where a ring opens, '-'
    where one closes, '±' both, 'o' the observer (the rendering of
    output/ascii.rs:88-104), from the sectors' compact rings or from ready
    (opening, closing) pairs."""
    cells = ["."] * geometry.dem_size
    if pairs is None:
        pairs = viewshed_rings(geometry, pov_id, sectors)
    for opening, closing in pairs:
        cells[opening] = "+" if cells[opening] == "." else "±"
        cells[closing] = "-" if cells[closing] == "." else "±"
        if opening == closing:
            cells[opening] = "±"
    cells[pov_id] = "o"
    w = geometry.dem_width
    return [" ".join(cells[r * w:(r + 1) * w]) for r in range(w)]

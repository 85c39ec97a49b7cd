"""Synthetic surveys for the fold tests: an orthogonal swath in a frame
rotated to an azimuth (a along it, c 90 degrees clockwise of it), X
records whose patches are clipped at the survey's edges.  This build's own
generator of the geometry the reference's fold tests describe."""
import math

import numpy as np

RL, ST, SL, SP = 10, 81, 20, 36          # receiver lines, stations; source lines, points
X0, Y0 = 512_000.0, 4_300_000.0


def at(az, a, c, x0=X0, y0=Y0):
    """Frame position (a along the azimuth, c clockwise of it) -> (x, y)."""
    t = az * (math.pi / 180.0)
    return x0 + a * math.sin(t) + c * math.cos(t), y0 + a * math.cos(t) - c * math.sin(t)


def receiver(i, k):
    return 50.0 * k, 200.0 * i


def source(m, q):
    return 25 + 200.0 * m, 25 + 50.0 * q


def swath_pairs():
    """Every (m, q, i, k) source-receiver pair of the swath: the 4 receiver
    lines nearest the source (2 on each side) x 32 stations (16 on each
    side), clipped at the edges."""
    out = []
    for m in range(SL):
        for q in range(SP):
            c = source(m, q)[1]
            i0 = int(math.floor(c / 200))
            for i in range(max(0, i0 - 1), min(RL - 1, i0 + 2) + 1):
                for k in range(max(0, 4 * m - 15), min(ST - 1, 4 * m + 16) + 1):
                    out.append((m, q, i, k))
    return out


def swath_grid(az):
    """The swath's natural grid: 25 x 25 bins centred on every midpoint,
    crosslines along the azimuth from (a, c) = (12.5, 12.5)."""
    from processseg_amd.fold import gridBinning
    x0, y0 = at(az, 12.5, 12.5)
    return gridBinning(x0, y0, az, 25.0, 25.0, inlines=72, crosslines=157)


def swath_keys(az, pairs=None):
    """[pairs, 10] key table (fold.KEY_COLUMNS): coordinates and |offset|
    for every pair, the midpoint given on every second one (NaN on the
    others: made from the coordinates)."""
    pairs = swath_pairs() if pairs is None else pairs
    k = np.full((len(pairs), 10), np.nan)
    for n, (m, q, i, kk) in enumerate(pairs):
        sx, sy = at(az, *source(m, q))
        rx, ry = at(az, *receiver(i, kk))
        k[n, 0:4] = sx, sy, rx, ry
        k[n, 6] = math.hypot(rx - sx, ry - sy)
        if (n + 1) % 2 == 0:
            k[n, 4:6] = (sx + rx) / 2, (sy + ry) / 2
    return k


def lattice_bins(pairs):
    """Each pair's bin on the natural grid from integers alone: the midpoint
    is at a = 12.5 + 25 (k + 4m), c = 12.5 + 25 (4i + q)."""
    return np.array([(4 * i + q) * 157 + (kk + 4 * m) for m, q, i, kk in pairs], np.int64)

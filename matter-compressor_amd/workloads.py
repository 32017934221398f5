"""Workload definitions shared by the benchmark and the tests.

Each workload is a list of chunks with a generator seed, an air fraction, a
quality and a max-error bound.  Per-chunk parameters that are drawn at random
come from a SplitMix64 stream, so they are identical on every machine.
"""
import math

MASK64 = (1 << 64) - 1

# Calibrated preset ladder (q, tau), same values as the reference's
# g_presets (src/mc_codec.c:504-513).
PRESETS = [(0.5, 1), (0.5, 2), (1.0, 4), (2.5, 8), (6.0, 16), (16.0, 32), (32.0, 64), (64.0, 128)]


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK64
    return x ^ (x >> 31)


def unit(seed, i, j):
    """Deterministic U[0,1) draw number (i, j) of stream `seed`."""
    r = splitmix64(splitmix64(seed) ^ splitmix64((i << 8) | j))
    return (r >> 11) * (1.0 / 9007199254740992.0)


def plain_chunks(n=64, q=6.0, air=0.419, first_seed=1):
    """One 256^3 chunk per seed (seeds first_seed..first_seed+n-1), one quality."""
    return [dict(seed=first_seed + i, air=air, q=q, tau=0, preset=None) for i in range(n)]


def mixed_chunks(n=256, first_seed=1001, param_seed=7):
    """Mixed-quality batch: air ~ U(0.30, 0.90), preset ~ U{0..7} per chunk."""
    out = []
    for i in range(n):
        air = 0.30 + 0.60 * unit(param_seed, i, 0)
        pi = min(7, int(math.floor(8.0 * unit(param_seed, i, 1))))
        q, tau = PRESETS[pi]
        out.append(dict(seed=first_seed + i, air=air, q=q, tau=tau, preset=pi))
    return out


def volume_chunks(nc=8, first_seed=20001, field_seed=13):
    """An nc^3-chunk volume (2048^3 voxels for nc=8): chunk (cz, cy, cx) has
    generator seed first_seed + index and air fraction 0.30 + 0.65 * v, v a
    U(0,1) field on the chunk grid smoothed with [1,2,1] along each axis
    (clamped at the edges) and min-max normalised, so air varies smoothly
    across the volume."""
    u = [[[unit(field_seed, (cz * nc + cy) * nc + cx, 0) for cx in range(nc)] for cy in range(nc)]
         for cz in range(nc)]

    def sm(f, axis):
        g = [[[0.0] * nc for _ in range(nc)] for _ in range(nc)]
        for z in range(nc):
            for y in range(nc):
                for x in range(nc):
                    idx = [z, y, x]
                    acc = 0.0
                    for d, w in ((-1, 1.0), (0, 2.0), (1, 1.0)):
                        j = list(idx)
                        j[axis] = min(nc - 1, max(0, j[axis] + d))
                        acc += w * f[j[0]][j[1]][j[2]]
                    g[z][y][x] = acc / 4.0
        return g

    f = sm(sm(sm(u, 2), 1), 0)
    flat = [f[z][y][x] for z in range(nc) for y in range(nc) for x in range(nc)]
    lo, hi = min(flat), max(flat)
    out = []
    for z in range(nc):
        for y in range(nc):
            for x in range(nc):
                v = (f[z][y][x] - lo) / (hi - lo) if hi > lo else 0.5
                out.append(dict(seed=first_seed + (z * nc + y) * nc + x, air=0.30 + 0.65 * v, q=6.0, tau=0,
                                preset=None, coord=(z, y, x)))
    return out


WORKLOADS = {
    # one chunk, CPU reference path
    "chunk_q6": lambda: plain_chunks(1),
    # 64 chunks of 256^3 at q=6 (1 GiB decoded), decoded one chunk per launch
    "plain64_q6": lambda: plain_chunks(64),
    # 256 chunks, mixed presets with max-error corrections
    "mixed256": lambda: mixed_chunks(256),
    # 2048^3 volume (8^3 chunks) at q=6 with a smooth air field, for serving
    "volume2048_q6": lambda: volume_chunks(8),
}

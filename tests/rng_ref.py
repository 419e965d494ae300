"""numpy restatement of the synthetic-input recipes used by the tests.

SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014): state += 0x9E3779B97F4A7C15, then two xor-shift-
multiply rounds.  Uniforms are u = (x >> 11) * 2^-53 in [0, 1).
Normals use Box-Muller on u1 = ((x >> 11) + 1) * 2^-53 in (0, 1] and u2 from the
next draw: z0 = sqrt(-2 ln u1) cos(2 pi u2), z1 = sqrt(-2 ln u1) sin(2 pi u2).
"""
import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix_u64(seed, n):
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + _G * np.arange(1, n + 1, dtype=np.uint64)
        z = s
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def splitmix_uniform(seed, n):
    return (splitmix_u64(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def splitmix_normal(seed, n):
    m = (n + 1) // 2
    u = splitmix_u64(seed, 2 * m) >> np.uint64(11)
    u1 = (u[0::2].astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = u[1::2].astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:n]


def rnd_matrix(n, seed):
    """The "rnd:N:SEED" test matrix: per row, the diagonal plus three column
    draws (u64 % n), sorted and de-duplicated; then one uniform u per kept
    column in ascending order, giving 4 + |2u - 1| on the diagonal and 2u - 1
    elsewhere.  Returns (row_ptr, col_idx, vals)."""
    draws = iter(splitmix_u64(seed, 7 * n).tolist())
    rp, ci, v = [0], [], []
    for i in range(n):
        cols = sorted({i, *(next(draws) % n for _ in range(3))})
        for j in cols:
            u = 2.0 * ((next(draws) >> 11) * 2.0 ** -53) - 1.0
            ci.append(j)
            v.append(4.0 + abs(u) if j == i else u)
        rp.append(len(ci))
    return (np.array(rp, np.int32), np.array(ci, np.int32), np.array(v))


def lcg_unit(n, seed):
    """Start vector of the analysis routines: a 32-bit LCG (multiplier 1664525,
    increment 1013904223), each state read as a signed int / 2^31, normalised."""
    v = np.empty(n)
    st = seed & 0xFFFFFFFF
    for i in range(n):
        st = (st * 1664525 + 1013904223) & 0xFFFFFFFF
        v[i] = (st - (1 << 32) if st >= (1 << 31) else st) / 2147483648.0
    return v / np.sqrt(np.sum(v * v))

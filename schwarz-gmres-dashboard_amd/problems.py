"""Synthetic test problems (host, numpy): the matrices and right-hand sides of
the benchmark configurations.

* ``poisson2d`` / ``poisson3d``: Dirichlet Laplacians with h = 1 scaling, rows in
  z-major, y, x order and neighbours in the order -nxy, -nx, -1, diag, +1, +nx,
  +nxy, as make_poisson_2d / make_poisson_3d (reference src/mesh/poisson.h:10-74).
* ``diffusion3d``: 7-point cell-centred finite-volume diffusion with harmonic-mean
  face coefficients 2 k_p k_q / (k_p + k_q); a Dirichlet boundary face adds k_p
  to the diagonal, so k == 1 gives exactly poisson3d.
* ``diffusion2d``: the 5-point analogue.
* ``block_jump_kappa``: k = 10^u, u uniform on {-3, 0, 3}, constant on cubic blocks.
* ``grf``: Gaussian random field with covariance sigma^2 exp(-r^2 / (2 l^2)) in
  unit-square coordinates (grid spacing 1/N), sampled by circulant embedding on
  a 2x-padded periodic grid: g = IFFT(sqrt(max(FFT(c), 0)) * FFT(w)), w white.
* Random numbers: SplitMix64, uniforms (x >> 11) * 2^-53, normals by
  Box-Muller on ((x >> 11) + 1) * 2^-53 and the next uniform.
"""
import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + _G * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed, n):
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def uniform_pm1(seed, n):
    return 2.0 * uniform01(seed, n) - 1.0


def normal(seed, n):
    m = (n + 1) // 2
    u = splitmix64(seed, 2 * m) >> np.uint64(11)
    u1 = (u[0::2].astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = u[1::2].astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(2.0 * np.pi * u2)
    out[1::2] = r * np.sin(2.0 * np.pi * u2)
    return out[:n]


def _assemble(shape, diag, offs):
    """CSR from per-row diagonal and neighbour lists in the reference order.

    offs: list of (flat offset, valid mask, value array) in column order."""
    n = int(np.prod(shape))
    cols, vals, masks = [], [], []
    idx = np.arange(n, dtype=np.int64)
    entries = []
    for off, mask, val in offs:
        entries.append((off, mask, val))
    # order: negative offsets (most negative first), diagonal, positive offsets
    neg = sorted([e for e in entries if e[0] < 0], key=lambda e: e[0])
    pos = sorted([e for e in entries if e[0] > 0], key=lambda e: e[0])
    ordered = neg + [(0, np.ones(n, bool), diag)] + pos
    counts = np.zeros(n, np.int64)
    for off, mask, val in ordered:
        counts += mask
    rp = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=rp[1:])
    nnz = int(rp[-1])
    ci = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    fill = rp[:-1].copy()
    for off, mask, val in ordered:
        sel = np.flatnonzero(mask)
        ci[fill[sel]] = (idx[sel] + off).astype(np.int32)
        v[fill[sel]] = val[sel] if np.ndim(val) else val
        fill[sel] += 1
    del cols, vals, masks
    return rp.astype(np.int32), ci, v


def poisson2d(nx, ny):
    n = nx * ny
    ix = np.arange(n) % nx
    iy = np.arange(n) // nx
    m1 = -np.ones(n)
    return _assemble((ny, nx), np.full(n, 4.0),
                     [(-nx, iy > 0, m1), (-1, ix > 0, m1), (1, ix < nx - 1, m1),
                      (nx, iy < ny - 1, m1)])


def poisson3d(nx, ny, nz):
    n = nx * ny * nz
    nxy = nx * ny
    i = np.arange(n)
    ix, iy, iz = i % nx, (i // nx) % ny, i // nxy
    m1 = -np.ones(n)
    return _assemble((nz, ny, nx), np.full(n, 6.0),
                     [(-nxy, iz > 0, m1), (-nx, iy > 0, m1), (-1, ix > 0, m1),
                      (1, ix < nx - 1, m1), (nx, iy < ny - 1, m1), (nxy, iz < nz - 1, m1)])


def _harm(a, b):
    return 2.0 * a * b / (a + b)


def diffusion3d(kappa):
    """kappa: array (nz, ny, nx) of cell coefficients."""
    nz, ny, nx = kappa.shape
    n = kappa.size
    k = kappa.ravel().astype(np.float64)
    nxy = nx * ny
    i = np.arange(n)
    ix, iy, iz = i % nx, (i // nx) % ny, i // nxy
    diag = np.zeros(n)
    offs = []
    for off, valid in ((-nxy, iz > 0), (-nx, iy > 0), (-1, ix > 0), (1, ix < nx - 1),
                       (nx, iy < ny - 1), (nxy, iz < nz - 1)):
        f = np.where(valid, _harm(k, k[np.clip(i + off, 0, n - 1)]), k)
        diag += f
        offs.append((off, valid, -f))
    return _assemble((nz, ny, nx), diag, offs)


def diffusion2d(kappa):
    ny, nx = kappa.shape
    n = kappa.size
    k = kappa.ravel().astype(np.float64)
    i = np.arange(n)
    ix, iy = i % nx, i // nx
    diag = np.zeros(n)
    offs = []
    for off, valid in ((-nx, iy > 0), (-1, ix > 0), (1, ix < nx - 1), (nx, iy < ny - 1)):
        f = np.where(valid, _harm(k, k[np.clip(i + off, 0, n - 1)]), k)
        diag += f
        offs.append((off, valid, -f))
    return _assemble((ny, nx), diag, offs)


def block_jump_kappa(n_side, block, seed):
    """k = 10^u, u in {-3, 0, 3} uniformly per block^3 cube (blocks in z, y, x order)."""
    nb = (n_side + block - 1) // block
    u = uniform01(seed, nb ** 3)
    expo = -3.0 + 3.0 * np.floor(3.0 * u)
    kb = (10.0 ** expo).reshape(nb, nb, nb)
    idx = np.arange(n_side) // block
    return kb[idx][:, idx][:, :, idx]


def grf(n_side, sigma, corr_len, seed):
    """Gaussian random field on an n_side^2 grid (see module docstring)."""
    M = 2 * n_side
    d = np.minimum(np.arange(M), M - np.arange(M)) / n_side
    r2 = d[:, None] ** 2 + d[None, :] ** 2
    c = sigma ** 2 * np.exp(-r2 / (2.0 * corr_len ** 2))
    lam = np.maximum(np.fft.fft2(c).real, 0.0)
    w = normal(seed, M * M).reshape(M, M)
    g = np.fft.ifft2(np.sqrt(lam) * np.fft.fft2(w)).real
    return g[:n_side, :n_side]


def batch_nparts(n):
    """Subdomains of one batched system: round(4 + 12 (n - 64^2) / (512^2 - 64^2))
    clamped to [4, 16]."""
    return int(min(16, max(4, round(4 + 12 * (n - 64 ** 2) / (512 ** 2 - 64 ** 2)))))


def batch_systems(nsys=64, base_seed=4000, side_range=(64, 512), sigma=1.0, corr_len=0.1):
    """Independent 2D diffusion systems of the batched workload.

    System i (seed s = base_seed + i): grid side = lo + floor(u * (hi - lo + 1))
    with u the first uniform of SplitMix64(s ^ 0x5EED); kappa = exp(g), g =
    grf(side, sigma, corr_len, s); b ~ N(0, 1) from SplitMix64(s + 10^6).
    Returns a list of (rp, ci, v, b)."""
    lo, hi = side_range
    out = []
    for i in range(nsys):
        s = base_seed + i
        side = lo + int(uniform01(s ^ 0x5EED, 1)[0] * (hi - lo + 1))
        rp, ci, v = diffusion2d(np.exp(grf(side, sigma, corr_len, s)))
        out.append((rp, ci, v, normal(s + 1_000_000, side * side)))
    return out

"""Seeded synthetic inputs shared by the tests and bench.py.

Workloads (named by what they are):
  plain_small        100 uniform points in [0,100)^2, EC mask 100x100 at cs 1,
                     density raster 1000x1000 at cs 0.1, all valid
  thomas_adaptive    Thomas process: parents uniform in [0,E)^2, Poisson(mu)
                     children with N(0, s^2) offsets, clipped; weights U[0.5,2];
                     mask = disk of radius 3900/8192*E minus rectangular holes
  mixture_search     50% uniform, 50% around 64 Gaussian centres (sigma 2% of
                     extent), all-valid mask
Seeds are 20260000 + workload number so every consumer sees the same bytes.
"""
import numpy as np

NODATA = -9999.0


def rng(seed):
    return np.random.default_rng(seed)


def uniform_points(n, extent=100.0, seed=20260001):
    r = rng(seed)
    x = r.uniform(0.0, extent, n).astype(np.float32)
    y = r.uniform(0.0, extent, n).astype(np.float32)
    return x, y


def clip_domain(v, extent):
    hi = np.nextafter(np.float32(extent), np.float32(0.0))
    return np.clip(v, np.float32(0.0), hi).astype(np.float32)


def thomas_points(n_parents, mu, sigma, extent, seed=20260002, weights=True):
    r = rng(seed)
    px = r.uniform(0.0, extent, n_parents)
    py = r.uniform(0.0, extent, n_parents)
    counts = r.poisson(mu, n_parents)
    cx = np.repeat(px, counts) + r.normal(0.0, sigma, counts.sum())
    cy = np.repeat(py, counts) + r.normal(0.0, sigma, counts.sum())
    x = clip_domain(cx.astype(np.float32), extent)
    y = clip_domain(cy.astype(np.float32), extent)
    w = r.uniform(0.5, 2.0, len(x)).astype(np.float32) if weights else np.ones(len(x), np.float32)
    return x, y, w


def mixture_points(n, extent, n_centres=64, sigma_frac=0.02, seed=20260003):
    r = rng(seed)
    nu = n // 2
    ux = r.uniform(0.0, extent, nu)
    uy = r.uniform(0.0, extent, nu)
    cxs = r.uniform(0.0, extent, n_centres)
    cys = r.uniform(0.0, extent, n_centres)
    k = r.integers(0, n_centres, n - nu)
    gx = cxs[k] + r.normal(0.0, sigma_frac * extent, n - nu)
    gy = cys[k] + r.normal(0.0, sigma_frac * extent, n - nu)
    x = clip_domain(np.concatenate([ux, gx]).astype(np.float32), extent)
    y = clip_domain(np.concatenate([uy, gy]).astype(np.float32), extent)
    return x, y


def disk_mask(ncols, nrows, radius_cells, n_holes=0, hole_min=50, hole_max=400, seed=20260002, nodata=NODATA):
    """Disk of valid cells (value 1.0) centred on the raster, minus holes."""
    r = rng(seed + 1000)
    rr, cc = np.ogrid[0:nrows, 0:ncols]
    cy, cx = (nrows - 1) / 2.0, (ncols - 1) / 2.0
    m = np.full((nrows, ncols), nodata, np.float32)
    m[(rr - cy) ** 2 + (cc - cx) ** 2 <= radius_cells ** 2] = 1.0
    for _ in range(n_holes):
        hw = int(r.integers(hole_min, hole_max + 1))
        hh = int(r.integers(hole_min, hole_max + 1))
        c0 = int(r.integers(0, max(1, ncols - hw)))
        r0 = int(r.integers(0, max(1, nrows - hh)))
        m[r0:r0 + hh, c0:c0 + hw] = nodata
    return m


def rect_mask(ncols, nrows, n_holes=0, hole_min=50, hole_max=400, seed=20260002, nodata=NODATA):
    """A centred rectangle of valid cells covering 40-90% of each side, minus holes."""
    r = rng(seed + 2000)
    m = np.full((nrows, ncols), nodata, np.float32)
    w = max(1, int(ncols * r.uniform(0.4, 0.9)))
    h = max(1, int(nrows * r.uniform(0.4, 0.9)))
    c0, r0 = (ncols - w) // 2, (nrows - h) // 2
    m[r0:r0 + h, c0:c0 + w] = 1.0
    for _ in range(n_holes):
        hw = int(r.integers(hole_min, hole_max + 1))
        hh = int(r.integers(hole_min, hole_max + 1))
        hc = int(r.integers(0, max(1, ncols - hw)))
        hr = int(r.integers(0, max(1, nrows - hh)))
        m[hr:hr + hh, hc:hc + hw] = nodata
    return m


def all_valid(ncols, nrows, value=0.0):
    return np.full((nrows, ncols), value, np.float32)


def mode0_grid(cs):
    """Mode-0 edge-correction mask geometry (tiled:429): int(100/cs) cells, origin (0,0)."""
    n = int(np.float32(100) / np.float32(cs))
    return n, n, 0.0, 0.0, float(np.float32(cs))


def mode0_density_tile(cs_den):
    """Mode-0 single density tile (tiled:604-605, 779): ceil(100/cs) cells,
    xll = 0, yll = 100 - rows*cs in float."""
    n = int(np.ceil(100.0 / np.float64(np.float32(cs_den))))
    cs = np.float32(cs_den)
    yll = np.float32(100.0) - np.float32(n) * cs
    return n, n, 0.0, float(yll), float(cs)


def plain_small_bandwidth(x, y):
    """The bandwidth named for the plain small workload:
    h = sqrt((sx^2 + sy^2)/2) * n^(-1/6) (sample standard deviations)."""
    n = len(x)
    sx2 = np.var(x.astype(np.float64), ddof=1)
    sy2 = np.var(y.astype(np.float64), ddof=1)
    return float(np.float32(np.sqrt((sx2 + sy2) / 2.0) * n ** (-1.0 / 6.0)))


def lognormal_clusters(n, extent, n_centres=1024, sigma=64.0, seed=20260004):
    """Cluster centres uniform in [0,E)^2 with log-normal (sigma 1.5)
    intensities normalised to multinomial counts; N(0, sigma^2) offsets."""
    r = rng(seed)
    cx = r.uniform(0.0, extent, n_centres)
    cy = r.uniform(0.0, extent, n_centres)
    lam = r.lognormal(0.0, 1.5, n_centres)
    counts = r.multinomial(n, lam / lam.sum())
    x = np.repeat(cx, counts) + r.normal(0.0, sigma, n)
    y = np.repeat(cy, counts) + r.normal(0.0, sigma, n)
    return clip_domain(x.astype(np.float32), extent), clip_domain(y.astype(np.float32), extent)

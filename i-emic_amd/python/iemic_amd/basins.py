"""Synthetic, seeded ocean basins for the benchmark and test workloads.

* box(n, m, l):   the flat-bottom North-Atlantic box of THCM's defaults
  (lon 286-350 E, lat 10-74 N, 4000 m, uniform levels, no land;
  THCM.C:2720-2727, topo.F90:177-187).
* global_basin(n, m, l, seed): lon 0-360 periodic, lat -81..81, 5000 m on
  stretched levels (qz = 2.25).  Land: a Gaussian random field with a 15
  degree correlation length, thresholded to about 30 % land, plus two
  meridional continents 20 degrees wide; the per-column depth (1..l levels)
  comes from the same field; only the largest connected ocean body is kept
  (a closed basin would leave its pressure level undetermined).

Forcing (THCM's external-field path, Wind Forcing Type 3 and Levitus T/S 2):
tau_x = -cos(2 pi (|lat| - 10) / 64) at v-point latitudes,
tatm = T* - 15 with T* = 15 cos(pi (|lat| - 42) / 64), and a salt flux
0.1 cos(pi (|lat| - 10) / 64) at cell-centre latitudes (|lat|: the northern
formulas mirrored about the equator).  State: rest velocities, T linear in
depth from T* - 15 at the surface to 4 - 15 at the bottom, S = 0, plus
N(0, 1e-3) noise on every unknown, surface w zeroed.

Masks use the reference's values: OCEAN 0, LAND 1, PERIO 3 on the periodic
border where both ends of the row are ocean (topo.F90:301-306).
"""
import numpy as np

OCEAN, LAND, PERIO = 0, 1, 3
CONFIG_FIELDS = ("n", "m", "l", "xmin_deg", "xmax_deg", "ymin_deg", "ymax_deg", "periodic",
                 "hdim", "qz", "alphaT", "alphaS", "ih", "vmix", "tap", "rho_mixing",
                 "coriolis_on", "tres", "sres", "iza", "ite", "its", "coupled_t", "coupled_s",
                 "forcing_type", "int_sign", "nic", "mic")


def default_config(**kw):
    """THCM::getDefaultInitParameters (THCM.C:2697-2765) plus overrides."""
    d = dict(n=16, m=16, l=16, xmin_deg=286.0, xmax_deg=350.0, ymin_deg=10.0, ymax_deg=74.0,
             periodic=0, hdim=4000.0, qz=1.0, alphaT=1.0e-4, alphaS=7.6e-4, ih=0, vmix=1,
             tap=1, rho_mixing=1, coriolis_on=1, tres=1, sres=1, iza=2, ite=1, its=1,
             coupled_t=0, coupled_s=0, forcing_type=0, int_sign=-1, nic=-1, mic=-1)
    d.update(kw)
    return d


def latitudes(cfg):
    dy = (cfg["ymax_deg"] - cfg["ymin_deg"]) / cfg["m"]
    y = cfg["ymin_deg"] + (np.arange(1, cfg["m"] + 1) - 0.5) * dy
    yv = cfg["ymin_deg"] + np.arange(1, cfg["m"] + 1) * dy
    return y, yv


def forcing(cfg):
    y, yv = latitudes(cfg)
    n = cfg["n"]
    ay, ayv = np.abs(y), np.abs(yv)
    taux = np.repeat((-np.cos(2 * np.pi * (ayv - 10.0) / 64.0))[:, None], n, 1)
    tstar = 15.0 * np.cos(np.pi * (ay - 42.0) / 64.0)
    tatm = np.repeat((tstar - 15.0)[:, None], n, 1)
    emip = np.repeat((0.1 * np.cos(np.pi * (ay - 10.0) / 64.0))[:, None], n, 1)
    z = np.zeros(n * cfg["m"])
    return dict(taux=taux.reshape(-1), tauy=z.copy(), tatm=tatm.reshape(-1),
                emip=emip.reshape(-1), spert=z.copy())


def seeded_state(cfg, seed, noise=1e-3):
    n, m, l = cfg["n"], cfg["m"], cfg["l"]
    y, _ = latitudes(cfg)
    tstar = 15.0 * np.cos(np.pi * (np.abs(y) - 42.0) / 64.0) - 15.0
    x = np.zeros((l, m, n, 6))
    for k in range(l):
        f = k / max(l - 1, 1)
        x[k, :, :, 4] = ((4.0 - 15.0) * (1 - f) + tstar * f)[:, None]
    rng = np.random.default_rng(seed)
    x += rng.normal(0.0, noise, x.shape)
    x[l - 1, :, :, 2] = 0.0
    return x.reshape(-1)


def box_mask(n, m, l):
    lm = np.full((l + 2, m + 2, n + 2), LAND, np.int32)
    lm[1:l + 1, 1:m + 1, 1:n + 1] = OCEAN
    return lm


def _smooth_periodic_x(f, sx, sy):
    """Gaussian smoothing, periodic in x (axis 1), reflecting in y (axis 0)."""
    m, n = f.shape
    kx = np.fft.fftfreq(n)
    g = np.fft.ifft(np.fft.fft(f, axis=1) * np.exp(-2 * (np.pi * kx * sx) ** 2)[None, :], axis=1).real
    ext = np.concatenate([g[::-1], g, g[::-1]], axis=0)
    ky = np.fft.fftfreq(3 * m)
    h = np.fft.ifft(np.fft.fft(ext, axis=0) * np.exp(-2 * (np.pi * ky * sy) ** 2)[:, None], axis=0).real
    return h[m:2 * m]


def _largest_ocean_component(lm, periodic):
    """Keep the largest 6-connected ocean body of the interior cells."""
    l2, m2, n2 = lm.shape
    ocean = lm[1:-1, 1:-1, 1:-1] == OCEAN
    lab = -np.ones(ocean.shape, np.int64)
    L, M, N = ocean.shape
    best, best_size, cur = -1, 0, 0
    for start in zip(*np.nonzero(ocean)):
        if lab[start] >= 0:
            continue
        stack = [start]
        lab[start] = cur
        size = 0
        while stack:
            k, j, i = stack.pop()
            size += 1
            for dk, dj, di in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                kk, jj, ii = k + dk, j + dj, i + di
                if periodic:
                    ii %= N
                if 0 <= kk < L and 0 <= jj < M and 0 <= ii < N and ocean[kk, jj, ii] and lab[kk, jj, ii] < 0:
                    lab[kk, jj, ii] = cur
                    stack.append((kk, jj, ii))
        if size > best_size:
            best, best_size = cur, size
        cur += 1
    keep = lab == best
    inner = lm[1:-1, 1:-1, 1:-1]
    inner[ocean & ~keep] = LAND
    return lm


def global_mask(n, m, l, seed=7, land_fraction=0.3, corr_deg=15.0, continent_deg=20.0):
    cfg = dict(n=n, m=m, ymin_deg=-81.0, ymax_deg=81.0)
    dx, dy = 360.0 / n, 162.0 / m
    rng = np.random.default_rng(seed)
    f = _smooth_periodic_x(rng.normal(size=(m, n)), corr_deg / dx, corr_deg / dy)
    f = (f - f.mean()) / f.std()
    lon = (np.arange(n) + 0.5) * dx
    lat, _ = latitudes(dict(cfg, m=m))
    land2d = f < np.quantile(f, land_fraction)
    for c0 in (70.0, 290.0):     # two meridional continents, open to the south
        band = np.abs(((lon - c0 + 180.0) % 360.0) - 180.0) <= continent_deg / 2
        rows = (lat >= -55.0) & (lat <= 75.0)
        land2d[np.ix_(rows, band)] = True
    # column depth in levels from the same field: deeper where f is larger
    q = np.clip((f - np.quantile(f, land_fraction)) / (f.max() - np.quantile(f, land_fraction)), 0, 1)
    depth = 1 + np.floor((l - 1) * np.sqrt(q) + 0.5).astype(int)
    lm = np.full((l + 2, m + 2, n + 2), LAND, np.int32)
    for j in range(m):
        for i in range(n):
            if not land2d[j, i]:
                d = min(l, depth[j, i])
                lm[l - d + 1:l + 1, j + 1, i + 1] = OCEAN
    lm = _largest_ocean_component(lm, True)
    both = (lm[:, :, 1] == OCEAN) & (lm[:, :, n] == OCEAN)
    lm[:, :, 0] = np.where(both, PERIO, LAND)
    lm[:, :, n + 1] = np.where(both, PERIO, LAND)
    return lm


def integral_row_location(lm, n, m, l):
    """Last surface ocean cell (0-based i, j) scanning from (n-1, m-1) back,
    for THCM's 'Integral row coordinate i/j' (THCM.C:663-698)."""
    for j in range(m - 1, -1, -1):
        for i in range(n - 1, -1, -1):
            if lm[l, j + 1, i + 1] == OCEAN:
                return i, j
    raise ValueError("no ocean at the surface")


def make_box(n=16, m=16, l=12, seed=11):
    """Seeded box workload (config of the 16x16x12 CPU case)."""
    cfg = default_config(n=n, m=m, l=l, sres=0, iza=3, ite=2, its=2)
    return cfg, box_mask(n, m, l).reshape(-1), forcing(cfg), seeded_state(cfg, seed), \
        dict(COMB=1.0, WIND=1.0, TEMP=1.0, SALT=1.0)


def make_global(n, m, l, mask_seed=7, noise_seed=11):
    """Seeded global workload (96x38x12 and 180x80x24 configurations)."""
    cfg = default_config(n=n, m=m, l=l, xmin_deg=0.0, xmax_deg=360.0, ymin_deg=-81.0,
                         ymax_deg=81.0, periodic=1, hdim=5000.0, qz=2.25, sres=0, iza=3, ite=2,
                         its=2)
    lm = global_mask(n, m, l, seed=mask_seed)
    cfg["nic"], cfg["mic"] = integral_row_location(lm, n, m, l)
    return cfg, lm.reshape(-1), forcing(cfg), seeded_state(cfg, noise_seed), \
        dict(COMB=1.0, WIND=1.0, TEMP=1.0, SALT=1.0)

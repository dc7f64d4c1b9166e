"""Seeded synthetic interferometer data (numpy, FP64).

Array layouts follow the w-gridder boundary: uvw float64 [n_rows, 3] (m),
rows time-major (row = t * n_baselines + b), freq float64 [n_chan] (Hz),
vis complex64 [n_rows, n_chan].

Earth-rotation synthesis: antenna ENU positions are rotated to local
equatorial XYZ at the site latitude, baselines b = X_j - X_i (i < j), and
for hour angle H and declination d

    u =  sin H bx + cos H by
    v = -sin d cos H bx + sin d sin H by + cos d bz
    w =  cos d cos H bx - cos d sin H by + sin d bz

Sky: point sources with S drawn from p(S) ~ S^-1.5 on [smin, smax] and
optional circular Gaussians; V = sum S g(u, v) exp(+2 pi i (u l + v m +
w (n - 1))) (the predict sign of WSClean's point-source generator,
gridding/msgridderdata.h:820-843).
"""
import numpy as np

C_LIGHT = 299792458.0
OMEGA_EARTH = 7.2921150e-5  # rad / s
ARCSEC = np.pi / 180.0 / 3600.0


def antennas_disc(rng, n_ant, radius_m):
    r = radius_m * np.sqrt(rng.random(n_ant))
    phi = 2 * np.pi * rng.random(n_ant)
    return np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(n_ant)], axis=1)


def antennas_logradial(rng, n_ant, radius_m, rmin_m=20.0):
    r = np.exp(rng.uniform(np.log(rmin_m), np.log(radius_m), n_ant))
    phi = 2 * np.pi * rng.random(n_ant)
    return np.stack([r * np.cos(phi), r * np.sin(phi), np.zeros(n_ant)], axis=1)


def earth_rotation_uvw(enu, n_times, dt_s, dec_rad, lat_rad=np.deg2rad(-30.0)):
    e, n, u = enu[:, 0], enu[:, 1], enu[:, 2]
    X = -np.sin(lat_rad) * n + np.cos(lat_rad) * u
    Y = e
    Z = np.cos(lat_rad) * n + np.sin(lat_rad) * u
    i, j = np.triu_indices(len(enu), k=1)
    bx, by, bz = X[j] - X[i], Y[j] - Y[i], Z[j] - Z[i]
    t = (np.arange(n_times) - (n_times - 1) / 2.0) * dt_s
    H = OMEGA_EARTH * t
    sH, cH = np.sin(H)[:, None], np.cos(H)[:, None]
    sd, cd = np.sin(dec_rad), np.cos(dec_rad)
    uu = sH * bx + cH * by
    vv = -sd * cH * bx + sd * sH * by + cd * bz
    ww = cd * cH * bx - cd * sH * by + sd * bz
    return np.stack([uu.ravel(), vv.ravel(), ww.ravel()], axis=1)


def power_law_fluxes(rng, n, smin=0.01, smax=1.0, index=-1.5):
    a = index + 1.0
    x = rng.random(n)
    return (smin ** a + x * (smax ** a - smin ** a)) ** (1.0 / a)


def point_sky(rng, n_src, npix, pixel_rad, on_pixel=False, margin=0.9):
    half = 0.5 * npix * pixel_rad * margin
    if on_pixel:
        xs = rng.integers(int(npix * (1 - margin) / 2) + 1, int(npix * (1 + margin) / 2) - 1, n_src)
        ys = rng.integers(int(npix * (1 - margin) / 2) + 1, int(npix * (1 + margin) / 2) - 1, n_src)
        l = (npix // 2 - xs) * pixel_rad
        m = (ys - npix // 2) * pixel_rad
    else:
        l = rng.uniform(-half, half, n_src)
        m = rng.uniform(-half, half, n_src)
        xs = ys = None
    return l, m, power_law_fluxes(rng, n_src), xs, ys


def predict_sources(uvw, freq, l, m, flux, gauss_sigma=None, chunk=4096):
    """Direct FP64 visibilities of point / circular Gaussian sources."""
    uvw = np.asarray(uvw, np.float64)
    n_rows = uvw.shape[0]
    vis = np.zeros((n_rows, len(freq)), np.complex128)
    nm1 = np.sqrt(1.0 - l * l - m * m) - 1.0
    for c, f in enumerate(freq):
        s = f / C_LIGHT
        for r0 in range(0, n_rows, chunk):
            u = uvw[r0:r0 + chunk, 0:1] * s
            v = uvw[r0:r0 + chunk, 1:2] * s
            w = uvw[r0:r0 + chunk, 2:3] * s
            ph = 2 * np.pi * (u * l[None, :] + v * m[None, :] + w * nm1[None, :])
            amp = flux[None, :]
            if gauss_sigma is not None:
                amp = amp * np.exp(-2 * np.pi ** 2 * gauss_sigma[None, :] ** 2 * (u * u + v * v))
            vis[r0:r0 + chunk, c] = (amp * np.exp(1j * ph)).sum(axis=1)
    return vis


def add_noise(rng, vis, sigma):
    return vis + sigma * (rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape))


def small_problem(seed=1001, n_ant=12, n_times=20, n_chan=2, npix=64, pixel_arcsec=20.0,
                  f0=150e6, df=5e6, dec_deg=-30.0, radius_m=1500.0, n_src=5, noise=0.0,
                  on_pixel=False, dt_s=60.0, sky_lm=(0.0, 0.0)):
    """A small seeded problem for parity tests: returns a dict.  The sky is
    centred on sky_lm (use the image's l/m shift to keep it in the field)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    enu = antennas_disc(rng, n_ant, radius_m)
    uvw = earth_rotation_uvw(enu, n_times, dt_s, np.deg2rad(dec_deg))
    freq = f0 + df * np.arange(n_chan)
    dx = pixel_arcsec * ARCSEC
    l, m, flux, xs, ys = point_sky(rng, n_src, npix, dx, on_pixel=on_pixel)
    l = l + sky_lm[0]
    m = m + sky_lm[1]
    vis = predict_sources(uvw, freq, l, m, flux)
    if noise > 0:
        vis = add_noise(rng, vis, noise)
    return dict(uvw=uvw, freq=freq, vis=vis.astype(np.complex64), npix=npix, dx=dx,
                l=l, m=m, flux=flux, xs=xs, ys=ys, seed=seed)


def large_wgridding_inputs(n_chan=64, npix=8192, seed=1002):
    """The large w-gridding workload: 128 antennas in a seeded 10 km
    log-radial layout (8128 baselines), 2880 x 10 s at dec -30 deg, n_chan
    channels over 120-180 MHz (64: 1.498e9 visibilities), 2 arcsec pixels,
    a sky of 500 power-law point sources and 5 circular Gaussians (10-60
    arcsec).  Returns (uvw [rows, 3] m, freq [n_chan] Hz, pixel size (rad),
    src [505, 5] = {l, m, n - 1, S, gaussian sigma}); the visibilities come
    from the device generator (gridder.synth_visibilities, noise 0.05)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    enu = antennas_logradial(rng, 128, 10000.0)
    uvw = earth_rotation_uvw(enu, 2880, 10.0, np.deg2rad(-30.0))
    freq = np.linspace(120e6, 180e6, 64)[:n_chan] if n_chan <= 64 else \
        np.linspace(120e6, 180e6, n_chan)
    dx = 2.0 * ARCSEC
    l, m, flux, _, _ = point_sky(rng, 500, npix, dx)
    gl, gm, gflux, _, _ = point_sky(rng, 5, npix, dx)
    gsig = rng.uniform(10.0, 60.0, 5) * ARCSEC
    src = np.zeros((505, 5))
    src[:500, 0], src[:500, 1], src[:500, 3] = l, m, flux
    src[500:, 0], src[500:, 1], src[500:, 3], src[500:, 4] = gl, gm, gflux, gsig
    src[:, 2] = np.sqrt(1.0 - src[:, 0] ** 2 - src[:, 1] ** 2) - 1.0
    return uvw, freq, dx, src


def batched_facets_workload(share=8, seed=1004, max_rows=None):
    """The batched-facets workload: a 12000^2 field at 2 arcsec split into a
    6 x 6 grid of 36 facets of uneven sizes (each side drawn from 1024 ...
    3072, a multiple of 16), each centred on its grid cell (its own l/m
    shift).  Facet f takes a log-uniform 1e6 ... 5e7 rows (capped at the
    track's 23.4e6, then at max_rows) of the 128-antenna 10 km, 2880 x 10 s
    track of large_wgridding_inputs, as a random contiguous range, at 16
    channels over 120-180 MHz; 1/share of those rows (the row shard of a
    `share`-GPU node).  Returns (uvw, freq, facets (dicts of constructor
    arguments), row_off, n_rows, src [500, 5])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rng_arr = np.random.Generator(np.random.PCG64(1002))
    enu = antennas_logradial(rng_arr, 128, 10000.0)
    uvw = earth_rotation_uvw(enu, 2880, 10.0, np.deg2rad(-30.0))
    freq = np.linspace(120e6, 180e6, 16)
    dx = 2.0 * ARCSEC
    R = uvw.shape[0]
    facets, row_off, n_rows = [], [], []
    for i in range(36):
        gx, gy = i % 6, i // 6
        nx = int(rng.integers(1024, 3073)) // 16 * 16
        ny = int(rng.integers(1024, 3073)) // 16 * 16
        rows = int(np.exp(rng.uniform(np.log(1e6), np.log(5e7))))
        rows = min(rows, R) // share
        if max_rows is not None:
            rows = min(rows, int(max_rows))
        row_off.append(int(rng.integers(0, R - rows + 1)))
        n_rows.append(rows)
        facets.append(dict(width=nx, height=ny, pixel_size_x=dx, pixel_size_y=dx,
                           l_shift=(6000 - (gx + 0.5) * 2000) * dx,
                           m_shift=((gy + 0.5) * 2000 - 6000) * dx, epsilon=1e-5))
    l, m, flux, _, _ = point_sky(rng, 500, 12000, dx)
    src = np.zeros((500, 5))
    src[:, 0], src[:, 1], src[:, 3] = l, m, flux
    src[:, 2] = np.sqrt(1.0 - l ** 2 - m ** 2) - 1.0
    return uvw, freq, facets, row_off, n_rows, src

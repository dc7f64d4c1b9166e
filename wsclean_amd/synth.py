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

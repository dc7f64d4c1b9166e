"""Generate the exponential-of-semicircle (ES) kernel parameter table.

For every even kernel support W (cells) and oversampling factor sigma this
script finds the shape parameter beta of

    phi(x) = exp(beta * (sqrt(1 - x^2) - 1)),   x = 2 s / W,  |s| <= W/2

that minimises the one-dimensional aliasing error of gridding on a grid
oversampled by sigma, and records that error.  The method is the one
published for the ES kernel (Barnett, Magland & af Klinteberg 2019,
arXiv:1808.06736; Arras et al. 2021, arXiv:2010.10122): with

    Phi(t) = int phi(2 s / W) exp(2 pi i s t) ds      (t in cycles per cell)

the relative RMS error of a gridded visibility seen at image position t
(|t| <= 1 / (2 sigma)) is

    err(t) = sqrt(sum_{k != 0} Phi(t + k)^2) / Phi(t).

The table stores max_t err(t).  Phi is integrated after the substitution
x = sin(theta), which makes the integrand analytic, with Gauss-Legendre
quadrature in theta.

Run:  python wsclean_amd/csrc/gen_es_table.py   (rewrites es_table.inc)
"""
import os

import numpy as np

SUPPORTS = list(range(4, 17, 2))
SIGMAS = [round(1.20 + 0.05 * i, 2) for i in range(17)]  # 1.20 .. 2.00
BETA_COARSE = np.linspace(1.2, 2.6, 29)                  # beta / W
NQUAD = 1536
KMAX = 32
NT = 48


def es_phi_hat(W, betas, tau, nquad=NQUAD):
    """Phi(tau) for every beta: array [len(betas), len(tau)]."""
    th, wq = np.polynomial.legendre.leggauss(nquad)
    th = th * (np.pi / 2)
    wq = wq * (np.pi / 2)
    x = np.sin(th)
    c = np.cos(th)
    # phi(sin th) * cos th * weight ; shape [nbeta, nquad]
    f = np.exp(np.outer(betas, c - 1.0)) * (c * wq)[None, :]
    # cos(pi W x tau) ; shape [nquad, ntau]
    cmat = np.cos(np.pi * W * np.outer(x, tau))
    return 0.5 * W * (f @ cmat)


def table():
    rows = []
    ks = np.array([k for k in range(-KMAX, KMAX + 1) if k != 0])
    for W in SUPPORTS:
        for sigma in SIGMAS:
            # the image edge t = 1/(2 sigma) is always a sample point
            t = np.linspace(0.0, 0.5 / sigma, NT)
            tau_alias = (t[None, :] + ks[:, None]).ravel()

            def errors(bf):
                betas = bf * W
                ph0 = es_phi_hat(W, betas, t)
                pha = es_phi_hat(W, betas, tau_alias).reshape(len(betas), len(ks), NT)
                return (np.sqrt((pha ** 2).sum(axis=1)) / np.abs(ph0)).max(axis=1)

            e0 = errors(BETA_COARSE)
            b0 = BETA_COARSE[int(np.argmin(e0))]
            fine = np.linspace(b0 - 0.05, b0 + 0.05, 21)
            e1 = errors(fine)
            ib = int(np.argmin(e1))
            beta = float(fine[ib] * W)
            # Amplification of grid-domain rounding noise by the correction
            # 1/Phi(t): RMS over |t| <= 1/(2 sigma) and its maximum (edge).
            ph = es_phi_hat(W, np.array([beta]), t)[0]
            amp = ph[0] / ph
            rows.append((W, sigma, beta, float(e1[ib]), float(np.sqrt(np.mean(amp ** 2))),
                         float(amp.max())))
    return rows


def main():
    rows = table()
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "es_table.inc")
    with open(out, "w") as f:
        f.write("// Generated by gen_es_table.py: {W, sigma, beta, 1-D RMS "
                "aliasing error, RMS and max correction amplification}.\n")
        for W, s, b, e, ar, am in rows:
            f.write("{%d, %.2f, %.6f, %.6e, %.6e, %.6e},\n" % (W, s, b, e, ar, am))
    print("wrote", out, len(rows), "rows")
    for W, s, b, e, ar, am in rows:
        if s in (1.2, 1.25, 1.5, 1.75, 2.0):
            print(W, s, round(b / W, 4), "%.2e" % e, "amp rms %.1f max %.1f" % (ar, am))


if __name__ == "__main__":
    main()

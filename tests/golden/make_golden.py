"""Write tests/golden/dft_small.npz.

The reference's own tests hold no golden vectors for the w-gridder
(ducc0 is absent and every system test needs downloaded Measurement Sets),
so the fixture is produced by the literal numpy restatement of
DirectMSGridder (oracle.dft_invert_reference_wrap, which follows
gridding/directmsgridder.cpp:56-132 statement by statement) and by a plain
numpy adjoint sum for the predict direction.  The test then checks the C
oracle against it.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402

C = 299792458.0


def main():
    rng = np.random.default_rng(1234)
    nx, ny = 12, 10
    dx, dy = 0.011, 0.013
    ls, ms = 0.004, -0.003
    uvw = rng.normal(0, 250.0, (9, 3))
    freq = np.array([120e6, 135e6, 150e6])
    vis = (rng.normal(size=(9, 3)) + 1j * rng.normal(size=(9, 3))).astype(np.complex64)
    vis[2, 1] = 0
    image = oracle.dft_invert_reference_wrap(nx, ny, dx, dy, ls, ms, uvw, freq, vis)
    model = np.zeros((ny, nx))
    model[2, 3] = 1.5
    model[7, 9] = -0.25
    model[5, 6] = 0.75
    pred = np.zeros((9, 3), np.complex128)
    for y in range(ny):
        for x in range(nx):
            if model[y, x] == 0:
                continue
            l = (nx // 2 - x) * dx + ls
            m = (y - ny // 2) * dy + ms
            nm1 = np.sqrt(1 - l * l - m * m) - 1
            for r in range(9):
                for c in range(3):
                    s = freq[c] / C
                    ph = 2 * np.pi * s * (uvw[r, 0] * l + uvw[r, 1] * m + uvw[r, 2] * nm1)
                    pred[r, c] += model[y, x] * np.exp(1j * ph)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dft_small.npz")
    np.savez(out, nx=nx, ny=ny, dx=dx, dy=dy, l_shift=ls, m_shift=ms, uvw=uvw, freq=freq,
             vis=vis, image=image, model=model, pred=pred)
    print("wrote", out)


if __name__ == "__main__":
    main()

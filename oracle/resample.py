"""Oracle for the image resampler: TEST INFRASTRUCTURE ONLY (used by tests/
as the checker; nothing in the product imports it).

WSClean resamples the small-inversion image to the requested image size
after inversion and the model back before prediction
(wgridder/wgriddingmsgridder.cpp:290-308, :326-336) with
schaapcommon::math::Resampler.  schaapcommon is absent from the reference
mount (external/schaapcommon is empty), so this restates the published
method -- Fourier interpolation: 2-D DFT of the input, spectrum cropped or
zero-padded to the output size, inverse DFT, scaled by 1 / (input pixels)
so that pixel values are samples of the same band-limited function
(output pixel m sits at input coordinate m * N_in / N_out in each axis).
At an even size's Nyquist bin the component is split evenly between the
+N/2 and -N/2 bins (upsampling), and the output's Nyquist bin collects
both input bins +-M/2 (downsampling): real inputs stay real, a cosine at
the Nyquist frequency keeps its amplitude, trigonometric polynomials below
both Nyquist limits are reproduced exactly, and upsampling then
downsampling returns the input.  The Nyquist treatment and the windowing
options of schaapcommon are parity unpinned.
"""
import numpy as np


def _axis_map(n_in, n_out):
    """For each output frequency bin: list of (input bin, weight)."""
    out = []
    for ko in range(n_out):
        f = ko if ko < (n_out + 1) // 2 else ko - n_out  # signed frequency
        terms = []
        lim = min(n_in, n_out)
        if n_out > n_in and n_in % 2 == 0 and abs(f) == n_in // 2:
            # the input's Nyquist bin, split over +-n_in/2
            terms.append((n_in // 2, 0.5))
        elif n_out < n_in and n_out % 2 == 0 and f == -(n_out // 2):
            # the output's Nyquist bin: both of the input's +-n_out/2
            terms.append((n_out // 2, 1.0))
            terms.append((n_in - n_out // 2, 1.0))
        elif (f >= 0 and f < (lim + 1) // 2) or (f < 0 and -f <= (lim - 1) // 2) or \
                (n_out == n_in):
            terms.append((f % n_in, 1.0))
        out.append(terms)
    return out


def _axis_matrix(n_in, n_out):
    m = np.zeros((n_out, n_in))
    for ko, terms in enumerate(_axis_map(n_in, n_out)):
        for i, w in terms:
            m[ko, i] += w
    return m


def resample(img, out_w, out_h):
    """img: [in_h][in_w] real; returns [out_h][out_w] float64."""
    img = np.asarray(img, np.float64)
    in_h, in_w = img.shape
    X = np.fft.fft2(img)
    Y = _axis_matrix(in_h, out_h) @ X @ _axis_matrix(in_w, out_w).T
    return np.real(np.fft.ifft2(Y)) * (out_w * out_h) / (in_w * in_h)

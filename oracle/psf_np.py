"""NumPy restatement of the 2-D PSF arithmetic (test infrastructure only).

resample_psf and pupil_to_intensity follow maomi/psf.py:154-203 step for step, with the same
library routines (scipy.ndimage.map_coordinates for the linear resampling, numpy's FFT).
Pinned in tests/test_psf_contracts.py by closed forms: resampling at the source pitch and shape
is the identity, a linear ramp is reproduced inside the support, the mass of a finer grid
scales with the pitch ratio; a zero-OPD pupil peaks at fft_size // 2 with the textbook
sinc^2 (Dirichlet) profile of a square aperture, an OPD tilt of m/n waves per sample moves the
peak by m pixels, and the result has unit sum.
"""

import numpy as np


def resample_coordinates(shape, pitch, src_pitch, src_shape):
    """Source-pixel coordinates of each output axis: (i - n // 2) step / old + size // 2."""
    return [(np.arange(n) - n // 2) * step / old + size // 2
            for n, step, old, size in zip(shape, pitch, src_pitch, src_shape)]


def resample_linear(intensity, pitch, src_pitch, shape):
    """Linear sampling with zero outside the support (psf.py:166-168) and the pre-normalisation
    mass estimate (psf.py:169); returns (unit-sum kernel, mass)."""
    from scipy.ndimage import map_coordinates
    coords = resample_coordinates(shape, pitch, src_pitch, np.shape(intensity))
    a = map_coordinates(np.asarray(intensity, float), np.array(np.meshgrid(*coords, indexing="ij")),
                        order=1, mode="constant", cval=0.0, prefilter=False)
    mass = float(a.sum() * np.prod(pitch) / np.prod(src_pitch))
    return a / a.sum(), mass


def pupil_to_intensity(opd_waves, fft_size=256, amplitude=None):
    """psf.py:179-203: pupil amp exp(2 pi i OPD) on its valid samples, embedded at
    n // 2 - s // 2, |fftshift(fft2(ifftshift(.)))|^2 normalised to unit sum."""
    opd = np.asarray(opd_waves, dtype=float)
    amp = np.ones_like(opd) if amplitude is None else np.asarray(amplitude, dtype=float)
    valid = np.isfinite(opd) & (amp > 0)
    pupil = np.zeros(opd.shape, dtype=np.complex128)
    pupil[valid] = amp[valid] * np.exp(2j * np.pi * opd[valid])
    n = int(fft_size)
    padded = np.zeros((n, n), dtype=np.complex128)
    s0, s1 = n // 2 - opd.shape[0] // 2, n // 2 - opd.shape[1] // 2
    padded[s0:s0 + opd.shape[0], s1:s1 + opd.shape[1]] = pupil
    inten = np.abs(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(padded)))) ** 2
    return inten / inten.sum()

"""Trigonometric interpolation of periodic grid fields from N to 2N per axis
(same box, fine node 2j on coarse node j), by zero-padding the spectrum:
the checker for the solver's coarse start (which has no counterpart in the
reference).  The Nyquist plane of each axis is split in halves between
+N/2 and -N/2, so a real field stays real and a field band-limited below
the coarse Nyquist frequency is reproduced exactly."""
import numpy as np


def upsample(t):
    """t: [..., N2, N1, N0] real, every N even -> [..., 2 N2, 2 N1, 2 N0]."""
    axes = (-3, -2, -1)
    F = np.fft.fftn(t, axes=axes)
    for ax in axes:
        n = F.shape[ax]
        h = n // 2
        lo = np.take(F, np.arange(0, h), axis=ax)
        nyq = 0.5 * np.take(F, [h], axis=ax)
        hi = np.take(F, np.arange(h + 1, n), axis=ax)
        zshape = list(F.shape)
        zshape[ax] = n - 1
        F = np.concatenate([lo, nyq, np.zeros(zshape, dtype=F.dtype), nyq, hi], axis=ax)
    # ifftn divides by the fine point count; the interpolant needs the coarse one
    return np.real(np.fft.ifftn(F, axes=axes)) * 8.0

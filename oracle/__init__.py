"""CPU oracle for the KRFFT / pseudo-spectral diffusion hot path.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy (FP64) restatement of
the reference algorithms (liuyxpp/CrystallographicFFT.jl), used as the checker
for the HIP library: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path never
routes through it.

Every function cites the reference file:line it restates.  How each function is
pinned (golden fixtures, closed forms, full-grid FFT identities, brute force) is
listed in the module docstrings and exercised by ``tests/test_oracle_*.py``.
"""
from .krfft import *  # noqa: F401,F403
from .scft import *   # noqa: F401,F403
from .asu import calc_asu, pack_asu, block_fft, block_ifft  # noqa: F401

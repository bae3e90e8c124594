"""The C CPU baseline (oracle/cstep.c) against the NumPy oracle's Strang step
(CPU only): FP64, 1e-12 relative, one and several threads."""
import numpy as np
import pytest

import oracle as O
from oracle import cstep
from conftest import group_setup


@pytest.fixture(scope="module", autouse=True)
def _built():
    cstep.build()


@pytest.mark.parametrize("sg,N", [(229, (16, 16, 16)), (47, (16, 16, 16)), (230, (24, 24, 24)),
                                  (225, (16, 16, 16)), (70, (16, 16, 16)), (221, (24, 24, 24))])
@pytest.mark.parametrize("threads", [1, 4])
def test_cstep_vs_oracle(sg, N, threads):
    _, _, _, R, t = group_setup(sg, N)
    if not np.all(O.auto_L(t) == 2):
        pytest.skip("the C baseline covers L = (2,2,2)")
    lat = 4.0 * np.eye(3)
    ds = 0.01
    Lv = O.auto_L(t)
    q = O.extract_subgrid(O.make_symmetric(R, t, N, seed=1), Lv) + 1.0
    w = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=2), Lv)
    cs = cstep.CStep(R, t, N, ds, lat, nthreads=threads)
    E = np.exp(-w * ds / 2.0)
    got = cs.step(q, E)
    spec = O.calc_spectral_asu(R, t, N)
    ref = O.strang_step(q, w, ds, lambda f: O.m2_scft(f, spec, R, t, N, ds, lat))
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))
    # pure diffusion = full-grid FFT . K . IFFT
    u = O.make_symmetric(R, t, N, seed=3)
    d = cs.step(O.extract_subgrid(u, Lv))
    full = O.fullgrid_reference(u, N, ds, lat, Lv)
    assert np.max(np.abs(d - full)) < 1e-12 * np.max(np.abs(full))

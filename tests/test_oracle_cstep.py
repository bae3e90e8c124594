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


@pytest.mark.parametrize("sg,N", [(229, (16, 16, 16)), (230, (16, 16, 16))])
def test_first_iteration_vs_numpy_sweep(sg, N):
    """cstep.first_iteration (C steps, q+ advanced alongside) equals the NumPy
    oracle's sweep / free_energy (all propagators stored) to 1e-12; a density
    pairing off by one contour step fails this by orders of magnitude."""
    _, _, _, R, t = group_setup(sg, N)
    lat = 4.0 * np.eye(3)
    Lv = O.auto_L(t)
    f, chiN, ns = 0.3, 20.0, 20
    wA = 2.0 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=4), Lv)
    wB = -1.5 * O.extract_subgrid(O.make_symmetric(R, t, N, seed=5), Lv)
    cs = cstep.CStep(R, t, N, 1.0 / ns, lat, nthreads=2)
    got = cstep.first_iteration(cs, wA, wB, f, chiN, ns)
    spec = O.calc_spectral_asu(R, t, N)
    table = O.m2_backward_table(spec, R, t, N)
    ref = O.sweep(wA, wB, f, ns, lambda x: O.m2_scft(x, spec, R, t, N, 1.0 / ns, lat, table))
    Fref = O.free_energy(wA, wB, ref["phiA"], ref["phiB"], ref["Q"], chiN)
    assert abs(got["Q"] - ref["Q"]) < 1e-12 * ref["Q"]
    assert abs(got["F"] - Fref) < 1e-12 * max(1.0, abs(Fref))
    for a, b in ((got["phiA"], ref["phiA"]), (got["phiB"], ref["phiB"]), (got["q"], ref["q"][ns]),
                 (got["qd"], ref["qd"][ns])):
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b))

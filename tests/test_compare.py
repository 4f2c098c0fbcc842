"""Host-side error measures (lbmpond_amd.compare)."""
import numpy as np

from lbmpond_amd.compare import CapDependence, dilate, state_errors


def brute_dilate(mask, shape, reach):
    nx, ny, nz = shape
    m = mask.reshape(nz, ny, nx)
    out = np.zeros_like(m)
    for z, y, x in zip(*np.nonzero(m)):
        for dz in range(-reach, reach + 1):
            for dy in range(-reach, reach + 1):
                for dx in range(-reach, reach + 1):
                    out[(z + dz) % nz, (y + dy) % ny, (x + dx) % nx] = True
    return out.reshape(-1)


def test_dilate_matches_brute_force_periodic():
    rng = np.random.default_rng(3)
    for shape, reach in (((9, 7, 5), 1), ((12, 10, 8), 2), ((16, 1, 1), 3), ((8, 6, 1), 2)):
        n = int(np.prod(shape))
        mask = np.zeros(n, bool)
        mask[rng.integers(0, n, 3)] = True
        assert np.array_equal(dilate(mask, shape, reach), brute_dilate(mask, shape, reach)), (shape, reach)


def test_cap_dependence_grows_by_reach_per_step():
    shape = (20, 20, 20)
    dep = CapDependence(shape, max_iter=100, reach=3)
    it = np.zeros(8000, np.int32)
    dep.update(it)
    assert not dep.mask.any() and dep.capped == [0]
    it[0] = 100
    dep.update(it)
    assert dep.mask.sum() == 1
    dep.update(np.zeros(8000, np.int32))
    assert dep.mask.sum() == 7 ** 3
    dep.update(np.zeros(8000, np.int32))
    assert dep.mask.sum() == 13 ** 3
    assert dep.capped == [0, 1, 0, 0]


def test_state_errors_velocity_as_vector_and_mask():
    n = 10
    ref = {"rho": np.full(n, 2.0), "ux": np.full(n, 0.5), "uy": np.zeros(n), "uz": np.zeros(n), "T": np.full(n, 0.1)}
    got = {k: v.copy() for k, v in ref.items()}
    got["uy"][3] = 1e-3  # measured against the flow speed 0.5, not against uy's own zero
    got["rho"][7] = 2.2
    e = state_errors(got, ref)
    assert np.isclose(e["u"], 2e-3) and np.isclose(e["rho"], 0.1) and e["T"] == 0.0
    mask = np.ones(n, bool)
    mask[7] = False
    e = state_errors(got, ref, mask)
    assert e["rho"] == 0.0 and np.isclose(e["u"], 2e-3)

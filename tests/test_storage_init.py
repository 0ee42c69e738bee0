"""Initial storage volume of subgrid cells (compute_initial_subgrid_volumes,
sfincs_subgrid.F90:952-956): a cell that starts wet has no storage for the water it
already holds. Pinned by its special cases: a dry cell (initial volume 0) keeps all its
storage, a cell holding more water than its storage keeps none, and in between the
storage left plus the water held equals the storage given."""
import numpy as np

from sfincs_amd import workloads as W


def _case(zsini):
    sc = W.subgrid_case(24, 30, dx=25.0, seed=107, t_end=600.0, nsrc=0, zsini=zsini)
    given = {}

    def storage(m):
        r = np.random.default_rng(108)
        v = r.uniform(0.0, 400.0, m.sizes()["np"]).astype(np.float32)
        given["v"] = v
        return v
    sc.storage = storage
    m = sc.build_model()
    return m, given["v"], m.table("storage_volume"), m.table("z_volume")


def test_dry_cells_keep_storage_and_wet_cells_lose_the_water_they_hold():
    m, given, left, vol0 = _case(0.0)
    dry = vol0 == 0.0
    assert dry.any() and (~dry).any()
    assert np.array_equal(left[dry], given[dry])
    full = vol0 >= given.astype(np.float64)
    assert full.any()
    assert np.all(left[full] == 0.0)
    part = ~dry & ~full
    assert part.any()
    # storage left + water held = storage given, to f32 rounding of the storage
    np.testing.assert_allclose(left[part].astype(np.float64) + vol0[part], given[part].astype(np.float64),
                               rtol=0, atol=np.spacing(given[part]).max())


def test_higher_initial_level_leaves_less_storage():
    _, given, low, _ = _case(-2.0)
    _, _, high, _ = _case(1.0)
    assert np.all(high <= low) and np.all(low <= given)
    assert np.sum(high) < np.sum(low)

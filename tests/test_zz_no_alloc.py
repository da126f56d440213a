"""fr_grid_execute_noalloc through the C ABI: the gridding pass with its
heavy-unit slice partials in the library's static per-GPU scratch instead of
plan-allocated memory.  It must give bit-identical grids to fr_grid_execute
(same kernels, same summation order) on consecutive calls whose second one
is smaller, and refuse a plan whose slices need more scratch than the array
holds."""
import numpy as np
import pytest

from fastrecon_amd import hip
from fastrecon_amd import synthetic as syn

pytestmark = pytest.mark.gpu
SCRATCH = 32 << 20


@pytest.fixture(scope="module")
def ctx():
    c = hip.Context(0)
    yield c
    c.close()


def make(ctx, spokes, readout, coils, image, seed, tune):
    N = syn.grid_size(image, 2.0)
    traj, dcf = syn.kooshball3d(spokes, readout)
    tn, _ = hip.normalize_traj(traj, N)
    M = tn.shape[0]
    s = hip.Samples(ctx, 3, M, coils)
    s.set_traj(tn)
    s.set_dcf(dcf)
    s.set_coils(syn.coil_data(coils, M, seed))
    lut, _ = hip.kb_lut(4.0, 2.0, 512)
    p = hip.Plan(ctx, s, N, 4.0, lut, tune)
    return s, p, N


def plan_within_scratch(ctx, spokes, readout, coils, image, seed, tune):
    """The smallest slice size (coarser slices need fewer partials) whose plan
    fits the scratch while still cutting heavy units into slices."""
    for split in (2048, 8192, 32768, 131072):
        s, p, N = make(ctx, spokes, readout, coils, image, seed, dict(tune, split_taps=split))
        info = p.info()
        need = info["scratch_bytes"] * ((coils + 31) // 32)
        if info["heavy_units"] > 0 and 0 < need <= SCRATCH:
            return s, p, N, info, need
        p.close()
        s.close()
    pytest.fail("no slice size gives a plan with slices within the scratch")


@pytest.mark.parametrize("tune", [{}, dict(accumulate=5), dict(tile_writes=1)])
def test_noalloc_equals_execute_on_two_calls(ctx, tune):
    # second call smaller (fewer samples, coils and grid): the scratch keeps
    # the first call's partials, which the second must not read
    for spokes, readout, coils, image, seed in ((1500, 96, 40, 48, 3), (600, 64, 12, 32, 4)):
        s, p, N, info, need = plan_within_scratch(ctx, spokes, readout, coils, image, seed, tune)
        g1 = hip.Grids(ctx, 3, N, coils)
        g2 = hip.Grids(ctx, 3, N, coils)
        hip.grid_execute(p, s, g1)
        hip.grid_execute_noalloc(p, s, g2)
        a, b = g1.download_all(), g2.download_all()
        print(f"[noalloc {tune}] {spokes}x{readout} C={coils} heavy={info['heavy_units']} "
              f"scratch {need} B, bitwise {np.array_equal(a.view(np.int32), b.view(np.int32))}", flush=True)
        np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))
        for h in (g1, g2, p, s):
            h.close()


def test_noalloc_refuses_plans_beyond_the_scratch(ctx):
    s, p, N = make(ctx, 6000, 128, 64, 64, 5, dict(split_taps=16))
    need = p.info()["scratch_bytes"] * 2
    g = hip.Grids(ctx, 3, N, 64)
    if need <= SCRATCH:
        pytest.skip(f"this plan needs only {need} bytes of scratch")
    with pytest.raises(hip.HipError, match="scratch"):
        hip.grid_execute_noalloc(p, s, g)
    for h in (g, p, s):
        h.close()

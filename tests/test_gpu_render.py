"""GPU checks of the render epilogue (iterate_and_render + download_render, mod.rs:302-335,
390-432): the frame is bit-exact against the oracle's colour pack of the state the step wrote,
and rendering does not change the step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _wide_state(cr, seed):
    """Values below 0 and above 2/3 too, so both clamps and every byte value occur."""
    rng = np.random.default_rng(seed)
    st = cr["state"].astype(np.float32)
    return (st + rng.uniform(-0.3, 0.6, st.shape)).astype(np.float32)


@pytest.mark.parametrize("n,C", [(64, 3), (256, 3), (1024, 3), (128, 5), (64, 1), (64, 2)])
def test_render_frame_bit_exact(cn, orc, n, C):
    cr = cn.baseline_creature(n, num_channels=C, radius=max(4, n // 8), seed=n + C)
    cr["state"] = _wide_state(cr, n)
    a, b = cn.make_engine(cr), cn.make_engine(cr)
    assert not a.download_render().any()              # zero frame before the first render
    for _ in range(2):
        a.iterate_and_render()
        b.iterate()
    st = a.download_all_channels().reshape(C, n, n)
    assert np.array_equal(st, b.download_all_channels().reshape(C, n, n))   # render changes no state
    rgb = a.download_render().reshape(n, n, 3)
    want = orc.render_rgb(st)
    assert np.array_equal(rgb, want), int((rgb != want).sum())
    assert len(np.unique(rgb[..., 0])) > 10           # a frame with many shades, not a constant
    a.iterate()                                        # a plain step keeps the last frame
    assert np.array_equal(a.download_render().reshape(n, n, 3), rgb)
    a.close()
    b.close()


def test_render_batch_of_mixed_sizes(cn, orc):
    crs = [cn.baseline_creature(n, num_channels=C, radius=6, seed=i) for i, (n, C) in
           enumerate([(64, 3), (128, 2), (64, 4)])]
    for i, cr in enumerate(crs):
        cr["state"] = _wide_state(cr, i)
    batch = cn.CreatureBatch(crs)
    for i, cr in enumerate(crs):
        batch.load(i, cr)
    batch.iterate_and_render()
    batch.sync()
    for i, cr in enumerate(crs):
        C, n = cr["num_channels"], cr["state"].shape[1]
        st = batch.creature_download(i).reshape(C, n, n)
        assert np.array_equal(batch.creature_download_render(i), orc.render_rgb(st))
    batch.close()

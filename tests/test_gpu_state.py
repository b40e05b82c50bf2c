"""Device state copies (atus_snapshot_*, atus_copy_components): the chirp
scan's restarts and the hand-over of one run's state to a batch must move
the exact bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from atus2_amd import hip as H
    H.load()
    return H


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_snapshot_restores_the_exact_state(hip):
    shape = (32, 64)
    with hip.Context(shape, [0.5, 0.5], [2 * np.pi / 16, 2 * np.pi / 32], 2) as ctx:
        a = [_rand(shape, s) for s in (1, 2)]
        for c in range(2):
            ctx.upload(c, a[c])
        ctx.snapshot_save(2)
        ctx.set_alpha([0.5, 0.5])
        ctx.set_g([1.0, 0.5, 0.5, 1.0])
        ctx.run_block(hip.STEP_BRAGG_AD, 1e-3, 5, 0.0, hip.laser(amp0=3.0, laser_k=1.0))
        assert not np.array_equal(ctx.download(0), a[0])
        ctx.snapshot_restore(2)
        for c in range(2):
            assert np.array_equal(ctx.download(c), a[c])
        # a second restore from the same slot gives the same bits again
        ctx.upload(0, np.zeros(shape, np.complex128))
        ctx.snapshot_restore(2)
        assert np.array_equal(ctx.download(0), a[0])
        with pytest.raises(hip.AtusError, match="holds no snapshot"):
            ctx.snapshot_restore(1)
        with pytest.raises(hip.AtusError, match="0..3"):
            ctx.snapshot_save(4)


def test_copy_components_into_batch_runs(hip):
    shape = (256,)
    dx, dk = [0.25], [2 * np.pi / 64]
    with hip.Context(shape, dx, dk, 2) as S, hip.BatchContext(shape, dx, dk, 3) as B:
        a = [_rand(shape, s) for s in (5, 6)]
        for c in range(2):
            S.upload(c, a[c])
        for r in range(3):
            B.copy_components(2 * r, S, 0, 2)
        for r in range(3):
            for c in range(2):
                assert np.array_equal(B.download(2 * r + c), a[c])
        # and back: run 2's second component into component 0
        B.upload(5, a[0] * 2)
        S.copy_components(0, B, 5, 1)
        assert np.array_equal(S.download(0), a[0] * 2)
        with pytest.raises(hip.AtusError, match="out of bounds"):
            S.copy_components(1, B, 0, 2)
    with hip.Context((128,), dx, dk, 1) as T, hip.Context(shape, dx, dk, 1) as U:
        with pytest.raises(hip.AtusError, match="grid sizes"):
            T.copy_components(0, U, 0, 1)

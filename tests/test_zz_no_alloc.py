"""atus_batch_particle_numbers_noalloc, the allocation-free variant for
callers that cannot allocate device memory, called through the C-ABI and
compared with atus_batch_particle_numbers: the same kernels on the same
data, so the results must agree bit for bit.  Two calls in a row, the
second on a smaller batch, check that the shared device scratch carries
nothing over from the previous call."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from atus2_amd import hip as H
    H.load()
    return H


def _batch(hip, shape, nbatch, seed):
    dim = len(shape)
    L = 32.0
    dx = [L / n for n in shape]
    dk = [2 * np.pi / L] * dim
    B = hip.BatchContext(shape, dx, dk, nbatch)
    rng = np.random.default_rng(seed)
    for c in range(2 * nbatch):
        B.upload(c, rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    return B


def _noalloc(hip, B):
    out = np.empty((B.nbatch, 2))
    rc = hip._lib.atus_batch_particle_numbers_noalloc(B._h, out.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise hip.AtusError(hip._lib.atus_last_error().decode())
    return out


@pytest.mark.parametrize("first,second", [(((64, 128), 6), ((32,), 3)),
                                          (((16, 16, 32), 40), ((1024,), 5)),
                                          (((4096,), 700), ((2048,), 2))])
def test_noalloc_particle_numbers_equal_the_allocating_entry_point(hip, first, second):
    for k, (shape, nb) in enumerate((first, second)):
        print(f"noalloc particle numbers: shape {shape}, {nb} runs", flush=True)
        with _batch(hip, shape, nb, 7 + k) as B:
            ref = B.particle_numbers()
            got = _noalloc(hip, B)
            assert got.shape == ref.shape
            assert np.array_equal(got, ref), (shape, nb, np.max(np.abs(got - ref)))
            # and again: the scratch now holds this call's partials
            assert np.array_equal(_noalloc(hip, B), ref)


def test_noalloc_particle_numbers_refuse_batches_beyond_the_scratch(hip):
    # 8193 runs = 16386 components: 2 * 16386 doubles of partials and
    # results, past the 32768 of the device scratch
    with _batch(hip, (2,), 8193, 1) as B:
        with pytest.raises(hip.AtusError, match="device scratch"):
            _noalloc(hip, B)
        # the allocating entry point still serves it
        assert np.all(np.isfinite(B.particle_numbers()))

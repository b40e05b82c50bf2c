"""The Sobolev loop run on the device in chunks (atus_sob_iterate): the
same iterations as one body per call, bit for bit, and the same stop as the
reference's do-while (CSOB_Base.h:427-505: the first body whose residual is
<= epsilon is the last one that updates psi); the oracle's loop stops at
the same iteration."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from atus2_amd import hip as H
    H.load()
    return H


def _setup(hip, orc, shape, nw):
    dim = len(shape)
    g = orc.make_grid(shape, L=[12.0, 12.0, 12.0][:dim])
    xs = orc.coords_x(g)
    Vs = [0.5 * sum((1.0 + 0.25 * a + 0.1 * i) ** 2 * x ** 2 for a, x in enumerate(xs)) * np.ones(shape)
          for i in range(nw)]
    guesses = [(np.exp(-sum(x ** 2 for x in xs) / (4.0 + i)) * np.ones(shape)).astype(np.complex128)
               for i in range(nw)]
    gs = [20.0] if nw == 1 else [20.0, 5.0, 5.0, 15.0]
    ctxs = []
    for _ in range(2):
        ctx = hip.Context.from_grid(g, nw)
        for i in range(nw):
            ctx.upload(i, guesses[i])
            ctx.set_potential(i, Vs[i])
        ctx.set_g(gs)
        ctx.sob_setup([0.5] * nw, 0.02)
        ctx.sob_renormalize()
        ctxs.append(ctx)
    S = orc.SobSolver(g, guesses, Vs, np.reshape(gs, (nw, nw)), [0.5] * nw, stepsize=0.02)
    S.renormalize()
    return g, ctxs, S


@pytest.mark.parametrize("shape,nw", [((256,), 1), ((32, 32), 1), ((16, 32, 32), 1), ((16, 16, 32), 2)])
def test_device_loop_equals_single_bodies_and_stops_where_the_reference_does(hip, orc, shape, nw):
    g, (A, B), S = _setup(hip, orc, shape, nw)
    try:
        res, rows, states = [], [], []
        for _ in range(40):
            r, rn, Ns = A.sob_iter()
            res.append(r)
            rows.append([Ns[0], Ns[1] if nw > 1 else 0.0, float(rn), r])
            states.append([A.download(i) for i in range(nw)])
        # an epsilon between two residuals of the sequence, so the stop
        # falls inside a chunk
        k = next(k for k in range(20, 40) if res[k] < res[k - 1])
        eps = float(np.sqrt(res[k - 1] * res[k]))
        kstop = next(j for j in range(40) if res[j] <= eps)
        done, r, log = B.sob_iterate(40, eps)
        assert done == kstop + 1 and r == res[kstop]
        assert np.array_equal(log, np.array(rows[:done]))
        for i in range(nw):
            assert np.array_equal(B.download(i), states[kstop][i]), i
        # the oracle's loop stops at the same body
        j = 0
        while True:
            rr = S.iterate()
            j += 1
            if rr <= eps:
                break
            assert j < 40
        assert j == done
        # a new call runs at least one body, as a new do-while does
        done2, r2, _ = B.sob_iterate(3, eps)
        assert done2 >= 1
    finally:
        A.close()
        B.close()


def test_sob_run_in_chunks_matches_single_bodies(hip, orc):
    g, (A, B), S = _setup(hip, orc, (32, 32, 32), 1)
    try:
        its, r = B.sob_run(1e-9, 100000)
        n = 0
        while True:
            rr, _, _ = A.sob_iter()
            n += 1
            if rr <= 1e-9:
                break
        assert n == its and rr == r
        assert np.array_equal(A.download(0), B.download(0))
    finally:
        A.close()
        B.close()

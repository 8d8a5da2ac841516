"""GPU parity at the benchmark configurations' full sizes and on the inputs
of the reference's own test-case generator, through the C ABI, against the
CPU oracle.  Integer work: bit-exact.

  * c4 (256 images, sides 256..4096): the batched call with 1 and 2 sweep
    rounds and with the UQ queue, hipGraph replays, the autotuned 'auto'
    call and the learned schedule's batch form.
  * c5 (32768^2 u8): every queue variant with the plain schedule, and the
    learned schedule with the shipped network.
  * The reference's InitRandomTestCase distribution
    (IWPP MP/src/testCaseGenerator.cu:28-35: marker uniform, mask = marker
    plus a residue) through every queue variant.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import iwpp
import oracle
from iwpp import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def up(a, pitch=None):
    return iwpp.DeviceImage.from_numpy(a, pitch=pitch)


def oracle_all(cases, conn):
    # ctypes releases the GIL during the oracle's C call: one image per thread
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda c: oracle.reconstruct(c.marker, c.mask, conn), cases))


@pytest.fixture(scope="module")
def c4_cases():
    cases = W.c4()
    say("c4 generated", len(cases), sum(c.pixels for c in cases))
    wants = oracle_all(cases, 8)
    masks = [up(c.mask) for c in cases]
    return cases, wants, masks


def _check_batch(pairs, wants, cases, what):
    bad = [c.name for (dj, _), want, c in zip(pairs, wants, cases)
           if not np.array_equal(dj.numpy(), want)]
    assert not bad, (what, len(bad), bad[:5])


@pytest.mark.parametrize("queue,sweeps", [("mutex", 1), ("mutex", 2), ("uq", -1), ("dq", 1)])
def test_config4_full_batch(torch_cuda, c4_cases, queue, sweeps):
    cases, wants, masks = c4_cases
    say("c4 batch", queue, sweeps)
    pairs = [(up(c.marker), m) for c, m in zip(cases, masks)]
    st = iwpp.reconstruct_batch(pairs, 8, 0, queue, "plain",
                                iwpp.default_options(sweep_rounds=sweeps))
    say("c4 batch stats", st)
    _check_batch(pairs, wants, cases, (queue, sweeps, st))


def test_config4_full_graph(torch_cuda, c4_cases):
    cases, wants, masks = c4_cases
    pairs = [(up(c.marker), m) for c, m in zip(cases, masks)]
    inputs = [up(c.marker) for c in cases]
    with iwpp.Workspace() as ws:
        for sweeps in (1, 2):
            g = iwpp.BatchGraph(pairs, ws, 8, options=iwpp.default_options(sweep_rounds=sweeps))
            try:
                for rep in range(2):
                    say("c4 graph", sweeps, rep)
                    for (dj, _), src in zip(pairs, inputs):
                        dj.buf.copy_(src.buf)
                    g.launch()
                    st = g.stats()
                    _check_batch(pairs, wants, cases, ("graph", sweeps, rep, st))
            finally:
                g.close()


def test_config4_full_autotuned_and_learned(torch_cuda, c4_cases):
    cases, wants, masks = c4_cases
    pairs = [(up(c.marker), m) for c, m in zip(cases, masks)]
    say("c4 autotune")
    best = iwpp.autotune(pairs, 8, 1)
    say("c4 autotuned", best)
    try:
        st = iwpp.reconstruct_batch(pairs, 8, 0, "auto")
        assert st["sweep_rounds"] == best["sweep_rounds"], (st, best)
        _check_batch(pairs, wants, cases, ("auto", best, st))
    finally:
        iwpp.lib().iwpp_autotune_clear()
    from iwpp import learned
    learned.install()
    try:
        say("c4 learned batch")
        pairs = [(up(c.marker), m) for c, m in zip(cases, masks)]
        st = iwpp.reconstruct_batch(pairs, 8, 0, "mutex", "learned")
        assert st["schedule"] == iwpp.SCHEDULES["learned"], st
        _check_batch(pairs, wants, cases, ("learned", st))
    finally:
        iwpp.set_model(None)


def test_config5_full(torch_cuda):
    case = W.c5()
    want = oracle.reconstruct(case.marker, case.mask, case.connectivity)
    mask = up(case.mask)
    from iwpp import learned
    runs = [("mutex", "plain"), ("dq", "plain"), ("uq", "plain"), ("mutex", "learned"),
            ("uq", "learned")]
    for queue, sched in runs:
        say(case.name, queue, sched)
        if sched == "learned":
            learned.install()
        try:
            dj = up(case.marker)
            st = iwpp.reconstruct(dj, mask, case.connectivity, case.tile, queue, sched)
        finally:
            iwpp.set_model(None)
        say(case.name, queue, sched, st)
        got = dj.numpy()
        assert np.array_equal(got, want), (queue, sched, st, int((got != want).sum()))
        if sched == "learned":
            assert st["schedule"] == iwpp.SCHEDULES["learned"], st
        del dj
    # size-independent properties on the last device result
    assert (got <= case.mask).all() and (got >= np.minimum(case.marker, case.mask)).all()
    again = up(got)
    iwpp.reconstruct(again, mask, case.connectivity, case.tile, "mutex")
    assert np.array_equal(again.numpy(), got)  # idempotent


@pytest.mark.parametrize("queue", ["mutex", "dq", "uq"])
@pytest.mark.parametrize("conn", [4, 8])
def test_reference_random_case(torch_cuda, queue, conn):
    for (h, w), seed in (((4096, 4096), 1312), ((512, 1536), 7), ((1000, 777), 11)):
        J, I = iwpp.synth_random_case(h, w, seed)
        want = oracle.reconstruct(J, I, conn)
        for sweeps in (0, -1):
            say("random case", queue, conn, h, w, sweeps)
            dj, di = up(J), up(I)
            st = iwpp.reconstruct(dj, di, conn, 0, queue, "plain",
                                  iwpp.default_options(sweep_rounds=sweeps))
            got = dj.numpy()
            assert np.array_equal(got, want), (h, w, sweeps, st, int((got != want).sum()))

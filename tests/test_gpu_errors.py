"""Error paths of the drop-in: the return codes of ps_init and
ps_calc_forces (src/polesitter.h:1699-1767, 1840-1909) when an allocation
fails or the device reports a failure, forced from the host with
ps_hip_debug_fail (no kernel faults; every launched kernel ends normally).

Bar: every forced failure comes back from the call that met it with the
reference's code for it: PS_EALLOC from ps_init (its "allocation failed",
H:1758-1761; PS_EOOM there means the caller's buffer is too small,
H:1720-1722), PS_EOOM from a tick, PS_EDEVICE for a device failure; a failed ps_init returns no context; and
the next tick on the same context, from the same input, equals bit for bit a
tick on a context that never failed (sorted state, forces)."""
import numpy as np
import pytest

from gpu_helpers import log
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("codes", "id", "x", "y", "z", "m", "fx", "fy", "fz")
N = 20_000


@pytest.fixture(autouse=True)
def disarm():
    api.debug_fail(api.FAIL_OFF)
    yield
    api.debug_fail(api.FAIL_OFF)


@pytest.fixture(scope="module")
def state():
    x, y, z, m = W.plummer(N, seed=501)
    with api.Solver(N, theta=2.0) as s:  # also grows the prepare's workspace
        s.load(x, y, z, m)
        s.tick(api.MORTON_AVX2)
        return (x, y, z, m), s.result()


def failing_tick(s, what, k):
    """arm, tick, disarm; the code the failing call returned (0: none failed)"""
    api.debug_fail(what, k)
    try:
        s.tick(api.MORTON_AVX2)
        rc = api.PS_OK
    except api.PsError as err:
        rc = err.rc
    api.debug_fail(api.FAIL_OFF)
    return rc


def assert_recovers(s, inputs, want):
    s.load(*inputs)
    s.tick(api.MORTON_AVX2)
    got = s.result()
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f


def test_init_allocation_failures():
    c0, _ = api.alloc_stats()
    rc, ctx = api.ps_init(api.buff_size_for(N), N, 2.0)
    api.check("ps_init", rc)
    c1, _ = api.alloc_stats()
    api.check("ps_destroy", api.ps_destroy(ctx))
    k_init = c1 - c0
    log(f"ps_init makes {k_init} allocations")
    assert k_init > 0
    for k in range(k_init):
        api.debug_fail(api.FAIL_ALLOC, k)
        rc, ctx = api.ps_init(api.buff_size_for(N), N, 2.0)
        api.debug_fail(api.FAIL_OFF)
        assert rc == api.PS_EALLOC and ctx is None, k
    rc, ctx = api.ps_init(api.buff_size_for(N), N, 2.0)
    api.check("ps_init", rc)
    api.check("ps_destroy", api.ps_destroy(ctx))


def test_tick_allocation_failures(state):
    inputs, want = state
    with api.Solver(N, theta=2.0) as s:
        s.load(*inputs)
        c0, _ = api.alloc_stats()
        s.tick(api.MORTON_AVX2)
        k_tick = api.alloc_stats()[0] - c0
    log(f"a first tick makes {k_tick} allocations")
    assert k_tick > 0
    for k in range(k_tick):
        with api.Solver(N, theta=2.0) as s:
            s.load(*inputs)
            assert failing_tick(s, api.FAIL_ALLOC, k) == api.PS_EOOM, k
            assert_recovers(s, inputs, want)


def test_tick_device_failures(state):
    # every host wait of a tick, the first tick and a later one (the
    # incremental sort's count readback); the first arming past the tick's
    # waits fails nothing
    inputs, want = state
    for warm in (False, True):
        waits = None
        for k in range(16):
            with api.Solver(N, theta=2.0) as s:
                s.load(*inputs)
                if warm:
                    s.tick(api.MORTON_AVX2)
                    s.load(*inputs)
                rc = failing_tick(s, api.FAIL_WAIT, k)
                if rc == api.PS_OK:
                    waits = k
                    break
                assert rc == api.PS_EDEVICE, (warm, k)
                assert_recovers(s, inputs, want)
        log(f"{'later' if warm else 'first'} tick: {waits} host waits")
        assert waits is not None and waits >= 2


def test_async_wait_device_failure(state):
    inputs, want = state
    with api.Stream() as stream, api.Solver(N, theta=2.0) as s:
        s.load(*inputs)
        s.tick_async(stream, api.MORTON_AVX2)
        api.debug_fail(api.FAIL_WAIT, 0)
        rc = api.ps_calc_forces_wait(s.ctx)
        api.debug_fail(api.FAIL_OFF)
        assert rc == api.PS_EDEVICE
        assert api.ps_calc_forces_wait(s.ctx) == api.PS_OK  # nothing pending any more
        assert_recovers(s, inputs, want)


def test_batch_allocation_failures():
    systems = [W.plummer(n, seed=510 + k) for k, n in enumerate((3000, 7000, 12000))]
    n = sum(len(t[0]) for t in systems)
    with api.Batch(n, len(systems), theta=2.0, graph=True) as b:
        b.load(systems)
        b.tick(api.MORTON_AVX2)
        want = b.result()
    c0, _ = api.alloc_stats()
    with api.Batch(n, len(systems), theta=2.0, graph=True) as b:
        c1, _ = api.alloc_stats()
        b.load(systems)
        b.tick(api.MORTON_AVX2)
        k_tick = api.alloc_stats()[0] - c1
    k_create = c1 - c0
    log(f"ps_hip_batch_create makes {k_create} allocations, a first batched tick {k_tick}")
    assert k_create > 0
    for k in range(k_create):
        api.debug_fail(api.FAIL_ALLOC, k)
        with pytest.raises(api.PsError) as err:
            api.Batch(n, len(systems), theta=2.0, graph=True)
        api.debug_fail(api.FAIL_OFF)
        assert err.value.rc == api.PS_EOOM, k
    for k in range(k_tick):
        with api.Batch(n, len(systems), theta=2.0, graph=True) as b:
            b.load(systems)
            api.debug_fail(api.FAIL_ALLOC, k)
            try:
                b.tick(api.MORTON_AVX2)
                rc = api.PS_OK
            except api.PsError as err:
                rc = err.rc
            api.debug_fail(api.FAIL_OFF)
            assert rc == api.PS_EOOM, k
            for _ in range(2):  # the direct run, then the recorded graph
                b.restore()
                b.tick(api.MORTON_AVX2)
                got = b.result()
                for f in want:
                    assert np.array_equal(got[f], want[f]), (k, f)

"""Serving loop (fsm_server_*): continuous batching.  Each result must be the
same alone and in the mix: every request's 17 output vectors are compared
with the same event served alone (one request in flight), and must be
identical.  Also: admission refuses work beyond the HBM budget, a queued
request can be cancelled, and requests arriving while earlier launches run are
launched together."""
import time

import numpy as np
import pytest

import fsm_amd
import fsm_amd.workloads as W

pytestmark = pytest.mark.gpu


def _mix(n, seed):
    evs = W.batch_mix(n, seed=seed)
    evs.append(W.small_3d())
    return evs


def _same(a, b, scale):
    # the same kernels run alone and in the mix, so the results are identical
    for name in fsm_amd.OUTPUT_NAMES:
        assert np.array_equal(a[name], b[name]), name


def _alone(ctx, evs):
    """Each event served by itself: one request in flight at a time."""
    srv = fsm_amd.Server(ctx, max_batch_events=1, max_inflight=1)
    outs = []
    for p, e in evs:
        rid = srv.submit(p, e)
        assert srv.poll(2000000) == rid
        outs.append(srv.fetch(rid))
    srv.close()
    return outs


def test_results_alone_and_in_the_mix(gpu_ctx):
    evs = _mix(12, seed=21)
    alone = _alone(gpu_ctx, evs)
    srv = fsm_amd.Server(gpu_ctx, max_batch_events=5, max_inflight=2)
    ids = []
    for p, e in evs:
        rid = srv.submit(p, e)
        assert rid is not None
        ids.append(rid)
    done = set()
    t0 = time.time()
    while len(done) < len(ids):
        rid = srv.poll(200000)
        assert rid is not None, "no completion within 0.2 s"
        done.add(rid)
        assert time.time() - t0 < 60
    for rid, ref in zip(ids, alone):
        st, lat, ptr = srv.result(rid)
        assert st == 0 and lat > 0 and ptr
        got = srv.fetch(rid)
        _same(got, ref, np.abs(ref["e"]).max())
    stats = srv.stats()
    assert stats["completed"] == len(ids) and stats["launches"] >= 3 and stats["bytes_in_use"] == 0
    srv.close()


def _hold_stream(s):
    """Queue tens of milliseconds of ordinary work (matrix products) on s:
    device-input submits on s upload only after it, so the server's launches
    wait on the device while the next requests arrive on the host."""
    import torch

    with torch.cuda.stream(s):
        a = torch.full((4096, 4096), 1.0e-3, dtype=torch.float32, device="cuda")
        for _ in range(24):
            a = torch.mm(a, a).clamp_(-1.0, 1.0)
    return a


def test_arrivals_during_running_work_are_batched(gpu_ctx):
    """Requests that arrive while a launch is in flight queue up and leave
    together, up to the launch's device-time cap (one launch slot)."""
    import torch

    p, e = W.config("C1")
    srv = fsm_amd.Server(gpu_ctx, max_batch_events=64, max_inflight=1)
    s = torch.cuda.current_stream()
    d_e = torch.from_numpy(e).cuda()
    # first round fills the memory pools (the first allocations are slow)
    warm = [srv.submit_device(p, d_e.data_ptr(), e.size, stream=s) for _ in range(40)]
    for _ in warm:
        srv.release(srv.poll(5000000))
    launches0 = srv.stats()["launches"]
    keep = _hold_stream(s)
    ids = [srv.submit_device(p, d_e.data_ptr(), e.size, stream=s) for _ in range(40)]
    got = 0
    while got < len(ids):
        if srv.poll(2000000) is not None:
            got += 1
    st = srv.stats()
    # the first request launches alone; the other 39 arrive while it waits
    assert st["launches"] - launches0 < len(ids) // 2, st
    ref = _alone(gpu_ctx, [(p, e)])[0]
    for rid in ids:
        _same(srv.fetch(rid), ref, np.abs(ref["e"]).max())
    del keep
    srv.close()


def test_2d_request_does_not_wait_behind_3d_launches(gpu_ctx):
    """Slot 0 takes only 2-D requests: with two launch slots, a 2-D request
    submitted after three 3-D requests completes before the last of them."""
    import torch

    p, e = W.config("C1")
    p3, e3 = W.c2_like(101, seed=5)
    srv = fsm_amd.Server(gpu_ctx, max_inflight=2)
    s = torch.cuda.current_stream()
    d2, d3 = torch.from_numpy(e).cuda(), torch.from_numpy(e3).cuda()
    warm = [srv.submit_device(p3, d3.data_ptr(), e3.size, stream=s), srv.submit_device(p, d2.data_ptr(), e.size, stream=s)]
    for _ in warm:
        srv.release(srv.poll(5000000))
    keep = _hold_stream(s)
    three = [srv.submit_device(p3, d3.data_ptr(), e3.size, stream=s) for _ in range(3)]
    two = srv.submit_device(p, d2.data_ptr(), e.size, stream=s)
    order = []
    while len(order) < 4:
        rid = srv.poll(2000000)
        assert rid is not None
        order.append(rid)
    assert order.index(two) < order.index(three[-1]), order
    for rid in order:
        srv.release(rid)
    del keep
    srv.close()


def test_at_most_two_3d_launches_in_flight(gpu_ctx):
    """With five launch slots, three take only 2-D requests: four 3-D requests
    submitted together run two at a time (the third completes only after one
    of the first two), and a 2-D request submitted after them completes before
    the last two."""
    import torch

    p, e = W.config("C1")
    p3, e3 = W.c2_like(41, seed=6)
    srv = fsm_amd.Server(gpu_ctx, max_inflight=5)
    s = torch.cuda.current_stream()
    d2, d3 = torch.from_numpy(e).cuda(), torch.from_numpy(e3).cuda()
    warm = [srv.submit_device(p3, d3.data_ptr(), e3.size, stream=s), srv.submit_device(p, d2.data_ptr(), e.size, stream=s)]
    for _ in warm:
        srv.release(srv.poll(5000000))
    keep = _hold_stream(s)
    four = [srv.submit_device(p3, d3.data_ptr(), e3.size, stream=s) for _ in range(4)]
    two = srv.submit_device(p, d2.data_ptr(), e.size, stream=s)
    order = []
    while len(order) < 5:
        rid = srv.poll(2000000)
        assert rid is not None
        order.append(rid)
    pos = {rid: k for k, rid in enumerate(order)}
    assert pos[four[2]] > min(pos[four[0]], pos[four[1]]), order
    assert pos[four[3]] > min(pos[four[0]], pos[four[1]]), order
    assert pos[two] < max(pos[four[2]], pos[four[3]]), order
    for rid in order:
        srv.release(rid)
    del keep
    srv.close()


def test_budget_admission_and_cancel(gpu_ctx):
    p, e = W.config("C1")
    D = p.DIM_X * p.DIM_Y
    per_event = D * (4 + 17 * 8)  # input + outputs, workspace comes on top
    srv = fsm_amd.Server(gpu_ctx, hbm_budget_bytes=3 * per_event, max_batch_events=1, max_inflight=1)
    a = srv.submit(p, e)
    b = srv.submit(p, e)
    assert a is not None
    # the third cannot fit the budget once workspace is counted
    c = srv.submit(p, e)
    assert c is None or b is None
    # cancel whatever is still queued
    if b is not None:
        r = srv.cancel(b)
        assert r in (0, fsm_amd.FSM_ERR_BUSY)
    seen = set()
    while len(seen) < (2 if b is not None else 1):
        rid = srv.poll(500000)
        assert rid is not None
        seen.add(rid)
    if b is not None:
        st, _, _ = srv.result(b)
        assert st in (0, fsm_amd.FSM_ERR_CANCELLED)
        srv.release(b)
    assert srv.result(a)[0] == 0
    srv.release(a)
    assert srv.stats()["bytes_in_use"] == 0
    srv.close()


def test_device_inputs_match_host_inputs(gpu_ctx):
    """fsm_server_submit_device (profiles already in HBM) gives the same
    results as the host-input submit, for 2-D, 3-D and baryon requests."""
    import torch

    pool = W.serving_mix(6, seed=3)
    pool.append((W.with_baryon(W.small_3d()[0]), W.small_3d()[1], W.baryon_density(W.small_3d()[1])))
    srv = fsm_amd.Server(gpu_ctx, max_batch_events=4, max_inflight=2)
    s = torch.cuda.current_stream()
    for p, e, nb in pool:
        rid = srv.submit(p, e, nb)
        assert srv.poll(5000000) == rid
        ref = srv.fetch(rid) if nb is None else srv.fetch_baryon(rid)
        de = torch.from_numpy(e).cuda()
        dn = None if nb is None else torch.from_numpy(nb).cuda()
        rid = srv.submit_device(p, de.data_ptr(), e.size, None if dn is None else dn.data_ptr(), stream=s)
        assert srv.poll(5000000) == rid
        got = srv.fetch(rid) if nb is None else srv.fetch_baryon(rid)
        if nb is None:
            _same(got, ref, 0)
        else:
            _same(got[0], ref[0], 0)
            assert np.array_equal(got[1], ref[1])
            for a, b in zip(got[2], ref[2]):
                assert np.array_equal(a, b)
    srv.close()

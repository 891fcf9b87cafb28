"""GPU tests of the served step (faces_server_*): a request's result is the same
bit for bit whether it runs alone or inside a mix of concurrent requests, and
equals the oracle; admission respects the footprint cap."""
from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

import faces_amd as fa  # noqa: E402
from faces_amd.batch import Server  # noqa: E402


def field(torch, shape, seed, kind=fa.FILL_RAND):
    d = torch.empty(fa.field_doubles(*shape), dtype=torch.float64, device="cuda")
    fa.fill(d, *shape, kind=kind, seed=seed)
    return d


def oracle_one(shape, seed, periodic=True, compute=True, kind=fa.FILL_RAND):
    u = oracle.new_field(*shape, kind, seed=seed)
    oracle.share_sweep(u, *shape, periodic=periodic, compute=compute)
    return u


def mix_shapes(count, seed):
    out = []
    for g in range(count):
        h = [oracle.hash64(seed, (g << 3) + k) for k in range(4)]
        out.append(((4, 5, 6, 7, 8)[h[3] % 5], 1 + h[0] % 9, 1 + h[1] % 9, 1 + h[2] % 9))
    return out


def test_alone_equals_in_mix(torch_cuda):
    print("case serve alone vs mix", flush=True)
    probes = [((7, 6, 5, 4), 11), ((4, 3, 3, 3), 12), ((8, 2, 5, 1), 13)]
    srv = Server(footprint_cap=4e9, max_batch=64, nstreams=2)
    alone = []
    for shape, seed in probes:
        d = field(torch_cuda, shape, seed)
        torch_cuda.cuda.synchronize()
        srv.wait(srv.submit(d, *shape))
        alone.append(d.cpu().numpy())
    # the same requests submitted among 120 others, all at once
    others = mix_shapes(120, seed=99)
    fields = [field(torch_cuda, s, 1000 + i) for i, s in enumerate(others)]
    mixed = [field(torch_cuda, shape, seed) for shape, seed in probes]
    torch_cuda.cuda.synchronize()
    tickets = []
    for i, s in enumerate(others):
        tickets.append(srv.submit(fields[i], *s))
        if i % 40 == 20:
            j = i // 40
            tickets.append(srv.submit(mixed[j], *probes[j][0]))
    for t in tickets:
        assert srv.wait(t) >= 0
    st = srv.stats()
    assert st["completed"] >= 123 and st["launches"] >= 2
    for j, (shape, seed) in enumerate(probes):
        got = mixed[j].cpu().numpy()
        assert np.array_equal(got.view(np.uint64), alone[j].view(np.uint64)), shape
        r = oracle.compare(got, oracle_one(shape, seed), *shape)
        assert r["ndiff"] == 0, (shape, r)
    for i in range(0, 120, 17):
        r = oracle.compare(fields[i].cpu().numpy(), oracle_one(others[i], 1000 + i), *others[i])
        assert r["ndiff"] == 0, (others[i], r)
    srv.close()


def test_admission_cap(torch_cuda):
    """A request larger than the cap is refused; with a cap that holds about
    two requests at a time, every request still completes and the admitted
    footprint never exceeds the cap."""
    print("case serve admission", flush=True)
    shape = (6, 4, 4, 4)
    nbytes = 8 * fa.field_doubles(*shape)
    srv = Server(footprint_cap=2.5 * nbytes, max_batch=16, nstreams=3)
    big = field(torch_cuda, (6, 8, 8, 8), 1)
    with pytest.raises(fa.FacesError) as e:
        srv.submit(big, 6, 8, 8, 8)
    assert e.value.status == fa.ERR_ADMISSION
    fields = [field(torch_cuda, shape, 300 + i) for i in range(12)]
    torch_cuda.cuda.synchronize()
    ts = [srv.submit(d, *shape) for d in fields]
    for t in ts:
        srv.wait(t)
    st = srv.stats()
    assert st["peak_bytes"] <= 2.5 * nbytes and st["completed"] == 12
    for i, d in enumerate(fields):
        r = oracle.compare(d.cpu().numpy(), oracle_one(shape, 300 + i), *shape)
        assert r["ndiff"] == 0, (i, r)
    srv.close()


@pytest.mark.parametrize("boundary,compute", [(fa.OPEN, True), (fa.PERIODIC, False)])
def test_other_modes(torch_cuda, boundary, compute):
    print("case serve other modes", boundary, compute, flush=True)
    shapes = [(4, 3, 2, 2), (13, 2, 1, 1), (7, 1, 1, 1)]
    srv = Server(footprint_cap=1e9, max_batch=8, nstreams=2)
    fields = [field(torch_cuda, s, 400 + i) for i, s in enumerate(shapes)]
    torch_cuda.cuda.synchronize()
    ts = [srv.submit(d, *s, boundary=boundary, compute=compute) for d, s in zip(fields, shapes)]
    for t in ts:
        srv.wait(t)
    for i, s in enumerate(shapes):
        want = oracle_one(s, 400 + i, periodic=boundary == fa.PERIODIC, compute=compute)
        r = oracle.compare(fields[i].cpu().numpy(), want, *s)
        assert r["ndiff"] == 0, (s, r)
    srv.close()


def test_run_trace(torch_cuda):
    """Timed arrivals (Poisson, 20k requests/s) through run_trace: latencies
    are reported and results match the oracle."""
    print("case serve trace", flush=True)
    rng = np.random.default_rng(5)
    nreq = 200
    arrivals = np.cumsum(rng.exponential(50.0, size=nreq)).tolist()
    shapes = mix_shapes(nreq, seed=17)
    fields = [field(torch_cuda, s, 2000 + i) for i, s in enumerate(shapes)]
    torch_cuda.cuda.synchronize()
    srv = Server(footprint_cap=8e9, max_batch=64, nstreams=2)
    lat, st = srv.run_trace(arrivals, shapes, fields)
    assert len(lat) == nreq and min(lat) > 0
    assert st["p50_us"] <= st["p99_us"] <= st["max_us"]
    assert st["launches"] >= 1 and st["req_per_s"] > 0
    for i in range(0, nreq, 23):
        r = oracle.compare(fields[i].cpu().numpy(), oracle_one(shapes[i], 2000 + i), *shapes[i])
        assert r["ndiff"] == 0, (shapes[i], r)
    srv.close()


def test_wait_unknown_ticket(torch_cuda):
    srv = Server(footprint_cap=1e9, max_batch=4, nstreams=1)
    with pytest.raises(fa.FacesError):
        srv.wait(12345, timeout_s=0.1)
    srv.close()


def test_ticket_waited_once(torch_cuda):
    """A ticket is released by its wait: waiting again is an error, and of
    two threads waiting on one ticket exactly one gets the result."""
    import threading
    print("case serve ticket waited once", flush=True)
    shape = (5, 3, 3, 3)
    srv = Server(footprint_cap=1e9, max_batch=4, nstreams=1)
    d = field(torch_cuda, shape, 700)
    torch_cuda.cuda.synchronize()
    t = srv.submit(d, *shape)
    srv.wait(t)
    with pytest.raises(fa.FacesError) as e:
        srv.wait(t, timeout_s=0.1)
    assert e.value.status == fa.ERR_ARG
    fields = [field(torch_cuda, shape, 710 + i) for i in range(8)]
    torch_cuda.cuda.synchronize()
    for i, d in enumerate(fields):
        t = srv.submit(d, *shape)
        outcome = []

        def waiter():
            try:
                srv.wait(t, timeout_s=30.0)
                outcome.append("ok")
            except fa.FacesError as x:
                outcome.append(x.status)

        th = [threading.Thread(target=waiter) for _ in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert sorted(outcome, key=str) == sorted(["ok", fa.ERR_ARG], key=str), outcome
        r = oracle.compare(d.cpu().numpy(), oracle_one(shape, 710 + i), *shape)
        assert r["ndiff"] == 0, (i, r)
    srv.close()


def test_submit_from_several_threads(torch_cuda):
    """Four host threads submit and wait for their own requests at once; every
    result is bit-exact and the server's counts add up."""
    import threading
    print("case serve threads", flush=True)
    shapes = mix_shapes(48, seed=23)
    fields = [field(torch_cuda, s, 800 + i) for i, s in enumerate(shapes)]
    torch_cuda.cuda.synchronize()
    srv = Server(footprint_cap=8e9, max_batch=16, nstreams=2)
    errors = []

    def worker(k):
        try:
            ts = [srv.submit(fields[i], *shapes[i]) for i in range(k, len(shapes), 4)]
            for t in ts:
                srv.wait(t, timeout_s=60.0)
        except Exception as x:  # reported below, in the main thread
            errors.append(repr(x))

    th = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    st = srv.stats()
    assert st["admitted"] == len(shapes) and st["completed"] == len(shapes)
    for i, s in enumerate(shapes):
        r = oracle.compare(fields[i].cpu().numpy(), oracle_one(s, 800 + i), *s)
        assert r["ndiff"] == 0, (s, r)
    srv.close()


def test_large_requests_launch_alone(torch_cuda):
    """A request of 64 MiB or more gets a launch of its own; the small ones
    around it are stepped in other launches, all bit-exact."""
    print("case serve large", flush=True)
    big = (7, 30, 30, 30)  # 74 MB
    shapes = [(4, 3, 3, 3), big, (6, 2, 3, 4), (5, 4, 4, 4), big, (8, 2, 2, 2)]
    fields = [field(torch_cuda, s, 500 + i) for i, s in enumerate(shapes)]
    torch_cuda.cuda.synchronize()
    srv = Server(footprint_cap=8e9, max_batch=64, nstreams=1)
    ts = [srv.submit(d, *s) for d, s in zip(fields, shapes)]
    for t in ts:
        srv.wait(t)
    assert srv.stats()["launches"] >= 3
    for i, s in enumerate(shapes):
        r = oracle.compare(fields[i].cpu().numpy(), oracle_one(s, 500 + i), *s)
        assert r["ndiff"] == 0, (s, r)
    srv.close()


def test_served_workload_full_size(torch_cuda):
    """The served workload at full size: 600 requests, Poisson arrivals at
    20k requests/s, mx, my, mz uniform in [2,24], n from {4,6,8}, 5% of them
    48^3 n=7, random FP64 [-1,1), footprint cap 200 GB, two launch slots.
    Every large request and every 8th small one equals the oracle bit for
    bit; three probes (two small, one large) give the same bits alone and
    inside the trace."""
    from bench import LARGE_SHAPE, served_workload
    print("case serve full-size workload", flush=True)
    nreq = 600
    arrivals, shapes = served_workload(nreq)
    large = [i for i, s in enumerate(shapes) if s == LARGE_SHAPE]
    small = [i for i, s in enumerate(shapes) if s != LARGE_SHAPE]
    assert len(large) >= 15, len(large)
    assert all(s[0] in (4, 6, 8) and all(2 <= m <= 24 for m in s[1:]) for s in (shapes[i] for i in small))
    probes = [small[len(small) // 4], large[len(large) // 2], small[3 * len(small) // 4]]
    seed = lambda i: 9000 + i  # noqa: E731
    srv = Server(footprint_cap=200e9, max_batch=256, nstreams=2)
    alone = []
    for i in probes:
        d = field(torch_cuda, shapes[i], seed(i))
        torch_cuda.cuda.synchronize()
        srv.wait(srv.submit(d, *shapes[i]))
        alone.append(d.cpu().numpy())
        del d
    fields = [field(torch_cuda, s, seed(i)) for i, s in enumerate(shapes)]
    torch_cuda.cuda.synchronize()
    lat, st = srv.run_trace(arrivals, shapes, fields)
    assert len(lat) == nreq and min(lat) > 0
    assert st["p50_us"] <= st["p99_us"] <= st["max_us"]
    assert srv.stats()["peak_bytes"] <= 200e9
    print(f"served {nreq} requests ({len(large)} large): p50 {st['p50_us']:.1f} us, "
          f"p99 {st['p99_us']:.1f} us, {st['launches']:.0f} launches", flush=True)
    srv.close()
    for j, i in enumerate(probes):
        got = fields[i].cpu().numpy()
        assert np.array_equal(got.view(np.uint64), alone[j].view(np.uint64)), (i, shapes[i])
    checked = large + small[::8]
    for i in checked:
        r = oracle.compare(fields[i].cpu().numpy(), oracle_one(shapes[i], seed(i)), *shapes[i])
        assert r["ndiff"] == 0, (i, shapes[i], r)
    print(f"checked {len(checked)} requests against the oracle", flush=True)

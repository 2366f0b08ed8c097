"""Dispatch order of the serving loop (host logic only, no GPU).

'fifo' (the default) keeps arrival order; 'sjf' takes the shortest job first by
work = particles + 4 * fields * dims^3, except that a request that has waited longer
than max_wait_s goes first (oldest first); 'hrrn' takes the highest ratio
(wait + service) / service.
"""
import collections
import time

import numpy as np

from pylians3_amd.serving import Request, Server


class _Srv:
    pass


def _item(n, dims, age, now):
    r = Request(np.zeros((n, 3), np.float32), dims, 100.0)
    r.t_submit = now - age
    return (r, None)


def test_shortest_job_first_with_aging():
    s = _Srv()
    s.policy, s.max_wait_s = 'sjf', 2.0
    now = time.perf_counter()
    s.queue = collections.deque([_item(1_000_000, 256, 0.5, now), _item(1000, 64, 0.1, now),
                                 _item(50_000, 128, 0.2, now)])
    assert Server._pick(s) == 1
    assert s.queue[1][0].work() < s.queue[2][0].work() < s.queue[0][0].work()
    s.queue[0][0].t_submit = now - 5.0          # waited too long: served first
    assert Server._pick(s) == 0
    s.queue[2][0].t_submit = now - 9.0          # the oldest of the overdue ones
    assert Server._pick(s) == 2


def test_fifo_policy_keeps_arrival_order():
    s = _Srv()
    s.policy, s.max_wait_s = 'fifo', 2.0
    now = time.perf_counter()
    s.queue = collections.deque([_item(1_000_000, 256, 0.0, now), _item(10, 16, 0.0, now)])
    assert Server._pick(s) == 0


def test_hrrn_policy_ratio():
    s = _Srv()
    s.policy, s.max_wait_s = 'hrrn', 2.0
    now = time.perf_counter()
    big, small = _item(10_000_000, 512, 0.0, now), _item(1000, 64, 0.0, now)
    s.queue = collections.deque([big, small])
    # equal waits: the short job has the higher ratio
    big[0].t_submit = small[0].t_submit = now - 0.001
    assert Server._pick(s) == 1
    # the long job has waited many times its own service estimate
    big[0].t_submit = now - 20 * big[0].service_s()
    assert Server._pick(s) == 0


def test_default_policy_is_arrival_order():
    import inspect
    assert inspect.signature(Server.__init__).parameters['policy'].default == 'fifo'


def test_warm_refuses_a_busy_server():
    import threading
    import pytest
    s = _Srv()
    s.cv = threading.Condition()
    s.ctxs = [None, None]
    s.free_slots = [0]                       # one slot running
    s.queue = collections.deque()
    with pytest.raises(RuntimeError):
        Server.warm(s, None)
    s.free_slots = [0, 1]
    s.queue = collections.deque([_item(10, 16, 0.0, time.perf_counter())])
    with pytest.raises(RuntimeError):
        Server.warm(s, None)

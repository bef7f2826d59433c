"""GPU tests of the continuous-batching server: requests join and leave the
running ensembles, and every result equals running the request alone."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_requests(n, seed):
    from lbswim.serving import poisson_requests
    return poisson_requests(n, rate=1000.0, seed=seed, sizes=(20, 40), steps=(15, 120), swimmers=(16, 300),
                            tracers=(0, 2000))


def test_results_same_alone_and_in_the_mix():
    from lbswim.serving import Server, estimate_bytes, run_alone

    class AdmitTogether(Server):
        """Admits only once every request being prepared is built, so the
        requests submitted together run together whatever the build times
        (a short request may otherwise finish before the next is built)."""

        def admit_ready(self):
            for _, fut in self.preparing.values():
                fut.result()
            super().admit_ready()

    reqs = small_requests(10, seed=7)
    # room for about three requests at a time: the rest queue and are admitted as others finish
    budget = 3 * max(estimate_bytes(r) for r in reqs)
    say("serving mix: %d requests, budget %d bytes, storages %s" % (len(reqs), budget, [r.storage for r in reqs]))
    srv = AdmitTogether(hbm_budget=budget, chunk=16)
    peak = []
    stats = srv.run(reqs, realtime=False, on_chunk=lambda s: peak.append(len(s.active)))
    say("stats %s, peak concurrent %d" % (stats, max(peak)))
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert max(peak) >= 2  # requests really overlapped on the device
    for r in reqs:
        alone = run_alone(r)
        for k, v in alone.items():
            assert np.array_equal(r.result[k], v), (r.rid, k)


def test_rejection_and_cancellation():
    from lbswim.serving import Request, Server, estimate_bytes, run_alone
    reqs = small_requests(5, seed=8)
    huge = Request(rid=99, size=96, steps=10, swimmers=10, tracers=0, seed=5)
    budget = max(estimate_bytes(r) for r in reqs) * 4
    assert estimate_bytes(huge) > budget
    say("serving: rejection and cancellation")
    srv = Server(hbm_budget=budget, chunk=8)
    victim = reqs[0].rid

    def cancel_once(s):
        if victim in s.active and victim not in s.cancelled:
            s.cancel(victim)

    stats = srv.run(reqs + [huge], realtime=False, on_chunk=cancel_once)
    assert stats["rejected"] == 1 and srv.rejected == [99]
    assert srv.cancelled == [victim]
    assert stats["completed"] == len(reqs) - 1
    for r in reqs[1:]:
        alone = run_alone(r)
        for k, v in alone.items():
            assert np.array_equal(r.result[k], v), (r.rid, k)

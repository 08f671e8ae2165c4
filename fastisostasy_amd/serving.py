"""Continuous-batching request server over resident LV-ELVA domains.

A thin mirror of the ``fi_server_*`` entry points of ``include/fastiso.h``
plus the open-loop driver that measures per-request latency (arrival to
result) for a Poisson request stream.  Scheduling, grouping and the batched
launches run in the native library; this module only feeds requests and reads
completions.
"""
import ctypes
import sys
import time

import numpy as np

from . import _lib
from ._lib import check, load


class Server:
    """fi_server: resident domains, a request queue per domain, one batched
    update per launch group and round."""

    def __init__(self, hbm_budget_bytes, max_batch=64):
        self.h = ctypes.c_void_p()
        check(load().fi_server_create(float(hbm_budget_bytes), int(max_batch), ctypes.byref(self.h)))
        self._bufs = []

    def admit(self, iso, bsl):
        """Make an initialised domain resident; raises FiError (FI_E_CAPACITY)
        when its device footprint does not fit the remaining budget."""
        slot = ctypes.c_int()
        check(load().fi_server_admit(self.h, iso.h, bsl.h, ctypes.byref(slot)))
        return slot.value

    def release(self, slot):
        check(load().fi_server_release(self.h, int(slot)))

    def submit(self, slot, time_target, tag, H_ice=None):
        if H_ice is None:
            ptr = None
        else:
            a = np.ascontiguousarray(H_ice, dtype=np.float64)
            ptr = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(load().fi_server_submit(self.h, int(slot), ptr, float(time_target), int(tag)))

    def run(self):
        """One scheduling round; returns the number of requests served."""
        n = ctypes.c_int()
        check(load().fi_server_run(self.h, ctypes.byref(n)))
        return n.value

    def poll(self, wait=False):
        """Continuous scheduling (fi_server_poll): follow up finished work and
        launch idle domains' requests on idle streams; returns the number of
        requests retired by this call."""
        n = ctypes.c_int()
        check(load().fi_server_poll(self.h, 1 if wait else 0, ctypes.byref(n)))
        return n.value

    def pending(self):
        n = ctypes.c_int()
        check(load().fi_server_pending(self.h, ctypes.byref(n)))
        return n.value

    def completed(self, max_tags=1 << 16):
        buf = (ctypes.c_longlong * max_tags)()
        n = ctypes.c_int()
        check(load().fi_server_completed(self.h, buf, max_tags, ctypes.byref(n)))
        return list(buf[:n.value])

    def failed(self, max_tags=1 << 16):
        """Tags of requests whose update ran but failed its instability / NaN
        checks (they are not re-run)."""
        buf = (ctypes.c_longlong * max_tags)()
        n = ctypes.c_int()
        check(load().fi_server_failed(self.h, buf, max_tags, ctypes.byref(n)))
        return list(buf[:n.value])

    def stats(self):
        r, g, s = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        used = ctypes.c_double()
        check(load().fi_server_stats(self.h, ctypes.byref(r), ctypes.byref(g), ctypes.byref(s), ctypes.byref(used)))
        return dict(rounds=r.value, groups=g.value, served=s.value, hbm_used=used.value)

    def close(self):
        if self.h:
            load().fi_server_destroy(self.h)
            self.h = ctypes.c_void_p()


def closed_loop_rate(server, slots, t_now, seconds, step=10.0, continuous=False):
    """Throughput (requests/s) with every domain always holding one pending
    request.  t_now: per-slot current model times (updated in place).  With
    `continuous`, fi_server_poll schedules and a domain's next request is
    submitted as soon as its last one completes (the work still in flight at
    the end is drained outside the timed span)."""
    if continuous:
        owner = {}
        tag = 0
        for k, s in enumerate(slots):
            t_now[k] += step
            server.submit(s, t_now[k], tag)
            owner[tag] = k
            tag += 1
        served = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            server.poll()
            for done in server.completed():
                served += 1
                k = owner.pop(done)
                t_now[k] += step
                server.submit(slots[k], t_now[k], tag)
                owner[tag] = k
                tag += 1
        wall = time.perf_counter() - t0
        while server.pending():
            server.poll(wait=True)
        server.completed()
        return served / wall
    tag = 0
    served = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for k, s in enumerate(slots):
            t_now[k] += step
            server.submit(s, t_now[k], tag)
            tag += 1
        while server.pending():
            served += server.run()
        server.completed()
    return served / (time.perf_counter() - t0)


def open_loop(server, slots, t_now, arrivals, domains, step=10.0, progress=0, continuous=False):
    """Serve a request stream with arrival times `arrivals` (seconds from the
    start) on domain indices `domains`; returns per-request latencies
    (seconds, arrival to completion) and the wall time.  With `progress` > 0
    a line goes to stderr every `progress` completed requests.  With
    `continuous`, fi_server_poll schedules instead of rounds."""
    n = len(arrivals)
    lat = np.full(n, np.nan)
    i = 0
    done = 0
    next_report = progress if progress > 0 else n + 1
    t0 = time.perf_counter()
    while done < n:
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            k = int(domains[i])
            t_now[k] += step
            server.submit(slots[k], t_now[k], i)
            i += 1
        if server.pending() == 0:
            # idle until the next arrival
            if i < n:
                wait = arrivals[i] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
            continue
        if continuous:
            server.poll()
        else:
            server.run()
        t_end = time.perf_counter() - t0
        for tag in server.completed():
            lat[tag] = t_end - arrivals[tag]
            done += 1
        if done >= next_report:
            print(f"# open loop: {done}/{n} requests served in {t_end:.1f} s", file=sys.stderr, flush=True)
            next_report += progress
    return lat, time.perf_counter() - t0

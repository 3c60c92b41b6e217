"""Request loop over one MI355X: continuous batching of split-operator
evolutions (1D and 2D) and 2D Born-series scattering solves.

The reference runs one solver per caller, synchronously, with a private
Device each (src/EvolverImpl.h:25); this loop keeps many solvers resident on
the device at once instead:

* Requests arrive over time.  A request is admitted (its grids built on the
  GPU and its job created) as soon as its HBM footprint fits the budget,
  first come first served.
* Admitted 2D evolutions of every grid size share one multi-size group
  (qsim_hip_split_group_*) on a stream of its own: each advance is one
  launch per sweep for all of them, replayed from a captured HIP graph, and
  each job runs its whole step count as one sweep program spread over as
  many advances as it needs.  New jobs join and finished jobs leave between
  advances.
* 1D evolutions join a lane of their length: a lane advances all its jobs
  in one launch per chunk (qsim_hip_split_batch_steps) on its own stream;
  jobs join and leave between chunks.
* Born solves run on worker threads, each with a stream of its own; each
  solve advances a fixed number of iterations per turn.
* A 1D job advances in chunks of a fixed length set by its kind and size
  (`quantum`), a 2D job in one program of its own step count, never by what
  else is on the device, and the kernels compute every job independently.
  So a result is bitwise the same alone and in any mix.

Footprints (bytes per grid point), which bound what a job allocates:
  evolution FP32-mixed: psi 16 + potential staging 16 + phase tables 32 + FP32 state 8 = 72
  evolution FP64 fused: 64; FP64 unfused (lengths without a fused kernel): 192
  Born FP32: V 16 + iterate 8 + work 8 = 32; Born FP64: 48 (+ small per-axis tables)
"""
import collections
import ctypes
import math
import queue
import threading
import time

import numpy as np

from . import _native as N
from .workloads import MT19937_64, SEED

KINDS = ("evo1d", "evo2d", "born2d")


# ------------------------------------------------------------------ requests
class Request:
    """One solve: kind, points per axis n, work (steps, or the iteration cap
    of a Born solve), arrival time (s) and the physical parameters."""
    __slots__ = ("rid", "kind", "n", "work", "arrival", "params")

    def __init__(self, rid, kind, n, work, arrival, params):
        self.rid, self.kind, self.n, self.work = rid, kind, int(n), int(work)
        self.arrival, self.params = float(arrival), params

    def points(self):
        return self.n if self.kind == "evo1d" else self.n * self.n

    def __repr__(self):
        return "Request(%d, %s, n=%d, work=%d, t=%.4f)" % (self.rid, self.kind, self.n,
                                                            self.work, self.arrival)


def make_requests(count=2000, rate=400.0, seed=SEED + 5, mix=(0.6, 0.3, 0.1)):
    """Seeded stream with Poisson arrivals at `rate` requests/s.

    * 1D evolutions (share mix[0]): N in {1024..16384}, steps U{100..2000}.
      Barrier h exp(-x^2/2) with h ~ U(0.3,0.9) on [-100,100), packet at -30
      with sigma 5 and k ~ U(0.8,1.2), dt 0.5 (the 1D tunnelling set-up).
    * 2D evolutions (mix[1]): n in {256,512,1024}, steps U{50..500}.
      16 Gaussian bumps plus the absorbing layer on [-150,150)^2, packet
      sigma 8 at (-60,0), k (0.8,0), dt 0.2 (the 2D benchmark set-up, with
      the request's own bumps).
    * 2D Born scattering (mix[2]): n in {256,512,1024} on [-80,80)^2.  8 bumps
      a ~ U(0,0.3), w ~ U(1.5,4), centres in [-10,10]^2; incident along x
      with k ~ U(1,3), E = k^2/2.  Vellekoop preconditioner, slow 0.5; up to
      200 iterations, stopping once the relative update is below 1e-6."""
    g = MT19937_64(seed)
    t = 0.0
    out = []
    c1, c2 = mix[0], mix[0] + mix[1]
    for i in range(count):
        t += -math.log(1.0 - g.uniform()) / rate
        u = g.uniform()
        if u < c1:
            n = 1024 << int(g.uniform() * 5)
            steps = 100 + int(g.uniform() * 1901)
            params = dict(h=g.uniform(0.3, 0.9), k=g.uniform(0.8, 1.2), center=-30.0, sigma=5.0,
                          dt=0.5, lo=-100.0, hi=100.0)
            out.append(Request(i, "evo1d", n, steps, t, params))
        elif u < c2:
            n = (256, 512, 1024)[int(g.uniform() * 3)]
            steps = 50 + int(g.uniform() * 451)
            bumps = []
            for _ in range(16):  # drawn in the order (cx, cy, h, w)
                cx, cy = g.uniform(-40, 40), g.uniform(-40, 40)
                h, w = g.uniform(0, 1), g.uniform(2, 6)
                bumps.append((h, w, cx, cy))
            params = dict(bumps=bumps, dt=0.2, lo=-150.0, hi=150.0, center=(-60.0, 0.0),
                          k=(0.8, 0.0), sigma=8.0, absorb=(20.0, 0.1))
            out.append(Request(i, "evo2d", n, steps, t, params))
        else:
            n = (256, 512, 1024)[int(g.uniform() * 3)]
            bumps = []
            for _ in range(8):
                a, w = g.uniform(0, 0.3), g.uniform(1.5, 4)
                bumps.append((a, w, g.uniform(-10, 10), g.uniform(-10, 10)))
            k = g.uniform(1, 3)
            params = dict(bumps=bumps, energy=0.5 * k * k, lo=-80.0, hi=80.0, tol=1e-6,
                          slow=0.5)
            out.append(Request(i, "born2d", n, 200, t, params))
    return out


def footprint(req, precision):
    """HBM bytes a request's job holds at most (see the module notes)."""
    pts = req.points()
    if req.kind == "born2d":
        return pts * (32 if precision == 32 else 48) + (1 << 20)
    if precision == 32:
        return pts * 72 + (1 << 20)
    return pts * (64 if fused(req, precision) else 192) + (1 << 20)


def fused(req, precision):
    """whether the split job runs the fused kernels (batchable); FP64 lines
    beyond the fused lengths take the unfused order on the four-step FFT"""
    if req.kind == "born2d" or precision == 32:
        return True
    return req.n <= (8192 if req.kind == "evo1d" else 4096)


def default_quantum(req):
    """Steps (or iterations) per turn: about a millisecond of device time,
    fixed by the request's own kind and size."""
    if req.kind == "evo1d":
        return 200
    if req.kind == "evo2d":
        return {256: 100, 512: 50}.get(req.n, 25)
    # Born iterations: the library checks the tolerance every 4 iterations
    # inside one call, so a turn only bounds how long a worker holds its
    # stream; every turn boundary costs a finishing half-iteration and a
    # residual read-back
    return {256: 200, 512: 100}.get(req.n, 50)


# --------------------------------------------------------------- device jobs
def _arr(v, t=ctypes.c_double):
    return (t * len(v))(*v)


class _Staging:
    """Reusable potential buffers per size: a split job reads its potential
    only while it is created, so one pair per size serves every request."""

    def __init__(self):
        self.bufs = {}

    def get(self, pts):
        if pts not in self.bufs:
            self.bufs[pts] = (N.DeviceArray(pts, np.float64), N.DeviceArray(pts, np.float64))
        return self.bufs[pts]

    def close(self):
        for a, b in self.bufs.values():
            a.close()
            b.close()
        self.bufs = {}


class EvoJob:
    """A split-operator evolution resident on the device."""

    def __init__(self, req, precision, staging):
        self.req, self.precision = req, precision
        p = req.params
        pts = req.points()
        self.psi = N.DeviceArray(pts, np.complex128)
        vre, vim = staging.get(pts)
        d = N.SplitDesc()
        if req.kind == "evo1d":
            dims, lo, hi = [req.n], [p["lo"]], [p["hi"]]
            N.call("qsim_hip_gen_packet", 1, _arr(dims, ctypes.c_size_t), _arr(lo), _arr(hi),
                   _arr([p["center"]]), _arr([p["k"]]), p["sigma"], 1, self.psi.ptr, None)
            N.call("qsim_hip_gen_potential", 1, _arr(dims, ctypes.c_size_t), _arr(lo), _arr(hi),
                   _arr([p["h"], 1.0, 0.0]), 1, 0.0, 0.0, vre.ptr, vim.ptr, None, None)
            d.dim, d.nx, d.ny = 1, req.n, 1
            d.x0, d.x1, d.y0, d.y1 = p["lo"], p["hi"], 0.0, 1.0
        else:
            dims, lo, hi = [req.n, req.n], [p["lo"]] * 2, [p["hi"]] * 2
            flat = [x for b in p["bumps"] for x in b]
            N.call("qsim_hip_gen_packet", 2, _arr(dims, ctypes.c_size_t), _arr(lo), _arr(hi),
                   _arr(p["center"]), _arr(p["k"]), p["sigma"], 1, self.psi.ptr, None)
            N.call("qsim_hip_gen_potential", 2, _arr(dims, ctypes.c_size_t), _arr(lo), _arr(hi),
                   _arr(flat), len(p["bumps"]), p["absorb"][0], p["absorb"][1], vre.ptr,
                   vim.ptr, None, None)
            d.dim, d.nx, d.ny = 2, req.n, req.n
            d.x0, d.x1, d.y0, d.y1 = p["lo"], p["hi"], p["lo"], p["hi"]
        d.dt_re, d.dt_im = p["dt"], 0.0
        d.mass = d.hbar = 1.0
        d.method, d.boundary, d.precision = 2, 0, precision
        h = ctypes.c_void_p()
        N.call("qsim_hip_split_create", ctypes.byref(h), ctypes.byref(d), vre.ptr, vim.ptr,
               self.psi.ptr)
        self.job = h.value
        self.fused = fused(req, precision)
        self.done = 0
        self.dv = ((p["hi"] - p["lo"]) / req.n) ** (1 if req.kind == "evo1d" else 2)

    def next_chunk(self, q):
        return min(q, self.req.work - self.done)

    def finished(self):
        return self.done >= self.req.work

    def result(self, keep):
        """norm after the evolution (and psi when kept); reads psi on the
        null stream, which the lanes' non-blocking streams do not wait on."""
        out = {"norm": N.reduce1("qsim_hip_norm2", self.psi.ptr, self.psi.size) * self.dv}
        if keep:
            out["psi"] = self.psi.download()
        return out

    def close(self):
        if self.job:
            N.call("qsim_hip_split_destroy", self.job)
            self.job = None
        self.psi.close()


# Born cross sections are reported toward these directions (angle from the
# incident x axis): forward, 30 degrees, 90 degrees.
BORN_ANGLES = (0.0, math.pi / 6, math.pi / 2)


class BornJob:
    """A 2D Born-series solve resident on the device."""

    def __init__(self, req, precision):
        self.req, self.precision = req, precision
        p = req.params
        pts = req.points()
        self.V = N.DeviceArray(pts, np.complex128)
        dims, lo, hi = [req.n, req.n], [p["lo"]] * 2, [p["hi"]] * 2
        flat = [x for b in p["bumps"] for x in b]
        N.call("qsim_hip_gen_potential", 2, _arr(dims, ctypes.c_size_t), _arr(lo), _arr(hi),
               _arr(flat), len(p["bumps"]), 0.0, 0.0, None, None, self.V.ptr, None)
        d = N.BornDesc()
        d.dim = 2
        d.n[0], d.n[1], d.n[2] = req.n, req.n, 1
        d.lo[0], d.lo[1], d.lo[2] = p["lo"], p["lo"], 0.0
        d.hi[0], d.hi[1], d.hi[2] = p["hi"], p["hi"], 1.0
        d.direction[0], d.direction[1], d.direction[2] = 1.0, 0.0, 0.0
        d.mass = d.hbar = 1.0
        d.preconditional, d.preconditioner, d.slow = 1, 1, p["slow"]
        d.epsilon, d.precision = 0.0, precision
        h = ctypes.c_void_p()
        E = ctypes.c_double(p["energy"])
        N.call("qsim_hip_born_create", ctypes.byref(h), ctypes.byref(d), self.V.ptr,
               ctypes.byref(E), 1)
        self.job = h.value
        self.done = 0
        self.resid = float("inf")
        self.tol = p["tol"]

    def next_chunk(self, q):
        return min(q, self.req.work - self.done)

    def finished(self):
        return self.done >= self.req.work or self.resid < self.tol

    def advance(self, q, stream):
        it, r = ctypes.c_int(), ctypes.c_double()
        N.call("qsim_hip_born_iterate", self.job, self.next_chunk(q), self.tol, ctypes.byref(it),
               ctypes.byref(r), stream)
        self.done += it.value
        self.resid = r.value

    def result(self, keep, stream=None):
        out = {"iterations": self.done, "residual": self.resid, "dsigma": []}
        for th in BORN_ANGLES:
            ds, amp = ctypes.c_double(), N.c128()
            N.call("qsim_hip_born_xsection", self.job, 0, math.cos(th), math.sin(th), 0.0,
                   ctypes.byref(ds), ctypes.byref(amp), stream)
            out["dsigma"].append(ds.value)
        if keep:
            buf = N.DeviceArray(self.req.points(), np.complex128)
            N.call("qsim_hip_born_get_dpsi", self.job, 0, buf.ptr, stream)
            N.synchronize(stream)
            out["dpsi"] = buf.download()
            buf.close()
        return out

    def close(self):
        if self.job:
            N.call("qsim_hip_born_destroy", self.job)
            self.job = None
        self.V.close()


class _Lane:
    """Same-shape split jobs advanced together on one stream."""

    CAPACITY = 4096

    def __init__(self, quantum):
        self.quantum = quantum
        p = ctypes.c_void_p()
        N.call("qsim_hip_stream_create", ctypes.byref(p))
        self.stream = p.value
        b = ctypes.c_void_p()
        N.call("qsim_hip_split_batch_create", ctypes.byref(b), self.CAPACITY)
        self.batch = b.value
        self.event = N.Event()
        self.jobs = []
        self.inflight = None

    def add(self, job):
        self.jobs.append(job)

    def launch(self):
        groups = collections.OrderedDict()
        for j in self.jobs:
            groups.setdefault(j.next_chunk(self.quantum), []).append(j)
        self.inflight = []
        for chunk, js in groups.items():
            batched = [j for j in js if j.fused]
            for j in js:
                if not j.fused:  # unfused order: one job at a time, same stream
                    N.call("qsim_hip_split_steps", j.job, chunk, self.stream)
            for k in range(0, len(batched), self.CAPACITY):
                part = batched[k:k + self.CAPACITY]
                arr = (ctypes.c_void_p * len(part))(*[j.job for j in part])
                N.call("qsim_hip_split_batch_steps", self.batch, arr, len(part), chunk,
                       self.stream)
            self.inflight.append((js, chunk))
        self.event.record(self.stream)

    def poll(self):
        """jobs whose last chunk completed and are finished (removed)."""
        if self.inflight is None or not self.event.query():
            return None
        done = []
        for js, chunk in self.inflight:
            for j in js:
                j.done += chunk
                if j.finished():
                    done.append(j)
        self.inflight = None
        if done:
            ids = set(id(j) for j in done)
            self.jobs = [j for j in self.jobs if id(j) not in ids]
        return done

    def close(self):
        if self.inflight is not None:
            self.event.synchronize()
        N.call("qsim_hip_split_batch_destroy", self.batch)
        N.call("qsim_hip_stream_destroy", self.stream)


def group_ok(req, precision):
    """2D evolutions whose grid a multi-size group takes (axes 64..2048)."""
    return req.kind == "evo2d" and 64 <= req.n <= 2048 and fused(req, precision)


class _GroupLane:
    """Every 2D evolution, whatever its grid size, in one multi-size group on
    one stream.  A job joins with its whole step count and runs it as one
    sweep program spread over the group's advances (PASSES passes each,
    replayed from a captured HIP graph), so its result is bitwise that of
    the job run alone in one call, whatever shares the launches; jobs join
    and leave between advances."""

    PASSES = 20
    CAPACITY = 4096

    def __init__(self, precision, method=2):
        p = ctypes.c_void_p()
        N.call("qsim_hip_stream_create", ctypes.byref(p))
        self.stream = p.value
        g = ctypes.c_void_p()
        N.call("qsim_hip_split_group_create", ctypes.byref(g), self.CAPACITY, precision, method)
        self.group = g.value
        self.event = N.Event()
        self.jobs = []
        self.inflight = None

    def add(self, job):
        N.call("qsim_hip_split_group_add", self.group, job.job, job.req.work)
        self.jobs.append(job)

    def launch(self):
        N.call("qsim_hip_split_group_advance", self.group, self.PASSES, 1, self.stream)
        self.event.record(self.stream)
        self.inflight = True

    def poll(self):
        """jobs finished by the completed advance (removed from the group)."""
        if self.inflight is None or not self.event.query():
            return None
        self.inflight = None
        done, keep = [], []
        fin = ctypes.c_int()
        for j in self.jobs:
            N.call("qsim_hip_split_group_status", self.group, j.job, ctypes.byref(fin), None)
            (done if fin.value else keep).append(j)
        for j in done:
            N.call("qsim_hip_split_group_remove", self.group, j.job)
            j.done = j.req.work
        self.jobs = keep
        return done

    def capture_failures(self):
        """advances that ran as direct launches because their graph capture
        was refused (the same kernels and results)"""
        n = ctypes.c_longlong()
        N.call("qsim_hip_split_group_info", self.group, None, None, None, ctypes.byref(n))
        return n.value

    def close(self):
        if self.inflight is not None:
            self.event.synchronize()
        self.failed_captures = self.capture_failures()
        N.call("qsim_hip_split_group_destroy", self.group)
        N.call("qsim_hip_stream_destroy", self.stream)


# -------------------------------------------------------------------- server
class Server:
    """Continuous-batching request loop on device 0.

    precision: 32 (FP32-mixed) or 64.  hbm_budget: bytes of job footprints
    admitted at once.  realtime: requests become visible at their arrival
    times (otherwise all at t = 0).  keep: request ids whose full fields
    (psi / dpsi) are returned."""

    def __init__(self, precision=32, hbm_budget=32 << 30, realtime=True, keep=(),
                 quantum=default_quantum, born_workers=4):
        self.precision = precision
        self.budget = int(hbm_budget)
        self.realtime = realtime
        self.keep = set(keep)
        self.quantum = quantum
        self.born_workers = int(born_workers)
        self._lock = threading.Lock()
        self._used = 0
        self.peak_used = 0
        self.records = {}
        self.results = {}
        self.errors = {}
        self._ndone = 0
        # finished jobs not yet freed: a free waits for the whole device, so
        # it happens when the lanes are idle or memory is needed
        self._graveyard = []
        # an exception that ended a Born worker thread
        self._worker_error = None
        self.capture_failures = 0

    # footprint accounting (shared with the Born worker)
    def _reserve(self, fp):
        with self._lock:
            if self._used + fp > self.budget:
                return False
            self._used += fp
            self.peak_used = max(self.peak_used, self._used)
            return True

    def _release(self, fp):
        with self._lock:
            self._used -= fp

    def _finish(self, rid, t0):
        with self._lock:
            self.records[rid]["done"] = time.perf_counter() - t0
            self._ndone += 1

    def _bury(self, job):
        with self._lock:
            self._graveyard.append(job)

    def _flush(self):
        """free every finished job (main thread)"""
        with self._lock:
            jobs, self._graveyard = self._graveyard, []
        for j in jobs:
            j.close()
            self._release(footprint(j.req, self.precision))
        return len(jobs)

    def _born_worker(self, inbox, t0):
        """Born solves on a stream of this worker's own, each advanced a
        fixed number of iterations per turn; new requests join between
        turns, one per turn, so several workers share the arrivals."""
        active = []
        stream = None
        try:
            p = ctypes.c_void_p()
            N.call("qsim_hip_stream_create", ctypes.byref(p))
            stream = p.value
            while True:
                try:
                    item = inbox.get(timeout=0.002) if not active else inbox.get_nowait()
                except queue.Empty:
                    item = None
                if item == "stop":
                    if not active:
                        break
                    inbox.put("stop")
                    item = None
                if item is not None:
                    try:
                        active.append(BornJob(item, self.precision))
                    except Exception as e:  # reported, never retried
                        self.errors[item.rid] = str(e)
                        self._finish(item.rid, t0)
                        self._release(footprint(item, self.precision))
                for job in list(active):
                    try:
                        job.advance(self.quantum(job.req), stream)
                        if job.finished():
                            self.results[job.req.rid] = job.result(job.req.rid in self.keep,
                                                                   stream)
                        else:
                            continue
                    except Exception as e:  # this request fails; the others go on
                        self.errors[job.req.rid] = str(e)
                    self._finish(job.req.rid, t0)
                    self._bury(job)
                    active.remove(job)
        except Exception as e:  # seen by the main loop, which stops the run
            self._worker_error = e
        finally:
            for job in active:
                self._bury(job)
            if stream is not None:
                N.call("qsim_hip_stream_destroy", stream)

    def run(self, requests):
        """Serve `requests`; returns the statistics (see `summary`)."""
        N.init(0)
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        pending = collections.deque(reqs)
        waiting = collections.deque()
        lanes = {}
        staging = _Staging()
        inbox = queue.Queue()
        t0 = time.perf_counter()
        workers = [threading.Thread(target=self._born_worker, args=(inbox, t0), daemon=True)
                   for _ in range(max(1, self.born_workers))]
        for w in workers:
            w.start()
        total = len(reqs)

        try:
            while self._ndone < total:
                if self._worker_error is not None:
                    raise RuntimeError("Born worker stopped: %r" % (self._worker_error,))
                now = time.perf_counter() - t0
                while pending and (not self.realtime or pending[0].arrival <= now):
                    waiting.append(pending.popleft())
                # admission, first come first served, by footprint
                while waiting:
                    r = waiting[0]
                    fp = footprint(r, self.precision)
                    if fp > self.budget:
                        waiting.popleft()
                        self.errors[r.rid] = "footprint %d B exceeds the HBM budget" % fp
                        self.records[r.rid] = {"arrival": r.arrival, "admitted": now}
                        self._finish(r.rid, t0)
                        continue
                    if not self._reserve(fp):
                        if self._flush():
                            continue
                        break
                    waiting.popleft()
                    self.records[r.rid] = {"arrival": r.arrival if self.realtime else 0.0,
                                           "admitted": time.perf_counter() - t0}
                    if r.kind == "born2d":
                        inbox.put(r)
                        continue
                    try:
                        job = EvoJob(r, self.precision, staging)
                    except Exception as e:
                        self.errors[r.rid] = str(e)
                        self._finish(r.rid, t0)
                        self._release(fp)
                        continue
                    if group_ok(r, self.precision):
                        key = "evo2d-group"
                        if key not in lanes:
                            lanes[key] = _GroupLane(self.precision)
                    else:
                        key = (r.kind, r.n)
                        if key not in lanes:
                            lanes[key] = _Lane(self.quantum(r))
                    lanes[key].add(job)
                busy = False
                for lane in lanes.values():
                    done = lane.poll()
                    if done:
                        for j in done:
                            self.results[j.req.rid] = j.result(j.req.rid in self.keep)
                            self._finish(j.req.rid, t0)
                            self._bury(j)
                    if lane.inflight is None and lane.jobs:
                        lane.launch()
                    busy = busy or lane.inflight is not None
                if not busy and self._graveyard:
                    self._flush()
                if not busy:
                    if pending and not waiting and inbox.empty():
                        wait = pending[0].arrival - (time.perf_counter() - t0)
                        time.sleep(min(max(wait, 0.0), 0.01))
                    else:
                        time.sleep(0.0002)
                else:
                    time.sleep(0.00005)
        finally:
            for _ in workers:
                inbox.put("stop")
            for w in workers:
                w.join()
            for lane in lanes.values():
                lane.close()
                self.capture_failures += getattr(lane, "failed_captures", 0)
            self._flush()
            staging.close()
        self.makespan = time.perf_counter() - t0
        return self.summary(reqs)

    def summary(self, reqs):
        lat = {k: [] for k in KINDS}
        qd = []
        for r in reqs:
            rec = self.records[r.rid]
            if r.rid in self.errors:
                continue
            lat[r.kind].append(rec["done"] - rec["arrival"])
            qd.append(rec["admitted"] - rec["arrival"])
        allv = sorted(v for k in KINDS for v in lat[k])

        def pct(v, p):
            return float(np.percentile(v, p)) if v else None

        first = min(self.records[r.rid]["arrival"] for r in reqs)
        last = max(self.records[r.rid]["done"] for r in reqs)
        return {
            "requests": len(reqs), "completed": len(allv), "errors": len(self.errors),
            "solves_per_s": len(allv) / (last - first) if last > first else None,
            "makespan_s": self.makespan,
            "latency_p50_s": pct(allv, 50), "latency_p99_s": pct(allv, 99),
            "queue_delay_p50_s": pct(qd, 50), "queue_delay_p99_s": pct(qd, 99),
            "per_kind": {k: {"count": len(lat[k]), "p50_s": pct(lat[k], 50),
                             "p99_s": pct(lat[k], 99)} for k in KINDS},
            "peak_footprint_bytes": self.peak_used, "hbm_budget_bytes": self.budget,
            "precision": self.precision, "graph_capture_failures": self.capture_failures,
        }


def solve_alone(req, precision=32, keep=True, iterations=None):
    """One request served by itself (arrival at t = 0): its result record.
    iterations: run a Born solve for exactly that many iterations (no
    tolerance stop), e.g. to match another precision's run."""
    params, work = req.params, req.work
    if iterations is not None:
        params, work = dict(req.params, tol=0.0), iterations
    r = Request(req.rid, req.kind, req.n, work, 0.0, params)
    s = Server(precision=precision, realtime=False, keep=(req.rid,) if keep else ())
    s.run([r])
    if req.rid in s.errors:
        raise RuntimeError(s.errors[req.rid])
    return s.results[req.rid]


def compare(kind, got, ref):
    """error of an FP32-mixed result against its FP64 twin: psi relative L2
    (evolutions), or for Born the largest cross-section difference over the
    angles relative to the largest cross section (the forward peak).  Side
    angles can sit 1e-13 below the peak, beneath what a complex64 iterate
    resolves, so an error relative to each angle's own value would measure
    that floor, not the solve."""
    if kind == "born2d":
        peak = max(abs(b) for b in ref["dsigma"])
        return max(abs(a - b) for a, b in zip(got["dsigma"], ref["dsigma"])) / peak
    a, b = got["psi"], ref["psi"]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def spot_check(requests, results, every=50):
    """Every `every`-th request re-solved alone in FP64; per-kind errors."""
    errs = {k: [] for k in KINDS}
    for r in requests[::every]:
        if r.rid not in results:
            continue
        got = results[r.rid]
        # a Born twin runs as many iterations as the FP32 solve did, so the
        # difference is the precision's, not the stopping point's
        its = got["iterations"] if r.kind == "born2d" else None
        ref = solve_alone(r, precision=64, iterations=its)
        errs[r.kind].append(compare(r.kind, got, ref))
    return {k: {"checked": len(v), "max": max(v) if v else None,
                "median": float(np.median(v)) if v else None} for k, v in errs.items()}

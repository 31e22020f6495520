"""Continuous-batching shot server on one GPU.

Shot requests (``Solver`` objects, or callables that build one) arrive
while others run.  The server

* admits requests in arrival order against an HBM budget (every admitted
  shot keeps its wavefield slots, models, shot record and the library's
  per-point arrays resident) and a cap on concurrent shots;
* prepares every admitted shot on its own worker thread and HIP stream
  (device-resident inputs, the library's persistent plan,
  ``simwave_hip_plan_*``);
* advances the shots that can be batched (stride 0 on the fused z-marching
  kernels: ``DeviceForward.batchable``) together: one runner thread per
  variant library calls ``simwave_hip_plans_run_<variant>`` on its own
  stream, which enqueues for every time step one step kernel and one
  receiver kernel covering all running shots of a radius, so shots join
  the running batch as they are admitted and leave it when their last step
  is enqueued;
* runs every other shot on its worker's stream in chunks of time steps (a
  stride-0 shot's chunks replay a hipGraph its plan captured once), so the
  kernels of different shots and batches overlap on the device;
* frees a finished shot's memory at once, which admits the next waiting
  request, and resolves the request's future with ``(u_full, shot_record)``
  exactly as ``Solver.forward()`` returns them;
* records per-request latency (arrival to result on the host).

A batched shot runs the same per-point operations on the same inputs as
alone, so every result is bit-identical to ``Solver.forward()``.  Device
memory, pinned staging, streams and events are the HIP library's own
(``simwave_amd.hipmem``); the server needs no other GPU runtime.
"""
import collections
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor

import numpy as np

__all__ = ["ShotServer", "ServeStats", "percentile", "shot_bytes"]


def percentile(values, q):
    """q-th percentile (linear interpolation, numpy's default)."""
    if not len(values):
        return float("nan")
    return float(np.percentile(np.asarray(values, dtype=np.float64), q))


def shot_bytes(solver):
    """HBM an admitted shot holds: wavefield slots, models, shot record and
    the library's per-point arrays (q, damping term, (dt^2/s)/den, density
    copy and, for variable density from space order 12, the density
    derivative sums)."""
    if hasattr(solver, "device_bytes"):           # test doubles
        return float(solver.device_bytes)
    sm = solver.space_model
    npts = int(np.prod(sm.extended_shape))
    esz = np.dtype(sm.dtype).itemsize
    slots = solver.num_snapshots + 2
    models = 3 if sm.density_model is not None else 2
    # variable density from space order 12: the density derivative sums
    sums = len(sm.extended_shape) if (sm.density_model is not None
                                       and sm.space_order >= 12) else 0
    rec = solver.time_model.timesteps * solver.receivers.count
    return float((slots + models + 4 + sums) * npts * esz + rec * esz)


class ServeStats:
    """Per-request timings of a server: arrival, admission, completion
    (time.perf_counter seconds) and interior grid-point updates."""

    def __init__(self):
        self.records = []

    def add(self, **rec):
        self.records.append(rec)

    @property
    def latencies(self):
        return [r["done"] - r["arrival"] for r in self.records]

    def summary(self):
        if not self.records:
            return {}
        lat = self.latencies
        first = min(r["arrival"] for r in self.records)
        last = max(r["done"] for r in self.records)
        updates = sum(r["updates"] for r in self.records)
        return {
            "requests": len(self.records),
            "updates": int(updates),
            "makespan_s": last - first,
            "gpts": updates / (last - first) / 1e9 if last > first else float("nan"),
            "latency_p50_s": percentile(lat, 50),
            "latency_p99_s": percentile(lat, 99),
            "queue_p50_s": percentile([r["admitted"] - r["arrival"] for r in self.records], 50),
            "max_active": max(r["active_at_admission"] for r in self.records),
            # mean per-request phases on its worker: argument preparation
            # and uploads (split into building the device-resident shot and
            # creating its plan), enqueue-to-completion on the device, result
            # fetch
            "prep_mean_s": float(np.mean([r.get("prep", np.nan) for r in self.records])),
            "prep_build_mean_s": float(np.mean([r.get("build", np.nan) for r in self.records])),
            "prep_plan_mean_s": float(np.mean([r.get("plan", np.nan) for r in self.records])),
            "device_mean_s": float(np.mean([r.get("device", np.nan) for r in self.records])),
            "fetch_mean_s": float(np.mean([r.get("fetch", np.nan) for r in self.records])),
            "batched": int(sum(1 for r in self.records if r.get("batched"))),
        }


class _Request:
    __slots__ = ("solver", "arrival", "future", "admitted", "keep_result", "seq", "build")

    def __init__(self, solver, arrival, seq=0):
        self.seq = seq
        self.solver = solver
        self.arrival = arrival
        self.future = Future()
        self.admitted = None
        self.keep_result = True
        self.build = None       # Future of the Solver a callable request builds

    def needs_build(self):
        s = self.solver
        return callable(s) and not hasattr(s, "space_model") and not hasattr(s, "device_bytes")


class ShotServer:
    """Admits and runs forward shots concurrently on one device.

    Parameters
    ----------
    hbm_budget : float
        Bytes of device memory the admitted shots may hold together.
    max_active : int
        Shots running at once (one worker thread and stream each).
    chunk_steps : int
        Time steps a worker enqueues per plan call (a multiple of 48 lets
        every full chunk of a stride-0 shot replay the plan's graph).
    plan_flags : int, optional
        Flags of the plans (default: FLAG_GRAPH, captured-graph replay).
    batch : bool
        Advance batchable shots together (one runner per variant library).
    device : str
    shot_factory : callable, optional
        ``solver -> shot`` with the ``DeviceForward`` interface (``steps``,
        ``updates_per_shot``, ``plan``, ``run_steps``, ``results``,
        ``release_plan``); replaceable for tests, as is ``stream_factory``.
    build_ahead : int
        Requests given as callables are built (their Solver: models and
        frontend arrays, on the host) by a pool of this many threads, for
        the first `build_ahead` waiting requests at a time, so that building
        overlaps the admission and runs of earlier shots; 0 builds each on
        the scheduler thread when it reaches the head of the queue.
    """

    def __init__(self, hbm_budget=200e9, max_active=16, chunk_steps=48, device="cuda",
                 shot_factory=None, stream_factory=None, plan_flags=None, batch=True,
                 build_ahead=4):
        self.hbm_budget = float(hbm_budget)
        self.max_active = int(max_active)
        self.chunk_steps = int(chunk_steps)
        if plan_flags is None:
            from .device import FLAG_GRAPH
            plan_flags = FLAG_GRAPH
        self.plan_flags = int(plan_flags)
        self.device = device
        self._shot_factory = shot_factory or self._device_shot
        self._stream_factory = stream_factory or self._new_stream
        self._cv = threading.Condition()
        self._waiting = collections.deque()
        self._active = 0
        self._used = 0.0
        self._closed = False
        self._workers = []
        self._idle_streams = []       # worker streams that finished a shot
        self.admissions = []          # submission numbers (0, 1, ...) in admission order
        self._submitted = 0
        self.batch = bool(batch)
        self._runners = {}            # variant -> _BatchRunner
        self.build_ahead = max(0, int(build_ahead))
        self._builders = (ThreadPoolExecutor(max_workers=self.build_ahead,
                                             thread_name_prefix="shot-build")
                          if self.build_ahead else None)
        self.stats = ServeStats()
        self._sched = threading.Thread(target=self._schedule, name="shot-scheduler",
                                       daemon=True)
        self._sched.start()

    # ------------------------------------------------------------ public
    def submit(self, solver, arrival=None, keep_result=True):
        """Queues a shot; returns a Future of (u_full, shot_record).  With
        keep_result False the results are still fetched (and timed) but the
        Future holds None, for callers that only measure the stream."""
        req = _Request(solver, time.perf_counter() if arrival is None else arrival)
        req.keep_result = bool(keep_result)
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            req.seq = self._submitted
            self._submitted += 1
            self._waiting.append(req)
            self._cv.notify_all()
        return req.future

    def close(self, wait=True):
        """Stops admitting new requests; with wait, drains the queue."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            self._sched.join()
            for t in list(self._workers):
                t.join()
        if self._builders is not None:
            self._builders.shutdown(wait=wait)
        with self._cv:
            runners = list(self._runners.values())
        for r in runners:
            r.close(wait)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close(wait=True)

    # ------------------------------------------------------------ internals
    def _device_shot(self, solver, stream=None):
        from .device import DeviceForward
        return DeviceForward(solver, device=self.device, stream=stream)

    def _new_stream(self):
        from . import hipmem
        return hipmem.Stream(hipmem.default_library())

    def _start_builds(self):
        """Starts building the first `build_ahead` waiting requests that are
        callables (called with the lock held; submitting does not block)."""
        if self._builders is None:
            return
        for i, req in enumerate(self._waiting):
            if i >= self.build_ahead:
                break
            if req.build is None and req.needs_build():
                req.build = self._builders.submit(req.solver)

    def _schedule(self):
        while True:
            with self._cv:
                while True:
                    self._start_builds()
                    if self._waiting and self._active < self.max_active:
                        break
                    if self._closed and not self._waiting:
                        return
                    self._cv.wait()
                req = self._waiting[0]
            try:
                if req.build is not None:
                    req.solver, req.build = req.build.result(), None
                elif req.needs_build():
                    req.solver = req.solver()
                need = shot_bytes(req.solver)
            except Exception as exc:          # a bad request fails alone
                with self._cv:
                    self._waiting.popleft()
                req.future.set_exception(exc)
                continue
            with self._cv:
                if need > self.hbm_budget:
                    self._waiting.popleft()
                    req.future.set_exception(MemoryError(
                        "shot needs %.3g GB, more than the %.3g GB budget"
                        % (need / 1e9, self.hbm_budget / 1e9)))
                    continue
                while (self._used + need > self.hbm_budget and self._active > 0) or \
                        self._active >= self.max_active:
                    self._start_builds()
                    self._cv.wait()
                self._waiting.popleft()
                self.admissions.append(req.seq)
                self._used += need
                self._active += 1
                active_now = self._active
                req.admitted = time.perf_counter()
            t = threading.Thread(target=self._work, name="shot-worker", daemon=True,
                                 args=(req, need, active_now))
            self._workers = [w for w in self._workers if w.is_alive()]
            self._workers.append(t)
            t.start()

    def _work(self, req, need, active_now):
        shot = None
        stream = None
        try:
            with self._cv:
                stream = self._idle_streams.pop() if self._idle_streams else None
            if stream is None:
                stream = self._stream_factory()
            t0 = time.perf_counter()
            if self._shot_factory == self._device_shot:
                # built on the worker's stream, which its memory is ordered on
                shot = self._shot_factory(req.solver, stream)
            else:
                shot = self._shot_factory(req.solver)
            tb = time.perf_counter()
            shot.plan(stream, flags=self.plan_flags)
            t1 = time.perf_counter()
            batched = self.batch and getattr(shot, "batchable", False)
            if batched:
                # the variant's runner enqueues the steps; this stream then
                # waits for the plan's last work before the results
                self._runner(shot).run(shot)
                shot.run_steps(0, stream)
            else:
                left = shot.steps
                while left > 0:
                    left = shot.run_steps(self.chunk_steps, stream)
            stream.synchronize()
            t2 = time.perf_counter()
            result = shot.results(stream)
            shot.release_plan()
            done = time.perf_counter()
            self.stats.add(arrival=req.arrival, admitted=req.admitted, done=done,
                           updates=shot.updates_per_shot, active_at_admission=active_now,
                           prep=t1 - t0, build=tb - t0, plan=t1 - tb, device=t2 - t1,
                           fetch=done - t2, batched=batched)
            req.future.set_result(result if req.keep_result else None)
        except Exception as exc:
            if shot is not None:
                try:
                    shot.release_plan()
                except Exception:
                    pass
            req.future.set_exception(exc)
        finally:
            shot = None
            with self._cv:
                if stream is not None:
                    self._idle_streams.append(stream)
                self._used -= need
                self._active -= 1
                self._cv.notify_all()


    def _runner(self, shot):
        with self._cv:
            r = self._runners.get(shot.variant)
            if r is None:
                r = _BatchRunner(self._stream_factory(), self.chunk_steps,
                                 type(shot).run_batched)
                self._runners[shot.variant] = r
            return r


class _BatchRunner:
    """Advances the batchable shots of one variant library on one stream:
    each round enqueues up to `chunk` steps of every running shot in one
    batched call; a shot whose last step is enqueued leaves the set and its
    waiter is released (the device work is ordered before its results by
    the plan's event).  At most two rounds are in flight on the device."""

    def __init__(self, stream, chunk, run_batched):
        self.stream = stream
        self.chunk = chunk
        self._run_batched = run_batched
        self._cv = threading.Condition()
        self._shots = {}              # shot -> [threading.Event, exception]
        self._closed = False
        self._thread = threading.Thread(target=self._loop, name="batch-runner", daemon=True)
        self._thread.start()

    def run(self, shot):
        """Adds `shot` to the running set; returns once its last step is
        enqueued (raises what the batched call raised)."""
        slot = [threading.Event(), None]
        with self._cv:
            if self._closed:
                raise RuntimeError("runner is closed")
            self._shots[shot] = slot
            self._cv.notify_all()
        slot[0].wait()
        if slot[1] is not None:
            raise slot[1]

    def close(self, wait=True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            self._thread.join()

    def _loop(self):
        inflight = collections.deque()
        record = getattr(self.stream, "record_event", None)   # absent in test doubles
        while True:
            with self._cv:
                while not self._shots and not self._closed:
                    self._cv.wait()
                if not self._shots:
                    return
                items = list(self._shots.items())
            shots = [s for s, _ in items]
            try:
                left = self._run_batched(shots, self.chunk, self.stream)
                err = None
            except Exception as exc:      # every shot of the round fails
                left, err = [0] * len(shots), exc
            if record is not None:
                inflight.append(record())
                if len(inflight) > 2:
                    inflight.popleft().synchronize()
            with self._cv:
                for (shot, slot), n in zip(items, left):
                    if n <= 0:
                        del self._shots[shot]
                        slot[1] = err
                        slot[0].set()

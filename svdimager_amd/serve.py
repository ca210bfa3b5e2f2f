"""Streaming snapshot imaging: requests that arrive over time are admitted against a device
memory budget, grouped by image size into batched calls, and timed from arrival to image-ready.

The reference images one snapshot per process (``main``, SVDIntegratedImager.cu:1768-1853) and
sizes its slab batches from free device memory (``choose_slab_batch_size``,
SVDIntegratedImager.cu:1470-1496).  This module is the request loop above that call for a stream
of device-resident snapshots of mixed image sizes:

* admission: a request costs its visibilities (the snapshot's device bytes) plus its output
  image, held from admission until its image is ready.  It is admitted, in arrival order, when
  the context's workspace plus the bytes already held plus its cost fit the budget.  The
  workspace is sized in advance (``reserve``: one pass over the largest shapes) and only grows,
  so admission never waits on an allocation inside the library;
* batching: admitted requests wait per (image size, cell).  Whenever fewer than
  ``max_inflight`` batched calls are in flight, the key of the oldest waiting request is taken,
  with up to ``max_batch`` requests of that key, and enqueued with
  ``svdi_image_snapshot_batch_async`` (which returns once the work is enqueued);
* completion: an event is recorded after each batched call; a request is done when the host
  sees its event complete, and its latency is that time minus its arrival time.

Everything runs on the calling thread: ``poll`` retires, admits and launches; ``run_poisson``
drives it with Poisson arrivals in wall-clock time.
"""
from __future__ import annotations

import collections
import dataclasses
import math
import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import imager
from ._lib import GRIDDER_MFMA, OUTPUT_DIRTY


@dataclasses.dataclass
class Request:
    rid: int
    snapshot: object                 # imager.DeviceSnapshot (or any object with nbytes / num_samples)
    image_size: int
    cell: float
    arrival: float                   # host clock, seconds
    admitted: float = math.nan
    launched: float = math.nan
    done: float = math.nan
    image: object = None
    plan: Optional[Dict] = None

    @property
    def key(self) -> Tuple[int, float]:
        return (int(self.image_size), float(self.cell))

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def request_bytes(req: Request) -> int:
    """Device bytes a request holds while admitted: its visibilities and its float32 image."""
    return int(req.snapshot.nbytes) + 4 * int(req.image_size) ** 2


class _Batch:
    __slots__ = ("handle", "requests")

    def __init__(self, handle, requests):
        self.handle = handle
        self.requests = requests


class SnapshotService:
    """Admission, batching by image size and completion tracking on one library context.

    ``launcher(requests, key)`` enqueues one batched call and returns a handle whose ``query()``
    is true once the call's images are ready; the default enqueues
    ``svdi_image_snapshot_batch_async`` on the current torch stream and records an event.
    ``workspace_bytes()`` reports what the library keeps (``svdi_workspace_bytes`` by default).
    Both are parameters so that the scheduling can be exercised without a device.
    """

    def __init__(self, ctx, budget_bytes: int, max_batch: int = 8, max_inflight: int = 2,
                 keep_images: bool = False, clock: Callable[[], float] = time.perf_counter,
                 launcher=None, workspace_bytes: Optional[Callable[[], int]] = None, **options):
        if max_batch < 1 or max_inflight < 1:
            raise ValueError("max_batch and max_inflight must be at least 1")
        self.ctx = ctx
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.max_inflight = int(max_inflight)
        self.keep_images = keep_images
        self.clock = clock
        self.options = dict(options)
        self.options.setdefault("gridder", GRIDDER_MFMA)
        self.options.setdefault("output", OUTPUT_DIRTY)
        self._launcher = launcher if launcher is not None else self._launch_device
        self._ws_fn = workspace_bytes if workspace_bytes is not None else ctx.workspace_bytes
        self.workspace = int(self._ws_fn())
        self.waiting: collections.deque = collections.deque()      # arrived, not admitted
        self.ready: Dict[Tuple[int, float], collections.deque] = {}  # admitted, not launched
        self.inflight: collections.deque = collections.deque()
        self.held = 0
        self.peak_held = 0
        self.completed: List[Request] = []
        self.batch_sizes: List[int] = []

    # -------------------------------------------------------------------------- device side
    def _launch_device(self, reqs: List[Request], key: Tuple[int, float]):
        import torch
        N, cell = key
        snaps = [r.snapshot for r in reqs]
        dev = snaps[0].tensors["uvw"].device
        images = [torch.empty((N, N), dtype=torch.float32, device=dev) for _ in reqs]
        _, plans = imager.image_snapshot_batch(self.ctx, snaps, N, cell, images=images, wait=False,
                                               **self.options)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(dev))
        for r, im, p in zip(reqs, images, plans):
            r.image = im
            r.plan = p
        return ev

    def reserve(self, requests: Sequence[Request]) -> int:
        """Size the workspace for the shapes of `requests` in advance and return its bytes.

        A batched call images its i-th snapshot on lane i mod (lane count), each lane with its
        own workspace, so every distinct (snapshot, key) is imaged in a batch of max_batch
        copies of itself: every lane a served batch can use then holds that shape's workspace,
        and serving the same shapes allocates nothing more."""
        import torch
        seen = set()
        for r in requests:
            ident = (id(r.snapshot), r.key)
            if ident in seen:
                continue
            seen.add(ident)
            N, cell = r.key
            imager.image_snapshot_batch(self.ctx, [r.snapshot] * self.max_batch, N, cell,
                                        wait=True, **self.options)
        torch.cuda.synchronize()
        self.workspace = int(self._ws_fn())
        return self.workspace

    # -------------------------------------------------------------------------- the loop
    def submit(self, req: Request) -> None:
        if self.workspace + request_bytes(req) > self.budget:
            raise ValueError(f"request {req.rid} needs {request_bytes(req)} bytes beside a "
                             f"{self.workspace}-byte workspace: over the {self.budget}-byte budget")
        self.waiting.append(req)

    def busy(self) -> bool:
        return bool(self.waiting or self.inflight or any(self.ready.values()))

    def poll(self) -> int:
        """Retire finished batches, admit, launch; returns the number of requests retired."""
        n = self._retire()
        self._admit()
        self._launch()
        return n

    def _retire(self) -> int:
        n = 0
        while self.inflight and self.inflight[0].handle.query():
            b = self.inflight.popleft()
            t = self.clock()
            for r in b.requests:
                r.done = t
                self.held -= request_bytes(r)
                if not self.keep_images:
                    r.image = None
                self.completed.append(r)
                n += 1
        return n

    def _admit(self) -> None:
        while self.waiting:
            r = self.waiting[0]
            c = request_bytes(r)
            if self.workspace + self.held + c > self.budget:
                if self.held == 0:
                    # Nothing admitted will free memory: the workspace has grown past what the
                    # budget leaves for this request (a shape that reserve() did not see).
                    raise RuntimeError(
                        f"request {r.rid} ({c} bytes) cannot be admitted: the workspace is now "
                        f"{self.workspace} bytes of the {self.budget}-byte budget")
                break
            self.waiting.popleft()
            r.admitted = self.clock()
            self.held += c
            self.peak_held = max(self.peak_held, self.held)
            self.ready.setdefault(r.key, collections.deque()).append(r)

    def _launch(self) -> None:
        while len(self.inflight) < self.max_inflight:
            keys = [k for k, q in self.ready.items() if q]
            if not keys:
                return
            key = min(keys, key=lambda k: self.ready[k][0].arrival)
            q = self.ready[key]
            reqs = [q.popleft() for _ in range(min(len(q), self.max_batch))]
            t = self.clock()
            for r in reqs:
                r.launched = t
            handle = self._launcher(reqs, key)
            self.inflight.append(_Batch(handle, reqs))
            self.batch_sizes.append(len(reqs))
            self.workspace = int(self._ws_fn())


class NativeService:
    """The library's own request loop (svdi_service_*): admission, batching by key and completion
    tracking run on a thread inside libsvdimager; this object submits requests, allocates their
    images and collects the results.  Same interface as SnapshotService for run_poisson
    (``submit``, ``poll``, ``busy``, ``completed``, ``batch_sizes``, ``clock``, ``reserve``).

    The native loop starts at the first ``submit``, after ``reserve`` has sized the workspace;
    from then on the context belongs to it until ``close``.  A request's ``launched`` and
    ``done`` are its arrival plus the library's queue and latency times, each counted from the
    submission, plus the delay between the request's arrival and its submission."""

    def __init__(self, ctx, budget_bytes: int, max_batch: int = 8, max_inflight: int = 2,
                 keep_images: bool = False, clock: Callable[[], float] = time.perf_counter, **options):
        if max_batch < 1 or max_inflight < 1:
            raise ValueError("max_batch and max_inflight must be at least 1")
        self.ctx = ctx
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.max_inflight = int(max_inflight)
        self.keep_images = keep_images
        self.clock = clock
        self.options = dict(options)
        self.options.setdefault("gridder", GRIDDER_MFMA)
        self.options.setdefault("output", OUTPUT_DIRTY)
        self.completed: List[Request] = []
        self.batch_sizes: List[int] = []
        self._calls = set()
        self._pending: Dict[int, Tuple[Request, float]] = {}
        self._svc = None

    def reserve(self, requests: Sequence[Request]) -> int:
        """As SnapshotService.reserve: every (snapshot, key) imaged in a batch of max_batch copies
        of itself, so that every lane holds every shape's workspace; returns its bytes."""
        if self._svc is not None:
            raise RuntimeError("reserve before the first submit")
        import torch
        seen = set()
        for r in requests:
            ident = (id(r.snapshot), r.key)
            if ident in seen:
                continue
            seen.add(ident)
            N, cell = r.key
            imager.image_snapshot_batch(self.ctx, [r.snapshot] * self.max_batch, N, cell, wait=True, **self.options)
        torch.cuda.synchronize()
        return int(self.ctx.workspace_bytes())

    def _start(self) -> None:
        import ctypes
        from ._lib import ServiceConfig, check, default_options, lib
        cfg = ServiceConfig(budget_bytes=self.budget, max_batch=self.max_batch, max_inflight=self.max_inflight,
                            options=default_options(**self.options))
        h = ctypes.c_void_p()
        check(lib().svdi_service_create(self.ctx.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self._svc = h

    def submit(self, req: Request) -> None:
        import ctypes
        import torch
        from ._lib import check, lib
        if self._svc is None:
            self._start()
        N = int(req.image_size)
        dev = req.snapshot.tensors["uvw"].device
        req.image = torch.empty((N, N), dtype=torch.float32, device=dev)
        s = req.snapshot.struct()
        t = self.clock()
        self._pending[int(req.rid)] = (req, t)
        try:
            check(lib().svdi_service_submit(self._svc, int(req.rid), ctypes.byref(s), int(req.snapshot.nbytes),
                                            imager._ptr(req.image), N, float(req.cell)))
        except Exception:
            del self._pending[int(req.rid)]
            raise

    def busy(self) -> bool:
        return bool(self._pending)

    def poll(self, wait_ms: int = 0) -> int:
        """Collect the finished requests (waiting up to wait_ms for one); returns their number."""
        import ctypes
        from ._lib import SVDI_OK, ServiceResult, SvdiError, lib, check
        if self._svc is None:
            return 0
        buf = (ServiceResult * 64)()
        n = ctypes.c_size_t()
        check(lib().svdi_service_collect(self._svc, buf, 64, ctypes.byref(n), int(wait_ms)))
        for i in range(n.value):
            res = buf[i]
            req, t_sub = self._pending.pop(int(res.id))
            if res.status != SVDI_OK:
                raise SvdiError(res.status, f"request {res.id} failed in the service")
            lag = t_sub - req.arrival
            req.admitted = req.arrival + lag
            req.launched = req.arrival + lag + res.queue_ms * 1e-3
            req.done = req.arrival + lag + res.latency_ms * 1e-3
            if not self.keep_images:
                req.image = None
            self.completed.append(req)
            if res.call not in self._calls:
                self._calls.add(res.call)
                self.batch_sizes.append(int(res.batch))
        return n.value

    def stats(self) -> Dict[str, int]:
        import ctypes
        from ._lib import ServiceStats, check, lib
        st = ServiceStats()
        if self._svc is not None:
            check(lib().svdi_service_stats(self._svc, ctypes.byref(st)))
        return {name: int(getattr(st, name)) for name, _ in ServiceStats._fields_}

    def close(self) -> None:
        """Wait for every submitted request and stop the native loop."""
        from ._lib import check, lib
        if self._svc is not None:
            while self._pending:
                self.poll(wait_ms=100)
            check(lib().svdi_service_destroy(self._svc))
            self._svc = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(rate: float, duration: float, rng: np.random.Generator) -> np.ndarray:
    """Arrival times in [0, duration) of a Poisson process of `rate` per second."""
    if rate <= 0 or duration <= 0:
        return np.zeros(0)
    n_max = int(rate * duration + 10 * math.sqrt(rate * duration) + 10)
    t = np.cumsum(rng.exponential(1.0 / rate, n_max))
    while t[-1] < duration:                       # practically never: extend the draw
        t = np.concatenate([t, t[-1] + np.cumsum(rng.exponential(1.0 / rate, n_max))])
    return t[t < duration]


def run_poisson(service, pool: Sequence[Tuple[object, int, float]], rate: float,
                duration: float, seed: int = 0, drain_limit: float = 60.0) -> Dict:
    """Requests drawn uniformly from `pool` ((snapshot, image size, cell) entries) arrive at
    `rate` per second for `duration` seconds of wall-clock time; the loop runs until every
    request is done and returns the latency and throughput figures, with the requests under
    ``"requests"``.  Raises if the work has not drained `drain_limit` seconds after the last
    arrival."""
    rng = np.random.default_rng(seed)
    times = poisson_arrivals(rate, duration, rng)
    picks = rng.integers(0, len(pool), len(times))
    clock = service.clock
    done_before = len(service.completed)
    reqs: List[Request] = []
    t0 = clock()
    i = 0
    while i < len(times) or service.busy():
        now = clock() - t0
        while i < len(times) and times[i] <= now:
            snap, N, cell = pool[picks[i]]
            r = Request(i, snap, int(N), float(cell), arrival=t0 + float(times[i]))
            service.submit(r)
            reqs.append(r)
            i += 1
        service.poll()
        if now > duration + drain_limit:
            raise RuntimeError(f"requests not drained {drain_limit} s after the last arrival")
    assert len(service.completed) - done_before == len(reqs)
    lat = np.array([r.latency for r in reqs]) * 1e3
    wait = np.array([r.launched - r.arrival for r in reqs]) * 1e3
    end = max((r.done for r in reqs), default=t0)
    span = end - t0
    samples = sum(int(r.snapshot.num_samples) for r in reqs)
    out = {
        "offered_rate": rate, "duration_s": duration, "requests": reqs, "count": len(reqs),
        "snapshots_per_s": len(reqs) / span if span > 0 else 0.0,
        "mvis_per_s": samples / span / 1e6 if span > 0 else 0.0,
        "latency_ms": {"p50": _pct(lat, 50), "p99": _pct(lat, 99), "mean": _mean(lat),
                       "max": float(lat.max()) if len(lat) else math.nan},
        "queue_ms": {"p50": _pct(wait, 50), "p99": _pct(wait, 99)},
        "mean_batch": float(np.mean(service.batch_sizes)) if service.batch_sizes else 0.0,
    }
    return out


def _pct(a: np.ndarray, q: float) -> float:
    return float(np.percentile(a, q)) if len(a) else math.nan


def _mean(a: np.ndarray) -> float:
    return float(a.mean()) if len(a) else math.nan

"""Ragged continuous-batching serving loop for GFN1 energy + forces.

Requests (one molecule each, arbitrary size) arrive over time.  The server

* sorts waiting requests into size classes by orbital count, so that one
  launch holds molecules of similar cost (a launch finishes when its largest
  molecule does, and the SCF kernel's LDS plan depends on the largest n);
* admits a micro-batch from the class whose oldest request has waited
  longest, as soon as a stream slot is free and the device-memory footprint
  of everything in flight stays under the HBM budget;
* launches the whole pipeline (integrals, persistent per-molecule SCF,
  analytic forces) for the micro-batch on its own HIP stream, so launches
  overlap on the device: one workgroup per molecule, and workgroups of a
  newer launch fill compute units that molecules of an older launch have
  released by converging.  One stream per hardware queue but the default
  stream's (GPU_MAX_HW_QUEUES - 1, three with HIP's default of four): two
  streams on one queue run their work one after the other;
* keeps each molecule's SCF state (potential, Anderson history, orbitals,
  density) resident in that launch's device buffers until its results are
  copied out, then frees the buffers and backfills with waiting requests;
* serves a request above the largest size class (1024 orbitals, the
  per-molecule kernel's limit) on the whole-GPU path, one at a time, from a
  worker thread on one of the streams, while the loop goes on.

Every request's result is computed by exactly the same kernels whatever its
neighbours are; ``tests/test_gpu_serving.py`` checks alone == in the mix.
"""

from __future__ import annotations

import functools
import os
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .engine import BatchLayout, DeviceTables, HostArena, Pipeline, ScfOptions
from .param import SpeciesTables, species_shells

DEFAULT_EDGES = (48, 96, 160, 256, 400, 1024)
DEFAULT_MAXB = (512, 512, 384, 256, 192, 128)
DEFAULT_CLASS_TIME = (2.0e-3, 4.7e-3, 12e-3, 22e-3, 57e-3, 143e-3)  # seconds per molecule


@dataclass
class Request:
    rid: int
    numbers: np.ndarray
    positions: np.ndarray  # bohr
    arrival: float = 0.0
    nao: int = 0
    energy: float | None = None
    forces: np.ndarray | None = None
    done: float | None = None
    status: int = 0
    iterations: int = 0

    @property
    def failed(self) -> bool:
        """The result is unusable: the overlap matrix was not positive
        definite (no SCF ran), the Fermi level or the final eigensolve did
        not converge, or non-finite values appeared.  Energy and forces are
        then NaN.  An SCF that stopped at maxiter still returns its last
        result (status bit MAXITER), as upstream warns and returns."""
        bad = _lib.MOL_OVERLAP | _lib.MOL_FERMI | _lib.MOL_NONFINITE | _lib.MOL_JACOBI
        return bool(self.status & bad)


@dataclass
class _Flight:
    pipe: Pipeline
    reqs: list
    stream: torch.cuda.Stream
    event: torch.cuda.Event
    bytes: int
    t_launch: float
    klass: int  # size class of the launch's requests


@dataclass
class ServerStats:
    completed: int = 0
    launches: int = 0
    admission_stalls: int = 0
    peak_bytes: int = 0
    latencies: list = field(default_factory=list)
    scf_busy_s: float = 0.0    # sum over molecules of the SCF kernel's per-molecule device time
    host_admit_s: float = 0.0  # host time spent building and launching micro-batches
    host_retire_s: float = 0.0  # host time spent copying results out
    spans: list = field(default_factory=list)  # (start_us, duration_us) per molecule, when timing
    large_runs: int = 0  # requests above the largest size class, served on the whole GPU
    launch_log: list = field(default_factory=list)  # per launch, when timing (see _retire)


def default_streams() -> int:
    """One stream per hardware queue of the process but the default
    stream's: GPU_MAX_HW_QUEUES (HIP's default 4) minus one, at most 15.
    The runtime reserves scratch memory per queue for the SCF kernel's
    stack (~2 KB per lane); with 24 queues its launches failed with
    HSA_STATUS_ERROR_OUT_OF_RESOURCES."""
    return min(15, max(1, int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) - 1))


_STREAMS: dict = {}


def _process_streams(device: torch.device, count: int) -> list:
    """The process's serving streams on `device`, created once and shared by
    every server.  Each new HIP stream beyond the hardware queues shares a
    queue with an earlier one, and a second server on fresh streams
    measured 300-500 ms p50 latency where the first one's streams gave
    ~45 ms (same load)."""
    key = device.index if device.index is not None else torch.cuda.current_device()
    have = _STREAMS.setdefault(key, [])
    while len(have) < count:
        have.append(torch.cuda.Stream(device=device))
    return have[:count]


@functools.lru_cache(maxsize=None)
def _orbital_table() -> np.ndarray:
    """Orbitals per element, indexed by atomic number (-1: unsupported)."""
    from . import constants as C

    tab = np.full(128, -1, dtype=np.int64)
    for z in C.SUPPORTED_ELEMENTS:
        tab[z] = sum(2 * l + 1 for l, *_ in species_shells(int(z)))
    return tab


def orbital_count(numbers: np.ndarray) -> int:
    per = _orbital_table()[np.asarray(numbers, dtype=np.int64)]
    if (per < 0).any():
        bad = int(np.asarray(numbers)[per < 0][0])
        species_shells(bad)  # raises the unsupported-element error
    return int(per.sum())


class RaggedServer:
    def __init__(self, device="cuda:0", opts: ScfOptions | None = None, streams: int | None = None,
                 edges=DEFAULT_EDGES, max_batch=DEFAULT_MAXB, hbm_budget_gb: float = 64.0,
                 elements=(1, 6, 7, 8), forces: bool = True, max_wait_s: float = 0.002,
                 wait_frac: float = 0.5, max_wait_cap_s: float = 0.1, xcd_rotate: bool = True,
                 stage_results: bool = False, timing: bool = False, merge_above: int | None = 1200,
                 merge_cap: int = 2400, merge_largest: bool = True, fused: bool = True):
        self.device = torch.device(device)
        self.opts = opts or ScfOptions()
        self.tables = SpeciesTables(elements)
        self.dtables = DeviceTables(self.tables, self.device)
        if streams is None:
            streams = default_streams()
        self.streams = _process_streams(self.device, streams)
        # one pinned host arena per stream (a stream holds one launch at a time)
        self.arenas = {id(s): HostArena() for s in self.streams}
        self.free_streams = deque(self.streams)
        self.edges = tuple(edges)
        self.max_batch = tuple(max_batch)
        self.queues = [deque() for _ in self.edges]
        self.inflight: list[_Flight] = []
        self.budget = int(hbm_budget_gb * 2**30)
        self.forces = forces
        self.max_wait = max_wait_s
        # a class waits for more requests up to wait_frac of its molecules'
        # mean SCF time (measured on retired launches; capped): a launch
        # holds its stream until its slowest molecule is done, so launches
        # of one or two large molecules every few milliseconds would tie up
        # every stream while most CUs idle
        self.wait_frac = wait_frac
        self.max_wait_cap = max_wait_cap_s
        # prior until launches of the class retire: the median per-molecule
        # SCF time of the default classes on MI355X (energy + forces, 1e-10
        # charge tolerance); other edges ~1.5 ms + 1.1 ns x (1.25 lower edge)^3
        if self.edges == DEFAULT_EDGES:
            self.class_time = list(DEFAULT_CLASS_TIME)
        else:
            lo = (0,) + self.edges[:-1]
            self.class_time = [1.5e-3 + 1.1e-9 * (1.25 * max(e, 20)) ** 3 for e in lo]
        self.xcd_rotate = xcd_rotate
        # stage_results: copy each launch's results to pinned host memory
        # behind its kernels (no device access at retirement).  Off by
        # default: those copies wait in the copy engines' queues behind the
        # launch's kernels and held up other streams' uploads (p50 latency
        # 46-139 ms at 4300/s against 11-19 ms with plain copies when the
        # launch is retired).
        self.stage_results = stage_results
        # saturation: with more than `merge_above` requests waiting, a launch
        # takes `merge_cap` of them whatever their class (the kernel starts
        # the largest molecules first), so that launches are few and long and
        # each one's tail overlaps the next one's start (None: class launches
        # only)
        self.merge_above = merge_above
        self.merge_cap = merge_cap
        # merged launches take the largest waiting molecules first (longest
        # job first: the backlog's longest SCFs start early instead of
        # running alone at its end) rather than the oldest
        self.merge_largest = merge_largest
        self.fused = fused
        self._xcd_next = 0
        self.stats = ServerStats()
        self.results: dict[int, Request] = {}
        self.timing = timing
        self.kernel_log: list[tuple[float, float]] = []  # (ms, algorithmic flops) per SCF launch
        # requests above the largest size class wait here for the whole-GPU
        # lane, which runs one at a time: (request, thread, result box, stream)
        self.large_queue: deque = deque()
        self.large_job = None

    # ------------------------------------------------------------------ intake
    def submit(self, req: Request) -> None:
        if req.nao == 0:
            req.nao = orbital_count(req.numbers)
        k = int(np.searchsorted(self.edges, req.nao))
        if k >= len(self.edges):
            # above the per-molecule kernel's classes: the whole-GPU lane
            self.large_queue.append(req)
            return
        self.queues[k].append(req)

    def submit_front(self, req: Request) -> None:
        """Put a request taken for a launch back at the head of its class."""
        self.queues[int(np.searchsorted(self.edges, req.nao))].appendleft(req)

    def pending(self) -> int:
        return sum(len(q) for q in self.queues) + len(self.large_queue)

    def busy(self) -> bool:
        """Anything waiting, in flight or running in the whole-GPU lane."""
        return bool(self.pending() or self.inflight or self.large_job is not None)

    # ---------------------------------------------------------- whole-GPU lane
    def _large_lane(self) -> list[Request]:
        """Requests above the largest size class run one at a time through
        the whole-GPU path (xtb_scf_large: grid kernels, an SCF loop on the
        host that reads a few values back per step) on a stream taken from
        the free ones, driven from a worker thread so that the loop keeps
        admitting and retiring micro-batches meanwhile (the library call
        releases the interpreter lock).  Returns the request it finished."""
        done = []
        job = self.large_job
        if job is not None and not job[1].is_alive():
            req, th, box, stream = job
            th.join()
            self.large_job = None
            self.free_streams.append(stream)
            if "error" in box:
                raise box["error"]
            rec = box["rec"]
            e = float(rec[_lib.OUT_ESCF] + rec[_lib.OUT_FENERGY] + rec[_lib.OUT_EREP])
            g = box["grad"]
            self._finish(req, e, int(rec[_lib.OUT_STATUS]), int(rec[_lib.OUT_ITER]),
                         None if g is None else -g, time.perf_counter())
            self.stats.completed += 1
            self.stats.large_runs += 1
            done.append(req)
        if self.large_job is None and self.large_queue and self.free_streams:
            req = self.large_queue.popleft()
            stream = self.free_streams.popleft()
            box: dict = {}
            th = threading.Thread(target=self._run_large, args=(req, stream, box), daemon=True)
            self.large_job = (req, th, box, stream)
            th.start()
        return done

    def _run_large(self, req: Request, stream, box: dict) -> None:
        try:
            with torch.cuda.stream(stream):
                lay = BatchLayout([req.numbers], self.tables, padded=False, chrg=np.zeros(1))
                pipe = Pipeline(lay, self.dtables, self.device, large=True)
                pos = torch.as_tensor(np.ascontiguousarray(req.positions), dtype=torch.float64).to(self.device)
                pipe.run(pos, self.opts, forces=self.forces, stream=stream)
                box["rec"] = pipe.records()[0].cpu().numpy()
                box["grad"] = pipe.gradient().cpu().numpy() if self.forces else None
        except BaseException as exc:  # re-raised by the loop's thread
            box["error"] = exc

    def _inflight_bytes(self) -> int:
        return sum(f.bytes for f in self.inflight)

    # --------------------------------------------------------------- admission
    def _pick_class(self, now: float, eager: bool = False) -> int | None:
        best, best_t = None, None
        for k, q in enumerate(self.queues):
            if not q:
                continue
            full = len(q) >= self.max_batch[k]
            waited = now - q[0].arrival >= self._wait_for(k)
            if not (eager or full or waited or not self.inflight):
                continue
            # oldest request first; among requests that arrived at the same
            # instant (a submitted batch) the larger class first, so that the
            # longest launches start earliest
            if best is None or q[0].arrival <= best_t:
                best, best_t = k, q[0].arrival
        return best

    def _wait_for(self, k: int) -> float:
        t = self.class_time[k]
        return min(self.max_wait_cap, max(self.max_wait, self.wait_frac * t))

    def _take_merged(self) -> list[Request]:
        """The oldest merge_cap waiting requests of all classes (ties: the
        larger class first), or with merge_largest the largest classes'
        requests first (oldest first within a class)."""
        take = []
        if self.merge_largest:
            for q in reversed(self.queues):
                while q and len(take) < self.merge_cap:
                    take.append(q.popleft())
            return take
        while len(take) < self.merge_cap:
            best, best_t = None, None
            for k, q in enumerate(self.queues):
                if q and (best is None or q[0].arrival <= best_t):
                    best, best_t = k, q[0].arrival
            if best is None:
                break
            q = self.queues[best]
            # the class's requests of that same arrival time, in one go
            while q and q[0].arrival == best_t and len(take) < self.merge_cap:
                take.append(q.popleft())
        return take

    def _admit(self, now: float, eager: bool = False) -> bool:
        if not self.free_streams:
            return False
        if self.merge_above is not None and sum(len(q) for q in self.queues) > self.merge_above:
            k = -1
            take = self._take_merged()
        else:
            k = self._pick_class(now, eager)
            if k is None:
                return False
            q = self.queues[k]
            take = [q.popleft() for _ in range(min(self.max_batch[k], len(q)))]
        lay = BatchLayout([r.numbers for r in take], self.tables, padded=False,
                          chrg=np.zeros(len(take)))
        nbytes = 8 * lay.work_doubles()
        if self.inflight and self._inflight_bytes() + nbytes > self.budget:
            for r in reversed(take):
                self.submit_front(r)
            self.stats.admission_stalls += 1
            return False
        stream = self.free_streams.popleft()
        th = time.perf_counter()
        arena = self.arenas[id(stream)]
        arena.reserve(HostArena.need(lay))
        with torch.cuda.stream(stream):
            pipe = Pipeline(lay, self.dtables, self.device, arena=arena)
            nat = int(lay.nat.sum())
            hpos = arena.take(24 * nat).view(torch.float64)
            np.concatenate([r.positions for r in take], out=hpos.numpy().reshape(nat, 3))
            pos = hpos.view(nat, 3).to(self.device, non_blocking=True)
            # launches start on the XCD after the previous launch's last one
            # integrals, SCF and forces in one launch: under a busy loop every
            # launch waits for CUs, so a batch's stages as separate kernels
            # held its stream idle between them
            pipe.run(pos, self.opts, forces=self.forces, stream=stream, timing=self.timing,
                     wg_skip=self._xcd_next if self.xcd_rotate else 0, fused=self.fused)
            self._xcd_next = (self._xcd_next + len(take)) % 8
            if self.stage_results:
                pipe.stage_results(stream)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.stats.host_admit_s += time.perf_counter() - th
        self.inflight.append(_Flight(pipe, take, stream, ev, nbytes, th, k))
        self.stats.launches += 1
        self.stats.peak_bytes = max(self.stats.peak_bytes, self._inflight_bytes())
        return True

    # -------------------------------------------------------------- retirement
    def _retire(self, f: _Flight) -> list[Request]:
        f.event.synchronize()
        th = time.perf_counter()
        pipe = f.pipe
        if self.timing:
            ev0, ev1 = pipe.scf_events
            self.kernel_log.append((ev0.elapsed_time(ev1), pipe.scf_flops()))
        if self.stage_results:
            rec = pipe.host_records()
            g = pipe.host_gradient() if self.forces else None
        else:
            with torch.cuda.stream(f.stream):
                rec = pipe.records().cpu().numpy()
                g = pipe.gradient().cpu().numpy() if self.forces else None
        e = rec[:, _lib.OUT_ESCF] + rec[:, _lib.OUT_FENERGY] + rec[:, _lib.OUT_EREP]
        fo = None if g is None else -g
        a0 = pipe.L.atom0
        now = time.perf_counter()
        for b, r in enumerate(f.reqs):
            self._finish(r, float(e[b]), int(rec[b, _lib.OUT_STATUS]), int(rec[b, _lib.OUT_ITER]),
                         None if fo is None else fo[a0[b]:a0[b] + len(r.numbers)], now)
        self.stats.completed += len(f.reqs)
        self.stats.scf_busy_s += float(rec[:, _lib.OUT_T_TOTAL].sum()) * 1e-6
        if f.klass >= 0:
            t = float(rec[:, _lib.OUT_T_TOTAL].mean()) * 1e-6
            self.class_time[f.klass] = 0.8 * self.class_time[f.klass] + 0.2 * t
        if self.timing:
            self.stats.spans += list(zip(rec[:, _lib.OUT_T_START], rec[:, _lib.OUT_T_TOTAL]))
            # host admit time, retire time, molecules, max n, device first start, last end (us)
            self.stats.launch_log.append((f.t_launch, th, len(f.reqs), int(pipe.L.n.max()),
                                          float(rec[:, _lib.OUT_T_START].min()),
                                          float((rec[:, _lib.OUT_T_START] + rec[:, _lib.OUT_T_TOTAL]).max())))
        self.stats.host_retire_s += time.perf_counter() - th
        self.free_streams.append(f.stream)
        return f.reqs

    def _finish(self, r: Request, energy: float, status: int, iterations: int, forces, now: float) -> None:
        r.energy = energy
        r.status = status
        r.iterations = iterations
        r.forces = forces
        if r.failed:
            r.energy = float("nan")
            if r.forces is not None:
                r.forces = np.full_like(r.forces, np.nan)
        r.done = now
        self.results[r.rid] = r
        self.stats.latencies.append(now - r.arrival)

    def poll(self, block: bool = False) -> list[Request]:
        out = self._large_lane()
        keep = []
        for f in self.inflight:
            if f.event.query():
                out += self._retire(f)
            else:
                keep.append(f)
        self.inflight = keep
        if block and not out and self.inflight:
            # wait for whichever launch finishes first
            while not out:
                time.sleep(2e-4)
                for f in self.inflight:
                    if f.event.query():
                        self.inflight.remove(f)
                        out += self._retire(f)
                        break
        return out

    def step(self, now: float | None = None, eager: bool = False) -> list[Request]:
        """Retire finished launches and admit micro-batches while streams
        and memory allow.  A class waits up to max_wait for more requests
        unless it is full, nothing is in flight, or `eager` (a batch that
        arrived whole has nothing more to wait for)."""
        now = time.perf_counter() if now is None else now
        done = self.poll()
        while self._admit(now, eager):
            pass
        return done

    def feed(self, reqs: list[Request], now: float | None = None) -> None:
        """Submit requests that arrive together now and admit what fits."""
        now = time.perf_counter() if now is None else now
        for r in reqs:
            r.arrival = now
            self.submit(r)
        self.step(now, eager=True)

    def _idle_wait(self) -> None:
        """Only the whole-GPU lane is running: wait a little for it."""
        if self.large_job is not None and not self.inflight:
            time.sleep(2e-4)

    def serve_until(self, reqs: list[Request]) -> None:
        """Run the loop (admitting anything waiting, in arrival order) until
        every request of `reqs` has its result."""
        want = [r.rid for r in reqs]
        while any(rid not in self.results for rid in want):
            self.step()
            if not self.free_streams or not self.pending():
                self.poll(block=True)
                self._idle_wait()

    def drain(self) -> None:
        while self.busy():
            self.step()
            if self.inflight and not self.free_streams:
                self.poll(block=True)
            self._idle_wait()

    # ------------------------------------------------------------------ driver
    def run_stream(self, reqs: list[Request], realtime: bool = True) -> dict:
        """Serve `reqs` (arrival offsets in seconds).  realtime=False submits
        everything at t0 (saturation throughput)."""
        t0 = time.perf_counter()
        pend = deque(sorted(reqs, key=lambda r: r.arrival))
        for r in pend:
            r.arrival = t0 + (r.arrival if realtime else 0.0)
        n = len(pend)
        while pend or self.busy():
            now = time.perf_counter()
            while pend and pend[0].arrival <= now:
                self.submit(pend.popleft())
            self.step(now)
            if not pend and not self.free_streams:
                self.poll(block=True)
            elif pend and not self.pending() and not self.inflight:
                time.sleep(max(0.0, min(pend[0].arrival - time.perf_counter(), 0.001)))
            elif not pend:
                self._idle_wait()
        t1 = time.perf_counter()
        lat = np.array([self.results[r.rid].done - self.results[r.rid].arrival for r in reqs])
        return {
            "molecules": n, "seconds": t1 - t0, "molecules_per_s": n / (t1 - t0),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "launches": self.stats.launches, "admission_stalls": self.stats.admission_stalls,
            "peak_gb": self.stats.peak_bytes / 2**30,
        }


def padded_dropin(reqs: list[Request], batch: int, device="cuda:0", opts: ScfOptions | None = None,
                  forces: bool = True) -> dict:
    """Plain padded-batch drop-in: requests in arrival order, `batch` at a
    time, each batch padded to its largest molecule and run to completion
    before the next (upstream's batch_mode=1 semantics)."""
    from . import workloads

    dev = torch.device(device)
    opts = opts or ScfOptions()
    tables = SpeciesTables((1, 6, 7, 8))
    dt = DeviceTables(tables, dev)
    t0 = time.perf_counter()
    lat = []
    for s in range(0, len(reqs), batch):
        chunk = reqs[s:s + batch]
        nums, pos = workloads.pad([(r.numbers, r.positions) for r in chunk])
        mask = nums > 0
        lay = BatchLayout([n[n > 0] for n in nums], tables, padded=True, chrg=np.zeros(len(chunk)))
        pipe = Pipeline(lay, dt, dev)
        p = torch.as_tensor(pos[mask], dtype=torch.float64, device=dev)
        pipe.run(p, opts, forces=forces, fused=True)
        e = pipe.energies().cpu().numpy()
        if forces:
            pipe.gradient().cpu()
        now = time.perf_counter()
        for r, ei in zip(chunk, e):
            r.energy = float(ei)
            lat.append(now - t0)
    t1 = time.perf_counter()
    return {"molecules": len(reqs), "seconds": t1 - t0, "molecules_per_s": len(reqs) / (t1 - t0)}

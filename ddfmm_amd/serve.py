"""Continuous-batching FMM matvec server on one MI355X.

The reference has no serving layer (it evaluates one density per process run,
tt.cpp:142-157); this module is the serving side of the evaluation pass:

* one resident ``Engine`` per geometry: octree, lists, translation operators and work
  buffers stay on the device after the first request (setup once, evaluate many);
* admission by HBM budget: an engine is created only if the resident total stays under
  the budget, evicting idle engines least-recently-used first; a request whose geometry
  cannot fit even alone is rejected with an error and does not disturb others;
* continuous batching: requests for the same geometry that are pending when its slot
  frees up are concatenated as RHS columns (up to the engine's slot capacity) into one
  evaluation; different geometries run on their own HIP streams and overlap on the GPU;
* latency = completion - arrival per request, reported as p50/p99 with the throughput.

Engines are pluggable (``make_engine``) so the scheduler can be exercised without a GPU.
"""
from __future__ import annotations

import collections
import concurrent.futures
import dataclasses
import time

import numpy as np

_COPY_POOL = concurrent.futures.ThreadPoolExecutor(max_workers=8)
_COPY_CHUNK = 1 << 23          # float64 elements (64 MiB) per copy task


def _par_copy(dst: np.ndarray, src: np.ndarray):
    """dst[:] = src for 1-D float64 arrays, in parallel chunks for large copies (numpy
    releases the GIL while copying); returns when the whole copy is done."""
    n = src.size
    if n <= _COPY_CHUNK:
        dst[:] = src
        return
    futs = [_COPY_POOL.submit(np.copyto, dst[i:i + _COPY_CHUNK], src[i:i + _COPY_CHUNK])
            for i in range(0, n, _COPY_CHUNK)]
    for f in futs:
        f.result()


@dataclasses.dataclass
class Request:
    rid: int
    geometry: str
    den: np.ndarray               # (nrhs, N) complex128
    arrival: float                # seconds from the start of the run
    result: np.ndarray | None = None
    error: str | None = None
    t_launch: float | None = None
    t_done: float | None = None
    batch: int | None = None

    @property
    def nrhs(self) -> int:
        return self.den.shape[0]

    @property
    def latency(self) -> float | None:
        return None if self.t_done is None else self.t_done - self.arrival


_TILE_CACHE: dict = {}


class GpuEngine:
    """One geometry resident on the device: a Wave3d context, a HIP stream, double-buffered
    pinned host and device staging (nrhs_slot x N complex each)."""

    def __init__(self, key, pos, wave_opts, ops, nrhs_slot, precision, kernel=0, numC=1, hip=None,
                 autotune=True, graph=True):
        from ._lib import Hip
        from .wave3d import Kernel3d, Mlib3d, Wave3d

        self.key = key
        self.hip = hip or Hip()
        self.N = len(pos)
        self.nrhs_slot = nrhs_slot
        ml = Mlib3d()
        ml.setup({"-mlib3d_ldname": ops[0], "-mlib3d_hdname": ops[1]})
        self.w = Wave3d()
        self.w.posptr, self.w.mlibptr, self.w.kernel = pos, ml, Kernel3d(kernel)
        self.w.geomprtn = np.zeros((numC,) * 3, dtype=np.int32)
        self.w.setup(wave_opts, nrhs_max=nrhs_slot, precision=precision)
        # GEMM tile heights: tuned once per (geometry, slot, precision, kernel) in this
        # process and reapplied when the geometry is admitted again after an eviction
        self.tiles = None
        if autotune:
            ck = (key, nrhs_slot, precision, kernel)
            self.tiles = _TILE_CACHE.get(ck)
            if self.tiles is None:
                self.tiles = self.w.autotune(nrhs_slot, reps=1)["tiles"]
                _TILE_CACHE[ck] = self.tiles
            else:
                self.w.set_tiles(self.tiles)
        if graph:
            self.w.set_graph(True)
        self.stream = self.hip.stream()
        nb = nrhs_slot * self.N * 16
        self.slots = []
        for _ in range(2):
            self.slots.append(dict(hden=self.hip.pinned(nb), hval=self.hip.pinned(nb), dden=self.hip.malloc(nb),
                                   dval=self.hip.malloc(nb), ev=self.hip.event(), busy=None))
        self.bytes = int(self.w.stats()["device_bytes"]) + 4 * nb

    @staticmethod
    def estimate_bytes(stats: dict, N: int, nrhs_slot: int) -> int:
        return int(stats["device_bytes"]) + 4 * nrhs_slot * N * 16

    def free_slot(self):
        for s in self.slots:
            if s["busy"] is None:
                return s
        return None

    def launch(self, slot, den: np.ndarray, token):
        self.launch_parts(slot, [den], token)

    def launch_parts(self, slot, parts, token):
        """Launch one evaluation whose RHS are the rows of the arrays in ``parts``, written
        straight into the pinned staging buffer (no concatenated copy)."""
        nb = sum(p.shape[0] for p in parts)
        nbytes = nb * self.N * 16
        host = np.ctypeslib.as_array((np.ctypeslib.ctypes.c_double * (2 * nb * self.N)).from_address(slot["hden"]))
        c = 0
        for p in parts:
            k = 2 * p.shape[0] * self.N
            _par_copy(host[c:c + k], np.ascontiguousarray(p).view(np.float64).reshape(-1))
            c += k
        self.hip.copy_async(slot["dden"], slot["hden"], nbytes, 1, self.stream)
        self.w.eval_device(slot["dden"], slot["dval"], nb, self.stream)
        self.hip.copy_async(slot["hval"], slot["dval"], nbytes, 2, self.stream)
        self.hip.record(slot["ev"], self.stream)
        slot["busy"] = (token, nb)

    def poll(self):
        """Yield (token, values (nb, N)) for every finished slot."""
        out = []
        for s in self.slots:
            if s["busy"] is not None and self.hip.done(s["ev"]):
                token, nb = s["busy"]
                arr = np.ctypeslib.as_array((np.ctypeslib.ctypes.c_double * (2 * nb * self.N)).from_address(s["hval"]))
                flat = np.empty(arr.size, dtype=np.float64)
                _par_copy(flat, arr)
                vals = flat.view(np.complex128).reshape(nb, self.N)
                s["busy"] = None
                out.append((token, vals))
        return out

    def idle(self) -> bool:
        return all(s["busy"] is None for s in self.slots)

    def close(self):
        self.hip.stream_sync(self.stream)
        for s in self.slots:
            self.hip.pinned_free(s["hden"])
            self.hip.pinned_free(s["hval"])
            self.hip.free(s["dden"])
            self.hip.free(s["dval"])
            self.hip.event_destroy(s["ev"])
        self.hip.stream_destroy(self.stream)
        self.w.close()


class Server:
    """Scheduler: admission by HBM budget, per-geometry continuous batching, latency stats.

    make_engine(key) -> engine with .bytes, .nrhs_slot, .free_slot(), .launch(slot, den, token),
    .poll(), .idle(), .close(); estimate(key) -> bytes the engine would take."""

    def __init__(self, make_engine, estimate, hbm_budget: int, clock=time.perf_counter):
        self.make_engine = make_engine
        self.estimate = estimate
        self.budget = int(hbm_budget)
        self.clock = clock
        self.engines: "collections.OrderedDict[str, object]" = collections.OrderedDict()
        self.pending: dict[str, collections.deque] = collections.defaultdict(collections.deque)
        self.inflight: dict[int, list[Request]] = {}
        self.batches = 0
        self.evictions = 0

    def resident_bytes(self) -> int:
        return sum(e.bytes for e in self.engines.values())

    def _admit_engine(self, key) -> bool:
        if key in self.engines:
            self.engines.move_to_end(key)
            return True
        need = self.estimate(key)
        if need > self.budget:
            raise MemoryError(f"geometry {key} needs {need} bytes, above the HBM budget {self.budget}")
        while self.resident_bytes() + need > self.budget:
            victim = next((k for k, e in self.engines.items() if e.idle() and not self.pending.get(k)), None)
            if victim is None:
                return False           # wait until memory frees up
            self.engines.pop(victim).close()
            self.evictions += 1
        self.engines[key] = self.make_engine(key)
        return True

    def _schedule(self, now):
        for key in list(self.pending):
            q = self.pending[key]
            if not q:
                continue
            try:
                if not self._admit_engine(key):
                    continue
            except MemoryError as e:
                while q:
                    r = q.popleft()
                    r.error = str(e)
                    r.t_done = now
                continue
            eng = self.engines[key]
            slot = eng.free_slot()
            if slot is None:
                continue
            batch, cols = [], 0
            while q and cols + q[0].nrhs <= eng.nrhs_slot:
                r = q.popleft()
                batch.append(r)
                cols += r.nrhs
            if not batch:                 # a single request larger than the slot
                r = q.popleft()
                r.error = f"request needs {r.nrhs} RHS, slot capacity is {eng.nrhs_slot}"
                r.t_done = now
                continue
            token = self.batches
            self.batches += 1
            for r in batch:
                r.t_launch = now
                r.batch = token
            self.inflight[token] = batch
            if hasattr(eng, "launch_parts"):      # writes the requests into its staging directly
                eng.launch_parts(slot, [r.den for r in batch], token)
            else:
                eng.launch(slot, np.concatenate([r.den for r in batch], axis=0), token)

    def _collect(self):
        for eng in self.engines.values():
            for token, vals in eng.poll():
                now = self.clock() - self.t0
                c = 0
                for r in self.inflight.pop(token):
                    r.result = vals[c:c + r.nrhs]
                    c += r.nrhs
                    r.t_done = now

    def run(self, requests: list[Request], idle_sleep: float = 1e-4) -> list[Request]:
        """Serve requests whose arrival times are relative to the start of this call."""
        order = sorted(requests, key=lambda r: r.arrival)
        self.t0 = self.clock()
        i = 0
        while i < len(order) or any(self.pending.values()) or self.inflight:
            now = self.clock() - self.t0
            while i < len(order) and order[i].arrival <= now:
                self.pending[order[i].geometry].append(order[i])
                i += 1
            self._schedule(now)
            self._collect()
            if i < len(order) and not self.inflight and not any(self.pending.values()):
                time.sleep(max(0.0, min(order[i].arrival - (self.clock() - self.t0), 0.01)))
            else:
                time.sleep(idle_sleep)
        return requests

    def close(self):
        for e in self.engines.values():
            e.close()
        self.engines.clear()


def latency_stats(reqs: list[Request]) -> dict:
    lat = np.array([r.latency for r in reqs if r.error is None and r.latency is not None])
    if len(lat) == 0:
        return {"p50_ms": None, "p99_ms": None, "served": 0}
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "served": int(len(lat))}


def poisson_requests(geoms: dict, qps: float, count: int, seed: int, max_rhs: int = 16) -> list[Request]:
    """Poisson arrivals at rate qps over the geometries (name -> N), 1..max_rhs complex-normal
    RHS per request, all seeded."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    names = sorted(geoms)
    for rid in range(count):
        t += rng.exponential(1.0 / qps)
        g = names[rng.integers(len(names))]
        nr = int(rng.integers(1, max_rhs + 1))
        x = rng.standard_normal((2, nr, geoms[g]))
        out.append(Request(rid, g, ((x[0] + 1j * x[1]) / np.sqrt(2.0)), t))
    return out

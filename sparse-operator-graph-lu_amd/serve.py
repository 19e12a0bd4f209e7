"""Serving loop for factor and solve requests on one GPU.

Requests arrive over time (a Poisson stream). Each carries a sparse matrix, or a
reference to an earlier matrix, and 1-16 right-hand sides. The server keeps
factors resident in HBM, so a request on a resident matrix is only a solve. It
admits new factorisations as slots and free HBM allow, and runs several
requests at once: every slot owns a device context with its own HIP stream, and
the C library releases the GIL while it works.

* Continuous batching: a request is admitted when a slot and memory free up, not
  when the previous batch ends. Requests queued for the same resident factor are
  solved together as one multi-right-hand-side solve.
* Admission by free HBM: a new factor is loaded only if the device's free
  memory, less a reserve, covers its pool, task graph and vectors. Otherwise
  least-recently-used idle factors are evicted first, and the request waits if
  that is not enough.
* Plans by sparsity pattern: matrices of one grid shape share a pattern, and the
  planner pool reuses the ordering and op list of an earlier plan of the same
  pattern (the same plan as one made from scratch), so only values are new.
* Every result is the same alone and in the mix: the factor and the solve of a
  request depend only on its own matrix and right-hand sides. The solve path is
  the same for the same number of columns; across batch widths it agrees to
  rounding, within 1e-13 relative.

The stream generator follows the serving workload: 2D 5-point (side U{32..512})
or 3D 7-point (side U{16..64}) variable-coefficient diffusion with log-normal
(sigma = 1) coefficients, 50/50; 1-16 RHS U[-1, 1]; and 30% of requests reuse an
earlier matrix.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import dataclasses
import sys
import threading
import time

import numpy as np

import sogl
import workloads as W

TILE = 64 * 64


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the stream starts
    matrix: tuple             # (dim, side, seed): the matrix key
    nrhs: int
    rhs_seed: int
    b: np.ndarray | None = None
    x: np.ndarray | None = None
    done: float = 0.0
    resident_hit: bool = False
    batched_with: int = 1


def make_stream(n_requests: int, rate_hz: float, seed: int = 3, reuse: float = 0.3,
                max_side_2d: int = 512, max_side_3d: int = 64) -> list[Request]:
    """The serving stream: Poisson arrivals (rate_hz) from counter RNG `seed`."""
    u = lambda k, off: (W.draws(seed, k, off).astype(np.float64) + 0.5) / 2.0 ** 64  # noqa: E731
    gaps = -np.log(u(n_requests, 0)) / rate_hz
    t = np.cumsum(gaps)
    coin, kind, side, nr, pick = (u(n_requests, (j + 1) * n_requests) for j in range(5))
    reqs, mats = [], []
    for i in range(n_requests):
        if mats and coin[i] < reuse:
            mk = mats[int(pick[i] * len(mats))]
        else:
            if kind[i] < 0.5:
                mk = (2, 32 + int(side[i] * (max_side_2d - 32 + 1)), 100 + i)
            else:
                mk = (3, 16 + int(side[i] * (max_side_3d - 16 + 1)), 100 + i)
            mats.append(mk)
        reqs.append(Request(i, float(t[i]), mk, 1 + int(nr[i] * 16), 50_000 + i))
    return reqs


def matrix_problem(key: tuple) -> W.Problem:
    dim, side, seed = key
    return W.diffusion_varcoef(side, dim, seed)


def request_rhs(r: Request, n: int) -> np.ndarray:
    b = W.uniform_pm1(r.rhs_seed, n * r.nrhs).reshape(r.nrhs, n).T.copy()
    return b


@dataclasses.dataclass
class Resident:
    key: tuple
    slot: int
    fid: int
    plan: sogl.Plan
    prob: W.Problem
    bytes: int
    last_use: float = 0.0
    busy: int = 0


class PatternPlans:
    """Plans by sparsity pattern: a new matrix whose pattern was planned before reuses
    that plan's ordering and op list (sogl_plan_create_like: the same plan, in a
    fraction of the time). Least recently used patterns are dropped first once the
    kept plans hold more than max_ops unit ops."""

    def __init__(self, max_ops: int = 60_000_000):
        self.max_ops = max_ops
        self.lock = threading.Lock()
        self.plans: collections.OrderedDict[int, sogl.Plan] = collections.OrderedDict()
        self.reused = 0
        self.planned = 0

    def plan(self, prob: W.Problem) -> sogl.Plan:
        h = sogl.pattern_hash(prob.n, prob.row, prob.col)
        with self.lock:
            like = self.plans.get(h)
            if like is not None:
                self.plans.move_to_end(h)
        pl = sogl.Plan(prob.n, prob.row, prob.col, prob.val, like=like)
        with self.lock:
            self.reused += pl.reused
            self.planned += 1
            if h not in self.plans:
                self.plans[h] = pl
                total = sum(p.info.n_ops for p in self.plans.values())
                while total > self.max_ops and len(self.plans) > 1:
                    _, old = self.plans.popitem(last=False)
                    total -= old.info.n_ops
        return pl


class Server:
    def __init__(self, slots: int = 4, precision: int = sogl.FP64, reserve_gb: float = 8.0,
                 exec_mode: int = sogl.EXEC_PERSISTENT, planners: int = 6, plan_cache_ops: int = 150_000_000):
        self.devs = [sogl.Device(0) for _ in range(slots)]
        for d in self.devs:  # many resident factors share the HBM: pools hold live tiles only
            d.set_pool_reuse(sogl.REUSE_ON)
        # host plans: made ahead by a planner pool as requests arrive, kept (least
        # recently used dropped first, by op count) so an evicted factor is not replanned
        self.planner = cf.ThreadPoolExecutor(max_workers=planners)
        self.plan_futs: collections.OrderedDict[tuple, cf.Future] = collections.OrderedDict()
        self.plan_cache_ops = plan_cache_ops
        self.patterns = PatternPlans()
        self.precision = precision
        self.exec_mode = exec_mode
        self.reserve = int(reserve_gb * 1e9)
        self.lock = threading.Condition()
        self.resident: dict[tuple, Resident] = {}
        self.slot_free = [True] * slots
        self.solve_q: dict[int, collections.deque] = {s: collections.deque() for s in range(slots)}
        self.pending: list[Request] = []          # arrived, not admitted (arrival order)
        self.admitted = 0
        self.evictions = 0
        self.committed = 0                        # bytes promised to loads in flight
        self.loading: set[tuple] = set()          # matrices being factored now
        self.completed = 0
        self.progress = 0                         # print a line every this many requests (0: never)

    def _plan_future(self, key: tuple) -> cf.Future:
        """The plan of a matrix, made by the planner pool (caller holds the lock)."""
        f = self.plan_futs.get(key)
        if f is None:
            f = self.plan_futs[key] = self.planner.submit(self._make_plan, key)
            # drop the oldest plans of matrices that are neither resident nor wanted
            total = sum(v.result()[1].info.n_ops for v in self.plan_futs.values() if v.done())
            for k in list(self.plan_futs):
                if total <= self.plan_cache_ops:
                    break
                v = self.plan_futs[k]
                if k in self.resident or k in self.loading or not v.done() or \
                        any(p.matrix == k for p in self.pending):
                    continue
                total -= v.result()[1].info.n_ops
                del self.plan_futs[k]
        else:
            self.plan_futs.move_to_end(key)
        return f

    def _make_plan(self, key: tuple):
        prob = matrix_problem(key)
        return prob, self.patterns.plan(prob)

    def close(self):
        self.planner.shutdown(wait=True)
        with self.lock:
            for r in list(self.resident.values()):
                self.devs[r.slot].release(r.fid)
            self.resident.clear()
        for d in self.devs:
            d.close()

    # ---- memory
    def _need_bytes(self, plan: sogl.Plan, nrhs: int) -> int:
        el = 8 if self.precision == sogl.FP64 else 4
        info = plan.info
        # with pool slot reuse the pool holds the factor tiles, the diagonal inverses, the
        # solve's two-row blocks (at most one L and one U block per pair of block rows)
        # and the zero tile, and the intermediates live at once: 11% more tiles for a 1024^2
        # 2D plan, 3% for a 64^3 3D plan, 2% for 512^2 and for the 1000^2 plan
        if info.symmetric:  # L, its transposes, L^-1 and theirs
            kept = 2 * info.n_l_tiles + 3 * info.block_rows + 1
        else:
            kept = info.n_l_tiles + info.n_u_tiles + 3 * info.block_rows + 1
        tiles = min(int(kept * 1.12) + 256, info.storage_count + 4 * info.block_rows + 1 + (kept if info.symmetric else 0))
        graph = plan.info.n_ops * 64
        vecs = 3 * plan.N * max(nrhs, 16) * 8 + 4 * plan.N * 8
        return tiles * TILE * el + graph + vecs

    def _free_bytes(self) -> int:
        free, _ = self.devs[0].mem_info()
        return free - self.reserve - self.committed

    def _evict_for(self, need: int) -> bool:
        """Evict idle resident factors, least recently used first, until need fits."""
        while self._free_bytes() < need:
            idle = [r for r in self.resident.values() if r.busy == 0]
            if not idle:
                return False
            v = min(idle, key=lambda r: r.last_use)
            self.devs[v.slot].release(v.fid)
            del self.resident[v.key]
            self.evictions += 1
        return True

    # ---- the run
    def run(self, reqs: list[Request], time_scale: float = 1.0, max_wall_s: float = 900.0) -> dict:
        """Serve the stream in real time (arrivals scaled by time_scale); returns metrics.
        Raises if the stream is not done within max_wall_s."""
        self.t0 = time.perf_counter()
        arrivals = sorted(reqs, key=lambda r: r.arrival)
        workers = [threading.Thread(target=self._worker, args=(s,), daemon=True) for s in range(len(self.devs))]
        self.stop = False
        self.error: BaseException | None = None
        for w in workers:
            w.start()
        for r in arrivals:
            delay = r.arrival * time_scale - (time.perf_counter() - self.t0)
            if delay > 0:
                time.sleep(delay)
            with self.lock:
                r.arrival = r.arrival * time_scale
                self.pending.append(r)
                if r.matrix not in self.resident:
                    self._plan_future(r.matrix)  # plan ahead on the host
                self.lock.notify_all()
        with self.lock:
            while self.error is None and (self.pending or any(self.solve_q.values()) or
                                          not all(self.slot_free)):
                self.lock.wait(0.05)
                if time.perf_counter() - self.t0 > max_wall_s:
                    self.error = TimeoutError(f"serving stream not done after {max_wall_s} s")
            self.stop = True
            self.lock.notify_all()
        for w in workers:
            w.join()
        if self.error is not None:
            raise self.error
        wall = time.perf_counter() - self.t0
        lat = np.array([r.done - r.arrival for r in reqs])
        return dict(requests=len(reqs), rhs=int(sum(r.nrhs for r in reqs)), wall_s=wall,
                    p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                    mean_ms=float(lat.mean() * 1e3), solves_per_s=sum(r.nrhs for r in reqs) / wall,
                    requests_per_s=len(reqs) / wall, resident_hits=sum(r.resident_hit for r in reqs),
                    evictions=self.evictions, max_batch=max(r.batched_with for r in reqs),
                    plans=self.patterns.planned, plans_reused=self.patterns.reused)

    def _take(self, s: int):
        """Next work for slot s (caller holds the lock): a batch of solves on one of its
        resident factors, or the oldest admissible pending request."""
        q = self.solve_q[s]
        if q:
            res, first = q.popleft()
            batch = [first]
            rest = collections.deque()
            while q:  # everything queued for the same factor goes in the same solve
                item = q.popleft()
                (batch.append(item[1]) if item[0] is res else rest.append(item))
            q.extend(rest)
            return ("solve", res, batch)
        for i, r in enumerate(self.pending):
            res = self.resident.get(r.matrix)
            if res is not None:  # resident: route to the owning slot's queue
                del self.pending[i]
                res.busy += 1
                r.resident_hit = True
                if res.slot == s:
                    return ("solve", res, [r])
                self.solve_q[res.slot].append((res, r))
                self.lock.notify_all()
                return None
            if any(p.matrix == r.matrix for p in self.pending[:i]) or \
                    any(k == r.matrix for k in self.loading):
                continue  # an earlier request loads this matrix: wait for it to be resident
            del self.pending[i]
            self.loading.add(r.matrix)
            return ("load", None, [r])
        return None

    def _worker(self, s: int):
        dev = self.devs[s]
        try:
            while True:
                with self.lock:
                    job = None
                    while job is None:
                        if self.stop or self.error is not None:
                            return
                        job = self._take(s)
                        if job is None:
                            self.lock.wait(0.02)
                    self.slot_free[s] = False
                kind, res, batch = job
                if kind == "load":
                    res = self._load(s, dev, batch[0])
                    if res is None:
                        continue
                self._solve(dev, res, batch)
                with self.lock:
                    res.busy -= len(batch)
                    res.last_use = time.perf_counter()
                    self.slot_free[s] = True
                    self.lock.notify_all()
        except BaseException as e:  # noqa: BLE001 - surfaced by run()
            with self.lock:
                self.error = e
                self.lock.notify_all()

    def _load(self, s: int, dev: sogl.Device, r: Request) -> Resident | None:
        with self.lock:
            fut = self._plan_future(r.matrix)
        prob, plan = fut.result()
        need = self._need_bytes(plan, r.nrhs)
        with self.lock:
            if not self._evict_for(need) and not self.resident and self.committed == 0:
                raise sogl.SoglError(3, f"request {r.rid} needs {need / 1e9:.1f} GB, more HBM than the "
                                        "device has free")
            if not self._evict_for(need):
                # not admitted yet: back to the head of the queue (plan kept), and this
                # slot serves its resident factors' solves, which is what frees memory
                self.pending.insert(0, r)
                self.loading.discard(r.matrix)
                self.slot_free[s] = True
                self.lock.notify_all()
                self.lock.wait(0.02)
                return None
            self.committed += need
        try:
            fid = dev.load(plan, self.precision)
            if self.precision != sogl.FP64:
                dev.attach_refine(fid, plan)
            dev.factor(fid, self.exec_mode)
        finally:
            with self.lock:
                self.committed -= need
        res = Resident(r.matrix, s, fid, plan, prob, need, time.perf_counter(), busy=1)
        with self.lock:
            self.resident[r.matrix] = res
            self.loading.discard(r.matrix)
            self.admitted += 1
            self.lock.notify_all()
        return res

    def _solve(self, dev: sogl.Device, res: Resident, batch: list[Request]):
        n = res.prob.n
        for r in batch:
            r.b = request_rhs(r, n)
        B = np.concatenate([r.b for r in batch], axis=1)
        if self.precision == sogl.FP64:
            X = res.plan.unpermute(dev.solve(res.fid, res.plan.permute_rhs(B)))
        else:
            X, _, _, _ = dev.solve_refined(res.fid, B, tol=1e-12, max_iter=30)
        X = X.reshape(n, -1)
        t = time.perf_counter() - self.t0
        c = 0
        for r in batch:
            r.x = X[:, c:c + r.nrhs].copy()
            c += r.nrhs
            r.done = t
            r.batched_with = len(batch)
        with self.lock:
            self.completed += len(batch)
            if self.progress and self.completed // self.progress != (self.completed - len(batch)) // self.progress:
                print(f"[serve] {self.completed} requests done at {t:.1f} s, {len(self.resident)} factors "
                      f"resident, {self.evictions} evicted", file=sys.stderr, flush=True)


def serve(reqs: list[Request], slots: int = 4, precision: int = sogl.FP64, time_scale: float = 1.0,
          exec_mode: int = sogl.EXEC_PERSISTENT, progress: int = 0) -> tuple[dict, list[Request]]:
    srv = Server(slots=slots, precision=precision, exec_mode=exec_mode)
    srv.progress = progress
    try:
        m = srv.run(reqs, time_scale)
    finally:
        srv.close()
    return m, reqs


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=20.0, help="mean arrivals per second")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--fp32", action="store_true", help="FP32 factors + FP64 refinement")
    ap.add_argument("--max-side-2d", type=int, default=512)
    ap.add_argument("--max-side-3d", type=int, default=64)
    args = ap.parse_args()
    reqs = make_stream(args.requests, args.rate, max_side_2d=args.max_side_2d, max_side_3d=args.max_side_3d)
    m, reqs = serve(reqs, args.slots, sogl.FP32 if args.fp32 else sogl.FP64, progress=25)
    worst = 0.0
    for r in reqs:
        prob = matrix_problem(r.matrix)
        worst = max(worst, sogl.oracle_residual(prob.n, prob.row, prob.col, prob.val, r.x, r.b))
    m.update(max_oracle_residual=worst, rate_hz=args.rate, slots=args.slots)
    print(json.dumps(m))


if __name__ == "__main__":
    main()

"""Continuous-batching request loop for simulation jobs on one GPU.

Jobs arrive (Poisson process), wait in a FIFO queue, and are admitted into free slots of a Batch whose
slot capacities are carved out of an HBM budget (size classes of 1024 .. 262144 cells).  All resident
jobs advance together, one engine step each per batch launch; a finished job's slot is refilled from the
queue at the next poll.  State stays in HBM from admission to completion; only the final state is copied
back.  Latency of a job = completion time - arrival time (wall clock).
"""
import time
from dataclasses import dataclass, field

import numpy as np

from . import fvs, ics

# HBM bytes per cell / face / CSR entry of a slot (FP64 state, see hfv_kernels.hpp layout)
BYTES_PER_CELL = 463
BYTES_PER_FACE = 125
BYTES_PER_CONN = 4


@dataclass
class Job:
    jid: int
    arrival: float
    prob: object
    kind: str
    admitted: float = -1.0
    finished: float = -1.0
    slot: int = -1
    steps: int = 0
    cell_updates: int = 0
    error_code: int = 0  # the simulation's hard error (HFV_ERR_*): it stopped at that step, no result
    result: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def served_jobs(count=40, rate=20.0, seed=5, side_min=32, side_max=512, sedov_t_end=0.05, vortex_t_end=1.0,
                precision=fvs.FP32_FLUX, threads=0):
    """The served workload: Poisson arrivals at `rate` jobs/s of 2D jobs, 50/50 Sedov blast and
    isentropic vortex (gamma 1.4), generator counts side^2 with side log-uniform in [side_min, side_max]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for j in range(count):
        t += rng.exponential(1.0 / rate)
        side = int(round(np.exp(rng.uniform(np.log(side_min), np.log(side_max)))))
        sub = int(rng.integers(0, 1 << 30))
        if rng.random() < 0.5:
            p = ics.sedov_2d(side=side, seed=sub, t_end=sedov_t_end, precision=precision, threads=threads, gamma=1.4)
            kind = "sedov"
        else:
            p = ics.isentropic_vortex_2d(side=side, seed=sub, t_end=vortex_t_end, gamma=1.4, precision=precision,
                                         threads=threads)
            kind = "vortex"
        jobs.append(Job(j, t, p, kind))
    return jobs


def slot_plan(hbm_budget_bytes, max_cells=262144 + 4096, classes=(1024, 4096, 16384, 65536, 262144), max_slots=256,
              face_per_cell=3.2, conn_per_cell=6.4, headroom=1.12):
    """Slot capacities: an equal share of the budget per size class."""
    share = hbm_budget_bytes / len(classes)
    per_class = []
    for c in classes:
        cells = int(c * headroom)
        faces = int(cells * face_per_cell)
        conn = int(cells * conn_per_cell)
        b = cells * BYTES_PER_CELL + faces * BYTES_PER_FACE + conn * BYTES_PER_CONN
        # each class gets an equal share of the budget and of the slot count, and at least one slot
        per_class.append([(cells, faces, conn), max(1, min(int(share // b), max_slots // len(classes)))])
    caps = [cap for cap, k in per_class for _ in range(k)]
    return [np.array(x, dtype=np.int64) for x in zip(*caps)]


class Server:
    def __init__(self, hbm_budget_bytes=24 << 30, launches_per_poll=16, **plan_kw):
        cc, fc, kc = slot_plan(hbm_budget_bytes, **plan_kw)
        self.batch = fvs.Batch(cc, fc, kc)
        self.caps = (cc, fc, kc)
        self.free = list(range(len(cc)))
        self.launches_per_poll = launches_per_poll
        self.resident = {}
        self.launches = 0
        self.samples = []  # (time, cell updates of every admitted job so far, jobs queued) at each poll

    def _fits(self, slot, p):
        cc, fc, kc = self.caps
        return p.n <= cc[slot] and p.mesh.n_faces <= fc[slot] and p.mesh.conn.size <= kc[slot]

    def _admit(self, job, now):
        best = None
        for s in self.free:
            if self._fits(s, job.prob) and (best is None or self.caps[0][s] < self.caps[0][best]):
                best = s
        if best is None:
            return False
        self.batch.load(best, job.prob)
        self.free.remove(best)
        job.slot = best
        job.admitted = now
        self.resident[best] = job
        return True

    def serve(self, jobs, fetch=("prim",), clock=time.perf_counter, progress=None, progress_every=5.0):
        """Run the request loop over `jobs` (sorted by arrival); returns the jobs with their timings.
        `progress(text)` is called about every `progress_every` seconds with the loop's counts."""
        for j in jobs:
            if not any(self._fits(s, j.prob) for s in range(len(self.caps[0]))):
                raise ValueError(f"job {j.jid} ({j.prob.n} cells) fits no slot of this server")
        pending = sorted(jobs, key=lambda j: j.arrival)
        queue = []
        t0 = clock()
        done = []
        pi = 0
        last = 0.0
        while len(done) < len(jobs):
            now = clock() - t0
            if progress and now - last >= progress_every:
                last = now
                progress(f"t={now:.1f}s done={len(done)} resident={len(self.resident)} queued={len(queue)} "
                         f"launches={self.launches}")
            while pi < len(pending) and pending[pi].arrival <= now:
                queue.append(pending[pi])
                pi += 1
            while queue and self._admit(queue[0], now):
                queue.pop(0)
            if not self.resident:
                if pi < len(pending):
                    time.sleep(max(0.0, pending[pi].arrival - (clock() - t0)))
                continue
            self.launches += self.batch.run(self.launches_per_poll)
            now = clock() - t0
            updates = sum(j.cell_updates for j in done)
            for s, job in list(self.resident.items()):
                info = self.batch.engine_info(s)
                updates += info["cell_updates"]
                if info["done"]:
                    job.finished = clock() - t0
                    job.steps = info["step_count"]
                    job.cell_updates = info["cell_updates"]
                    job.error_code = info["error_code"]
                    if fetch and not job.error_code:
                        job.result = self.batch.get_state(s, list(fetch))
                    self.batch.unload(s)
                    self.free.append(s)
                    del self.resident[s]
                    done.append(job)
            self.samples.append((now, updates, len(queue)))
        self.wall = clock() - t0
        return done

    def saturated_rate(self):
        """Cell updates per second while jobs waited in the queue (every slot they fit was busy): the
        server's throughput at saturation, without the drain at the end of a finite stream; None when the
        queue never held a job at two polls."""
        busy = [(t, u) for t, u, q in self.samples if q > 0]
        if len(busy) < 2 or busy[-1][0] <= busy[0][0]:
            return None
        return (busy[-1][1] - busy[0][1]) / (busy[-1][0] - busy[0][0])


def latency_summary(jobs, wall):
    """Latency percentiles and throughput over the jobs that completed; jobs that stopped on a hard
    error are counted apart (failed_jobs) and contribute neither latency nor cell updates."""
    ok = [j for j in jobs if not j.error_code]
    lat = np.array([j.latency for j in ok]) if ok else np.array([np.nan])
    cu = sum(j.cell_updates for j in ok)
    return {"jobs": len(ok), "failed_jobs": len(jobs) - len(ok),
            "failed": [(j.jid, j.error_code) for j in jobs if j.error_code],
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "wall_s": wall, "cell_updates": int(cu),
            "cell_updates_per_s": cu / wall if wall > 0 else 0.0}

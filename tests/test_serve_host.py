"""Host logic of the serving loop (CPU): latency percentiles and throughput exclude failed jobs, and the
saturated rate counts only the polls at which jobs waited in the queue."""
from types import SimpleNamespace

import numpy as np

from hydro_rs_amd import serve


def _job(jid, arrival, finished, updates, err=0):
    j = serve.Job(jid, arrival, None, "sedov")
    j.finished, j.cell_updates, j.error_code = finished, updates, err
    return j


def test_latency_summary_counts_failed_jobs_apart():
    jobs = [_job(0, 0.0, 1.0, 100), _job(1, 0.5, 2.5, 300), _job(2, 1.0, 1.5, 50, err=-4)]
    s = serve.latency_summary(jobs, wall=4.0)
    assert s["jobs"] == 2 and s["failed_jobs"] == 1 and s["failed"] == [(2, -4)]
    assert s["cell_updates"] == 400 and s["cell_updates_per_s"] == 100.0
    assert s["p50_latency_s"] == np.percentile([1.0, 2.0], 50)


def test_saturated_rate_uses_queued_polls_only():
    samples = [(0.0, 0, 0), (1.0, 100, 3), (2.0, 300, 2), (3.0, 500, 1), (4.0, 550, 0), (9.0, 600, 0)]
    rate = serve.Server.saturated_rate(SimpleNamespace(samples=samples))
    assert rate == (500 - 100) / (3.0 - 1.0)
    assert serve.Server.saturated_rate(SimpleNamespace(samples=[(0.0, 0, 0), (1.0, 5, 1)])) is None

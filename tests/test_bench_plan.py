"""bench.py's launch arithmetic and host facts (no GPU): the timed region covers exactly the
requested KMC steps for every (steps, chunk) the driver or a user may pass."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("steps,chunk", [(20, 5000), (5, 5000), (1, 1), (100000, 5000),
                                         (12345, 5000), (5000, 5000), (5001, 5000), (7, 3)])
def test_launch_plan_covers_exactly_the_steps(bench, steps, chunk):
    plan = bench.launch_plan(steps, chunk)
    assert sum(c * k for c, k in plan) == steps
    assert all(c > 0 and k > 0 for c, k in plan)
    assert all(c <= min(steps, chunk) for c, _ in plan)
    assert len(plan) <= 2
    # the first entry is the graph part: full launches of the clamped chunk
    assert plan[0][0] == min(steps, chunk)


def test_launch_plan_rejects_empty(bench):
    for steps, chunk in [(0, 10), (10, 0), (-1, 5)]:
        with pytest.raises(ValueError):
            bench.launch_plan(steps, chunk)


def test_driver_command_plan(bench):
    """The driver's `--steps 20 --warmup 5` with the default chunk: one 20-step launch timed,
    one 5-step warm-up launch."""
    assert bench.launch_plan(20, 5000) == [(20, 1)]
    assert bench.launch_plan(5, 5000) == [(5, 1)]


def test_host_cores_reports_its_sources(bench, monkeypatch):
    aff = len(os.sched_getaffinity(0))
    monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
    cores, host = bench.host_cores()
    assert cores == aff and host["affinity_cpus"] == aff and host["nproc"] == os.cpu_count()
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    cores, host = bench.host_cores()
    assert cores == 1 and host["omp_num_threads"] == "1"


def test_graph_only_for_repeated_launches(bench):
    assert not bench.uses_graph(bench.launch_plan(20, 5000))
    assert bench.uses_graph(bench.launch_plan(100000, 5000))
    assert not bench.uses_graph(bench.launch_plan(100000, 5000), eager=True)
    assert not bench.uses_graph(bench.launch_plan(7000, 5000))  # one full launch + remainder

"""The uneven batch's memory-aware packing (UnevenBatch.Packer and
footprint_bytes) on the CPU: simulated workers with random run times and
held-memory outcomes."""
import random
import threading
import time

import pytest

from kalnajs_log_spirals_amd.UnevenBatch import Packer, config_for, draw_batch, footprint_bytes

GB = 1 << 30


def simulate(sizes, budget, workers, held_of=None, seed=0):
    """Run the packer with `workers` threads; returns (order started, max in
    use, per-item worker).  held_of(i) is what a context holds after item i
    (default: its resident estimate)."""
    rng = random.Random(seed)
    p = Packer(sizes, budget, workers)
    lock = threading.Lock()
    state = {"use": 0, "max": 0, "order": [], "who": {}, "use_at_start": {}}
    held = [0] * workers
    durations = [rng.uniform(0.001, 0.01) for _ in sizes]

    def work(w):
        while True:
            i = p.next(w)
            if i is None:
                return
            if i == Packer.RELEASE:
                with lock:
                    state["use"] -= held[w]
                    held[w] = 0
                p.released(w)
                continue
            r, t = sizes[i]
            with lock:
                state["use_at_start"][i] = state["use"]
                grow = max(held[w], r) - held[w]
                state["use"] += grow + t
                held[w] += grow
                state["max"] = max(state["max"], state["use"])
                state["order"].append(i)
                state["who"][i] = w
            time.sleep(durations[i])
            h = held_of(i) if held_of else r
            with lock:
                state["use"] -= t
                state["use"] += max(h, held[w]) - held[w]
                held[w] = max(h, held[w])
            p.done(w, held[w])

    ts = [threading.Thread(target=work, args=(w,)) for w in range(workers)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=30)
        assert not t.is_alive(), "packer deadlocked"
    simulate.use_at_start = state["use_at_start"]
    return state["order"], state["max"], state["who"]


def test_every_draw_once_largest_first_within_budget():
    rng = random.Random(1)
    sizes = [(rng.randint(1, 120) * GB, rng.randint(0, 8) * GB) for _ in range(60)]
    budget = 245 * GB
    order, peak, _ = simulate(sizes, budget, 4)
    assert sorted(order) == list(range(len(sizes)))
    assert peak <= budget
    # the first item is the largest
    assert sum(sizes[order[0]]) == max(sum(s) for s in sizes)


def test_a_context_holding_memory_is_released_when_nothing_fits():
    # two workers: after the 200 GB item the 150 GB items need the worker
    # holding 200 GB to let go
    sizes = [(200 * GB, 0), (150 * GB, 0), (150 * GB, 0), (10 * GB, 0)]
    order, peak, who = simulate(sizes, 250 * GB, 2)
    assert sorted(order) == [0, 1, 2, 3]
    assert peak <= 250 * GB


def test_an_item_above_the_budget_runs_alone():
    """It starts once every other context is released (nothing held) and
    after the items that fit."""
    sizes = [(300 * GB, 5 * GB), (20 * GB, 0), (20 * GB, 0)]
    order, peak, _ = simulate(sizes, 250 * GB, 3)
    assert sorted(order) == [0, 1, 2]
    assert order[-1] == 0
    assert simulate.use_at_start[0] == 0


def test_one_worker_is_the_sequential_order():
    sizes = [(s * GB, 0) for s in (5, 50, 20, 80, 1)]
    order, _, _ = simulate(sizes, 250 * GB, 1)
    assert order == [3, 1, 2, 0, 4]


def test_footprint_estimate_of_the_default_sized_draws():
    """The largest draw of the 48 (NR, Ne, N_alpha, l_max) holds W (104 GB),
    its real basis and bf16 planes: ~180 GB of the estimate's 202 GB; every
    draw fits 85% of a 288 GB device on its own."""
    fps = [footprint_bytes(config_for(d), "mixed") for d in draw_batch(48, 0)]
    big = max(fps, key=sum)
    assert 1.8e11 < big[0] < 2.2e11
    assert all(sum(f) < 0.85 * 288e9 for f in fps)

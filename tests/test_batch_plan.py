"""The batched schedules' task queue (vct_batch_plan, host only): the
invariants the persistent queue's waits rely on, checked on the 64-scene
workload's shapes, and a model of the queue showing that every task runs
whatever the number of resident workgroups and the order in which runnable
ones finish. A worker takes the next task only while it runs and waits only on
counters of its scene's previous phase, so the queue drains if and only if
every dependency's tasks come earlier in the queue and each counter's target
equals the number of tasks that advance it; the model fails on a planted
off-by-one in a target."""
import ctypes as C
import random

import numpy as np
import pytest

from vk_voxel_cone_tracing_amd import _abi as A
from vk_voxel_cone_tracing_amd import scenes

VERTEX, RASTER, MIP, TRACE = 0, 1, 2, 3


def _plan(shapes, chunk=256, brick=8):
    arr = (A.BatchSceneShape * len(shapes))(*[A.BatchSceneShape(t, r, w, h) for t, r, w, h in shapes])
    tun = A.BatchTuning(chunk, brick, 0, 0)
    n = C.c_size_t()
    A.call("vct_batch_plan", arr, len(shapes), C.byref(tun), None, 0, C.byref(n), None)
    tasks = np.zeros((n.value, 4), np.uint32)
    need = np.zeros((len(shapes), A.BATCH_PHASES), np.uint32)
    A.call("vct_batch_plan", arr, len(shapes), C.byref(tun), tasks.ctypes.data, n.value, C.byref(n), need.ctypes.data)
    return tasks, need


def _phase(task):
    kind, level = task[0] >> 24, task[3]
    return {VERTEX: 0, RASTER: 1, MIP: 1 + level}.get(kind)  # a tile advances no counter


def _dep(task, levels):
    kind, level = task[0] >> 24, task[3]
    return {VERTEX: None, RASTER: 0, MIP: level, TRACE: levels}[kind]


def _levels(res):
    return int(np.log2(res)) + 1


SHAPES = [(t, r, 512, 512) for t, r in scenes.batch_sizes(3, 64)]


@pytest.mark.parametrize("chunk,brick", [(256, 8), (64, 4), (128, 8)])
def test_plan_covers_the_work_and_counts_every_phase(chunk, brick):
    tasks, need = _plan(SHAPES, chunk, brick)
    scene = tasks[:, 0] & 0xFFFFFF
    kind = tasks[:, 0] >> 24
    for s, (tris, res, w, h) in enumerate(SHAPES):
        mine = tasks[scene == s]
        for k, step in ((VERTEX, 256), (RASTER, chunk)):
            rng = sorted((int(a), int(b)) for _, a, b, _ in mine[mine[:, 0] >> 24 == k])
            assert rng[0][0] == 0 and rng[-1][1] == tris
            assert all(b0 == a1 for (_, b0), (a1, _) in zip(rng, rng[1:]))
            assert all(b - a <= step for a, b in rng)
            assert need[s, 0 if k == VERTEX else 1] == len(rng)
        L = _levels(res)
        for m in range(1, L):
            n = res >> m
            be = min(n, brick)
            bricks = sorted(int(a) for _, a, _, lv in mine[(mine[:, 0] >> 24 == MIP)] if lv == m)
            assert bricks == list(range(((n + be - 1) // be) ** 3))
            assert need[s, 1 + m] == len(bricks)
        assert all(need[s, p] == 0 for p in range(1 + L, A.BATCH_PHASES))
        tiles = sorted(int(a) for _, a, _, _ in mine[mine[:, 0] >> 24 == TRACE])
        assert tiles == list(range(((w + 15) // 16) * ((h + 15) // 16)))
    assert set(np.unique(kind)) == {VERTEX, RASTER, MIP, TRACE}


def test_every_dependency_precedes_its_waiters():
    tasks, _ = _plan(SHAPES)
    last = {}
    for i, t in enumerate(tasks):
        p = _phase(t)
        if p is not None:
            last[(int(t[0] & 0xFFFFFF), p)] = i
    for i, t in enumerate(tasks):
        s = int(t[0] & 0xFFFFFF)
        d = _dep(t, _levels(SHAPES[s][1]))
        if d is not None and (s, d) in last:
            assert last[(s, d)] < i


def _drain(tasks, need, levels, workers, seed):
    """Workers pull tasks in queue order; a worker whose dependency counter is
    below its target waits. Each round, every runnable worker finishes, in a
    random order (any interleaving of independent finishes is possible on
    the device). Returns the number of tasks finished before the queue
    drained or stalled."""
    rnd = random.Random(seed)
    done = np.zeros_like(need)
    head, finished = 0, 0
    held = []
    while True:
        while len(held) < workers and head < len(tasks):
            held.append(head)
            head += 1
        runnable, waiting = [], []
        for ti in held:
            t = tasks[ti]
            s = int(t[0] & 0xFFFFFF)
            d = _dep(t, levels[s])
            (runnable if d is None or done[s, d] >= need[s, d] else waiting).append(ti)
        if not runnable:
            return finished
        rnd.shuffle(runnable)
        for ti in runnable:
            p = _phase(tasks[ti])
            if p is not None:
                done[int(tasks[ti][0] & 0xFFFFFF), p] += 1
            finished += 1
        held = waiting


@pytest.mark.parametrize("workers", [1, 3, 64, 768])
def test_queue_drains_for_any_number_of_resident_workgroups(workers):
    shapes = [(t, r, 64, 48) for t, r, _, _ in SHAPES[:16]]  # few tiles: the model runs in Python
    tasks, need = _plan(shapes)
    levels = [_levels(r) for _, r, _, _ in shapes]
    assert _drain(tasks, need, levels, workers, seed=workers) == len(tasks)


def test_model_detects_an_unreachable_target():
    shapes = [(t, r, 64, 48) for t, r, _, _ in SHAPES[:4]]
    tasks, need = _plan(shapes)
    levels = [_levels(r) for _, r, _, _ in shapes]
    bad = need.copy()
    bad[2, 1] += 1  # one raster task more than the queue holds
    assert _drain(tasks, bad, levels, 64, seed=1) < len(tasks)


def test_plan_rejects_bad_shapes():
    arr = (A.BatchSceneShape * 1)(A.BatchSceneShape(10, 100, 8, 8))
    n = C.c_size_t()
    assert A.lib().vct_batch_plan(arr, 1, C.byref(A.BatchTuning(256, 8, 0, 0)), None, 0, C.byref(n), None) == 1
    arr = (A.BatchSceneShape * 1)(A.BatchSceneShape(10, 64, 8, 8))
    assert A.lib().vct_batch_plan(arr, 1, C.byref(A.BatchTuning(100, 8, 0, 0)), None, 0, C.byref(n), None) == 1

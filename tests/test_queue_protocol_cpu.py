"""The multi-clip queue kernel's work protocol (k_sad_queue in
vid.stab_amd/csrc/hip/vsg_search.hip), simulated on the CPU.

Workgroups are interleaved generators that stop at every atomic step, run
under random and adversarial schedules: some start only after others have
finished (workgroups the GPU cannot keep resident), any may be delayed
between any two atomics.  Each workgroup does what the kernel does:

  loop: first = atomicAdd(queue[0], chunk); stop when first >= units;
        for each unit of [first, first + chunk): atomicMin(best[job], key),
        before = atomicAdd(done[job], 1); the unit that makes done reach the
        job's tile count reads best, resets best / done and refines the job
        (one-tile jobs refine at once, with no atomics)
  count out: if atomicAdd(queue[1], 1) == members - 1: queue[0] = queue[1] = 0

Several launches run in a row on the same counters.  Every launch must end,
every unit run exactly once, every job be refined exactly once with the
minimum key of its units, and every counter be back at zero.  A variant
with the reset moved to the last unit of the queue (planted on the CPU, in
the simulation only) must be caught.
"""
import random

import pytest

NEVER = 2**64 - 1


def make_work(rng, njobs):
    tiles = [rng.choice([1, 1, 2, 3, 4]) for _ in range(njobs)]
    units = [(j, t) for j in range(njobs) for t in range(tiles[j])]
    keys = {u: rng.getrandbits(40) for u in units}
    return tiles, units, keys


class Mem:
    def __init__(self, njobs):
        self.queue = [0, 0]
        self.best = [NEVER] * njobs
        self.done = [0] * njobs


def workgroup(mem, b, nb, units, tiles, keys, chunk, log, reset_by_last_unit=False):
    total = len(units)
    while True:
        first = mem.queue[0]
        mem.queue[0] += chunk            # atomicAdd
        yield
        if first >= total:
            break
        for u in range(first, min(first + chunk, total)):
            job, tile = units[u]
            log["ran"].append(u)
            if tiles[job] == 1:
                log["refined"].append((job, keys[(job, tile)]))
            else:
                mem.best[job] = min(mem.best[job], keys[(job, tile)])   # atomicMin
                yield
                before = mem.done[job]
                mem.done[job] += 1       # atomicAdd
                yield
                if before == tiles[job] - 1:
                    key = mem.best[job]
                    mem.best[job] = NEVER
                    mem.done[job] = 0
                    log["refined"].append((job, key))
            if reset_by_last_unit and u == total - 1:
                mem.queue[0] = 0         # the planted mistake: re-opens the queue
                yield
    members = nb
    before = mem.queue[1]
    mem.queue[1] += 1                    # atomicAdd
    yield
    if before == members - 1:
        mem.queue[0] = 0
        mem.queue[1] = 0


def run_launch(mem, rng, nb, resident, units, tiles, keys, chunk, schedule, reset_by_last_unit=False,
               step_limit=200000):
    log = {"ran": [], "refined": []}
    waiting = list(range(nb))
    if schedule == "reversed":
        waiting.reverse()
    running = {}
    steps = 0
    while waiting or running:
        while waiting and len(running) < resident:
            b = waiting.pop(0)
            running[b] = workgroup(mem, b, nb, units, tiles, keys, chunk, log, reset_by_last_unit)
        if schedule == "round_robin":
            b = min(running)
        else:
            b = rng.choice(sorted(running))
        try:
            next(running[b])
        except StopIteration:
            del running[b]
        steps += 1
        if steps > step_limit:
            return None
    return log


def check_launch(log, units, tiles, keys, mem):
    assert log is not None, "the launch did not end"
    assert sorted(log["ran"]) == list(range(len(units))), "a unit ran twice or never"
    refined = sorted(log["refined"])
    want = sorted((j, min(keys[(j, t)] for t in range(tiles[j]))) for j in range(len(tiles)))
    assert refined == want, "a job was refined twice, never, or with a wrong minimum"
    assert mem.queue == [0, 0] and all(d == 0 for d in mem.done) and all(b == NEVER for b in mem.best)


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("schedule", ["random", "reversed", "round_robin"])
def test_every_launch_ends_and_runs_each_unit_once(seed, schedule):
    rng = random.Random(1000 * seed + len(schedule))
    njobs = rng.randint(1, 40)
    nb = rng.randint(1, 24)
    resident = rng.randint(1, nb)   # fewer resident than launched: late starters
    chunk = rng.randint(1, 5)
    mem = Mem(njobs)
    for launch in range(3):
        tiles, units, keys = make_work(rng, njobs)
        log = run_launch(mem, rng, nb, resident, units, tiles, keys, chunk, schedule)
        check_launch(log, units, tiles, keys, mem)


def test_a_reset_by_the_last_unit_is_caught():
    """the simulation's teeth: resetting the pull counter when the queue's last
    unit is done (instead of when the last workgroup counts out) lets a late
    workgroup pull units again"""
    caught = 0
    for seed in range(40):
        rng = random.Random(seed)
        njobs, nb, resident, chunk = 8, 6, 3, 2
        mem = Mem(njobs)
        tiles, units, keys = make_work(rng, njobs)
        log = run_launch(mem, rng, nb, resident, units, tiles, keys, chunk, "random", reset_by_last_unit=True,
                         step_limit=20000)
        try:
            check_launch(log, units, tiles, keys, mem)
        except AssertionError:
            caught += 1
    assert caught > 0

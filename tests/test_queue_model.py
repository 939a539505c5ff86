"""The work-queue protocol of the multi-pass launch (k_pass_multi in st_step.hip),
modelled on the CPU and run under random schedules.

Per queue, items i = 0 .. npass * nbq - 1 stand for (pass k = i // nbq, brick i % nbq).
Each workgroup claims its first two items, then loops: claim the item after next and
read done[] of the next item's brick, wait for (k - 1, b) unless the value read for this
item says it is published -- publishing its own finished item first -- stage the brick,
step its parcels, publish the item it finished before (every thread has drained by
then), flush; at the end it publishes its last item.  The scheduler picks any workgroup that can move
at each step, with per-workgroup speeds drawn at random, so slow and fast workgroups
interleave in many ways.

Checked: no schedule deadlocks; every item runs exactly once; an item starts only after
the item it depends on was published; done[b] ends at npass for every brick.  A planted
mistake (no publish before a wait) must deadlock in some schedule, so the model can
tell a protocol that waits on itself from one that does not.
"""
import random

import pytest


class Deadlock(Exception):
    pass


def run_queue(nbq, npass, nwg, seed, publish_before_wait=True, read_ahead=True):
    rng = random.Random(seed)
    total = nbq * npass
    ctr = [0]
    done = [0] * nbq
    ran = [0] * total

    def claim():
        v = ctr[0]
        ctr[0] += 1
        return v

    def wg():
        """One workgroup as a generator: yields 'run' for a step of work, or a predicate
        it waits on."""
        item = claim()
        nxt = claim()
        ready = 0
        prev = None  # (brick, value) finished but not published
        while item < total:
            k, b = divmod(item, nbq)
            after = claim()
            ahead = done[nxt % nbq] if (read_ahead and nxt < total) else 0
            yield "run"
            if k > 0 and ready < k:
                if publish_before_wait and prev is not None:
                    done[prev[0]] = prev[1]
                    prev = None
                yield lambda b=b, k=k: done[b] >= k
            assert k == 0 or done[b] >= k, "started before its dependency was published"
            yield "run"  # staging
            yield "run"  # parcel loop
            if prev is not None:  # stepped(): every thread drained
                done[prev[0]] = prev[1]
            yield "run"  # flush
            ran[item] += 1
            prev = (b, k + 1)
            item, nxt, ready = nxt, after, ahead
        yield "run"
        if prev is not None:
            done[prev[0]] = prev[1]

    procs = [wg() for _ in range(nwg)]
    speed = [rng.uniform(0.05, 1.0) for _ in range(nwg)]
    state = [next(p, None) for p in procs]
    while True:
        live = [i for i, s in enumerate(state) if s is not None]
        if not live:
            break
        movable = [i for i in live if state[i] == "run" or state[i]()]
        if not movable:
            raise Deadlock(f"{len(live)} workgroups blocked, done={done}")
        w = rng.choices(movable, weights=[speed[i] for i in movable])[0]
        state[w] = next(procs[w], None)
    return ran, done


@pytest.mark.parametrize("nbq,npass,nwg", [(1, 10, 4), (2, 10, 8), (3, 7, 96), (64, 10, 96), (512, 10, 96),
                                           (5, 3, 1), (40, 10, 13)])
def test_every_schedule_completes(nbq, npass, nwg):
    for seed in range(40 if nbq * npass < 1000 else 4):
        ran, done = run_queue(nbq, npass, nwg, seed)
        assert all(r == 1 for r in ran), seed
        assert all(d == npass for d in done), seed


def test_stale_read_ahead_only_slows():
    # reading nothing ahead (every item takes the waiting path) is still correct
    for seed in range(20):
        ran, done = run_queue(3, 6, 8, seed, read_ahead=False)
        assert all(r == 1 for r in ran) and all(d == 6 for d in done)


def test_planted_mistake_deadlocks():
    # a workgroup that waits while holding its own unpublished brick can wait on itself
    hits = 0
    for seed in range(40):
        try:
            run_queue(1, 4, 3, seed, publish_before_wait=False)
        except Deadlock:
            hits += 1
    assert hits > 0

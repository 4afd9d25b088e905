"""CPU checks of the tile-QR persistent launch's task list (nla_tile_qr_schedule).

The launch's workgroups claim tasks in list order and each task waits on progress counters
(mirrored here from the kernel: PANEL(k, b) on tile i waits for panel_prog[k][b-1] >= i+1, or
for b = 0 on strip_prog[k-1][s] >= i+1 for the strip groups of tile column k; WALK(k, s) on
row i waits for panel_prog[k][last] >= i+1 and strip_prog[k-1][s] >= i+1).  These tests run
the list with P simulated workers under random interleavings and check that

* every run finishes (no deadlock for any worker count), and
* the tile operations happen in an order the flat-tree tile QR allows: every tile gets the
  steps' updates in step order, a panel tile is factored only after all earlier steps updated
  it, its block stages run in order and after the tile above's same stage, and an update of
  step k uses only fully factored panel tiles of column k.

The second property is the algorithm's own data dependence, checked independently of the
wait rules, so a schedule or wait rule that lets a tile be read before it is final fails."""
import random

import pytest

import nextla_hip as nl

IB, TW = 32, 64


def ops_of(task, mt, nt, nb):
    """The kernel's row loop of one task: (waits, publish, op) per row."""
    typ, k, s = task
    spt, nblk = nb // TW, nb // IB
    out = []
    if typ == "P":
        b = s
        for i in range(k, mt):
            if b > 0:
                waits = [("P", k, b - 1, i + 1)]
            else:
                waits = [("S", k - 1, g, i + 1) for g in range(k * spt, (k + 1) * spt)] if k > 0 else []
            out.append((waits, ("P", k, b, i + 1), ("panel", k, i, b)))
    else:
        waits = [("P", k, nblk - 1, k + 1)] + ([("S", k - 1, s, k + 1)] if k > 0 else [])
        out.append((waits, None, ("unmqr", k, s)))
        for i in range(k + 1, mt):
            waits = [("P", k, nblk - 1, i + 1)] + ([("S", k - 1, s, i + 1)] if k > 0 else [])
            out.append((waits, ("S", k, s, i + 1), ("tsmqr", k, i, s)))
    return out


def simulate(tasks, mt, nt, nb, P, seed):
    rnd = random.Random(seed)
    prog = {}
    log = []
    nxt = 0
    workers = [None] * P
    done = 0
    while done < len(tasks):
        moved = False
        order = list(range(P))
        rnd.shuffle(order)
        for w in order:
            if workers[w] is None:
                if nxt < len(tasks):
                    workers[w] = [ops_of(tasks[nxt], mt, nt, nb), 0]
                    nxt += 1
                    moved = True
                continue
            rows, idx = workers[w]
            waits, pub, op = rows[idx]
            if all(prog.get(x[:3], 0) >= x[3] for x in waits):
                log.append(op)
                if pub:
                    prog[pub[:3]] = pub[3]
                workers[w][1] += 1
                moved = True
                if workers[w][1] == len(rows):
                    workers[w] = None
                    done += 1
        if not moved:
            raise AssertionError(f"deadlock: mt={mt} nt={nt} P={P} seed={seed}")
    return log


def check_dependences(log, mt, nt, nb):
    spt, nblk = nb // TW, nb // IB
    steps = {}   # (tile row, strip group) -> steps applied
    blocks = {}  # (k, i) -> block stages done on panel tile (i, k)
    for op in log:
        if op[0] == "panel":
            _, k, i, b = op
            for g in range(k * spt, (k + 1) * spt):
                assert steps.get((i, g), 0) == k, op
            assert blocks.get((k, i), 0) == b, op
            if i > k:
                assert blocks.get((k, i - 1), 0) > b, op
            blocks[(k, i)] = b + 1
            if b == nblk - 1:
                for g in range(k * spt, (k + 1) * spt):
                    steps[(i, g)] = k + 1
        elif op[0] == "unmqr":
            _, k, s = op
            assert steps.get((k, s), 0) == k and blocks.get((k, k), 0) == nblk, op
            steps[(k, s)] = k + 0.5   # row k's strip stays in the walker until the walk ends
        else:
            _, k, i, s = op
            assert steps.get((i, s), 0) == k and blocks.get((k, i), 0) == nblk, op
            steps[(i, s)] = k + 1
    # every tile got every step: a panel tile (i >= j) ends factored by step j, a tile above
    # the diagonal (i < j) ends with steps 0..i-1 applied and row i's own unmqr
    for i in range(mt):
        for s in range(nt * spt):
            j = s // spt
            assert steps.get((i, s), 0) == (j + 1 if i >= j else i + 0.5), (i, s, steps.get((i, s), 0))


@pytest.mark.parametrize("m,n,nb", [(512, 512, 128), (1024, 1024, 128), (1024, 256, 64), (256, 1024, 64),
                                    (2048, 512, 128), (512, 512, 64)])
def test_schedule_is_deadlock_free_and_respects_dependences(m, n, nb):
    tasks = nl.tile_qr_schedule(m, n, nb)
    mt, nt = m // nb, n // nb
    kt, spt, nblk = min(mt, nt), nb // TW, nb // IB
    # every panel stage once, every (k, strip group right of k) walker once
    want = {("P", k, b) for k in range(kt) for b in range(nblk)}
    want |= {("W", k, s) for k in range(kt) for s in range((k + 1) * spt, nt * spt)}
    assert len(tasks) == len(want) and set(tasks) == want
    for P in (1, 2, 3, 7, 64):
        for seed in range(2):
            log = simulate(tasks, mt, nt, nb, P, seed)
            check_dependences(log, mt, nt, nb)


def test_schedule_runs_panels_ahead():
    """Lookahead: at 16384 x 16384 (nb 128) the panel of step 8 is claimed before the walkers
    of step 0 on the far columns, so it need not wait for those to finish."""
    tasks = nl.tile_qr_schedule(16384, 16384, 128)
    pos = {t: i for i, t in enumerate(tasks)}
    assert pos[("P", 8, 0)] < pos[("W", 0, 2 * 100)]
    assert pos[("P", 1, 0)] > pos[("W", 0, 3)]


def test_schedule_rejects_unsupported_shapes():
    with pytest.raises(nl.ArgumentError, match="multiples of nb"):
        nl.tile_qr_schedule(500, 512, 128)
    with pytest.raises(nl.ArgumentError, match="nb must be 64 or 128"):
        nl.tile_qr_schedule(512, 512, 96)


def test_checks_catch_a_bad_order_and_a_missing_wait():
    """The two checks above are not vacuous: a panel claimed before the walkers it waits on
    deadlocks one worker, and a panel stage that skips its wait breaks the data order."""
    m = n = 512
    nb = 128
    mt = nt = m // nb
    tasks = nl.tile_qr_schedule(m, n, nb)
    bad = [t for t in tasks if t != ("P", 1, 0)]
    bad.insert(bad.index(("P", 0, 3)) + 1, ("P", 1, 0))
    with pytest.raises(AssertionError, match="deadlock"):
        simulate(bad, mt, nt, nb, 1, 0)
    global ops_of
    real = ops_of

    def no_stage_wait(task, mt_, nt_, nb_):
        rows = real(task, mt_, nt_, nb_)
        return [([] if task[0] == "P" and task[2] > 0 else w, p, o) for (w, p, o) in rows]

    ops_of = no_stage_wait
    try:
        failed = False
        for seed in range(20):
            try:
                check_dependences(simulate(tasks, mt, nt, nb, 8, seed), mt, nt, nb)
            except AssertionError:
                failed = True
                break
        assert failed
    finally:
        ops_of = real

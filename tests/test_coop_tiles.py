"""Schedule of the cooperative exact step's column tiles, checked on the CPU.

k_step_exact_coop gives work-group g the frame columns [a, e) and runs s
steps between grid barriers on an LDS tile of the columns [a - s, e + s):
sub-step k updates pressures on [a - (s-1-k), e + (s-k)) and velocities on
[a - (s-1-k), e + (s-1-k)), clipped to the interior.  This model replays
that schedule with a time level per tile value and asserts, for every
update, that each input it reads (FDTDEngineCPU.cpp:316-410: a pressure
reads its own P and V, the left Vx and the lower Vy; a velocity reads its
own V, its own, the right and the upper new P) lies inside the tile and is
of the right step; that cells carrying source coefficients (the excitation
column and the one left of it) are only updated by work-groups that run the
source; that a source work-group's excitation-cell reads are current; and
that the own columns partition the interior.  The partition and source rules
restate coop_chunk / coop_columns / coop_owner / coop_runs_source /
coop_source_ok of csrc/vs_common.h.
"""
import random

import pytest


def chunk(W, G):
    return (W - 2 + G - 1) // G


def columns(W, G, g):
    a = 1 + g * chunk(W, G)
    return a, min(a + chunk(W, G), W - 1)


def owner(W, G, c):
    return (min(max(c, 1), W - 2) - 1) // chunk(W, G)


def runs_source(W, G, s, g, ex):
    a, e = columns(W, G, g)
    if a >= e:
        return False
    lo, hi = a - (s - 1), e + (s - 1)
    return g == owner(W, G, ex) or lo <= ex < hi or lo <= ex - 1 < hi


def source_ok(W, G, s, ex):
    if s < 2:
        return True
    for g in range(G):
        if runs_source(W, G, s, g, ex):
            a, e = columns(W, G, g)
            if ex < a - 1 or ex >= e + 2:
                return False
    return True


def replay(H, W, G, s, ex, ey, lx, ly):
    """Replays one period of every length 1..s for every work-group."""
    for g in range(G):
        a, e = columns(W, G, g)
        L, R = max(0, a - s), min(W, e + s)
        src = runs_source(W, G, s, g, ex)
        for sp in range(1, s + 1):
            lvP = {(r, c): 0 for r in range(H) for c in range(L, R)}
            lvV = dict(lvP)

            def ring(r, c):
                return r in (0, H - 1) or c in (0, W - 1)

            def need(lv, r, c, want, what):
                assert L <= c < R and 0 <= r < H, (what, "outside the tile", g, r, c)
                assert ring(r, c) or lv[(r, c)] == want, (what, "stale", g, sp, r, c, lv[(r, c)], want)

            for k in range(sp):
                if src and k >= 1:
                    need(lvP, ey, ex, k, "excitation P")
                pl, pr = max(1, a - (sp - 1 - k)), min(W - 1, e + (sp - k))
                done = []
                for r in range(1, H - 1):
                    for c in range(pl, pr):
                        need(lvP, r, c, k, "P")
                        need(lvV, r, c, k, "V")
                        need(lvV, r, c - 1, k, "left Vx")
                        need(lvV, r + 1, c, k, "lower Vy")
                        done.append((r, c))
                for rc in done:
                    lvP[rc] = k + 1
                if g == owner(W, G, lx):
                    need(lvP, ly, lx, k + 1, "probe")
                vr = min(W - 1, e + (sp - 1 - k))
                done = []
                for r in range(1, H - 1):
                    for c in range(pl, vr):
                        need(lvV, r, c, k, "V")
                        need(lvP, r, c, k + 1, "new P")
                        need(lvP, r, c + 1, k + 1, "right P")
                        need(lvP, r - 1, c, k + 1, "upper P")
                        if c in (ex - 1, ex):
                            assert src, ("source cell updated without the source", g, c)
                        done.append((r, c))
                for rc in done:
                    lvV[rc] = k + 1
            for r in range(1, H - 1):
                for c in range(a, e):
                    need(lvP, r, c, sp, "stored P")
                    need(lvV, r, c, sp, "stored V")
    own = [0] * W
    for g in range(G):
        a, e = columns(W, G, g)
        for c in range(a, e):
            own[c] += 1
    assert all(own[c] == 1 for c in range(1, W - 1))


def plan(W, G, s, ex):
    G = min(G, W - 2)
    G = (W - 2 + chunk(W, G) - 1) // chunk(W, G)   # no empty work-groups
    while s > 1 and not source_ok(W, G, s, ex):
        s -= 1
    return G, s


@pytest.mark.parametrize("G,s", [(2, 1), (11, 2), (16, 4), (42, 8), (24, 1), (48, 16)])
def test_vowel_tube_frame(G, s):
    # the vowel tube's 250-column frame (excitation column 7, listener
    # column 237), with 12 of its 47 rows: rows do not enter the schedule
    G, s = plan(250, G, s, 7)
    replay(12, 250, G, s, 7, 6, 237, 5)


def test_random_frames():
    rng = random.Random(3)
    for _ in range(60):
        H, W = rng.randint(4, 12), rng.randint(6, 60)
        G0, s0 = rng.randint(2, 16), rng.randint(1, 8)
        ex = rng.choice([1, 2, W // 2, W - 2, rng.randint(1, W - 2)])
        G, s = plan(W, G0, s0, ex)
        if G < 2:
            continue
        replay(H, W, G, s, ex, rng.randint(1, H - 2), rng.randint(1, W - 2), rng.randint(1, H - 2))


def test_source_rule_lowers_steps_next_to_a_boundary():
    # a boundary two columns right of the excitation column: a neighbour's
    # halo reaches the source cells, so 4 steps per barrier are not allowed
    W, G = 34, 4              # chunk 8: columns [1, 9), [9, 17), ...
    assert not source_ok(W, G, 4, 7)
    assert source_ok(W, G, 1, 7)
    assert plan(W, G, 4, 7)[1] < 4

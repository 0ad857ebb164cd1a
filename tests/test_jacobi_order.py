"""Pair orderings of the parallel one-sided Jacobi sweeps, replayed on the CPU.

Each schedule below restates the column moves of a kernel's sweep (which
column sits in which register slot of which lane group or pair slot between
rounds) and checks that every pair of the KP columns meets exactly once per
sweep, in KP - 1 rounds:

* the fused batched kernel with two pairs per lane group (KP = 64,
  Wave::jacobi, NP = 2: renames for bit 0, lane-group moves for higher bits);
* the fused batched kernel with one pair per lane group (KP = 32, NP = 1);
* the 256-thread workgroup form for KP = 128 (jacobi_regs128: 64 pair slots,
  moves along slot bits 0-3 inside a wavefront, bits 4-5 through LDS).
"""
import itertools


def _check(rounds, kp):
    seen = {}
    for pairs in rounds:
        cols = [c for p in pairs for c in p]
        assert len(cols) == len(set(cols)) == kp, "a round pairs every column once"
        for a, b in pairs:
            key = (min(a, b), max(a, b))
            seen[key] = seen.get(key, 0) + 1
    assert len(rounds) == kp - 1
    assert set(seen) == set(itertools.combinations(range(kp), 2))
    assert max(seen.values()) == 1


def _slot_schedule(nslot):
    """One pair (top, bottom) per slot; slot s starts with columns 2s, 2s+1."""
    top = [2 * s for s in range(nslot)]
    bot = [2 * s + 1 for s in range(nslot)]
    rounds = []
    g = nslot
    while g >= 1:
        for r in range(g):
            rounds.append(list(zip(top, bot)))
            if r + 1 < g:
                m = (r + 1) & -(r + 1)
                bot = [bot[s ^ m] for s in range(nslot)]
        if g > 1:
            d = g >> 1
            nt, nb = top[:], bot[:]
            for s in range(nslot):
                if s & d == 0:
                    nb[s] = top[s ^ d]
                    nt[s ^ d] = bot[s]
            top, bot = nt, nb
        g >>= 1
    return rounds


def _wave_np2_schedule(kp=64):
    """Two pairs per lane group: group c (kp / 4 of them) holds column slots
    J = 0..3, first column J * (kp / 4) + c; rounds pair (X0, X1), (X2, X3)
    or (X0, X3), (X2, X1).  kp = 64: the fused kernel's 16 lane groups;
    kp = 128: the workgroup form's 32 groups of eight threads."""
    ng = kp // 4
    x = [[j * ng + c for j in range(4)] for c in range(ng)]
    rounds = []

    def rnd(t0, b0, t1, b1):
        rounds.append([(x[c][t0], x[c][b0]) for c in range(ng)]
                      + [(x[c][t1], x[c][b1]) for c in range(ng)])

    def move(j, v):
        old = [row[j] for row in x]
        for c in range(ng):
            x[c][j] = old[c ^ v]

    def xswap(t, b, d):
        old = [row[:] for row in x]
        for c in range(ng):
            if c & d == 0:
                x[c][b] = old[c ^ d][t]
            else:
                x[c][t] = old[c ^ d][b]

    g = kp // 2
    while g >= 4:
        for r in range(0, g, 4):
            rnd(0, 1, 2, 3)
            rnd(0, 3, 2, 1)
            move(1, 1)
            move(3, 1)
            rnd(0, 3, 2, 1)
            rnd(0, 1, 2, 3)
            if r + 4 < g:
                v = ((r + 4) & -(r + 4)) >> 1
                move(1, v)
                move(3, v)
        d = g >> 2
        xswap(0, 1, d)
        xswap(2, 3, d)
        g >>= 1
    rnd(0, 1, 2, 3)
    rnd(0, 3, 2, 1)
    rnd(0, 2, 3, 1)
    return rounds


def test_wave_two_pairs_per_lane_group_kp64():
    _check(_wave_np2_schedule(), 64)


def test_wave_one_pair_per_lane_group_kp32():
    _check(_slot_schedule(16), 32)


def test_workgroup_register_jacobi_kp128():
    _check(_slot_schedule(64), 128)


def test_workgroup_two_pairs_per_group_kp128():
    _check(_wave_np2_schedule(128), 128)


def test_workgroup_cross_wave_moves_are_six_a_sweep():
    """Moves and swaps along slot bits 4 and 5 (16 slots a wavefront) are the
    only ones that need LDS and a barrier."""
    n, g = 0, 64
    while g >= 1:
        n += sum(1 for r in range(g - 1) if ((r + 1) & -(r + 1)) >= 16)
        n += 1 if g > 1 and (g >> 1) >= 16 else 0
        g >>= 1
    assert n == 6

"""The oracle's structural plasticity (ProcStructuralPlasticity restated).

Pins, on small networks whose state is written by hand:
  * the prune test of Prj::prune (Prj.cpp:112-138): the age test
    ti < simstep - t-th, the weight test with its closed bounds (w >= wp,
    w <= wn keep a row; NaN does not), rows with ii < 0 untouched, the state
    reset of remove_conn (Prj.cpp:78-107: pi, ei, ti, pij, eij, zj2 zeroed;
    zi, zi2, wij, tij kept) plus the queue, and the fan-out credit of one per
    removed row (two rows from one source: two);
  * the growth of STAGE 3-5 (ProcStructuralPlasitcity.cpp:354-576) against a
    brute-force checker of its invariants: new rows only in open hypercolumns
    (e-trace sum above 0.5 and a free slot), in their lowest free slots, from
    active units (count above int(0.2 * period * 100 * dt)) not already
    connected to that hypercolumn, no source twice in one hypercolumn, no unit
    more new rows than its fan-out, fan-out debited once per new row,
    counters reset, delay 1; a case with a single possible outcome, checked
    exactly;
  * the draws: words equal Philox4x32-10 of (block, step, lane, 1), shuffles
    are permutations, and the geometric delay follows the distance (d-norm 0
    gives delays 1 or 2 around the mean of 1 ms, a far hypercolumn longer
    ones).
"""

import numpy as np
import pytest

import oracle as O
from gsbn_amd import configs


def _net(H=3, M=2, slots=None, fanout=None, seed=0, **sp):
    spec = configs.recall_net(H, M, 2, 20, 0, 10, 0.5, name="sp")
    if slots is not None:
        spec.projs[0].slot_num = float(slots)
    spec.pops[0].fanout = fanout
    spec.conn = [[]]
    spec.structural = configs.StructuralSpec(seed=seed, **sp)
    return spec, O.OracleNet(spec)


def _set(net, **vals):
    for k, v in vals.items():
        net.write(k, np.asarray(v))


def test_prune_rules_and_state_reset():
    spec, net = _net(H=2, M=2, slots=3, t_th=10, wp=0.5, wn=-0.5)
    # 2 HCUs x 3 slots: rows 0..5; sources are units 0..3
    ii = np.array([0, 1, -1, 2, 3, 0], np.int32)
    ti = np.array([95, 89, 0, 95, 95, 95], np.int32)
    wij = np.array([[0.1, 0.5],       # row 0: w == wp keeps it
                    [9.0, 9.0],       # row 1: too old (89 < 100 - 10)
                    [0.0, 0.0],       # row 2: free slot, untouched
                    [np.nan, 0.2],    # row 3: nothing at or beyond the bounds
                    [-0.5, 0.0],      # row 4: w == wn keeps it
                    [0.49, -0.49]],   # row 5: weak, removed; source 0 again
                   np.float32)
    rng = np.random.default_rng(0)
    other = {n: rng.random(12).astype(np.float32) for n in ("pij", "eij", "zi2", "zj2", "tij")}
    other["tij"] = rng.integers(1, 90, 12).astype(np.int32)
    rows = {n: rng.random(6).astype(np.float32) + 0.1 for n in ("pi", "ei", "zi")}
    _set(net, ii_0=ii, di_0=np.array([1, 1, 0, 2, 1, 3], np.int32), ti_0=ti, wij_0=wij.ravel(),
         qi_0=np.full(6, 1, np.int32), **{".siq_0": np.ones(6, np.int8)},
         **{n + "_0": v for n, v in other.items()}, **{n + "_0": v for n, v in rows.items()})
    fan0 = net.sp.pops[0]["fanout"].copy()
    net.sp.prune(100)
    gone = np.array([False, True, False, True, False, True])
    assert np.array_equal(net.vector("ii_0"), np.where(gone, -1, ii))
    assert np.array_equal(net.vector("di_0"), np.where(gone, 0, [1, 1, 0, 2, 1, 3]))
    assert np.array_equal(net.vector("ti_0"), np.where(gone, 0, ti))
    for n in ("pi", "ei"):
        assert np.array_equal(net.vector(n + "_0"), np.where(gone, 0, rows[n]))
    assert np.array_equal(net.vector("zi_0"), rows["zi"])
    for n in ("qi", ".siq"):
        assert np.array_equal(net.vector(n + "_0"), np.where(gone, 0, 1))
    g2 = np.repeat(gone, 2)
    for n in ("pij", "eij", "zj2"):
        assert np.array_equal(net.vector(n + "_0"), np.where(g2, 0, other[n]))
    for n in ("zi2", "tij"):
        assert np.array_equal(net.vector(n + "_0"), other[n])
    assert np.array_equal(net.vector("wij_0").view(np.uint32), wij.ravel().view(np.uint32))
    want = fan0.copy()
    want[1] += 1  # row 1
    want[2] += 1  # row 3
    want[0] += 1  # row 5
    assert np.array_equal(net.sp.pops[0]["fanout"], want)


def _check_growth(before_ii, after_ii, after_di, H, C, open_h, active, fan_before, fan_after):
    new = (before_ii < 0) & (after_ii >= 0)
    assert np.all(after_ii[before_ii >= 0] == before_ii[before_ii >= 0])
    assert np.all(after_di[new] == 1)
    per_unit = np.zeros_like(fan_before)
    for h in range(H):
        row_b, row_a = before_ii[h * C:(h + 1) * C], after_ii[h * C:(h + 1) * C]
        added = np.nonzero((row_b < 0) & (row_a >= 0))[0]
        if h not in open_h:
            assert added.size == 0
            continue
        free = np.nonzero(row_b < 0)[0]
        assert np.array_equal(added, free[:added.size])  # lowest free slots
        srcs = row_a[row_a >= 0]
        assert len(set(srcs.tolist())) == srcs.size    # no source twice
        for u in row_a[added]:
            assert u in active
            per_unit[u] += 1
    assert np.all(per_unit <= np.maximum(fan_before, 0))
    assert np.array_equal(fan_after, fan_before - per_unit)
    return int(new.sum())


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_growth_invariants(seed):
    H, M, C = 4, 3, 6
    spec, net = _net(H=H, M=M, slots=C, fanout=2, seed=seed, period=20)
    rng = np.random.default_rng(seed)
    ii = np.full(H * C, -1, np.int32)
    for h in range(H):
        k = rng.integers(0, C)
        ii[h * C:h * C + k] = rng.choice(H * M, size=k, replace=False)
    # HCU 1's e-traces sum to 0.5 exactly: not above the threshold, closed
    ej = np.full(H * M, 0.25, np.float32)
    ej[1 * M:2 * M] = [0.25, 0.125, 0.125]
    cnt = rng.integers(0, 3, H * M).astype(np.int32)   # threshold int(0.4) = 0
    _set(net, ii_0=ii, di_0=np.where(ii >= 0, 1, 0).astype(np.int32), ej_0=ej,
         **{".counter_0": cnt})
    fan0 = net.sp.pops[0]["fanout"].copy()
    open_h = [h for h in range(H) if h != 1 and (ii[h * C:(h + 1) * C] < 0).any()]
    net.sp.grow(20, 0, 1, int(0.2 * 20 * 100 * float(np.float32(spec.dt))), np.float32(0.5))
    added = _check_growth(ii, net.vector("ii_0"), net.vector("di_0"), H, C, open_h,
                          set(np.nonzero(cnt > 0)[0].tolist()), fan0, net.sp.pops[0]["fanout"])
    assert added > 0
    assert not net.vector(".counter_0").any()


def test_growth_single_outcome():
    # one open HCU with one free slot, one active unit not yet connected
    H, M, C = 2, 2, 3
    spec, net = _net(H=H, M=M, slots=C, fanout=1, period=20)
    ii = np.array([0, -1, 1, 0, 1, 2], np.int32)
    _set(net, ii_0=ii, di_0=np.where(ii >= 0, 1, 0).astype(np.int32),
         ej_0=np.array([0.3, 0.3, 0.0, 0.0], np.float32),
         **{".counter_0": np.array([5, 0, 0, 9], np.int32)})
    net.sp.grow(20, 0, 1, 0, np.float32(0.5))
    # unit 0 is connected to HCU 0 already; unit 3 is the only candidate
    assert np.array_equal(net.vector("ii_0"), [0, 3, 1, 0, 1, 2])
    assert net.sp.pops[0]["fanout"][3] == 0 and net.sp.pops[0]["fanout"][0] == 1


def test_draw_words_and_shuffle():
    d = O.SpDraws(12345, 77, 1)
    words = [d.word() for _ in range(8)]
    key = np.array([12345, 0], np.uint32)
    want = list(O.philox(np.array([0, 77, 1, 1], np.uint32), key)) + \
        list(O.philox(np.array([1, 77, 1, 1], np.uint32), key))
    assert words == [int(w) for w in want]
    for n in (1, 2, 7, 50):
        v = list(range(n))
        O.SpDraws(3, n, 0).shuffle(v)
        assert sorted(v) == list(range(n))
    b = [O.SpDraws(9, 1, 0).below(5) for _ in range(3)]
    assert all(0 <= x < 5 for x in b)


def test_geometric_delays_follow_distance():
    spec, net = _net(H=4, M=2, slots=8, enable_geometry=1, d_norm=0.0)
    sp = net.sp
    q = sp.prjs[0]
    d = O.SpDraws(1, 1, 1)
    near = [sp._delay(d, q, 0, 0) for _ in range(200)]
    assert set(near) <= {1, 2} and 1 in near and 2 in near
    spec2, net2 = _net(H=4, M=2, slots=8, enable_geometry=1, d_norm=0.75, v_cond=0.2)
    net2.sp.pops[0]["shape"] = [4]
    q2 = net2.sp.prjs[0]
    d2 = O.SpDraws(1, 1, 1)
    # hypercolumn 3 from 0: distance 3, mean 0.75 * 3 / 0.2 + 1 = 12.25 ms
    far = [net2.sp._delay(d2, q2, 0, 3) for _ in range(200)]
    assert 11 <= np.mean(far) <= 14 and min(far) >= 9

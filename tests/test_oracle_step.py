"""Step-level pins of the oracle's restatement of ProcUpdLazy.

Each test checks a property the reference's algorithm fixes, computed
independently of the oracle's code:
  * trace evolution of single synapses driven by scripted spike trains
    against RK4 integration of the continuous cascade with kicks,
  * the delay-queue latency (a presynaptic spike at step s updates its row at
    step s + di + 2; Proj.cpp:645-664 with ProcExchangeSpike's one-step lag),
  * the WTA soft-max against a float64 evaluation of its formula,
  * the postsynaptic Euler recurrence against its matrix power,
  * ProcCheck's argmax / criterion special cases,
  * the table, float-expression and double variants against each other.
"""

import numpy as np
import pytest

import oracle as O
from gsbn_amd import configs


def tiny_spec(H=2, M=2, steps=200, ext=None, prn=1.0, delay=1, tauzi=0.01, tauzj=0.01,
              taue=0.02, taup=1.0, snoise=0.0, wtagain=4.0):
    stim = np.full((1, H), configs.UNCLAMPED, np.uint32)
    mask = np.zeros((1, H), np.float32)
    modes = [configs.ModeSpec(configs.step_time(0, 0.001), configs.step_time(steps, 0.001),
                              time_step=steps, begin_lgidx_id=0, begin_wmask_id=0, prn=prn,
                              lgidx_step=0, wmask_step=0)]
    pop = configs.PopSpec(hcu=H, mcu=M, snoise=snoise, wtagain=wtagain)
    proj = configs.ProjSpec(src=0, dest=0, tauzi=tauzi, tauzj=tauzj, taue=taue, taup=taup,
                            slot_num=float((H - 1) * M))
    spec = configs.NetSpec(dt=0.001, pops=[pop], projs=[proj], modes=modes, stim=stim, mask=mask,
                           conn=[configs.dense_conn(H, M, delay)], name="tiny")
    if ext is not None:
        spec.ext_spikes = {0: ext}
    return spec


def rk4_synapse(kzi, kzj, ke, kp, kicks_i, kicks_j, t_end, zi0=0.0, zj0=0.0, e0=0.0, p0=0.0,
                h=0.01):
    """Integrate z_i, z_j, e, p with dz/dt=-k z (+kicks), de/dt=ke(zi zj - e),
    dp/dt = kp (e - p) from t=0 to t_end; kicks are (time, amount) applied at
    the start of their time instant."""
    y = np.array([zi0, zj0, e0, p0], dtype=np.float64)

    def f(y):
        zi, zj, e, p = y
        return np.array([-kzi * zi, -kzj * zj, ke * (zi * zj - e), kp * (e - p)])

    events = sorted([(t, 0, a) for t, a in kicks_i] + [(t, 1, a) for t, a in kicks_j])
    t = 0.0
    for te, which, amount in events + [(t_end, -1, 0.0)]:
        n = int(round((te - t) / h))
        for _ in range(n):
            k1 = f(y)
            k2 = f(y + h / 2 * k1)
            k3 = f(y + h / 2 * k2)
            k4 = f(y + h * k3)
            y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        t = te
        if which >= 0:
            y[which] += amount
    return y


def test_synapse_traces_against_rk4():
    # unit 0 (HCU 0) is presynaptic to slot 0 of HCU 1; unit 2 (HCU 1, MCU 0) is postsynaptic
    pre_steps = [3, 10, 11, 40]
    post_steps = [6, 11, 25, 41]
    ext = {}
    for s in pre_steps:
        ext.setdefault(s, []).append(0)
    for s in post_steps:
        ext.setdefault(s, []).append(2)
    T = 80
    spec = tiny_spec(ext=ext, steps=T)
    net = O.OracleNet(spec)
    for _ in range(T):
        net.step()
    q = spec.projs[0]
    kzi, kzj, ke, kp = (np.float32(0.001) / np.float32(v) for v in (q.tauzi, q.tauzj, q.taue, q.taup))
    kfti = 1 / (np.float32(100) * np.float32(q.tauzi))
    kftj = 1 / (np.float32(100) * np.float32(q.tauzj))
    C, M = 2, 2
    # slot of source unit 0 in HCU 1 is k = 0; synapse (h=1, k=0, j=0)
    idx = (1 * C + 0) * M + 0
    tij = int(net.vector("tij_0")[idx])
    # row kicks: pre spike at s reaches the row update at s + di + 2 (di = 1)
    kicks_i = [(s + 3, float(kfti)) for s in pre_steps if s + 3 <= tij]
    kicks_j = [(s, float(kftj)) for s in post_steps if s <= tij]
    y = rk4_synapse(float(kzi), float(kzj), float(ke), float(kp), kicks_i, kicks_j, tij,
                    p0=1.0 / (C * M))
    got = [net.vector(n)[idx] for n in ("zi2_0", "zj2_0", "eij_0", "pij_0")]
    np.testing.assert_allclose(got, y, rtol=2e-5, atol=1e-9)
    # row trace of slot (1,0): pi, ei, zi with kicks at the row updates only
    r = 1 * C + 0
    ti = int(net.vector("ti_0")[r])
    yi = rk4_synapse(float(kzi), 0.0, float(ke), float(kp),
                     [(s + 3, float(kfti)) for s in pre_steps if s + 3 <= ti], [], ti,
                     zj0=1.0, p0=1.0 / C)
    got_i = [net.vector(n)[r] for n in ("zi_0", "ei_0", "pi_0")]
    np.testing.assert_allclose(got_i, [yi[0], yi[2], yi[3]], rtol=2e-5, atol=1e-9)


@pytest.mark.parametrize("delay", [1, 2, 5])
def test_queue_latency(delay):
    # one presynaptic spike of unit 1 at step 10: its rows (one per other HCU)
    # are active exactly at step 10 + delay + 2
    spec = tiny_spec(H=3, M=2, ext={10: [1]}, steps=40, delay=delay)
    net = O.OracleNet(spec)
    active = {}
    for _ in range(40):
        net.step()
        s = net.extgen.simstep
        ssi = net.vector(".ssi_0")
        if len(ssi):
            active[s] = sorted(ssi.tolist())
    ii = net.vector("ii_0")
    rows = sorted(int(r) for r in np.nonzero(ii == 1)[0])
    assert active == {10 + delay + 2: rows}


def test_wta_softmax_formula():
    # random dsup; compare act with a float64 evaluation of Pop.cpp:263-303
    spec = tiny_spec(H=3, M=7, steps=5, snoise=0.5)
    net = O.OracleNet(spec)
    for _ in range(3):
        net.step()
        dsup = net.vector("dsup_0").astype(np.float64).reshape(3, 7)
        act = net.vector("act_0").astype(np.float64).reshape(3, 7)
        w = 4.0
        for h in range(3):
            mx = dsup[h].max()
            a = np.exp(w * (dsup[h] - mx))
            if np.exp(w * mx) < 1:
                a *= np.exp(w * mx)
            if a.sum() > 1:
                a /= a.sum()
            np.testing.assert_allclose(act[h], a, rtol=1e-5)


def test_jxx_euler_matrix_power():
    # no spikes after step 5: (pj, ej, zj) follow the linear recurrence
    spec = tiny_spec(H=2, M=2, ext={5: [0, 1, 2, 3]}, steps=30)
    net = O.OracleNet(spec)
    for _ in range(5):
        net.step()
    s0 = np.array([net.vector(n)[0] for n in ("pj_0", "ej_0", "zj_0")], np.float64)
    for _ in range(20):
        net.step()
    q = spec.projs[0]
    kp, ke, kz = 0.001 / q.taup, 0.001 / q.taue, 0.001 / q.tauzj
    A = np.array([[1 - kp, kp, 0], [0, 1 - ke, ke], [0, 0, 1 - kz]])
    # order of updates: pj uses old ej, ej uses old zj, then zj decays
    want = np.linalg.matrix_power(A, 20) @ s0
    got = [net.vector(n)[0] for n in ("pj_0", "ej_0", "zj_0")]
    np.testing.assert_allclose(got, want, rtol=1e-5)
    bj = net.vector("bj_0")[0]
    np.testing.assert_allclose(bj, np.log(got[0] + 0.001), rtol=1e-6)


def test_variants_agree_with_exact_double():
    # the float table path, the reference's float expression and the exact
    # double path on the same scripted spikes
    rng = np.random.default_rng(3)
    ext = {s: sorted(rng.choice(12, size=rng.integers(0, 4), replace=False).tolist())
           for s in range(1, 300)}
    outs = {}
    for arith in (O.ARITH_TABLE, O.ARITH_REF32):
        spec = tiny_spec(H=3, M=4, ext=ext, steps=300)
        net = O.OracleNet(spec, arith=arith)
        for _ in range(300):
            net.step()
        outs[arith] = {n: net.vector(n) for n in ("pij_0", "eij_0", "zi2_0", "zj2_0", "wij_0", "pi_0")}
    # exact double path in replay with the same spike trains
    spec = tiny_spec(H=3, M=4, ext=ext, steps=300)
    ref = O.OracleNet(spec, arith=O.ARITH_TABLE)
    f64 = O.OracleNet(spec, arith=O.ARITH_F64)
    L = O.lib()
    for _ in range(300):
        ref.step()
        eg = ref.extgen
        L.gor_set_globals(f64.h, eg.simstep, eg.prn, eg.old_prn, eg.cycle_flag, 0, eg.wmask_idx)
        if eg.cycle_flag == 1:
            sj = ref.vector("spike_0")
            L.gor_proj_step_replay(f64.h, 0, O._p(sj))
            L.gor_exchange_spikes(f64.h)
    for n in ("pij_0", "eij_0", "zi2_0", "zj2_0", "pi_0"):
        exact = f64.vector(n)
        for arith in outs:
            np.testing.assert_allclose(outs[arith][n], exact, rtol=1e-4, atol=1e-7)


def test_check_criterion_cases():
    spec = configs.config0()
    chk = O.Check(spec, threshold=0)
    m = chk.modes[0]
    H, M = 10, 10
    # counts: ties resolve to the first maximum; an all-zero HCU reports -1
    spikes = np.zeros(H * M, np.int8)
    exp = spec.stim[m["lgexp"]]
    for h in range(H):
        spikes[h * M + int(exp[h])] = 1
    chk.update(1, m["begin"] + 1, [spikes])
    chk.update(1, m["end"], [spikes])
    assert chk.results[-1]["ok"] and chk.results[-1]["cnt"] == [1] * H
    chk2 = O.Check(spec, threshold=0)
    spikes2 = spikes.copy()
    spikes2[0 * M:(0 + 1) * M] = 0
    chk2.update(1, m["begin"] + 1, [spikes2])
    chk2.update(1, m["end"], [spikes2])
    r = chk2.results[-1]
    assert not r["ok"] and r["idx"][0] == -1 and r["cnt"][0] == 0
    # threshold 2: one spike is not enough
    chk3 = O.Check(spec, threshold=2)
    chk3.update(1, m["begin"] + 1, [spikes])
    chk3.update(1, m["end"], [spikes])
    assert not chk3.results[-1]["ok"]


def test_extgen_time_truncation_quirk():
    # ProcExtGen.cpp:22 truncates float(begin_time)/float(dt): 0.005 s at 1 ms is step 4,
    # so the workload generator writes times half a step past the boundary
    assert int(np.float32(0.005) / np.float32(0.001)) == 4
    for k in range(0, 20000, 7):
        assert int(np.float32(configs.step_time(k, 0.001)) / np.float32(0.001)) == k


def test_conn_slot_fill_matches_first_free_scan():
    # the oracle's cursor fill against a literal restatement of
    # ProcInitConn::add_row (scan for the first ii < 0, drop when full), on a
    # map with repeated destination lines and more sources than slots
    rng = np.random.default_rng(9)
    H, M, C = 4, 3, 7
    rows = []
    for _ in range(10):
        h = int(rng.integers(0, H))
        pairs = [(int(rng.integers(0, H * M)), int(rng.integers(1, 4)))
                 for _ in range(int(rng.integers(0, 5)))]
        rows.append((h, pairs))
    spec = tiny_spec(H=H, M=M, steps=3)
    spec.projs[0].slot_num = float(C)
    spec.conn = [rows]
    net = O.OracleNet(spec)
    want_ii = np.full(H * C, -1, np.int32)
    want_di = np.zeros(H * C, np.int32)
    for h, pairs in rows:
        for s, d in pairs:
            for k in range(C):
                if want_ii[h * C + k] < 0:
                    want_ii[h * C + k] = s
                    want_di[h * C + k] = d
                    break
    assert (net.vector("ii_0") == want_ii).all()
    assert (net.vector("di_0") == want_di).all()
    assert (want_ii >= 0).sum() < sum(len(p) for _, p in rows)  # some sources were dropped

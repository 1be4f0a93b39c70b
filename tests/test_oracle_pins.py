"""Independent pins of the oracle functions that lockstep parity alone does
not fix (product and oracle were written from the same reading, so a shared
misreading would pass product == oracle).

Every expected value here is computed from the reference's equations by
code of its own (float64 numpy / scipy), never by the oracle:

* the full catch-up sweep on a learning-rate switch (Proj.cpp:209-308,
  538-588): target step simstep-1, rate kp = taupdt * old-prn, Wij
  recomputed only when kp != 0 and the gap is positive, tij stamped when the
  gap is <= 0; checked synapse by synapse against scipy.linalg.expm of the
  3x3 trace cascade, and over a train -> recall -> train schedule against
  RK4 integration of the continuous cascade with the learning rate switching
  at simstep-1;
* support integration (Pop.cpp:229-261) and spike sampling
  (Pop.cpp:305-323) from the oracle's own inputs of that step;
* ProcExtGen's mode chunking and per-step globals (ProcExtGen.cpp:17-54,
  113-153) for the reference's shipped data/10x10_rnd.prototxt and for a
  hand-made schedule with gaps and ragged chunks, written out by hand.

test_pins_catch_planted_mistakes shows that each pin fails when the oracle
is made to deviate in the way the pin is meant to exclude (old/new prn,
t/t-1, Wij at kp = 0, pi before/after its catch-up, a dropped bj term, a
dropped maxfq*dt factor).
"""

import os

import numpy as np
import pytest
import scipy.linalg

import oracle as O
from gsbn_amd import configs

DT = 0.001
F = np.float32
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rates(q, dt=DT):
    """Per-step rates as the procedure computes them in float (Proj.cpp:38-52)."""
    f = lambda tau: float(F(dt) / F(tau))  # noqa: E731
    kfti = float(F(1) / (F(q.maxfq) * F(q.tauzi)))
    kftj = float(F(1) / (F(q.maxfq) * F(q.tauzj)))
    return dict(kzi=f(q.tauzi), kzj=f(q.tauzj), ke=f(q.taue), taupdt=f(q.taup), kfti=kfti,
                kftj=kftj, eps=float(F(dt) / F(q.taup)))


def cascade(kz, ke, kp, gap, z, e, p):
    """(z, e, p) after `gap` steps of dz=-kz z, de=ke (z-e), dp=kp (e-p)."""
    A = np.array([[-kz, 0.0, 0.0], [ke, -ke, 0.0], [0.0, kp, -kp]], np.float64)
    return scipy.linalg.expm(A * gap) @ np.array([z, e, p], np.float64)


def schedule_spec(prns, seg, H=3, M=3, ext=None, clamp=False, wmask=0.0, snoise=0.0,
                  seed=0):
    """One population with a dense recurrent projection (no self-HCU
    connections), consecutive modes of `seg` steps with learning rates
    `prns`; optionally one clamped pattern and recurrent support."""
    stim = np.array([[h % M for h in range(H)] if clamp else [configs.UNCLAMPED] * H],
                    np.uint32)
    mask = np.full((1, H), wmask, np.float32)
    modes = [configs.ModeSpec(configs.step_time(i * seg, DT), configs.step_time((i + 1) * seg, DT),
                              time_step=seg, begin_lgidx_id=0, begin_wmask_id=0, prn=prn,
                              lgidx_step=0, wmask_step=0)
             for i, prn in enumerate(prns)]
    pop = configs.PopSpec(hcu=H, mcu=M, snoise=snoise)
    proj = configs.ProjSpec(src=0, dest=0, tauzi=0.01, tauzj=0.01, taue=0.02, taup=1.0,
                            slot_num=float((H - 1) * M))
    spec = configs.NetSpec(dt=DT, pops=[pop], projs=[proj], modes=modes, stim=stim, mask=mask,
                           conn=[configs.dense_conn(H, M, 1)], seed=seed, name="pins")
    if ext is not None:
        spec.ext_spikes = {0: ext}
    return spec


SYN = ("pij_0", "eij_0", "zi2_0", "zj2_0", "tij_0", "wij_0")
ROW = ("pi_0", "ei_0", "zi_0", "ti_0")


def state(net, names=SYN + ROW + ("pj_0", "ii_0")):
    return {n: net.vector(n).copy() for n in names}


# ---------------------------------------------------------------- the sweep

SEG = 40
SWEEP_PRNS = (1.0, 0.0, 0.5, 1.0)  # sweeps at steps 1, 41 (kp = taupdt), 81 (kp = 0), 121


def sweep_ext(H, M, seed=4):
    """Random spikes, none at t-3 or t of a sweep step t (no row or column
    update at t, so the state after step t is the sweep's alone), and
    pre/post spikes at t-4 / t-1 (rows and columns updated at t-1, which the
    sweep only stamps).  The column updated at t-1 alternates between two
    units, so the weights a column update left stale before a kp = 0 window
    (a column update does not recompute Wij, Proj.cpp:481-536) are still
    stale, and not stamped, when the window ends."""
    rng = np.random.default_rng(seed)
    sweeps = [SEG * i + 1 for i in range(1, len(SWEEP_PRNS))]
    post = [H * M - 1, M + 1]
    ext = {}
    for s in range(1, SEG * len(SWEEP_PRNS)):
        if any(s in (t - 3, t) for t in sweeps):
            continue
        u = set(int(v) for v in np.nonzero(rng.random(H * M) < 0.08)[0])
        if any(s == t - 4 for t in sweeps):
            u |= {0}
        for n, t in enumerate(sweeps):
            if s == t - 1:
                u = (u - set(post)) | {post[n % 2]}
        if u:
            ext[s] = sorted(u)
    return ext, sweeps


def sweep_mismatches(mistake=None):
    """Run the schedule; at every sweep step compare the state after it with
    the float64 expectation from the state before it.  Returns a list of
    (step, what) for every disagreement."""
    H, M = 3, 3
    ext, sweeps = sweep_ext(H, M)
    spec = schedule_spec(SWEEP_PRNS, SEG, H=H, M=M, ext=ext)
    q = spec.projs[0]
    r = rates(q)
    C = (H - 1) * M
    eps, eps2 = r["eps"], float(F(r["eps"]) * F(r["eps"]))
    wgain = float(q.wgain)
    net = O.OracleNet(spec)
    net.plant(mistake)
    bad = []
    w_after_kp0 = None
    prev = None
    while True:
        if net.extgen.simstep + 1 in sweeps:
            prev = state(net)
        f = net.step()
        if f < 0:
            break
        t = net.extgen.simstep
        if t not in sweeps:
            continue
        old_prn, new_prn = net.extgen.old_prn, net.extgen.prn
        assert old_prn != new_prn
        kp = float(F(r["taupdt"]) * F(old_prn))
        now = state(net)
        target = t - 1
        # rows
        pi_new = prev["pi_0"].astype(np.float64)
        for i in range(H * C):
            ti = int(prev["ti_0"][i])
            gap = target - ti
            if gap > 0:
                z, e, p = cascade(r["kzi"], r["ke"], kp, gap, prev["zi_0"][i], prev["ei_0"][i],
                                  prev["pi_0"][i])
                want = (p, e, z, target)
            else:
                want = (prev["pi_0"][i], prev["ei_0"][i], prev["zi_0"][i], ti)
            pi_new[i] = want[0]
            got = [now[n][i] for n in ROW]
            if got[3] != want[3] or not np.allclose(got[:3], want[:3], rtol=2e-6, atol=1e-30):
                bad.append((t, "row %d: %s vs %s" % (i, got, want)))
        # synapses
        for idx in range(H * C * M):
            i, j = divmod(idx, M)
            tij = int(prev["tij_0"][idx])
            gap = target - tij
            zi2, zj2 = float(prev["zi2_0"][idx]), float(prev["zj2_0"][idx])
            if gap > 0:
                z, e, p = cascade(r["kzi"] + r["kzj"], r["ke"], kp, gap, zi2 * zj2,
                                  prev["eij_0"][idx], prev["pij_0"][idx])
                zi2n, zj2n = zi2 * np.exp(-r["kzi"] * gap), zj2 * np.exp(-r["kzj"] * gap)
                want = [p, e, zi2n, zj2n]
                if kp != 0.0:
                    pj = float(prev["pj_0"][i // C * M + j])
                    w = wgain * np.log((p + eps2) / ((pi_new[i] + eps) * (pj + eps)))
                else:
                    w = float(prev["wij_0"][idx])
            else:
                want = [prev[n][idx] for n in ("pij_0", "eij_0", "zi2_0", "zj2_0")]
                w = float(prev["wij_0"][idx])
            got = [now[n][idx] for n in ("pij_0", "eij_0", "zi2_0", "zj2_0")]
            if int(now["tij_0"][idx]) != target:
                bad.append((t, "tij %d: %d vs %d" % (idx, now["tij_0"][idx], target)))
            if not np.allclose(got, want, rtol=2e-6, atol=1e-30):
                bad.append((t, "syn %d: %s vs %s" % (idx, got, want)))
            if not np.isclose(now["wij_0"][idx], w, rtol=2e-6, atol=2e-6):
                bad.append((t, "wij %d: %r vs %r" % (idx, now["wij_0"][idx], w)))
        if old_prn == 0.0 and w_after_kp0 is not None:
            # Wij is frozen through a kp = 0 window, the sweep that ends it included
            if (now["wij_0"] != w_after_kp0).any():
                bad.append((t, "wij changed through the kp = 0 window"))
        w_after_kp0 = now["wij_0"].copy() if new_prn == 0.0 else None
    net.close()
    return bad


def test_sweep_against_matrix_exponential():
    bad = sweep_mismatches()
    assert not bad, bad[:5]


def test_sweep_schedule_reaches_both_branches():
    # the schedule exercises gaps > 0 and gaps <= 0 at the sweeps with kp != 0
    H, M = 3, 3
    ext, sweeps = sweep_ext(H, M)
    spec = schedule_spec(SWEEP_PRNS, SEG, H=H, M=M, ext=ext)
    net = O.OracleNet(spec)
    seen = {}
    while True:
        if net.extgen.simstep + 1 in sweeps:
            tij = net.vector("tij_0")
            seen[net.extgen.simstep + 1] = (int((tij < net.extgen.simstep).sum()),
                                           int((tij >= net.extgen.simstep).sum()))
        if net.step() < 0:
            break
    for t in sweeps:
        assert seen[t][0] > 0 and seen[t][1] > 0, (t, seen[t])


def rk4_piecewise(kzi, kzj, ke, kp_of, kicks_i, kicks_j, t_end, p0, h=0.01):
    """RK4 of (zi, zj, e, p) with a learning rate kp_of(t) that is constant
    on each unit step; kicks (time, amount) at the start of their instant."""
    y = np.array([0.0, 0.0, 0.0, p0], np.float64)

    def f(y, kp):
        zi, zj, e, p = y
        return np.array([-kzi * zi, -kzj * zj, ke * (zi * zj - e), kp * (e - p)])

    events = sorted([(t, 0, a) for t, a in kicks_i] + [(t, 1, a) for t, a in kicks_j])
    t = 0
    for te, which, amount in events + [(t_end, -1, 0.0)]:
        while t < te:
            kp = kp_of(t + 1)  # the rate of the step (t, t+1]
            for _ in range(int(round(1 / h))):
                k1 = f(y, kp)
                k2 = f(y + h / 2 * k1, kp)
                k3 = f(y + h / 2 * k2, kp)
                k4 = f(y + h * k3, kp)
                y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
            t += 1
        if which >= 0:
            y[which] += amount
    return y


def test_train_recall_train_against_rk4():
    # one synapse through prn 1 -> 0 -> 0.25 (sweeps at 41 and 81): the
    # sweep integrates to simstep - 1 with the old rate, so in continuous
    # time the rate switches at the end of step 40 resp. 80
    pre = [3, 10, 20, 33, 45, 52, 70, 85, 90, 101, 108]
    post = [6, 11, 25, 37, 46, 60, 77, 88, 96, 105, 112]
    ext = {}
    for s in pre:
        ext.setdefault(s, []).append(0)
    for s in post:
        ext.setdefault(s, []).append(2)
    prns = (1.0, 0.0, 0.25)
    spec = schedule_spec(prns, SEG, H=2, M=2, ext=ext)
    net = O.OracleNet(spec)
    while net.step() >= 0:
        pass
    q = spec.projs[0]
    r = rates(q)

    def kp_of(step):
        return float(F(r["taupdt"]) * F(prns[min((step - 1) // SEG, len(prns) - 1)]))

    C, M = 2, 2
    idx = (1 * C + 0) * M + 0  # HCU 1, slot 0 (source unit 0), MCU 0 (unit 2)
    tij = int(net.vector("tij_0")[idx])
    assert tij >= 112
    kicks_i = [(s + 3, r["kfti"]) for s in pre if s + 3 <= tij]
    kicks_j = [(s, r["kftj"]) for s in post if s <= tij]
    y = rk4_piecewise(r["kzi"], r["kzj"], r["ke"], kp_of, kicks_i, kicks_j, tij, 1.0 / (C * M))
    got = [net.vector(n)[idx] for n in ("zi2_0", "zj2_0", "eij_0", "pij_0")]
    np.testing.assert_allclose(got, y, rtol=2e-5, atol=1e-9)
    # and the same run with the rate switching at simstep instead: pij differs
    def kp_late(step):
        return kp_of(step - 1 if step > 1 else step)
    y_late = rk4_piecewise(r["kzi"], r["kzj"], r["ke"], kp_late, kicks_i, kicks_j, tij,
                           1.0 / (C * M))
    assert abs(y_late[3] - y[3]) > 1e-3 * abs(y[3])
    net.close()


# ------------------------------------------------- support and spike sampling

def recall_pop_spec(H, M):
    """A clamped pattern, 80 training steps (wmask 0), then 80 recall steps
    with the recurrent support on (wmask 1) and a bias that keeps units
    spiking against the negative bj."""
    spec = schedule_spec((1.0, 0.0), 80, H=H, M=M, clamp=True, snoise=0.3, seed=7)
    spec.mask = np.stack([np.zeros(H, np.float32), np.ones(H, np.float32)])
    spec.modes[1].begin_wmask_id = 1
    spec.pops[0].lgbias = 1.5
    return spec


def pop_mismatches(mistake=None, steps=160):
    """Support integration and spike sampling of every unit at every step,
    from the oracle's inputs of that step (the previous step's epsc and bj,
    this step's draws), in float64 resp. float32 as the comparison needs."""
    H, M = 3, 4
    spec = recall_pop_spec(H, M)
    p = spec.pops[0]
    taumdt = float(F(DT) / F(p.taum))
    maxfqdt = F(p.maxfq) * F(DT)
    net = O.OracleNet(spec)
    net.plant(mistake)
    bad = []
    spiked = 0
    for _ in range(steps):
        before = {n: net.vector(n).astype(np.float64) for n in ("dsup_0", "epsc_0", "bj_0")}
        cnt0 = net.vector(".counter_0")
        if net.step() < 0:
            break
        t = net.extgen.simstep
        eg = net.extgen
        lginp = eg.lginp.astype(np.float64)
        wmask = spec.mask[eg.wmask_idx].astype(np.float64)
        noise = net.vector(".rnd_normal_0").astype(np.float64)
        wsup = before["bj_0"] + before["epsc_0"]
        h = np.arange(H * M) // M
        sup = p.lgbias + p.igain * lginp + noise + p.wgain * wmask[h] * wsup
        want = before["dsup_0"] + (sup - before["dsup_0"]) * taumdt
        got = net.vector("dsup_0")
        if not np.allclose(got, want, rtol=1e-5, atol=1e-6):
            bad.append((t, "dsup", np.max(np.abs(got - want))))
        u = net.vector(".rnd_uniform01_0")
        act = net.vector("act_0")
        spk = (u < act * maxfqdt).astype(np.int8)
        if (net.vector("spike_0") != spk).any():
            bad.append((t, "spike"))
        if (net.vector(".counter_0") - cnt0 != spk).any():
            bad.append((t, "counter"))
        spiked += int(spk.sum())
    net.close()
    assert spiked > 0
    return bad


def test_support_and_spikes_from_their_inputs():
    bad = pop_mismatches()
    assert not bad, bad[:5]


def test_support_uses_recurrent_terms():
    # the pin is not vacuous: in the recall mode bj and epsc are non-zero,
    # wmask = 1, and units spike
    spec = recall_pop_spec(3, 4)
    net = O.OracleNet(spec)
    spikes = 0
    for _ in range(160):
        net.step()
        if net.extgen.simstep > 80:
            spikes += int(net.vector("spike_0").sum())
            assert net.extgen.wmask_idx == 1
    assert np.abs(net.vector("bj_0")).max() > 0.1
    assert np.abs(net.vector("epsc_0")).max() > 0.01
    assert spikes > 0
    net.close()


# ---------------------------------------------------------------- ExtGen

def extgen_trace(spec, steps):
    eg = O.ExtGen(spec)
    out = []
    for _ in range(steps):
        eg.update()
        out.append((eg.simstep, eg.cycle_flag, eg.prn, eg.old_prn, eg.wmask_idx, eg.old_lgidx,
                    eg.lginp.copy()))
    return eg, out


def test_extgen_shipped_10x10_schedule():
    # data/10x10_rnd.prototxt:10-56: mode 1 = 0..5 s, 500-step chunks,
    # stimulus rows 0, 1, ..., 9, wmask row 0, prn 1, expected rows 0..9
    # (lgexp_step is not given: 1 by gsbn.proto:27); mode 2 = 5..10 s,
    # 500-step chunks, stimulus rows 100..109, wmask row 1, prn 0, expected
    # rows 0..9.  Written out by hand from ProcExtGen.cpp:17-54, 113-153.
    import refconfig
    with open(os.path.join(GOLDEN, "10x10_rnd.prototxt")) as f:
        text = f.read()
    with open(os.path.join(GOLDEN, "stimuli_10x10_100_-1_rnd.bin"), "rb") as f:
        spec = refconfig.spec_from_prototxt(text, f.read())
    eg, tr = extgen_trace(spec, 10002)
    assert [(m["begin"], m["end"], m["lgidx"], m["lgexp"], m["wmask"], m["prn"])
            for m in eg.modes] == (
        [(500 * c, 500 * (c + 1), c, c, 0, 1.0) for c in range(10)] +
        [(5000 + 500 * c, 5500 + 500 * c, 100 + c, c, 1, 0.0) for c in range(10)])
    eps = np.float32(0.001)
    for simstep, flag, prn, old_prn, wm, lgidx, lginp in tr:
        if simstep > 10000:
            assert flag == -1
            continue
        assert flag == 1
        assert prn == (1.0 if simstep <= 5000 else 0.0)
        assert old_prn == (0.0 if simstep == 1 else 1.0 if simstep <= 5001 else 0.0)
        assert wm == (0 if simstep <= 5000 else 1)
        want_row = (simstep - 1) // 500 if simstep <= 5000 else 100 + (simstep - 5001) // 500
        assert lgidx == want_row
        pat = spec.stim[want_row]
        want = np.array([np.log(1 + eps) if (pat[h] == j) else np.log(eps)
                         for h in range(10) for j in range(10)], np.float32)
        np.testing.assert_allclose(lginp, want, rtol=1e-6)
    # ProcCheck windows are the prn == 0 chunks (ProcCheck.cpp:35-62)
    chk = O.Check(spec)
    assert [(m["begin"], m["end"], m["lgexp"]) for m in chk.modes] == \
        [(5000 + 500 * c, 5500 + 500 * c, c) for c in range(10)]


def test_extgen_gaps_and_ragged_chunks():
    # modes (times in s): [0.0035, 0.0205) chunks of 7 steps; gap; [0.0305,
    # 0.0405) prn 0 chunks of 4; the float truncation int(t / dt) gives steps
    # 3..20 and 30..40 (0.0035/0.001 = 3.4999998 -> 3)
    modes = [configs.ModeSpec(0.0035, 0.0205, time_step=7, begin_lgidx_id=2, lgidx_step=3,
                              begin_wmask_id=1, wmask_step=2, prn=1.0),
             configs.ModeSpec(0.0305, 0.0405, time_step=4, begin_lgidx_id=0, lgidx_step=1,
                              begin_wmask_id=0, wmask_step=0, prn=0.0)]
    assert int(F(0.0035) / F(0.001)) == 3 and int(F(0.0205) / F(0.001)) == 20
    assert int(F(0.0305) / F(0.001)) == 30 and int(F(0.0405) / F(0.001)) == 40
    stim = np.zeros((20, 2), np.uint32)
    spec = configs.NetSpec(dt=0.001, pops=[configs.PopSpec(hcu=2, mcu=2)], projs=[],
                           modes=modes, stim=stim, mask=np.zeros((20, 2), np.float32), conn=[])
    eg, tr = extgen_trace(spec, 42)
    assert [(m["begin"], m["end"], m["lgidx"], m["wmask"]) for m in eg.modes] == [
        (3, 10, 2, 1), (10, 17, 5, 3), (17, 20, 8, 5),
        (30, 34, 0, 0), (34, 38, 1, 0), (38, 40, 2, 0)]
    want = {}
    for s in range(1, 43):
        if s <= 3:
            want[s] = (0, None)
        elif s <= 10:
            want[s] = (1, (1.0, 1, 2))
        elif s <= 17:
            want[s] = (1, (1.0, 3, 5))
        elif s <= 20:
            want[s] = (1, (1.0, 5, 8))
        elif s <= 30:
            want[s] = (0, None)
        elif s <= 34:
            want[s] = (1, (0.0, 0, 0))
        elif s <= 38:
            want[s] = (1, (0.0, 0, 1))
        elif s <= 40:
            want[s] = (1, (0.0, 0, 2))
        else:
            want[s] = (-1, None)
    prev_prn = 0.0
    for simstep, flag, prn, old_prn, wm, lgidx, _ in tr:
        wf, wv = want[simstep]
        assert flag == wf, simstep
        if wv is not None:
            assert (prn, wm, lgidx) == wv, simstep
            assert old_prn == prev_prn, simstep
            prev_prn = prn


# ------------------------------------------------------- tests of the tests

@pytest.mark.parametrize("mistake", ["sweep_new_prn", "sweep_target_t", "sweep_w_kp0",
                                     "sweep_old_pi"])
def test_sweep_pin_catches_planted_mistakes(mistake):
    assert sweep_mismatches(mistake), "the sweep pin passed with the mistake %s" % mistake


@pytest.mark.parametrize("mistake", ["sup_no_bj", "spike_no_maxfq"])
def test_pop_pin_catches_planted_mistakes(mistake):
    assert pop_mismatches(mistake), "the support/spike pin passed with the mistake %s" % mistake

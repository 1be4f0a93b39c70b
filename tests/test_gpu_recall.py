"""Frozen-weight batched recall on the GPU (-m gpu).

* the support product on the matrix cores against numpy: exactly for
  integer weights (every partial sum is exact in float), within the float
  summation bound for random bfloat16 weights, including ragged tile edges;
* the exact mode against the oracle: for several cues, the oracle runs one
  cue in recall mode with the same frozen weights and seed, and every step's
  spikes must be identical;
* the bfloat16 matrix-core mode against the exact mode: recall outcomes agree.
"""

import ctypes
import tempfile

import numpy as np
import pytest
import torch

import oracle as O
from gsbn_amd import Solver, configs, lib
from gsbn_amd import recall as R

pytestmark = pytest.mark.gpu


def _gemm(S, WT, E, beta):
    L = lib()
    R._setup(L)
    B, Kp = S.shape
    N = E.shape[1]
    dev = torch.device("cuda:0")
    Bp = (B + 255) // 256 * 256
    Np = (N + 255) // 256 * 256
    s = torch.zeros((Bp, Kp), dtype=torch.int16, device=dev)
    w = torch.zeros((Np, Kp), dtype=torch.int16, device=dev)
    s[:B] = torch.from_numpy(S.view(np.int16)).to(dev)
    w[:N] = torch.from_numpy(WT.view(np.int16)).to(dev)
    e = torch.from_numpy(E.copy()).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.gsbn_spike_gemm(B, N, Kp, s.data_ptr(), w.data_ptr(), e.data_ptr(), beta, stream)
    assert rc == 0
    torch.cuda.synchronize()
    return e.cpu().numpy()


def _bf16_bits(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(torch.bfloat16).view(
        torch.int16).numpy().view(np.uint16)


@pytest.mark.parametrize("B,N,K", [(128, 128, 64), (200, 300, 500), (256, 1000, 960)])
def test_spike_gemm_integer_weights_exact(B, N, K):
    rng = np.random.default_rng(B + N + K)
    Kp = (K + 63) // 64 * 64
    S = np.zeros((B, Kp), np.uint16)
    S[:, :K] = np.where(rng.random((B, K)) < 0.05, 0x3F80, 0)
    W = rng.integers(-8, 9, size=(N, K)).astype(np.float32)
    WT = np.zeros((N, Kp), np.uint16)
    WT[:, :K] = _bf16_bits(W)
    E = (rng.integers(-1024, 1024, size=(B, N)) / 1024.0).astype(np.float32)
    got = _gemm(S, WT, E, 0.5)
    want = (0.5 * E.astype(np.float64) + (S[:, :K] != 0).astype(np.float64) @ W.T.astype(np.float64))
    assert np.array_equal(got, want.astype(np.float32))


def test_spike_gemm_random_bf16_weights():
    rng = np.random.default_rng(7)
    B, N, K = 300, 700, 1300
    Kp = (K + 63) // 64 * 64
    S = np.zeros((B, Kp), np.uint16)
    spk = rng.random((B, K)) < 0.02
    S[:, :K] = np.where(spk, 0x3F80, 0)
    Wf = rng.normal(0, 1, size=(N, K)).astype(np.float32)
    WT = np.zeros((N, Kp), np.uint16)
    WT[:, :K] = _bf16_bits(Wf)
    Wb = (WT[:, :K].astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    E = rng.normal(0, 1, size=(B, N)).astype(np.float32)
    beta = np.float32(0.95)
    got = _gemm(S, WT, E, float(beta)).astype(np.float64)
    prod = spk.astype(np.float64) @ Wb.T
    want = np.float64(beta) * E.astype(np.float64) + prod
    bound = 2 ** -22 * (np.abs(E) + spk.astype(np.float64) @ np.abs(Wb).T + 1e-30) * (K + 2)
    assert np.all(np.abs(got - want) <= bound)


def _train(tmp_path, spec):
    """Run the solver through training into the first recall step (after the
    sweep of the learning-rate switch) and return it."""
    s = Solver(configs.write(spec, str(tmp_path)))
    configs.apply_conn_arrays(s, spec)
    for _ in range(100000):
        s.step(1)
        if s.glob("cycle-flag") == 1 and s.glob("prn", "f") == 0.0:
            break
    return s


@pytest.fixture(scope="module")
def trained10(tmp_path_factory):
    spec = configs.config0()
    s = _train(tmp_path_factory.mktemp("train10"), spec)
    yield spec, s
    s.close()


def test_recall_exact_matches_oracle(trained10):
    spec, s = trained10
    pats = spec.stim[:10].astype(np.int64)
    B, T = 6, 60
    clamp, wmask, want = R.cues(pats, B, 0.4, seed=3)
    rc = R.Recall.from_solver(s, spec, B, R.EXACT, seed=1234)
    idx, cnt, ras, _ = rc.run(clamp, wmask, T, count_from=0, raster=True)
    ii, di = s.vector("ii_0"), s.vector("di_0")
    wij, bj = s.vector("wij_0"), s.vector("bj_0")
    for b in range(B):
        ospec = R.oracle_recall_spec(spec, clamp[b], wmask[b], T, 1234 + b, ii, di)
        net = O.OracleNet(ospec)
        net.write("wij_0", wij)
        net.write("bj_0", bj)
        counts = np.zeros(spec.pops[0].hcu * spec.pops[0].mcu, np.int64)
        for t in range(T):
            assert net.step() == 1
            sp = net.vector("spike_0")
            assert np.array_equal(sp, ras[t, b]), "cue %d diverges at step %d" % (b, t + 1)
            counts += sp
        H, M = spec.pops[0].hcu, spec.pops[0].mcu
        c = counts.reshape(H, M)
        for h in range(H):
            mi, mc = -1, 0
            for j in range(M):
                if c[h, j] > mc:
                    mi, mc = j, int(c[h, j])
            assert (idx[b, h], cnt[b, h]) == (mi, mc)
    rc.close()


def test_recall_mfma_agrees_with_exact(trained10):
    spec, s = trained10
    pats = spec.stim[:10].astype(np.int64)
    B, T = 512, 60
    clamp, wmask, want = R.cues(pats, B, 0.4, seed=5)
    out = {}
    for mode in (R.EXACT, R.MFMA):
        rc = R.Recall.from_solver(s, spec, B, mode, seed=99)
        idx, cnt, _, ms = rc.run(clamp, wmask, T, count_from=T // 2, time_support=True)
        out[mode] = idx
        rc.close()
    ok_exact = R.recall_correct(out[R.EXACT], want)
    ok_mfma = R.recall_correct(out[R.MFMA], want)
    agree = np.mean(ok_exact == ok_mfma)
    winners = np.mean(out[R.EXACT] == out[R.MFMA])
    print("recall exact %.3f mfma %.3f outcome agreement %.3f winner agreement %.3f"
          % (ok_exact.mean(), ok_mfma.mean(), agree, winners))
    assert agree >= 0.9 and winners >= 0.9


def test_spike_gemm_full_recall_shape_sampled_tiles():
    # the support product at the frozen-recall shape of the 100 x 100
    # network: (8192 cues x 10^4 units) 0/1 spikes times (10^4 x 10^4) bf16
    # weights, checked on sampled 64 x 64 output tiles (corners, middle,
    # ragged edge) against float64 within the float summation bound
    rng = np.random.default_rng(21)
    B, N, K = 8192, 10000, 10000
    Kp = (K + 63) // 64 * 64
    spk = rng.random((B, K)) < 0.01
    S = np.zeros((B, Kp), np.uint16)
    S[:, :K] = np.where(spk, 0x3F80, 0)
    WT = np.zeros((N, Kp), np.uint16)
    WT[:, :K] = _bf16_bits(rng.normal(0, 0.5, size=(N, K)).astype(np.float32))
    E = rng.normal(0, 1, size=(B, N)).astype(np.float32)
    beta = np.float32(0.95)
    got = _gemm(S, WT, E, float(beta))
    rows = [0, 4096, B - 64]
    cols = [0, 5000, N - 64]
    for r0 in rows:
        sp = spk[r0:r0 + 64].astype(np.float64)
        for c0 in cols:
            Wb = (WT[c0:c0 + 64, :K].astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            want = np.float64(beta) * E[r0:r0 + 64, c0:c0 + 64].astype(np.float64) + sp @ Wb.T
            bound = 2 ** -22 * (np.abs(E[r0:r0 + 64, c0:c0 + 64]) + sp @ np.abs(Wb).T
                                + 1e-30) * (K + 2)
            err = np.abs(got[r0:r0 + 64, c0:c0 + 64].astype(np.float64) - want)
            assert np.all(err <= bound), (r0, c0, float(err.max()))


@pytest.fixture(scope="module")
def trained100(tmp_path_factory):
    spec = configs.config1()
    s = _train(tmp_path_factory.mktemp("train100"), spec)
    yield spec, s
    s.close()


def test_recall_8192_cues_exact_sampled_against_oracle(trained100):
    # the frozen-recall workload at its own size: the 100 x 100 network
    # trained through its 6000 training steps, 8192 cues (40 % of the HCUs
    # clamped) x 50 lock-step steps in the exact mode; four sampled cues
    # (first, second, middle, last) against single-cue oracle runs with the
    # same frozen weights and draws: spikes of every step, winners, counts
    spec, s = trained100
    pats = spec.stim[:100].astype(np.int64)
    B, T = 8192, 50
    clamp, wmask, want = R.cues(pats, B, 0.4, seed=11)
    rc = R.Recall.from_solver(s, spec, B, R.EXACT, seed=77)
    idx, cnt, ras, _ = rc.run(clamp, wmask, T, count_from=0, raster=True)
    rc.close()
    ii, di = s.vector("ii_0"), s.vector("di_0")
    wij, bj = s.vector("wij_0"), s.vector("bj_0")
    H, M = spec.pops[0].hcu, spec.pops[0].mcu
    spiking = 0
    for b in (0, 1, B // 2, B - 1):
        ospec = R.oracle_recall_spec(spec, clamp[b], wmask[b], T, 77 + b, ii, di)
        net = O.OracleNet(ospec)
        net.write("wij_0", wij)
        net.write("bj_0", bj)
        counts = np.zeros(H * M, np.int64)
        for t in range(T):
            assert net.step() == 1
            sp = net.vector("spike_0")
            assert np.array_equal(sp, ras[t, b]), "cue %d diverges at step %d" % (b, t + 1)
            counts += sp
        spiking += int(counts.sum())
        c = counts.reshape(H, M)
        for h in range(H):
            mi, mc = -1, 0
            for j in range(M):
                if c[h, j] > mc:
                    mi, mc = j, int(c[h, j])
            assert (idx[b, h], cnt[b, h]) == (mi, mc)
        net.close()
    assert spiking > 0
    del ras


@pytest.mark.parametrize("M", [40, 100, 200])
def test_recall_mfma_population_step_matches_exact(M):
    # with integer weights every support sum is exact in float and in the
    # bfloat16 matrix product, and kepsc = 0 keeps the decay exact, so the
    # mfma mode's support equals the exact mode's; its one-wave-per-(HCU,
    # cue) population step (1, 2 and 4 units per lane) must then give the
    # exact mode's spikes, winners and counts bit for bit.  37 cues: the last
    # workgroup of four (HCU, cue) pairs is partly empty.
    H, B, T = 6, 37, 30
    C = (H - 1) * M
    rng = np.random.default_rng(M)
    ii = np.concatenate([np.array([u for u in range(H * M) if u // M != h], np.int32)
                         for h in range(H)])
    di = np.ones(H * C, np.int32)
    wij = rng.integers(-2, 3, size=H * C * M).astype(np.float32)
    bj = rng.uniform(-1.0, 0.0, size=H * M).astype(np.float32)
    params = dict(taumdt=0.1, wtagain=4.0, maxfqdt=0.1, igain=1.0, wgain=0.05, lgbias=0.0,
                  snoise=0.3, kepsc=0.0, lg_on=float(np.log(np.float32(1.001))),
                  lg_off=float(np.log(np.float32(0.001))))
    clamp = np.where(rng.random((B, H)) < 0.5, rng.integers(0, M, (B, H)), -1).astype(np.int16)
    wmask = rng.random((B, H)).astype(np.float32)
    out = {}
    for mode in (R.EXACT, R.MFMA):
        rc = R.Recall(H, M, C, ii, di, wij, bj, params, B, mode, seed=5)
        out[mode] = rc.run(clamp, wmask, T, count_from=0, raster=True)
        rc.close()
    (ie, ce, re_, _), (im, cm, rm, _) = out[R.EXACT], out[R.MFMA]
    spikes = int(re_.sum())
    assert 0 < spikes < re_.size
    for t in range(T):
        assert np.array_equal(re_[t], rm[t]), "step %d differs" % (t + 1)
    assert np.array_equal(ie, im) and np.array_equal(ce, cm)

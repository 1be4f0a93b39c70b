"""The kernel-level C ABI with caller-owned buffers against the oracle (-m gpu).

gsbn_pop_update (the population update of Pop.cpp:220-409 and the WTA of
Pop.cu:73-79) is called through ctypes on device buffers the test
allocates, outside any solver, and compared bit for bit with the oracle's
population step (gor_updlazy_step on a network of one population and no
projection) on the same inputs: a selected stimulus row and mask row,
random dsup and counters, the per-HCU and the population-seeded WTA maximum
(whose workspace the caller passes), external spikes, ragged sizes, three
steps in a row with the state carried in the caller's buffers.
"""

import ctypes

import numpy as np
import pytest

import oracle as O
from gsbn_amd import lib

from abi import PopArgs

pytestmark = pytest.mark.gpu

DT = 0.001
POP = dict(taum=0.01, wtagain=4.0, maxfq=100.0, igain=1.0, wgain=1.0, lgbias=0.5, snoise=0.3)
SEED = 0x5EED0000F00D


def _oracle(H, M, wta_pop_max, lginp, wmask, dsup, counter, ext):
    L = O.lib()
    cfg = O.PopCfg(hcu=H, mcu=M, **POP)
    n = L.gor_net_new(DT, 1, ctypes.byref(cfg), 0, None, SEED, wta_pop_max, 0)  # table arithmetic
    assert n
    L.gor_set_inputs(n, O._p(lginp), lginp.shape[0], O._p(wmask), wmask.shape[0])
    assert L.gor_vector_write(n, b"dsup_0", O._p(dsup), dsup.size) == 0
    assert L.gor_vector_write(n, b".counter_0", O._p(counter), counter.size) == 0
    if ext is not None:
        off, units = ext
        L.gor_set_ext_spikes(n, 0, off.size - 1, O._p(off), O._p(units))
    return n


def _oracle_vec(n, name, dtype, count):
    L = O.lib()
    out = np.empty(count, dtype)
    t = ctypes.c_int()
    assert L.gor_vector(n, name.encode(), ctypes.byref(t), O._p(out), count) == count
    return out


@pytest.mark.parametrize("H,M,wta_pop_max,with_ext", [
    (12, 37, 0, False), (12, 37, 1, False), (5, 130, 0, True), (3, 1, 0, False), (40, 64, 1, True)])
def test_pop_update_caller_buffers_match_oracle(H, M, wta_pop_max, with_ext):
    import torch
    print("kernel ABI case H=%d M=%d wta_pop_max=%d ext=%d" % (H, M, wta_pop_max, with_ext),
          flush=True)
    Lg = lib()
    Lg.gsbn_pop_update.argtypes = [ctypes.POINTER(PopArgs), ctypes.c_void_p]
    Lg.gsbn_pop_update.restype = ctypes.c_int
    rng = np.random.default_rng(H * 1000 + M)
    N = H * M
    lginp = np.where(rng.random((3, N)) < 0.2, np.log(1.001), np.log(0.001)).astype(np.float32)
    wmask = rng.random((2, H)).astype(np.float32)
    dsup = rng.uniform(-2.0, 2.0, N).astype(np.float32)
    counter = rng.integers(0, 9, N).astype(np.int32)
    ext = None
    if with_ext:
        # steps 0..9; every other step forces a few units to spike
        off, units = [0], []
        for s in range(10):
            if s % 2:
                units += sorted(rng.choice(N, size=min(3, N), replace=False).tolist())
            off.append(len(units))
        ext = (np.array(off, np.int32), np.array(units, np.int32))
    ora = _oracle(H, M, wta_pop_max, lginp, wmask, dsup, counter, ext)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    g = {"lginp": t(lginp), "wmask": t(wmask), "dsup": t(dsup), "counter": t(counter),
         "act": torch.zeros(N, dtype=torch.float32, device=dev),
         "spike": torch.zeros(N, dtype=torch.int8, device=dev),
         "rnd_u": torch.zeros(N, dtype=torch.float32, device=dev),
         "rnd_n": torch.zeros(N, dtype=torch.float32, device=dev),
         "spk_list": torch.zeros(N, dtype=torch.int32, device=dev),
         "spk_cnt": torch.zeros(H, dtype=torch.int32, device=dev),
         "scratch": torch.zeros(1, dtype=torch.float32, device=dev)}
    if ext is not None:
        g["ext_off"], g["ext_units"] = t(ext[0]), t(ext[1])
    stream = torch.cuda.current_stream().cuda_stream
    L = O.lib()
    for k, (simstep, lrow, wrow) in enumerate(((3, 1, 1), (4, 2, 0), (5, 0, 1))):
        L.gor_set_globals(ora, simstep, 1.0, 1.0, 1, lrow, wrow)
        L.gor_updlazy_step(ora)
        a = PopArgs()
        a.pop_id, a.H, a.M, a.simstep, a.exec = 0, H, M, simstep, simstep
        a.seed = SEED
        a.taumdt = np.float32(DT) / np.float32(POP["taum"])
        a.wtagain, a.igain, a.wgain = POP["wtagain"], POP["igain"], POP["wgain"]
        a.maxfqdt = np.float32(POP["maxfq"]) * np.float32(DT)
        a.lgbias, a.snoise, a.prn = POP["lgbias"], POP["snoise"], 1.0
        a.wta_pop_max = wta_pop_max
        a.lginp = g["lginp"].data_ptr() + lrow * N * 4
        a.wmask = g["wmask"].data_ptr() + wrow * H * 4
        for name in ("dsup", "act", "spike", "counter", "rnd_u", "rnd_n", "spk_list", "spk_cnt"):
            setattr(a, name, g[name].data_ptr())
        a.scratch_dsup0 = g["scratch"].data_ptr()
        if ext is not None:
            a.ext_off, a.ext_units, a.ext_nsteps = (g["ext_off"].data_ptr(),
                                                   g["ext_units"].data_ptr(), ext[0].size - 1)
        a.nproj_in = a.nproj_out = 0
        a.do_jxx = 1
        assert Lg.gsbn_pop_update(ctypes.byref(a), stream) == 0
        torch.cuda.synchronize()
        for name, key, dtype in (("dsup_0", "dsup", np.float32), ("act_0", "act", np.float32),
                                 ("spike_0", "spike", np.int8), (".counter_0", "counter", np.int32),
                                 (".rnd_uniform01_0", "rnd_u", np.float32),
                                 (".rnd_normal_0", "rnd_n", np.float32)):
            want = _oracle_vec(ora, name, dtype, N)
            got = g[key].cpu().numpy()
            assert got.tobytes() == want.tobytes(), (k, name, np.nonzero(got != want)[0][:5])
        if ext is not None:
            # the forced units of this step spike, and no other unit does
            forced = ext[1][ext[0][simstep]:ext[0][simstep + 1]]
            assert np.array_equal(np.nonzero(g["spike"].cpu().numpy())[0], np.sort(forced)), k
        # the ascending per-HCU spike lists (.ssj order, Proj.cpp:632-643)
        spk = g["spike"].cpu().numpy().reshape(H, M)
        cnt = g["spk_cnt"].cpu().numpy()
        lst = g["spk_list"].cpu().numpy().reshape(H, M)
        for h in range(H):
            js = np.nonzero(spk[h])[0]
            assert cnt[h] == js.size and np.array_equal(lst[h, :js.size], js), (k, h)
    L.gor_net_free(ora)

"""The population update for callers that cannot allocate
(gsbn_pop_update_noalloc) against the existing entry point (gsbn_pop_update),
both called through the C ABI with caller-owned device buffers.

* CPU: the ctypes mirror of gsbn_pop_args used here has the C layout (every
  field offset and the size, from gcc on the header);
* GPU: two calls in a row, the second with a smaller population, each bit
  for bit against gsbn_pop_update on identical inputs, with the WTA maximum
  seeded from the population's first unit (wta_pop_max), the case that uses
  the workspace; the seed is made to matter (results differ without it).
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from gsbn_amd import lib

from abi import Fanout, PopArgs, PopProj

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
vp = ctypes.c_void_p


def test_pop_args_mirror_matches_c_layout(tmp_path):
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "gsbn_amd.h"',
             "int main(void) {"]
    for cls, cname in ((PopProj, "gsbn_pop_proj"), (Fanout, "gsbn_fanout"),
                       (PopArgs, "gsbn_pop_args")):
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for name, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, name, cname, name))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = {}
    for line in out:
        if line:
            a, b, c = line.split()
            want[(a, b)] = int(c)
    for cls, cname in ((PopProj, "gsbn_pop_proj"), (Fanout, "gsbn_fanout"),
                       (PopArgs, "gsbn_pop_args")):
        assert want[(cname, "size")] == ctypes.sizeof(cls), cname
        for name, _ in cls._fields_:
            assert want[(cname, name)] == getattr(cls, name).offset, (cname, name)


def test_noalloc_symbol_exported():
    L = lib()
    assert hasattr(L, "gsbn_pop_update_noalloc")


def _population(torch, H, M, seed, dsup0_high):
    """Inputs of one population update: lginp, wmask, dsup, counter (the
    rest are outputs), on cuda:0."""
    rng = np.random.default_rng(seed)
    dev = torch.device("cuda:0")
    N = H * M
    dsup = rng.uniform(-1.0, 1.0, N).astype(np.float32)
    if dsup0_high:
        dsup[0] = 5.0  # the population's first unit seeds every HCU's maximum above zero
    lginp = np.where(rng.random(N) < 0.1, np.log(1.001), np.log(0.001)).astype(np.float32)
    wmask = rng.random(H).astype(np.float32)
    counter = rng.integers(0, 5, N).astype(np.int32)
    t = lambda a: torch.from_numpy(a).to(dev)  # noqa: E731
    return {"lginp": t(lginp), "wmask": t(wmask), "dsup": t(dsup), "counter": t(counter),
            "act": torch.zeros(N, dtype=torch.float32, device=dev),
            "spike": torch.zeros(N, dtype=torch.int8, device=dev),
            "rnd_u": torch.zeros(N, dtype=torch.float32, device=dev),
            "rnd_n": torch.zeros(N, dtype=torch.float32, device=dev),
            "spk_list": torch.full((N,), -7, dtype=torch.int32, device=dev),
            "spk_cnt": torch.full((H,), -7, dtype=torch.int32, device=dev)}


def _args(bufs, H, M, step, wta_pop_max, scratch=None):
    a = PopArgs()
    a.pop_id, a.H, a.M, a.simstep, a.exec = 3, H, M, step, step
    a.seed = 0x1234ABCD5678
    a.taumdt, a.wtagain, a.maxfqdt, a.igain, a.wgain = 0.1, 4.0, 0.1, 1.0, 1.0
    a.lgbias, a.snoise, a.prn = 0.0, 0.3, 1.0
    a.wta_pop_max = wta_pop_max
    for name in ("lginp", "wmask", "dsup", "act", "spike", "counter", "rnd_u", "rnd_n",
                 "spk_list", "spk_cnt"):
        setattr(a, name, bufs[name].data_ptr())
    a.scratch_dsup0 = scratch.data_ptr() if scratch is not None else None
    a.nproj_in = a.nproj_out = 0
    a.do_jxx = 1
    return a


OUTS = ("dsup", "act", "spike", "counter", "rnd_u", "rnd_n", "spk_list", "spk_cnt")


@pytest.mark.gpu
def test_pop_update_noalloc_matches_pop_update():
    import torch
    L = lib()
    L.gsbn_pop_update.argtypes = [ctypes.POINTER(PopArgs), vp]
    L.gsbn_pop_update.restype = ctypes.c_int
    L.gsbn_pop_update_noalloc.argtypes = [ctypes.POINTER(PopArgs), vp]
    L.gsbn_pop_update_noalloc.restype = ctypes.c_int
    stream = torch.cuda.current_stream().cuda_stream
    for H, M, seed, step in ((64, 100, 1, 17), (8, 40, 2, 18)):  # the second call is smaller
        print("noalloc case H=%d M=%d" % (H, M), flush=True)
        ref = _population(torch, H, M, seed, True)
        mine = {k: v.clone() for k, v in ref.items()}
        plain = {k: v.clone() for k, v in ref.items()}
        scratch = torch.full((1,), float("nan"), dtype=torch.float32, device="cuda:0")
        ra = _args(ref, H, M, step, 1, scratch)
        assert L.gsbn_pop_update(ctypes.byref(ra), stream) == 0
        ma = _args(mine, H, M, step, 1)
        assert L.gsbn_pop_update_noalloc(ctypes.byref(ma), stream) == 0
        pa = _args(plain, H, M, step, 0)
        assert L.gsbn_pop_update(ctypes.byref(pa), stream) == 0
        torch.cuda.synchronize()
        for k in OUTS:
            a, b = ref[k].cpu().numpy(), mine[k].cpu().numpy()
            assert a.tobytes() == b.tobytes(), (H, M, k)
        # the workspace value matters: the per-HCU maximum gives other results
        assert ref["act"].cpu().numpy().tobytes() != plain["act"].cpu().numpy().tobytes()
        assert np.isfinite(scratch.cpu().numpy()).all()
        assert int(ref["spk_cnt"].cpu().numpy().min()) >= 0

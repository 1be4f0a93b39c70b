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

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
vp, i32, u64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_float


class PopProj(ctypes.Structure):
    _fields_ = [("contrib", vp), ("pend_cnt", vp), ("C", i32), ("pj", vp), ("ej", vp),
                ("zj", vp), ("taupdt", f32), ("ke", f32), ("kzj", f32), ("kepsc", f32),
                ("kftj", f32), ("bgain", f32), ("eps", f32), ("fut", vp), ("W", i32),
                ("rows_list", vp), ("rows_cnt", vp), ("act_stamp", vp), ("dhist", vp),
                ("dcnt", vp), ("dfl", vp), ("walk_list", vp), ("walk_cnt", vp), ("walk_T", vp),
                ("dK", i32), ("dpad", i32)]


class Fanout(ctypes.Structure):
    _fields_ = [("fwd_off", vp), ("fwd_row", vp), ("fwd_del", vp), ("fut", vp), ("H", i32),
                ("C", i32), ("W", i32), ("pad", i32)]


class PopArgs(ctypes.Structure):
    _fields_ = [("pop_id", i32), ("H", i32), ("M", i32), ("simstep", i32), ("exec", i32),
                ("seed", u64), ("taumdt", f32), ("wtagain", f32), ("maxfqdt", f32),
                ("igain", f32), ("wgain", f32), ("lgbias", f32), ("snoise", f32), ("prn", f32),
                ("wta_pop_max", i32), ("lginp", vp), ("wmask", vp), ("dsup", vp), ("act", vp),
                ("spike", vp), ("counter", vp), ("rnd_u", vp), ("rnd_n", vp), ("epsc", vp),
                ("bj", vp), ("scratch_dsup0", vp), ("ext_off", vp), ("ext_units", vp),
                ("ext_nsteps", i32), ("spk_list", vp), ("spk_cnt", vp), ("nproj_out", i32),
                ("pad_out", i32), ("outs", vp), ("nproj_in", i32), ("do_jxx", i32),
                ("max_slots", i32), ("projs", vp), ("work", vp), ("proj0", PopProj),
                ("out0", Fanout)]


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

"""CPU checks of the boundary: the library loads and exports every entry
point include/gsbn_amd.h declares; the prototxt and StimRawData readers
accept the reference's own files; the workload writer round-trips, with the
ProcUpdLazy step-plan keys; bench.py's --plan maps onto them."""

import ctypes
import os
import re

import numpy as np
import pytest

import gsbn_amd
from gsbn_amd import configs

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def declared_functions():
    text = open(gsbn_amd.INCLUDE).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsbn_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_declared_symbols():
    L = gsbn_amd.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_reference_stimulus_fixture_decodes():
    # data/stimuli_10x10_100_-1_rnd.bin: 100 patterns + 100 cues with HCU 9
    # unclamped; mask rows: training (all 0) and recall (HCU 9 only)
    L = gsbn_amd.lib()
    L.gsbn_stim_read.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                 ctypes.c_void_p, ctypes.c_int64]
    dims = np.zeros(4, np.int32)
    data = np.zeros(4000, np.uint32)
    mask = np.zeros(100, np.float32)
    path = os.path.join(GOLDEN, "stimuli_10x10_100_-1_rnd.bin")
    rc = L.gsbn_stim_read(path.encode(), dims.ctypes.data, data.ctypes.data, data.size,
                          mask.ctypes.data, mask.size)
    assert rc == 0
    assert list(dims) == [200, 10, 2, 10]
    d = data[:2000].reshape(200, 10)
    assert (d[:100] < 10).all()
    assert (d[100:, 9] == 0x7FFFFFFF).all() and (d[100:, :9] == d[:100, :9]).all()
    m = mask[:20].reshape(2, 10)
    assert (m[0] == 0).all() and list(m[1]) == [0] * 9 + [1]
    # the Python decoder agrees
    pd, pm = configs.decode_stim(open(path, "rb").read())
    assert (pd == d).all() and (pm == m).all()


def test_reference_prototxt_parses(tmp_path):
    # data/10x10_rnd.prototxt with its stimulus path pointed at the fixture
    text = open(os.path.join(GOLDEN, "10x10_rnd.prototxt")).read()
    text = text.replace("../data/stimuli_10x10_100_-1_rnd.bin",
                        os.path.join(GOLDEN, "stimuli_10x10_100_-1_rnd.bin"))
    p = tmp_path / "net.prototxt"
    p.write_text(text)
    L = gsbn_amd.lib()
    L.gsbn_config_check.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    out = np.zeros(4, np.int32)
    assert L.gsbn_config_check(str(p).encode(), out.ctypes.data) == 0, L.gsbn_last_error()
    assert list(out) == [1, 1, 2, 7]


def test_bad_configs_fail_cleanly(tmp_path):
    L = gsbn_amd.lib()
    L.gsbn_config_check.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    out = np.zeros(4, np.int32)
    p = tmp_path / "bad.prototxt"
    p.write_text("gen_param { dt: 0.001 stim_file: \"/nonexistent.bin\" }\nnet_param { }\n")
    assert L.gsbn_config_check(str(p).encode(), out.ctypes.data) == -2
    assert b"File not found" in L.gsbn_last_error()
    p.write_text("gen_param { dt: 0.001 bogus: 3 }\nnet_param { }\n")
    assert L.gsbn_config_check(str(p).encode(), out.ctypes.data) == -2
    assert b"unknown field bogus" in L.gsbn_last_error()
    # truncated stimulus file
    s = tmp_path / "s.bin"
    s.write_bytes(configs.encode_stim(np.zeros((2, 3), np.uint32), np.zeros((1, 3), np.float32))[:-3])
    p.write_text("gen_param { dt: 0.001 stim_file: \"%s\" }\nnet_param { }\n" % s)
    assert L.gsbn_config_check(str(p).encode(), out.ctypes.data) == -2


def test_written_workload_parses(tmp_path):
    spec = configs.config0()
    path = configs.write(spec, str(tmp_path))
    L = gsbn_amd.lib()
    L.gsbn_config_check.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    out = np.zeros(4, np.int32)
    assert L.gsbn_config_check(path.encode(), out.ctypes.data) == 0, L.gsbn_last_error()
    assert list(out) == [1, 1, len(spec.modes), 5]
    d, m = configs.decode_stim(open(path.replace(".prototxt", "_stim.bin"), "rb").read())
    assert (d == spec.stim).all() and (m == spec.mask).all()


def test_solver_without_device_fails_loudly(tmp_path):
    # on a machine without a GPU the product raises instead of falling back
    import subprocess
    import sys
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    path = configs.write(configs.config0(), str(tmp_path))
    code = "import gsbn_amd; gsbn_amd.Solver(%r)" % path
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=os.path.dirname(HERE), timeout=120)
    assert r.returncode != 0 and "GsbnError" in r.stderr


@pytest.mark.parametrize("field,key,value", [("chained_step", "chained-step", 1),
                                             ("chained_step", "chained-step", -1),
                                             ("rowcol_graded", "rowcol-graded", 2),
                                             ("rowcol_wave", "rowcol-wave", 3),
                                             ("pop_wave", "pop-wave", 0),
                                             ("ovf_keep", "ovf-keep", 3),
                                             ("ovf_cap", "ovf-cap", 4096),
                                             ("fused_step", "fused-step", 1),
                                             ("rowcol_parts", "rowcol-parts", 15),
                                             ("resident_stimulus", "resident-stimulus", 0)])
def test_step_plan_keys_written(tmp_path, field, key, value):
    # the ProcUpdLazy plan options reach the prototxt, and the file still parses
    spec = configs.config0()
    setattr(spec, field, value)
    path = configs.write(spec, str(tmp_path))
    text = open(path).read()
    assert 'argi { key: "%s" val: %d }' % (key, value) in text
    L = gsbn_amd.lib()
    L.gsbn_config_check.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    out = np.zeros(4, np.int32)
    assert L.gsbn_config_check(path.encode(), out.ctypes.data) == 0, L.gsbn_last_error()


def test_bench_plan_option():
    # bench.py --plan maps onto the same spec fields
    import importlib.util
    spec_b = importlib.util.spec_from_file_location("bench_mod", os.path.join(os.path.dirname(HERE),
                                                                              "bench.py"))
    bench = importlib.util.module_from_spec(spec_b)
    spec_b.loader.exec_module(bench)
    s = configs.config0()
    bench.apply_plan(s, "auto")
    assert (s.fused_step, s.chained_step, s.rowcol_parts) == (0, 0, 0)
    bench.apply_plan(s, "chained")
    assert s.chained_step == 1
    s = configs.config0()
    bench.apply_plan(s, "split:15")
    assert s.rowcol_parts == 15 and s.chained_step == -1
    s = configs.config0()
    bench.apply_plan(s, "fused")
    assert s.fused_step == 1
    with pytest.raises(SystemExit):
        bench.apply_plan(configs.config0(), "nonsense")

"""ctypes mirrors of the kernel-level argument structs of include/gsbn_amd.h
(gsbn_pop_proj, gsbn_fanout, gsbn_pop_args), for tests that call the C ABI
with caller-owned device buffers.  tests/test_zz_no_alloc.py checks every
field offset and the sizes against gcc's layout of the header."""

import ctypes

vp, i32, u64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_float


class PopProj(ctypes.Structure):
    _fields_ = [("contrib", vp), ("pend_cnt", vp), ("C", i32), ("pj", vp), ("ej", vp),
                ("zj", vp), ("taupdt", f32), ("ke", f32), ("kzj", f32), ("kepsc", f32),
                ("kftj", f32), ("bgain", f32), ("eps", f32), ("fut", vp), ("W", i32),
                ("rows_list", vp), ("rows_cnt", vp), ("act_stamp", vp), ("dhist", vp),
                ("dcnt", vp), ("dfl", vp), ("walk_list", vp), ("walk_cnt", vp), ("walk_T", vp),
                ("dK", i32), ("bidx", i32)]


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

"""The reference's own shipped configuration, end to end on the GPU (-m gpu).

data/10x10_rnd.prototxt (with its stimulus file data/stimuli_10x10_100_-1_rnd.bin,
both kept byte for byte under tests/golden/) runs the procedure list
ProcExtGen -> ProcUpdLazy -> ProcExchangeSpike -> ProcStructuralPlasticity ->
ProcSpkRec -> ProcSnapshot -> ProcCheck for 10 000 steps: 10 patterns trained
for 500 steps each with structural plasticity growing every connection from
an empty map (no ProcInitConn), then 10 cues recalled for 500 steps each and
scored.  Parameters are the file's own: wtagain 1, snoise 0, taup 60 s,
tauzi/tauzj/taue 5 ms / 5.001 ms / 5.0002 ms (rates within 1e-4 of each
other), tauepsc 50 ms.

The product reads the file unchanged except for the two paths (stimulus file
and record directory).  The oracle reads it with its own parser
(tests/refconfig.py) and runs in lockstep: spikes compared at every step, the
whole Database state every 500 steps and at the end, the ProcCheck outcome,
the raster file ProcSpkRec wrote against the oracle's spikes, and every
snapshot file ProcSnapshot wrote against the oracle's vectors at that step.
"""

import os

import numpy as np
import pytest

import oracle as O
from gsbn_amd import Solver
from gsbn_amd import snapshot as S

import helpers
import refconfig

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
STIM = os.path.join(GOLDEN, "stimuli_10x10_100_-1_rnd.bin")


def _bits(a):
    return a.view(np.uint32) if a.dtype == np.float32 else a


def test_shipped_10x10_rnd_config_lockstep(tmp_path):
    text = open(os.path.join(GOLDEN, "10x10_rnd.prototxt")).read()
    rec = str(tmp_path / "rec")
    assert '"../data/stimuli_10x10_100_-1_rnd.bin"' in text
    assert '"../data/snapshot_10x10_rnd"' in text
    text = text.replace('"../data/stimuli_10x10_100_-1_rnd.bin"', '"%s"' % STIM)
    text = text.replace('"../data/snapshot_10x10_rnd"', '"%s"' % rec)
    path = tmp_path / "10x10_rnd.prototxt"
    path.write_text(text)
    with open(STIM, "rb") as f:
        spec = refconfig.spec_from_prototxt(text, f.read())
    gpu = Solver(str(path))
    ora = O.OracleNet(spec)
    names = helpers.state_names(spec)
    raster = []
    executed = 0
    while True:
        gpu.step(1)
        flag = ora.step()
        t = ora.extgen.simstep
        assert gpu.glob("simstep") == t and gpu.glob("cycle-flag") == flag
        if flag < 0:
            break
        if flag != 1:
            continue
        executed += 1
        so = ora.vector("spike_0")
        sg = gpu.vector("spike_0")
        assert (sg == so).all(), "spikes differ at step %d" % t
        sp = np.nonzero(so > 0)[0]
        if sp.size:
            raster.append(",".join([str(t)] + [str(int(i)) for i in sp]))
        if t % 500 == 0:
            bad = helpers.compare_state(gpu, ora, spec, names)
            assert not bad, (t, bad)
    assert executed == 10000
    bad = helpers.compare_state(gpu, ora, spec, names)
    assert not bad, bad
    # connections were grown (the map starts empty) and the network spiked
    ii = ora.vector("ii_0")
    assert (ii >= 0).sum() > 100 and len(raster) > 1000
    # ProcCheck: outcome of every recall window
    c, tot, flags = gpu.check()
    assert (c, tot) == (ora.check.correct, ora.check.total) and tot == 10
    assert list(flags) == [int(r["ok"]) for r in ora.check.results]
    gpu.close()
    # ProcSpkRec (period 1, offset 0): the raster of every executed step
    lines = open(os.path.join(rec, "ProcSpkRec", "spk_pop_0.csv")).read().splitlines()
    assert lines[0] == "10,10,0.001,100"
    assert lines[1:] == raster
    # ProcSnapshot (period 500, offset 500): 20 files; each holds the state of
    # its step, which a second oracle run reaches again
    snaps = sorted(os.listdir(os.path.join(rec, "ProcSnapshot")))
    assert len(snaps) == 20
    ora2 = O.OracleNet(spec)
    want = {S.snapshot_name(k, spec.dt): k for k in (1000, 5000, 10000)}
    assert set(want) <= set(snaps)
    while ora2.extgen.simstep < 10000:
        ora2.step()
        name = S.snapshot_name(ora2.extgen.simstep, spec.dt)
        if name not in want:
            continue
        _, prn, vecs = S.read_state(os.path.join(rec, "ProcSnapshot", name))
        assert prn == ora2.extgen.prn
        compared = 0
        for n in names:
            if n in vecs:
                assert np.array_equal(_bits(vecs[n][0]), _bits(ora2.vector(n))), (name, n)
                compared += 1
        assert compared >= 20
        assert np.array_equal(vecs["fanout_0"][0], ora2.sp.pops[0]["fanout"]), name
    ora2.close()
    ora.close()

"""Checkpoint / restart through the C ABI (lq_checkpoint_write / _read): a
run resumed from a checkpoint continues bit for bit as the run that wrote
it, in eager and graph stepping, including the contact state, the
displacement record and the motion tables that the reference's .dat/.xml
restart drops (loquat_io.cu:315-468)."""
import numpy as np
import pytest

from retainwall_code_amd import cases, sim
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

FIELDS = ["codes", "cellpos", "orig_id", "pos", "pos_ori", "vel", "idm", "normal", "str", "stn", "cvar", "cmag",
          "accad", "spin", "acestn"]


def bits(d):
    return {k: np.ascontiguousarray(v).view(np.int32) for k, v in d.items()}


@pytest.mark.parametrize("graph", [False, True])
def test_restart_is_bit_exact(tmp_path, graph):
    case = cases.config("small")  # rotating wall: contact forces and slip flags are live
    path = tmp_path / "run.ckpt"
    say(f"[checkpoint] case=small graph={graph}")
    a = sim.open_case(case)
    a.step(5)
    a.sync()
    a.checkpoint_write(path)
    a.step(7, graph=graph)
    a.sync()
    ref = bits(a.snapshot(FIELDS))
    t_ref = a.time()
    mv_ref = a.get_motion(len(case.motion))
    f_ref = a.wall_force(case.probe_box)
    a.close()
    assert np.any(ref["cmag"] != 0)  # contact state present
    b = sim.open_case(case)
    b.step(2)  # a different state first: the read must replace all of it
    b.checkpoint_read(path)
    b.step(7, graph=graph)
    b.sync()
    got = bits(b.snapshot(FIELDS))
    for k in FIELDS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    assert b.time() == t_ref
    mv = b.get_motion(len(case.motion))
    for x, y in zip(mv, mv_ref):
        assert bytes(x) == bytes(y)
    assert b.wall_force(case.probe_box) == f_ref
    b.close()


def test_restart_rejects_another_shape(tmp_path):
    path = tmp_path / "tiny.ckpt"
    a = sim.open_case(cases.config("tiny"))
    a.step(1)
    a.checkpoint_write(path)
    a.close()
    b = sim.open_case(cases.config("small"))
    with pytest.raises(sim.LqError, match="another shape"):
        b.checkpoint_read(path)
    bad = tmp_path / "junk.ckpt"
    bad.write_bytes(b"not a checkpoint" * 8)
    with pytest.raises(sim.LqError, match="not a checkpoint"):
        b.checkpoint_read(bad)
    b.close()

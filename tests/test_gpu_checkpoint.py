"""Exact checkpoint (vx_case_save / vx_case_load): a case restored from a
checkpoint continues bit for bit like the case that wrote it, on the direct
FP64 path (C0's cylinder, tail slots included) and on the mixed tree path
(whose diffusion carries its domain-scale target hint across steps); the
caller's blob comes back unchanged; a checkpoint of another case is refused
with error 2.  The reference restarts from %lf Kadr text (src/main.c:558-605),
which this replaces with an exact file."""
import ctypes as C

import numpy as np
import pytest

from vortex_amd import geometry as geo
from vortex_amd._lib import lib
from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
from vortex_amd.state import uniform_wake

pytestmark = pytest.mark.gpu
VINF = (1.0, 0.0)


def run_pair(prm, pan, start, k1, k2, path):
    a = Case(prm, pan)
    a.build_matrix()
    if start is not None:
        a.set_state(*start)
    for _ in range(k1):
        a.time_step(VINF)
    blob = b"step=%d" % k1
    a.save(path, blob)
    fa = np.array([a.time_step(VINF) for _ in range(k2)])
    sa = a.state()
    b = Case(prm, pan)
    b.build_matrix()
    assert b.load(path) == blob
    fb = np.array([b.time_step(VINF) for _ in range(k2)])
    sb = b.state()
    assert a.n == b.n and a.d_g == b.d_g
    assert np.array_equal(fa, fb), np.max(np.abs(fa - fb))
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    a.close()
    b.close()


def test_checkpoint_direct_fp64(tmp_path):
    pan = geo.circle(200, 0.5)
    prm = make_params(pan, dt=0.01, nu=1e-3, r_col_diff_sign=1.0, r_col_same_sign=0.3, birth_offset_rel=0.5,
                      precision=VX_PREC_F64)
    print("[ckpt] direct f64", flush=True)
    run_pair(prm, pan, None, 12, 8, tmp_path / "c0.ckpt")


def test_checkpoint_tree_mixed(tmp_path):
    pan = geo.circle(1000, 0.5)
    x, y, g = uniform_wake(200000, seed=4)
    prm = make_params(pan, dt=0.005, nu=1e-3, ve_size=float(np.pi / 1000), precision=VX_PREC_MIXED,
                      tree_min_n=65536)
    print("[ckpt] tree mixed", flush=True)
    run_pair(prm, pan, (x, y, g), 2, 3, tmp_path / "tree.ckpt")


def test_checkpoint_of_another_case_is_refused(tmp_path):
    pan = geo.circle(64, 0.5)
    a = Case(make_params(pan, precision=VX_PREC_F64), pan)
    a.build_matrix()
    a.time_step(VINF)
    path = str(tmp_path / "a.ckpt")
    a.save(path)
    other = Case(make_params(geo.circle(80, 0.5), precision=VX_PREC_F64), geo.circle(80, 0.5))
    got = C.c_int64()
    assert lib().vx_case_load(other._h, path.encode(), None, 0, C.byref(got)) == 2
    a.close()
    other.close()


CONF = """steps {steps}
saving_step 10
dt 0.01
inc_step 8192
pr_file profile.txt
viscosity 0.001
rho 1.0
v_inf_x 1.0
v_inf_y 0.0
v_inf_incr_steps 5
n_of_points 20
x_max 10.0
x_min -2.0
y_max 10
y_min -10
ve_size {eps}
n_col 2
r_col_diff_sign 1
r_col_same_sign 0.3
max_ve_g 0.00628162
rc_x 0.0
rc_y 0.0
matrix_load 0
birth_offset 0.5
checkpoint_step 10
{extra}"""


def test_cli_restart_from_checkpoint_continues_the_run(tmp_path):
    """The `vortex` CLI with checkpoint_step 10: a 20-step run, against a
    10-step run restarted from its step-10 checkpoint for 10 more steps (the
    free-stream ramp spans the first 5 steps).  The snapshots and the probe
    file written after the restart are byte for byte the same, and so is
    every force line."""
    import os
    import subprocess
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "vortex_amd", "vortex")
    pan = geo.circle(200, 0.5)
    eps = repr(float(pan["length"][0]))
    runs = {}
    for name, parts in (("whole", [(20, "")]),
                        ("split", [(10, ""), (10, "checkpoint_file output/checkpoints/Ckpt000010.bin\n")])):
        d = tmp_path / name
        d.mkdir()
        geo.write_profile(str(d / "profile.txt"), pan)
        for steps, extra in parts:
            (d / "run.conf").write_text(CONF.format(steps=steps, eps=eps, extra=extra))
            print(f"[ckpt] cli {name}: steps {steps} {extra.strip()}", flush=True)
            out = subprocess.run([cli, "-c", "run.conf"], cwd=str(d), capture_output=True, text=True, timeout=300)
            assert out.returncode == 0, out.stdout + out.stderr
        lines = {}
        for ln in (d / "output" / "F_p.txt").read_text().splitlines():
            lines[int(ln.split()[0])] = ln  # a restart logs its start step again
        runs[name] = (d, lines)
    (da, fa), (db, fb) = runs["whole"], runs["split"]
    assert sorted(fa) == list(range(21)) and fa == fb
    for rel in ("output/kadrs/Kadr000020.txt", "output/kadrs/Kadr000010.txt", "output/vels/contr_Vel000010.txt"):
        assert (da / rel).read_bytes() == (db / rel).read_bytes(), rel

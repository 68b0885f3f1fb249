"""Drop-in run surface: the `vortex` CLI (run files, profiles, snapshots,
force history) and the reference-named C entry points (vortex_compat.h).

CPU tests parse run files without touching a GPU (`vortex -c FILE -p`); the
GPU tests run the CLI and a C driver written against the reference's API and
compare them with the native C-ABI on the same case."""
import json
import os
import subprocess

import numpy as np
import pytest

from vortex_amd import geometry as geo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "vortex_amd", "vortex")
DRIVER = os.path.join(ROOT, "vortex_amd", "compat_driver")
DRIVER_NT = os.path.join(ROOT, "vortex_amd", "compat_driver_nt")  # built with NO_TREE (24-byte Vortex)
REF = "/root/reference"

RUN_CONF = """steps {steps}                     # number of steps
saving_step {save}
dt 0.01
inc_step 8192
pr_file profile.txt     # panels
viscosity 0.001
rho 1.0
v_inf_x 1.0
v_inf_y 0.0
v_inf_incr_steps 00
n_of_points 20
x_max 10.0
x_min -2.0
y_max 10
y_min -10
ve_size {eps}
n_col 2
h_col_x 0.1
h_col_y 0.1
r_col_diff_sign 0.3
r_col_same_sign 0.3
max_ve_g {maxg}
rc_x 0.0
rc_y 0.0
matrix_load 0
timings_file output/timings.txt
precision {prec}
unknown_key_is_ignored 42
"""


def write_case(tmp, P=64, steps=6, save=2, prec="f64"):
    pan = geo.circle(P, 0.5)
    geo.write_profile(os.path.join(tmp, "profile.txt"), pan)
    eps = float(pan["length"][0])
    with open(os.path.join(tmp, "run.conf"), "w") as f:
        f.write(RUN_CONF.format(steps=steps, save=save, eps=repr(eps), maxg=repr(2 * eps), prec=prec))
    return pan, eps


def parse(conf, cwd):
    out = subprocess.run([CLI, "-c", conf, "-p"], cwd=cwd, capture_output=True, text=True, timeout=60)
    return out.returncode, json.loads(out.stdout)


def test_cli_parses_run_file_and_profile(tmp_path):
    pan, eps = write_case(str(tmp_path))
    rc, js = parse("run.conf", str(tmp_path))
    assert rc == 0 and js["profile_ok"] == 1 and js["P"] == 64
    p = js["params"]
    assert js["steps"] == 6 and js["saving_step"] == 2 and js["v_inf_incr_steps"] == 0
    assert p["ve_size"] == eps and p["max_ve_g"] == 2 * eps and p["dt"] == 0.01
    assert (p["x_min"], p["x_max"], p["y_min"], p["y_max"]) == (-2.0, 10.0, -10.0, 10.0)
    assert (p["r_col_diff_sign"], p["r_col_same_sign"], p["n_col"]) == (0.3, 0.3, 2)
    bx = geo.profile_bbox(pan)
    assert (p["px_min"], p["px_max"], p["py_min"], p["py_max"]) == pytest.approx(bx, abs=1e-15)


def test_cli_usage_and_missing_files(tmp_path):
    out = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "Usage" in out.stdout
    out = subprocess.run([CLI, "-c", "nope.conf", "-p"], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert out.returncode == 1


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference run files not present")
@pytest.mark.parametrize("name", ["krug_1000.conf", "krug_1000_1.conf", "plas_2576.conf", "plas_2576_1.conf"])
def test_cli_parses_reference_run_files(name):
    """The reference's own run files: every key read as src/main.c:69-115 reads it."""
    rc, js = parse(name, REF)
    want = {}
    for line in open(os.path.join(REF, name), encoding="utf-8", errors="replace"):
        tok = line.split()
        if len(tok) >= 2 and not line.startswith("#"):
            want.setdefault(tok[0], tok[1])
    p = js["params"]
    assert js["pr_file"] == want["pr_file"] and js["steps"] == int(want["steps"])
    assert p["dt"] == float(want["dt"]) and p["ve_size"] == float(want["ve_size"])
    assert p["max_ve_g"] == float(want["max_ve_g"]) and p["n_col"] == int(want["n_col"])
    assert p["r_col_diff_sign"] == float(want["r_col_diff_sign"])
    assert js["kadr_file"] == want.get("kadr_file", "")
    # the krug profile ships with the reference; the plas run files name a missing one
    assert js["profile_ok"] == (1 if "krug" in name else 0)
    if "krug" in name:
        assert js["P"] == 1000


def write_points(path, x, y, g):
    with open(path, "w") as f:
        f.write(f"{len(x)}\n")
        for i in range(len(x)):
            f.write(f"{i} 0.008 {float(x[i])!r} {float(y[i])!r} 0 0 0 {float(g[i])!r}\n")


def tracer_points(n=40):
    th = np.linspace(0.3, 2 * np.pi - 0.3, n)
    r = np.linspace(0.51, 0.9, n)
    return -0.6 + 0.0 * th + r * np.cos(th) * 0.2, r * np.sin(th), np.ones(n)


def test_cli_second_component_requires_rel_t(tmp_path):
    """src/main.c:480-484: a second component without rel_t stops the run;
    steady_flow without one too (:506-509).  Both before any GPU work."""
    tmp = str(tmp_path)
    write_case(tmp)
    write_points(os.path.join(tmp, "pts.txt"), *tracer_points())
    with open(os.path.join(tmp, "run.conf"), "a") as f:
        f.write("second_points_file pts.txt\nrel_t 0\n")
    out = subprocess.run([CLI, "-c", "run.conf"], cwd=tmp, capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "rel_t is nessesary" in out.stdout
    write_case(tmp)
    with open(os.path.join(tmp, "run.conf"), "a") as f:
        f.write("steady_flow 1\n")
    out = subprocess.run([CLI, "-c", "run.conf"], cwd=tmp, capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "steady_flow and no second component" in out.stdout


# ---------------------------------------------------------------- GPU# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_cli_run_writes_reference_outputs(tmp_path):
    from vortex_amd.solver import Case, make_params
    tmp = str(tmp_path)
    pan, eps = write_case(tmp, steps=6, save=2)
    out = subprocess.run([CLI, "-c", "run.conf"], cwd=tmp, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    kd = os.path.join(tmp, "output", "kadrs")
    assert sorted(os.listdir(kd)) == ["Kadr000000.txt", "Kadr000002.txt", "Kadr000004.txt", "Kadr000006.txt"]
    fp = np.loadtxt(os.path.join(tmp, "output", "F_p.txt"))
    assert fp.shape == (7, 4) and np.all(fp[0, 1:] == 0)  # forces are written one step late
    assert os.path.exists(os.path.join(tmp, "output", "vels", "contr_Vel000002.txt"))
    assert len(open(os.path.join(tmp, "output", "timings.txt")).readlines()) == 7
    # the same case through the C-ABI
    c = Case(make_params(pan, dt=0.01, ve_size=eps, max_ve_g=2 * eps, r_col_diff_sign=0.3, r_col_same_sign=0.3),
             pan)
    c.build_matrix()
    f = [c.time_step((1.0, 0.0)) for _ in range(6)]
    assert np.allclose(fp[1:, 1], [q[1] for q in f], rtol=1e-6, atol=1e-6)
    x, y, g = c.state()
    snap = np.loadtxt(os.path.join(kd, "Kadr000006.txt"), skiprows=1)
    assert snap.shape[0] == c.n
    assert np.allclose(snap[:, 2], x, atol=1e-6) and np.allclose(snap[:, 7], g, atol=1e-6)


@pytest.mark.gpu
def test_cli_restart_from_snapshot(tmp_path):
    """A run restarted from Kadr000004 continues from step 4 (src/main.c:594-605)."""
    tmp = str(tmp_path)
    write_case(tmp, steps=6, save=2)
    assert subprocess.run([CLI, "-c", "run.conf"], cwd=tmp, capture_output=True, timeout=300).returncode == 0
    os.rename(os.path.join(tmp, "output", "kadrs", "Kadr000004.txt"), os.path.join(tmp, "Kadr000004.txt"))
    with open(os.path.join(tmp, "run.conf"), "a") as f:
        f.write("kadr_file Kadr000004.txt\nsteps 2\nmatrix_load 1\n")
    out = subprocess.run([CLI, "-c", "run.conf"], cwd=tmp, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout
    fp = np.loadtxt(os.path.join(tmp, "output", "F_p.txt"))
    assert list(fp[:, 0]) == [4, 5, 6]


@pytest.mark.gpu
@pytest.mark.parametrize("prec,driver", [("f64", DRIVER), ("mixed", DRIVER), ("f64", DRIVER_NT)])
def test_compat_driver_matches_native_api(tmp_path, prec, driver):
    """The compat driver (the reference's loop on the unit_1.cuh calls), built
    with the default 32-byte Vortex and with NO_TREE's 24-byte one, against
    the native API step by step."""
    from vortex_amd.solver import Case, make_params, VX_PREC_F64, VX_PREC_MIXED
    tmp = str(tmp_path)
    pan = geo.circle(64, 0.5)
    geo.write_profile(os.path.join(tmp, "profile.txt"), pan)
    steps = 8
    args = [driver, "profile.txt", str(steps)] + (["mixed"] if prec == "mixed" else [])
    out = subprocess.run(args, cwd=tmp, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rows = np.array([[float(v) for v in line.split()] for line in out.stdout.strip().splitlines()
                     if not line.startswith(("first", "second"))])
    eps = float(pan["length"][0])
    c = Case(make_params(pan, dt=0.01, ve_size=eps, max_ve_g=2 * eps, r_col_diff_sign=0.3, r_col_same_sign=0.3,
                         precision=VX_PREC_MIXED if prec == "mixed" else VX_PREC_F64), pan)
    c.build_matrix()
    for k in range(steps):
        f = c.time_step((1.0, 0.0))
        assert rows[k, 1] == c.n, k
        assert np.allclose(rows[k, 2:], f[1:], rtol=1e-12, atol=1e-12), (k, rows[k], f)
    assert os.path.exists(os.path.join(tmp, "M.txt"))  # matr_creation writes it, as the reference does


@pytest.mark.gpu
def test_cli_second_component_matches_native_api(tmp_path):
    """Inertial and passive tracer points (src/main.c:729-741): the CLI's point
    snapshots and velocity files against the same loop on the native API."""
    from vortex_amd.solver import Case, make_params, points_relax
    tmp = str(tmp_path)
    pan, eps = write_case(tmp, steps=4, save=2)
    px, py, pg = tracer_points()
    write_points(os.path.join(tmp, "pts.txt"), px, py, pg)
    with open(os.path.join(tmp, "run.conf"), "a") as f:
        f.write("second_points_file pts.txt\nrel_t 0.1\n")
    out = subprocess.run([CLI, "-c", "run.conf"], cwd=tmp, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    for sub in ("first", "second"):
        assert sorted(os.listdir(os.path.join(tmp, "output", sub))) == [
            "Kadr000000.txt", "Kadr000002.txt", "Kadr000004.txt"]
    assert os.path.exists(os.path.join(tmp, "output", "vels", "Vel000003_4.txt"))
    # the same loop on the native API
    c = Case(make_params(pan, dt=0.01, ve_size=eps, max_ve_g=2 * eps, r_col_diff_sign=0.3, r_col_same_sign=0.3),
             pan)
    c.build_matrix()
    fx, fy, fgv = px.copy(), py.copy(), pg.copy()
    sx, sy, sg = px.copy(), py.copy(), pg.copy()
    su = sv = None
    for k in range(4):
        c.vort_creation((1.0, 0.0))
        ue, ve = c.velocity_probe(sx, sy, (1.0, 0.0))
        if su is None:
            su, sv = ue.copy(), ve.copy()
        if k == 3:
            v4 = np.loadtxt(os.path.join(tmp, "output", "vels", "Vel000003_4.txt"), skiprows=1)
            assert np.allclose(v4[:, 3], ue, atol=2e-6) and np.allclose(v4[:, 1], sx, atol=2e-6)
        su, sv = points_relax(sg, ue, ve, su, sv, 0.1, 0.01)
        sx, sy, sg = c.points_move(sx, sy, sg, su, sv)
        uf, vf = c.velocity_probe(fx, fy, (1.0, 0.0))
        fx, fy, fgv = c.points_move(fx, fy, fgv, uf, vf)
        c.speed((1.0, 0.0))
        c.step()
    for sub, (x, y, g) in (("first", (fx, fy, fgv)), ("second", (sx, sy, sg))):
        snap = np.loadtxt(os.path.join(tmp, "output", sub, "Kadr000004.txt"), skiprows=1)
        assert np.allclose(snap[:, 2], x, atol=2e-6) and np.allclose(snap[:, 3], y, atol=2e-6)
        assert np.array_equal(snap[:, 7], g)
    assert not np.allclose(fx, sx)  # inertia separates the two copies


@pytest.mark.gpu
@pytest.mark.parametrize("driver", [DRIVER, DRIVER_NT])
def test_compat_second_component_matches_native_api(tmp_path, driver):
    """second_speed / first_speed through the reference-named entry points
    (32- and 24-byte Vortex) against the same loop on the native API (FP64)."""
    from vortex_amd.solver import Case, make_params, points_relax
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "output", "vels"))
    pan = geo.circle(64, 0.5)
    geo.write_profile(os.path.join(tmp, "profile.txt"), pan)
    px, py, pg = tracer_points(24)
    write_points(os.path.join(tmp, "pts.txt"), px, py, pg)
    steps = 5
    out = subprocess.run([driver, "profile.txt", str(steps), "f64", "pts.txt"], cwd=tmp, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    pts = {"first": [], "second": []}
    for line in out.stdout.splitlines():
        tok = line.split()
        if tok[0] in pts:
            pts[tok[0]].append([float(v) for v in tok[1:]])
    eps = float(pan["length"][0])
    c = Case(make_params(pan, dt=0.01, ve_size=eps, max_ve_g=2 * eps, r_col_diff_sign=0.3, r_col_same_sign=0.3),
             pan)
    c.build_matrix()
    fx, fy, fgv = px.copy(), py.copy(), pg.copy()
    sx, sy, sg = px.copy(), py.copy(), pg.copy()
    su = sv = None
    for k in range(steps):
        c.vort_creation((1.0, 0.0))
        ue, ve = c.velocity_probe(sx, sy, (1.0, 0.0))
        if su is None:
            su, sv = ue.copy(), ve.copy()
        su, sv = points_relax(sg, ue, ve, su, sv, 0.1, 0.01)
        sx, sy, sg = c.points_move(sx, sy, sg, su, sv)
        uf, vf = c.velocity_probe(fx, fy, (1.0, 0.0))
        fx, fy, fgv = c.points_move(fx, fy, fgv, uf, vf)
        c.speed((1.0, 0.0))
        c.step()
    for name, (x, y, g) in (("first", (fx, fy, fgv)), ("second", (sx, sy, sg))):
        a = np.array(pts[name])
        assert a.shape == (len(px), 3)
        assert np.allclose(a[:, 0], x, rtol=0, atol=1e-12) and np.allclose(a[:, 1], y, rtol=0, atol=1e-12)
        assert np.array_equal(a[:, 2], g)
    assert os.path.exists(os.path.join(tmp, "output", "vels", "Vel000004_4.txt"))

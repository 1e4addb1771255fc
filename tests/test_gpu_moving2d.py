"""2D moving meshes on the device: the step's Voronoi rebuild (Space::volume_calculation, space.rs:236-263)
and the device engine loop with drift (engine.rs:152-192), through the C-ABI.

The device rebuild reproduces the host tessellator (the stand-in for meshless_voronoi) bit for bit: faces,
their order, the CSR and the cell geometry.  Whole runs match the oracle, whose steps re-tessellate with
the host tessellator, within 1e-9 (FP64; bins and step counts exact) and L1(rho) < 1e-5 (FP32 flux).
"""
import dataclasses
import sys

import numpy as np
import pytest

from hydro_rs_amd import fvs, ics
from hydro_rs_amd.mesh import VoronoiMesh
from tests.helpers import (advected_wave, compare_states, oracle_sim, reflections, star_problem, wall_box,
                           wrapped_fraction)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")


def moving_sedov(boundary=fvs.PERIODIC, motion=fvs.FLUID, side=24, t_end=0.02, **cfg):
    """The 2D Sedov blast with moving generators; walls get the non-periodic initial mesh."""
    prob = ics.sedov_2d(side=side, t_end=t_end)
    prob = dataclasses.replace(prob, cfg=dataclasses.replace(prob.cfg, motion=motion, boundary=boundary, **cfg))
    if boundary != fvs.PERIODIC:
        prob = dataclasses.replace(prob, mesh=VoronoiMesh(2, prob.cfg.box, False, prob.loc))
    return prob


def host_mesh_of(dev, prob):
    loc = dev.get_state(["loc"])["loc"].reshape(3, prob.n).T.copy()
    return VoronoiMesh(2, prob.cfg.box, prob.cfg.boundary == fvs.PERIODIC, loc).as_dict()


@pytest.mark.parametrize("boundary", [fvs.PERIODIC, fvs.REFLECTIVE])
def test_device_rebuild_equals_host_tessellator(boundary):
    """After steps with every cell active (sync_all), the device mesh is the host tessellation of the
    device's generator positions, array for array."""
    prob = moving_sedov(boundary, sync_all=True)
    dev = ics.build_device_sim(prob)
    for steps in (1, 12):
        say(f"[rebuild] boundary={boundary} after {steps} steps")
        dev.run(steps if steps == 1 else steps - 1)
        got = dev.get_mesh()
        want = host_mesh_of(dev, prob)
        assert got["n_faces"] == want["n_faces"], (got["n_faces"], want["n_faces"])
        for k in ("area", "normal", "centroid", "shift", "left", "right", "offset", "count", "conn", "volume",
                  "cell_centroid", "search_radius"):
            a, b = np.asarray(got[k]), np.asarray(want[k])
            assert a.shape == b.shape and np.array_equal(a, b), (k, np.max(np.abs(a - b)) if a.shape == b.shape else a.shape)
    # the generators moved: the mesh is not the initial one
    assert not np.array_equal(got["area"][: prob.mesh.n_faces], prob.mesh.area[: got["n_faces"]])


@pytest.mark.parametrize("boundary", [fvs.PERIODIC, fvs.REFLECTIVE])
def test_device_rebuild_cell_of_40_edges(boundary):
    """A generator ringed by 40 others on a small circle: its 40-edge cell outgrows the first pass's 32
    edges and is built by the second pass (48 edges), as the host tessellates it."""
    prob = star_problem(moving_sedov(fvs.PERIODIC, side=8, sync_all=True), 40, boundary)
    assert prob.mesh.as_dict()["count"][0] == 40
    dev = ics.build_device_sim(prob)
    say(f"[rebuild star 40] n={prob.n} boundary={boundary}")
    dev.run(1)
    assert dev.engine_info()["error_code"] == 0
    got, want = dev.get_mesh(), host_mesh_of(dev, prob)
    assert got["n_faces"] == want["n_faces"], (got["n_faces"], want["n_faces"])
    for k in ("area", "normal", "centroid", "shift", "left", "right", "offset", "count", "conn", "volume",
              "cell_centroid", "search_radius"):
        a, b = np.asarray(got[k]), np.asarray(want[k])
        assert a.shape == b.shape and np.array_equal(a, b), k


def test_device_rebuild_cell_beyond_capacity_stops_with_mesh_error():
    """A 60-edge cell (more than the 48 a 2D cell holds): the rebuild stops the simulation with
    HFV_ERR_MESH (-22), as hfv_run documents."""
    prob = star_problem(moving_sedov(fvs.PERIODIC, side=8, sync_all=True), 60, fvs.PERIODIC)
    assert prob.mesh.as_dict()["count"][0] == 60
    dev = ics.build_device_sim(prob)
    say(f"[rebuild star 60] n={prob.n}")
    with pytest.raises(fvs.HfvError) as ex:
        dev.run(1)
    assert ex.value.code == -22
    assert dev.engine_info()["error_code"] == -22


@pytest.mark.parametrize("case", ["periodic_fluid", "reflective_fluid", "periodic_steer", "periodic_sync_all"])
def test_moving2d_full_run_parity(ora, case):
    """hfv_run (drift, device rebuild, step) against the oracle's run with host re-tessellation."""
    prob = {"periodic_fluid": lambda: moving_sedov(fvs.PERIODIC),
            "reflective_fluid": lambda: moving_sedov(fvs.REFLECTIVE),
            "periodic_steer": lambda: moving_sedov(fvs.PERIODIC, motion=fvs.STEER),
            "periodic_sync_all": lambda: moving_sedov(fvs.PERIODIC, sync_all=True)}[case]()
    osim = oracle_sim(prob)
    n_ref = osim.advance(1 << 40)
    ref = osim.get_fields()
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(1 << 40)
    assert dev.engine_info()["done"] and dev.engine_info()["error_code"] == 0
    errs = compare_states(dev.get_state(), ref, prob.n)
    say(f"[moving2d {case}] steps ref={n_ref} dev={n_dev} errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs
    moved = np.abs(ref["loc"][: prob.n] - prob.loc[:, 0])
    assert np.max(moved) > 1e-3  # generators moved by more than a tenth of a cell


def test_moving2d_fp32_flux_l1(ora):
    """The FP32-flux context on the moving mesh: L1(rho) against the FP64 oracle below 1e-5."""
    prob = moving_sedov(fvs.PERIODIC)
    ref = oracle_sim(prob)
    ref.advance(1 << 40)
    rho = ref.get_fields()["prim"][: prob.n]
    d32 = ics.build_device_sim(ics.with_precision(prob, fvs.FP32_FLUX))
    d32.run(1 << 40)
    assert d32.engine_info()["done"] and d32.engine_info()["error_code"] == 0
    l1 = float(np.mean(np.abs(d32.get_state(["prim"])["prim"][: prob.n] - rho)))
    say(f"[moving2d fp32] L1(rho) vs oracle {l1:.2e}")
    assert l1 < 1e-5


def test_moving2d_step_equals_run():
    """hfv_step (the caller's TimestepInfo, the caller's drift through hfv_set_state) takes the same device
    rebuild: a step sequence driven from the host equals hfv_run's first steps bit for bit."""
    prob = moving_sedov(fvs.PERIODIC, sync_all=True)
    a = ics.build_device_sim(prob)
    a.run(1)
    b = ics.build_device_sim(prob)
    tb = prob.cfg.t_end / float(1 << 57)
    b.step(b.timestep_info(0, 0, tb))
    sa, sb = a.get_state(), b.get_state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    ma, mb = a.get_mesh(), b.get_mesh()
    for k in ("area", "left", "right", "conn"):
        assert np.array_equal(ma[k], mb[k]), k


def test_moving2d_periodic_wrap_full_run_parity(ora):
    """Generators carried by a uniform flow through every face of the periodic box (k_drift's box_wrap,
    part.rs:449-455, then the rebuild's periodic images): the whole run matches the oracle's within 1e-9,
    bins and step counts exact, and at least a sixth of the generators wrapped on each axis."""
    vel = (1.0, 0.6)
    prob = advected_wave(2, 12, vel, 0.3)
    osim = oracle_sim(prob)
    n_ref = osim.advance(1 << 40)
    ref = osim.get_fields()
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(1 << 40)
    assert dev.engine_info()["done"] and dev.engine_info()["error_code"] == 0
    errs = compare_states(dev.get_state(), ref, prob.n)
    loc = dev.get_state(["loc"])["loc"].reshape(3, prob.n).T
    wrapped = wrapped_fraction(prob.loc, loc, vel)
    say(f"[wrap 2d] steps ref={n_ref} dev={n_dev} wrapped={wrapped} errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs
    assert min(wrapped) >= 1.0 / 6.0 - 1e-12


def test_moving2d_wall_reflection_parity(ora):
    """Generators drifting out through every reflective wall are reflected with their primitives, gradients
    and kept rate (k_drift; box_reflect, reflect_quantities, reflect_gradients) as in the oracle: 9 steps
    (80 reflected coordinates) within the whole-run tolerance, then the reference's negative-pressure
    panic at the same step with the same code."""
    prob = wall_box(2, 10)
    osim = oracle_sim(prob)
    assert reflections(osim, 9, 2) == 80
    dev = ics.build_device_sim(prob)
    assert dev.run(9) == 9
    errs = compare_states(dev.get_state(), osim.get_fields(), prob.n)
    say(f"[wall 2d] 9 steps: errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs
    with pytest.raises(RuntimeError) as oe:
        osim.advance(1 << 40)
    with pytest.raises(fvs.HfvError) as de:
        dev.run(1 << 40)
    assert de.value.code == int(str(oe.value).split()[-1]) == -4
    assert dev.engine_info()["step_count"] == osim.info()["step_count"]


@pytest.mark.parametrize("case", ["wall", "wrap"])
def test_moving2d_drift_through_box_fp32(ora, case):
    """The FP32-flux context on the same drifts (reflective walls for 9 steps, periodic wraps to the end):
    it sums each generator's deferred predicts with the rate it had before a reflection and extrapolates the
    reflected state afterwards; L1(rho) against the FP64 oracle < 1e-5."""
    if case == "wall":
        prob, steps = wall_box(2, 10, precision=fvs.FP32_FLUX), 9
    else:
        prob, steps = advected_wave(2, 12, (1.0, 0.6), 0.3, precision=fvs.FP32_FLUX), 1 << 40
    ref = oracle_sim(dataclasses.replace(prob, cfg=dataclasses.replace(prob.cfg, precision=fvs.FP64)))
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(steps)
    ref.advance(steps)
    assert dev.engine_info()["error_code"] == 0
    if case == "wall":  # the same step of the same timeline (the wraps run to t_end)
        assert dev.engine_info()["ti_current"] == ref.info()["ti_current"]
    rho = ref.get_fields()["prim"][: prob.n]
    l1 = float(np.mean(np.abs(dev.get_state(["prim"])["prim"][: prob.n] - rho)))
    say(f"[2d {case} fp32] {n_dev} steps, L1(rho) vs oracle {l1:.2e}")
    assert l1 < 1e-5

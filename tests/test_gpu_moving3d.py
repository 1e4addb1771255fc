"""3D moving meshes on the device: the step's Voronoi rebuild (Space::volume_calculation, space.rs:236-263)
in 3D and the device engine loop with drift (engine.rs:152-192), through the C-ABI.

The device rebuild reproduces the host tessellator's 3D cells (the stand-in for meshless_voronoi) bit for
bit: faces, their order, the CSR and the cell geometry.  Whole runs match the oracle, whose steps
re-tessellate with the host tessellator, within 1e-9 (FP64; timebins and step counts exact) and
L1(rho) < 1e-5 (FP32 flux).
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

MESH_KEYS = ("area", "normal", "centroid", "shift", "left", "right", "offset", "count", "conn", "volume",
             "cell_centroid", "search_radius")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if fvs.device_count() < 1:
        pytest.skip("no HIP device")


def moving_kh(boundary=fvs.PERIODIC, motion=fvs.FLUID, side=8, t_end=0.05, **cfg):
    """The 3D Kelvin-Helmholtz shear layer with moving generators (jittered lattice); walls get the
    non-periodic initial mesh."""
    prob = ics.kelvin_helmholtz_3d(side=side, t_end=t_end)
    prob = dataclasses.replace(prob, cfg=dataclasses.replace(prob.cfg, motion=motion, boundary=boundary, **cfg))
    if boundary != fvs.PERIODIC:
        prob = dataclasses.replace(prob, mesh=VoronoiMesh(3, prob.cfg.box, False, prob.loc))
    return prob


def host_mesh_of(dev, prob):
    loc = dev.get_state(["loc"])["loc"].reshape(3, prob.n).T.copy()
    return VoronoiMesh(3, prob.cfg.box, prob.cfg.boundary == fvs.PERIODIC, loc).as_dict()


def assert_same_mesh(got, want):
    assert got["n_faces"] == want["n_faces"], (got["n_faces"], want["n_faces"])
    for k in MESH_KEYS:
        a, b = np.asarray(got[k]), np.asarray(want[k])
        assert a.shape == b.shape, (k, a.shape, b.shape)
        assert np.array_equal(a, b), (k, np.max(np.abs(a.astype(np.float64) - b)))


@pytest.mark.parametrize("boundary", [fvs.PERIODIC, fvs.REFLECTIVE])
def test_device_rebuild3d_equals_host_tessellator(boundary):
    """After steps with every cell active (sync_all), the device mesh is the host tessellation of the
    device's generator positions, array for array."""
    prob = moving_kh(boundary, sync_all=True)
    dev = ics.build_device_sim(prob)
    for steps in (1, 6):
        say(f"[rebuild3d] boundary={boundary} after {steps} steps")
        dev.run(1 if steps == 1 else steps - 1)
        assert dev.engine_info()["error_code"] == 0
        assert_same_mesh(dev.get_mesh(), host_mesh_of(dev, prob))


def test_device_rebuild3d_random_generators():
    """Uniform random generators (no lattice: cells with many faces, uneven bins) on a periodic box."""
    rng = np.random.default_rng(11)
    prob = moving_kh(fvs.PERIODIC, sync_all=True, side=7)
    loc = rng.random((prob.n, 3))
    prob = dataclasses.replace(prob, loc=loc, mesh=VoronoiMesh(3, prob.cfg.box, True, loc))
    dev = ics.build_device_sim(prob)
    say(f"[rebuild3d random] n={prob.n}")
    dev.run(2)
    assert dev.engine_info()["error_code"] == 0
    assert_same_mesh(dev.get_mesh(), host_mesh_of(dev, prob))


@pytest.mark.parametrize("boundary", [fvs.PERIODIC, fvs.REFLECTIVE])
def test_device_rebuild3d_exact_lattice(boundary):
    """Generators on an exact cubic lattice, gas at rest (the generators stay put): equal distances break
    ties by index and image code, and bisectors parallel to the walls and to each other meet in no point
    (the clip's fallback to the removed vertex's position)."""
    prob = moving_kh(boundary, sync_all=True, side=6)
    side = 6
    g = (np.arange(side) + 0.5) / side
    loc = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    mesh = VoronoiMesh(3, prob.cfg.box, boundary == fvs.PERIODIC, loc)
    prob = dataclasses.replace(prob, loc=loc, mesh=mesh, vel=np.zeros_like(prob.vel),
                               mass=np.full(prob.n, 1.0 / prob.n), u_int=np.full(prob.n, 1.5))
    dev = ics.build_device_sim(prob)
    say(f"[rebuild3d lattice] n={prob.n} boundary={boundary}")
    dev.run(2)
    assert dev.engine_info()["error_code"] == 0
    got = dev.get_mesh()
    assert_same_mesh(got, host_mesh_of(dev, prob))
    assert np.array_equal(dev.get_state(["loc"])["loc"].reshape(3, prob.n).T, loc)  # at rest


@pytest.mark.parametrize("boundary", [fvs.PERIODIC, fvs.REFLECTIVE])
def test_device_rebuild3d_clustered_generators(boundary):
    """Half the generators in a cube of 1/125 of the box (60x denser than the rest): blocks of several hundred
    candidates, taken in bands of squared distance."""
    rng = np.random.default_rng(12)
    prob = moving_kh(boundary, sync_all=True, side=10)
    loc = np.concatenate([rng.random((prob.n // 2, 3)), 0.4 + 0.2 * rng.random((prob.n - prob.n // 2, 3))])
    prob = dataclasses.replace(prob, loc=loc,
                               mesh=VoronoiMesh(3, prob.cfg.box, boundary == fvs.PERIODIC, loc))
    dev = ics.build_device_sim(prob)
    say(f"[rebuild3d clustered] n={prob.n} boundary={boundary}")
    dev.run(1)
    assert dev.engine_info()["error_code"] == 0
    assert_same_mesh(dev.get_mesh(), host_mesh_of(dev, prob))


@pytest.mark.parametrize("boundary", [fvs.PERIODIC, fvs.REFLECTIVE])
def test_device_rebuild3d_cell_of_40_faces(boundary):
    """A generator ringed by 40 others on a small sphere: its cell has 40 faces (46 of the 64 planes a cell
    holds, with the box's 6 it starts from), and the rebuild gives the host tessellation."""
    prob = star_problem(moving_kh(fvs.PERIODIC, sync_all=True, side=6), 40, boundary)
    assert prob.mesh.as_dict()["count"][0] == 40
    dev = ics.build_device_sim(prob)
    say(f"[rebuild3d star 40] n={prob.n} boundary={boundary}")
    dev.run(1)
    assert dev.engine_info()["error_code"] == 0
    assert_same_mesh(dev.get_mesh(), host_mesh_of(dev, prob))


def test_device_rebuild3d_cell_beyond_capacity_stops_with_mesh_error():
    """A cell of 91 faces needs more planes than a 3D cell holds (64, the box's 6 included): the rebuild
    stops the simulation with HFV_ERR_MESH (-22), as hfv_run documents, and emits no partial cell."""
    prob = star_problem(moving_kh(fvs.PERIODIC, sync_all=True, side=6), 91, fvs.PERIODIC)
    assert prob.mesh.as_dict()["count"][0] == 91
    dev = ics.build_device_sim(prob)
    say(f"[rebuild3d star 91] n={prob.n}")
    with pytest.raises(fvs.HfvError) as ex:
        dev.run(1)
    assert ex.value.code == -22
    assert dev.engine_info()["error_code"] == -22


@pytest.mark.parametrize("case", ["periodic_fluid", "reflective_fluid", "periodic_steer", "periodic_sync_all"])
def test_moving3d_full_run_parity(ora, case):
    """hfv_run (drift, device rebuild of the active cells, step) against the oracle's run with host
    re-tessellation."""
    prob = {"periodic_fluid": lambda: moving_kh(fvs.PERIODIC),
            "reflective_fluid": lambda: moving_kh(fvs.REFLECTIVE),
            "periodic_steer": lambda: moving_kh(fvs.PERIODIC, motion=fvs.STEER),
            "periodic_sync_all": lambda: moving_kh(fvs.PERIODIC, sync_all=True)}[case]()
    osim = oracle_sim(prob)
    n_ref = osim.advance(1 << 40)
    ref = osim.get_fields()
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(1 << 40)
    assert dev.engine_info()["done"] and dev.engine_info()["error_code"] == 0
    errs = compare_states(dev.get_state(), ref, prob.n)
    say(f"[moving3d {case}] steps ref={n_ref} dev={n_dev} errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs
    moved = np.abs(ref["loc"][: prob.n] - prob.loc[:, 0])
    assert np.max(moved) > 1e-3


@pytest.mark.parametrize("boundary", [fvs.PERIODIC, fvs.REFLECTIVE])
def test_moving3d_random_generators_full_run_parity(ora, boundary):
    """A whole run from uniform random generators (irregular cells, blocks beyond the LDS candidate list)
    with steered motion, against the oracle's run."""
    rng = np.random.default_rng(21)
    prob = moving_kh(boundary, motion=fvs.STEER, side=8, t_end=0.03)
    loc = rng.random((prob.n, 3))
    prob = dataclasses.replace(prob, loc=loc,
                               mesh=VoronoiMesh(3, prob.cfg.box, boundary == fvs.PERIODIC, loc))
    osim = oracle_sim(prob)
    n_ref = osim.advance(1 << 40)
    ref = osim.get_fields()
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(1 << 40)
    assert dev.engine_info()["done"] and dev.engine_info()["error_code"] == 0
    errs = compare_states(dev.get_state(), ref, prob.n)
    say(f"[moving3d random {boundary}] steps ref={n_ref} dev={n_dev} errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs


def test_moving3d_fp32_flux_l1(ora):
    """The FP32-flux context on the moving 3D mesh (16^3 generators): L1(rho) against the FP64 oracle below
    1e-5 (measured 7e-7).  Generators move with the fluid velocity at the generator, which takes the
    limited gradient: where an FP32 roundoff flips a limiter decision at the shear layer, the generator
    velocity changes by a fraction of the velocity jump and the mesh follows, so on 8^3 generators the same
    run ends at L1 3e-5 (max |d rho| 2e-3), against 1.6e-8 with fixed generators."""
    prob = moving_kh(fvs.PERIODIC, side=16)
    ref = oracle_sim(prob)
    ref.advance(1 << 40)
    rho = ref.get_fields()["prim"][: prob.n]
    d32 = ics.build_device_sim(ics.with_precision(prob, fvs.FP32_FLUX))
    d32.run(1 << 40)
    assert d32.engine_info()["done"] and d32.engine_info()["error_code"] == 0
    l1 = float(np.mean(np.abs(d32.get_state(["prim"])["prim"][: prob.n] - rho)))
    say(f"[moving3d fp32] L1(rho) vs oracle {l1:.2e}")
    assert l1 < 1e-5


def test_moving3d_batch_member_equals_alone():
    """A 3D moving simulation in a batch (beside a 2D fixed one) is rebuilt in its slot and ends bit for bit
    where it ends alone."""
    moving = moving_kh(fvs.PERIODIC, t_end=0.02)
    alone = ics.build_device_sim(moving)
    alone.run(1 << 40)
    b = fvs.Batch.for_problems([ics.sedov_2d(side=16, t_end=0.01), moving], headroom=1.5)
    b.run(1 << 40)
    assert b.engine_info(1)["done"] and b.engine_info(1)["error_code"] == 0
    assert b.engine_info(1)["step_count"] == alone.engine_info()["step_count"]
    got, want = b.get_state(1), alone.get_state()
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_moving3d_periodic_wrap_full_run_parity(ora):
    """Generators carried by a uniform flow through every face of the periodic box (k_drift's box_wrap,
    part.rs:449-455, then the rebuild's periodic images): the whole run matches the oracle's within 1e-9,
    bins and step counts exact, and at least a sixth of the generators wrapped on each axis."""
    vel = (1.0, 0.6, -0.8)
    prob = advected_wave(3, 6, vel, 0.25)
    osim = oracle_sim(prob)
    n_ref = osim.advance(1 << 40)
    ref = osim.get_fields()
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(1 << 40)
    assert dev.engine_info()["done"] and dev.engine_info()["error_code"] == 0
    errs = compare_states(dev.get_state(), ref, prob.n)
    loc = dev.get_state(["loc"])["loc"].reshape(3, prob.n).T
    wrapped = wrapped_fraction(prob.loc, loc, vel)
    say(f"[wrap 3d] steps ref={n_ref} dev={n_dev} wrapped={wrapped} errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert n_dev == n_ref
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs
    assert min(wrapped) >= 1.0 / 6.0 - 1e-12


def test_moving3d_wall_reflection_parity(ora):
    """Generators drifting out through every reflective wall are reflected with their primitives, gradients
    and kept rate (k_drift; box_reflect, reflect_quantities, reflect_gradients) as in the oracle: 7 steps
    (364 reflected coordinates) within the whole-run tolerance, then the reference's negative-pressure
    panic at the same step with the same code."""
    prob = wall_box(3, 6)
    osim = oracle_sim(prob)
    assert reflections(osim, 7, 3) == 364
    dev = ics.build_device_sim(prob)
    assert dev.run(7) == 7
    errs = compare_states(dev.get_state(), osim.get_fields(), prob.n)
    say(f"[wall 3d] 7 steps: errs={ {k: f'{v:.1e}' for k, v in errs.items()} }")
    assert errs.pop("timebin_mismatch") == 0
    assert max(errs.values()) <= 1e-9, errs
    with pytest.raises(RuntimeError) as oe:
        osim.advance(1 << 40)
    with pytest.raises(fvs.HfvError) as de:
        dev.run(1 << 40)
    assert de.value.code == int(str(oe.value).split()[-1]) == -4
    assert dev.engine_info()["step_count"] == osim.info()["step_count"]


@pytest.mark.parametrize("case", ["wall", "wrap"])
def test_moving3d_drift_through_box_fp32(ora, case):
    """The FP32-flux context on the same drifts (reflective walls for 7 steps, periodic wraps to the end):
    it sums each generator's deferred predicts with the rate it had before a reflection and extrapolates the
    reflected state afterwards; L1(rho) against the FP64 oracle < 1e-5."""
    if case == "wall":
        prob, steps = wall_box(3, 6, precision=fvs.FP32_FLUX), 7
    else:
        prob, steps = advected_wave(3, 6, (1.0, 0.6, -0.8), 0.25, precision=fvs.FP32_FLUX), 1 << 40
    ref = oracle_sim(dataclasses.replace(prob, cfg=dataclasses.replace(prob.cfg, precision=fvs.FP64)))
    dev = ics.build_device_sim(prob)
    n_dev = dev.run(steps)
    ref.advance(steps)
    assert dev.engine_info()["error_code"] == 0
    if case == "wall":  # the same step of the same timeline (the wraps run to t_end)
        assert dev.engine_info()["ti_current"] == ref.info()["ti_current"]
    rho = ref.get_fields()["prim"][: prob.n]
    l1 = float(np.mean(np.abs(dev.get_state(["prim"])["prim"][: prob.n] - rho)))
    say(f"[3d {case} fp32] {n_dev} steps, L1(rho) vs oracle {l1:.2e}")
    assert l1 < 1e-5

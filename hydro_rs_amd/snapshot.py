"""Snapshots and initial conditions in the reference's on-disk layout.

Snapshot: Space::dump (src/space.rs:1283-1417).  Initial conditions: InitialConditions::from_hdf5
(src/initial_conditions.rs:47-117).

The reference writes HDF5 through the hdf5 crate; libhdf5 and h5py are not available here, so the
container is an .npz whose keys are the reference's HDF5 paths ("Header/Time", "PartType0/Coordinates",
"VoronoiFaces/Left", ...) with the same shapes and dtypes (usize -> uint64); a file that an HDF5
library writes with these paths and arrays holds the same data.  Header attributes are stored as
arrays, as the reference stores them (create_attr! of a slice).  The reference's names are kept as
they are, including the two on which its writer and reader disagree: a snapshot header has
"Numpart_Total" and the dataset "InternalEnergies", while from_hdf5 reads "NumPart_Total" and
"InternalEnergy" (so a reference snapshot is not an initial-conditions file, here as there).
"""
import numpy as np

from .fvs import PERIODIC, Config

ISOTHERMAL = 1


def snapshot_arrays(state, n, cfg: Config, time, faces=None):
    """The datasets of Space::dump for one simulation state (component-major state arrays as
    EulerEqnsFvs.get_state returns them).  faces: a mesh with area/centroid/normal/left/right, or None."""
    prim = np.asarray(state["prim"], np.float64).reshape(5, n)
    cons = np.asarray(state["cons"], np.float64).reshape(5, n)
    vec = lambda k: np.ascontiguousarray(np.asarray(state[k], np.float64).reshape(3, n).T)
    mass, mom, energy = cons[0], cons[1:4], cons[4]
    rho, vel, pressure = prim[0], prim[1:4], prim[4]
    with np.errstate(divide="ignore", invalid="ignore"):
        # Particle::internal_energy (part.rs:314-317): (E - 0.5 * p . v) / m, glam's dot order
        dot = (mom[0] * vel[0] + mom[1] * vel[1]) + mom[2] * vel[2]
        u = (energy - 0.5 * dot) / mass
        # GasLaw::gas_entropy_from_internal_energy (gas_law.rs:108-120)
        g = cfg.gamma
        u_eos = u if cfg.eos != ISOTHERMAL else np.full(n, cfg.isothermal_u)
        entropy = (g - 1.0) * u_eos * np.power(rho, 1.0 - g)
    out = {
        "Header/Time": np.array([float(time)]),
        "Header/BoxSize": np.asarray(cfg.box, np.float64),
        "Header/Numpart_Total": np.array([n, 0, 0, 0, 0], np.uint64),
        "Header/Dimension": np.array([cfg.dim], np.uint64),
        "PartType0/Coordinates": vec("loc"),
        "PartType0/Masses": mass.copy(),
        "PartType0/Densities": rho.copy(),
        "PartType0/Momenta": np.ascontiguousarray(mom.T),
        "PartType0/Velocities": np.ascontiguousarray(vel.T),
        "PartType0/Pressures": pressure.copy(),
        "PartType0/TotalEnergies": energy.copy(),
        "PartType0/InternalEnergies": u,
        "PartType0/Entropies": entropy,
        "PartType0/Volumes": np.asarray(state["volume"], np.float64).copy(),
        "PartType0/Centroids": vec("centroid"),
        "PartType0/ParticleVelocities": vec("v"),
        # a_grav is never set on the finite-volume path (no gravity solver is wired in the reference)
        "PartType0/GravitationalAccelerations": np.zeros((n, 3)),
        "PartType0/Timestep": np.asarray(state["dt"], np.float64).copy(),
    }
    if faces is not None:
        out.update({
            "VoronoiFaces/Area": np.asarray(faces.area, np.float64).copy(),
            "VoronoiFaces/Centroid": np.asarray(faces.centroid, np.float64).reshape(-1, 3).copy(),
            "VoronoiFaces/Normal": np.asarray(faces.normal, np.float64).reshape(-1, 3).copy(),
            "VoronoiFaces/Left": np.asarray(faces.left).astype(np.uint64),
            "VoronoiFaces/Right": np.asarray(faces.right).astype(np.int64),
        })
    return out


def write_snapshot(path, sim, cfg: Config, faces=None, slot=None):
    """Space::dump of a device simulation (EulerEqnsFvs, or slot `slot` of a Batch) at its current
    time: Time = ti_current * time_base (TimestepInfo::dt_from_dti, timeline.rs:141-143)."""
    fields = ["prim", "cons", "loc", "volume", "centroid", "v", "dt"]
    if slot is None:
        state, info, n = sim.get_state(fields), sim.engine_info(), sim.n
    else:
        state, info, n = sim.get_state(slot, fields), sim.engine_info(slot), sim.n_cells[slot]
    time = float(info["ti_current"]) * info["time_base"]
    arrays = snapshot_arrays(state, n, cfg, time, faces)
    np.savez(path, **arrays)
    return arrays


def write_ic(path, dim, box, coordinates, masses, velocities, internal_energy, periodic=None):
    """An initial-conditions file as InitialConditions::from_hdf5 reads it."""
    n = len(masses)
    arrays = {
        "Header/NumPart_Total": np.array([n, 0, 0, 0, 0], np.uint64),
        "Header/Dimension": np.array([dim], np.uint64),
        "Header/BoxSize": np.asarray(box, np.float64),
        "PartType0/Coordinates": np.asarray(coordinates, np.float64).reshape(n, 3),
        "PartType0/Masses": np.asarray(masses, np.float64),
        "PartType0/Velocities": np.asarray(velocities, np.float64).reshape(n, 3),
        "PartType0/InternalEnergy": np.asarray(internal_energy, np.float64),
    }
    if periodic is not None:
        arrays["Header/Periodic"] = np.array([bool(periodic)])
    np.savez(path, **arrays)


def read_ic(path):
    """InitialConditions::from_hdf5 (initial_conditions.rs:47-117): header (NumPart_Total, Dimension,
    optional Periodic, BoxSize padded with 1 beyond the dimension) and the particle arrays, with its
    length checks."""
    with np.load(path, allow_pickle=False) as z:
        num_parts = int(z["Header/NumPart_Total"][0])
        dim = int(z["Header/Dimension"][0])
        periodic = bool(z["Header/Periodic"][0]) if "Header/Periodic" in z.files else False
        box = [float(x) for x in z["Header/BoxSize"]]
        coordinates = np.asarray(z["PartType0/Coordinates"], np.float64).ravel()
        masses = np.asarray(z["PartType0/Masses"], np.float64).ravel()
        velocities = np.asarray(z["PartType0/Velocities"], np.float64).ravel()
        internal_energy = np.asarray(z["PartType0/InternalEnergy"], np.float64).ravel()
    if dim > 3:
        raise ValueError("Dimension must be <= 3!")
    for i in range(dim, 3):
        if len(box) < i + 1:
            box.append(1.0)
        else:
            box[i] = 1.0
    if coordinates.size != 3 * num_parts:
        raise ValueError("Incorrect length of coordinates vector!")
    if masses.size != num_parts:
        raise ValueError("Incorrect length of masses vector!")
    if velocities.size != 3 * num_parts:
        raise ValueError("Incorrect length of velocities vector!")
    if internal_energy.size != num_parts:
        raise ValueError("Incorrect length of internal energies vector!")
    return {"num_parts": num_parts, "dimension": dim, "periodic": periodic, "box_size": tuple(box[:3]),
            "coordinates": coordinates.reshape(num_parts, 3), "masses": masses,
            "velocities": velocities.reshape(num_parts, 3), "internal_energy": internal_energy}


def problem_from_ic(path, name="ic", threads=0, **cfg_kw):
    """A Problem (tessellated) from an initial-conditions file; cfg_kw: the remaining Config fields
    (gamma, cfl, solver, boundary, motion, t_end, ...); dim and box come from the file, and a periodic
    file sets the periodic boundary unless a boundary is given."""
    from .ics import Problem
    from .mesh import VoronoiMesh

    ic = read_ic(path)
    if ic["periodic"] and "boundary" not in cfg_kw:
        cfg_kw["boundary"] = PERIODIC
    cfg = Config(dim=ic["dimension"], box=ic["box_size"], **cfg_kw)
    mesh = VoronoiMesh(cfg.dim, cfg.box, cfg.boundary == PERIODIC, ic["coordinates"], n_threads=threads)
    return Problem(name, cfg, ic["coordinates"], ic["masses"], ic["velocities"], ic["internal_energy"], mesh)

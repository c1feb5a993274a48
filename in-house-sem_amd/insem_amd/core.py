"""Object model of the simulator, mirroring the reference's Python API
(python/insem/__init__.py:7-29, python/src/bindings.cpp:110-494): MacroMesh,
Material, source time functions, Source, Discretization, Simulation and
Solver.  The time loop runs on the MI355X through the C-ABI of
include/insem_hip.h.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import InsemError, check, check_count, check_ptr

MEDIA = {"acoustic": 0, "elastic": 1}
BOUNDARY_TYPES = {"free": 0, "rigid": 1, "absorbing": 2}
_STF_TYPES = {"ricker": 0, "gaussian": 1, "gaussian_derivative": 2, "samples": 3}
_SOURCE_TYPES = {"force": 0, "moment_tensor": 1, "monopole": 2, "dipole": 3}

real_dtype = np.float32


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


# ------------------------------------------------------------------ mesh


class MacroMesh:
    """Conforming coarse mesh (reference: mesh::MacroMesh, macro_mesh.hpp:42-106).

    cell_type 'quad4' or 'quad9'; coords 'cartesian' or 'axisymmetric';
    nodes (n, 2); cells (n_cells, 4 | 9); material (n_cells,);
    boundary (n, 3) rows of (cell, face, tag).
    """

    def __init__(self, cell_type: str, coords: str, nodes, cells, material, boundary=None):
        if cell_type not in ("quad4", "quad9"):
            raise InsemError(f"unknown cell type '{cell_type}'")
        if coords not in ("cartesian", "axisymmetric"):
            raise InsemError(f"unknown coordinate system '{coords}'")
        npc = 4 if cell_type == "quad4" else 9
        nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 2)
        cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, npc)
        material = np.ascontiguousarray(material, dtype=np.int32).reshape(-1)
        boundary = np.ascontiguousarray(
            np.zeros((0, 3)) if boundary is None else boundary, dtype=np.int32).reshape(-1, 3)
        lib = _lib.host()
        self._h = check_ptr(lib.insem_mesh_new(npc, int(coords == "axisymmetric"), _dptr(nodes), nodes.shape[0],
                                               _iptr(cells), cells.shape[0], _iptr(material), _iptr(boundary),
                                               boundary.shape[0]))
        self.cell_type = cell_type
        self.coords = coords
        self.nodes = nodes
        self.material = material
        info = np.zeros(6, dtype=np.int32)
        check(lib.insem_mesh_info(self._h, _iptr(info)))
        self.n_nodes, self.n_cells, self.n_edges = int(info[0]), int(info[1]), int(info[2])
        # After the axis-face rotation of axisymmetric meshes.
        self.cells = np.zeros((self.n_cells, npc), dtype=np.int32)
        self.boundary = np.zeros((int(info[5]), 3), dtype=np.int32)
        check(lib.insem_mesh_arrays(self._h, _iptr(self.cells), _iptr(self.boundary)))

    def map(self, cell: int, xi: float, eta: float) -> Tuple[float, float]:
        out = np.zeros(2)
        check(_lib.host().insem_mesh_map(self._h, cell, xi, eta, _dptr(out)))
        return float(out[0]), float(out[1])

    def is_axis_face(self, cell: int, face: int) -> bool:
        return bool(check_count(_lib.host().insem_mesh_is_axis_face(self._h, cell, face)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.host().insem_mesh_free(h)
            self._h = None

    def __repr__(self):
        return f"MacroMesh({self.cell_type}, {self.coords}, {self.n_nodes} nodes, {self.n_cells} cells)"


# ----------------------------------------------------------- description


@dataclass(frozen=True)
class Material:
    """Isotropic material (reference: description.hpp:133-146)."""

    medium: str
    vp: float
    vs: float
    rho: float

    @staticmethod
    def acoustic(vp: float, rho: float) -> "Material":
        return Material("acoustic", float(vp), 0.0, float(rho))

    @staticmethod
    def elastic(vp: float, vs: float, rho: float) -> "Material":
        return Material("elastic", float(vp), float(vs), float(rho))


@dataclass(frozen=True)
class SourceTimeFunction:
    """Source time function (reference: description.hpp:161-184, description.cpp:88-113)."""

    kind: str
    f0: float = 0.0
    t0: float = 0.0
    amplitude: float = 1.0
    samples: Tuple[float, ...] = ()
    sample_dt: float = 0.0

    def __call__(self, t):
        t = np.asarray(t, dtype=np.float64)
        if self.kind in ("ricker", "gaussian"):
            a = (np.pi * self.f0 * (t - self.t0)) ** 2
            return self.amplitude * ((1.0 - 2.0 * a) if self.kind == "ricker" else 1.0) * np.exp(-a)
        if self.kind == "gaussian_derivative":
            tau = t - self.t0
            c = (np.pi * self.f0) ** 2
            return self.amplitude * (-2.0 * c * tau) * np.exp(-c * tau * tau)
        s = np.asarray(self.samples)
        return self.amplitude * np.interp(t / self.sample_dt, np.arange(s.size), s, left=0.0, right=0.0)


def _stf(kind, f0, t0, amplitude):
    if f0 <= 0:
        raise InsemError("source time function frequency must be positive")
    return SourceTimeFunction(kind, float(f0), 1.2 / f0 if t0 is None else float(t0), float(amplitude))


def ricker(f0: float, t0: Optional[float] = None, amplitude: float = 1.0) -> SourceTimeFunction:
    return _stf("ricker", f0, t0, amplitude)


def gaussian(f0: float, t0: Optional[float] = None, amplitude: float = 1.0) -> SourceTimeFunction:
    return _stf("gaussian", f0, t0, amplitude)


def gaussian_derivative(f0: float, t0: Optional[float] = None, amplitude: float = 1.0) -> SourceTimeFunction:
    return _stf("gaussian_derivative", f0, t0, amplitude)


def sampled_stf(samples, dt: float, amplitude: float = 1.0) -> SourceTimeFunction:
    s = tuple(float(v) for v in np.asarray(samples, dtype=np.float64).ravel())
    if dt <= 0 or not s:
        raise InsemError("sampled source time function needs samples and a positive interval")
    return SourceTimeFunction("samples", 0.0, 0.0, float(amplitude), s, float(dt))


@dataclass(frozen=True)
class Source:
    """Point source (reference: description.hpp:202-221)."""

    kind: str
    position: Tuple[float, float]
    components: Tuple[float, float, float]
    stf: SourceTimeFunction

    @staticmethod
    def force(x, z, fx, fz, stf):
        return Source("force", (float(x), float(z)), (float(fx), float(fz), 0.0), stf)

    @staticmethod
    def moment_tensor(x, z, mxx, mzz, mxz, stf):
        return Source("moment_tensor", (float(x), float(z)), (float(mxx), float(mzz), float(mxz)), stf)

    @staticmethod
    def monopole(x, z, stf):
        return Source("monopole", (float(x), float(z)), (1.0, 0.0, 0.0), stf)

    @staticmethod
    def dipole(x, z, fx, fz, stf):
        return Source("dipole", (float(x), float(z)), (float(fx), float(fz), 0.0), stf)


@dataclass(frozen=True)
class Receiver:
    name: str
    position: Tuple[float, float]
    quantities: Tuple[str, ...]
    medium: str = ""


@dataclass(frozen=True)
class TimeConfig:
    dt: float = 0.0
    safety: float = 0.8
    n_steps: int = 0
    duration: float = 0.0
    record_every: int = 1
    energy_every: int = 0


@dataclass(frozen=True)
class Discretization:
    """Element order and refinement (reference: simulation.hpp:22-48)."""

    order: int = 4
    uniform_level: int = 0
    f_max: float = 0.0
    points_per_wavelength: float = 5.0
    max_level: int = 12
    regions: Tuple[Tuple[float, float, float, float, int], ...] = ()
    azimuthal_order: int = 0


# ---------------------------------------------------------- setup image


class SetupImage:
    """The plain-array state handed across the C-ABI (insem_hip_setup)."""

    def __init__(self, handle, owned: bool = True):
        self._h = handle
        self._owned = owned
        lib = _lib.host()
        self.view = check_ptr(lib.insem_image_view(handle))
        self.n_domains = check_count(lib.insem_image_n_domains(handle))
        self.domains = []
        for di in range(self.n_domains):
            info = np.zeros(10, dtype=np.int32)
            check(lib.insem_image_domain_info(handle, di, _iptr(info)))
            self.domains.append(dict(medium="acoustic" if info[0] == 0 else "elastic", ncomp=int(info[1]),
                                     n_global=int(info[2]), n_elements=int(info[3]), n_blocks=int(info[4]),
                                     n_dirichlet=int(info[5]), n_absorbing=int(info[6]), n_axis=int(info[7]),
                                     constrained=bool(info[8]), azimuthal_order=int(info[9])))
            blocks = []
            for bi in range(int(info[4])):
                b = np.zeros(3, dtype=np.int32)
                check(lib.insem_image_block_info(handle, di, bi, _iptr(b)))
                blocks.append(dict(kind=("rect", "affine", "curved", "axisym")[b[0]], n_elements=int(b[1]),
                                   azimuthal_order=int(b[2])))
            self.domains[-1]["blocks"] = blocks
        counts = np.zeros(8, dtype=np.int64)
        check(lib.insem_image_counts(handle, counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        self.n_points, self.n_dofs, self.n_coupling = int(counts[0]), int(counts[1]), int(counts[2])
        self.n_sources, self.n_receivers, self.ngll = int(counts[3]), int(counts[4]), int(counts[5])
        self.axisymmetric, self.azimuthal_order = bool(counts[6]), int(counts[7])
        td = np.zeros(3)
        ti = np.zeros(3, dtype=np.int32)
        check(lib.insem_image_time(handle, _dptr(td), _iptr(ti)))
        self.time = TimeConfig(dt=float(td[0]), safety=float(td[1]), duration=float(td[2]), n_steps=int(ti[0]),
                               record_every=int(ti[1]), energy_every=int(ti[2]))
        self.traces = []
        for t in range(check_count(lib.insem_image_n_traces(handle))):
            name = ctypes.create_string_buffer(256)
            q = ctypes.create_string_buffer(64)
            nc = ctypes.c_int32()
            check(lib.insem_image_trace_info(handle, t, name, 256, q, 64, ctypes.byref(nc)))
            self.traces.append((name.value.decode(), q.value.decode(), nc.value))

    def domain_index(self, medium: str) -> int:
        for i, d in enumerate(self.domains):
            if d["medium"] == medium:
                return i
        return -1

    def coordinates(self, di: int) -> np.ndarray:
        out = np.zeros((2, self.domains[di]["n_global"]))
        check(_lib.host().insem_image_coords(self._h, di, _dptr(out)))
        return out

    def mass(self, di: int) -> np.ndarray:
        """Lumped diagonal mass of a domain (as the time loop uses it, float)."""
        out = np.zeros(self.domains[di]["n_global"])
        check(_lib.host().insem_image_mass(self._h, di, _dptr(out)))
        return out

    def start_vectors(self) -> np.ndarray:
        n = check_count(_lib.host().insem_image_start_len(self._h))
        out = np.zeros(n, dtype=np.float32)
        check(_lib.host().insem_image_start_vectors(self._h, _fptr(out)))
        return out

    def stage_size(self) -> int:
        """Device bytes of this model's context (host work only, any thread:
        insem_hip_stage_size)."""
        n = ctypes.c_int64()
        check(_lib.host().insem_image_stage_size(self._h, ctypes.byref(n)))
        return int(n.value)

    def time_step(self, lambda_max: float) -> Tuple[float, int]:
        """(dt, n_steps) the model gets from lambda_max, as its Solver would
        (reference: solver.cpp:131-144)."""
        d = np.zeros(2)
        n = ctypes.c_int32()
        check(_lib.host().insem_image_model_time(self._h, float(lambda_max), _dptr(d), ctypes.byref(n)))
        return float(d[0]), int(n.value)

    def amplitudes(self, dt: float, n_steps: int) -> np.ndarray:
        out = np.zeros((self.n_sources, n_steps + 1), dtype=np.float32)
        check(_lib.host().insem_image_amplitudes(self._h, dt, n_steps, _fptr(out)))
        return out

    def close(self):
        if self._h and self._owned:
            _lib.host().insem_image_free(self._h)
        self._h = None

    def __del__(self):
        self.close()


# ----------------------------------------------------------------- solver


class Solver:
    """Time-stepping state on the MI355X (reference: solver::Solver,
    solver.hpp:91-153; bindings.cpp:330-397)."""

    def __init__(self, image: SetupImage):
        lib = _lib.host()
        handle = image._h
        image._owned = False  # the solver takes the image over
        self._handle = check_ptr(lib.insem_solver_new(handle))
        image._h = None
        self.image = SetupImage(lib.insem_solver_image(self._handle), owned=False)

    @property
    def _h(self):
        h = getattr(self, "_handle", None)
        if not h:
            raise InsemError("the solver has been closed")
        return h

    def _info(self):
        d = np.zeros(2)
        i = np.zeros(2, dtype=np.int32)
        check(_lib.host().insem_solver_info(self._h, _dptr(d), _iptr(i)))
        return d, i

    @property
    def dt(self) -> float:
        return float(self._info()[0][0])

    @property
    def stable_dt(self) -> float:
        return float(self._info()[0][1])

    @property
    def n_steps(self) -> int:
        return int(self._info()[1][0])

    @property
    def step(self) -> int:
        return int(self._info()[1][1])

    @property
    def time(self) -> float:
        return self.step * self.dt

    @property
    def context(self):
        """The insem_hip_ctx* of the C-ABI."""
        return _lib.host().insem_solver_context(self._h)

    def advance(self, n_steps: int) -> None:
        check(_lib.host().insem_solver_advance(self._h, int(n_steps)))

    def run(self) -> None:
        self.advance(self.n_steps - self.step)

    @property
    def media(self) -> List[str]:
        return [d["medium"] for d in self.image.domains]

    def _domain(self, medium: str) -> int:
        di = self.image.domain_index(medium)
        if di < 0:
            raise InsemError(f"the model has no {medium} region")
        return di

    def coordinates(self, medium: str) -> np.ndarray:
        return self.image.coordinates(self._domain(medium))

    def _field(self, di: int, which: int) -> np.ndarray:
        d = self.image.domains[di]
        out = np.zeros((d["ncomp"], d["n_global"]))
        check(_lib.host().insem_solver_field(self._h, di, which, _dptr(out)))
        return out

    def state(self, medium: str):
        di = self._domain(medium)
        return self._field(di, 0), self._field(di, 1), self._field(di, 2)

    def set_state(self, medium: str, u, v) -> None:
        if medium not in MEDIA:
            raise InsemError(f"unknown medium '{medium}' (expected 'acoustic' or 'elastic')")
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        v = np.ascontiguousarray(v, dtype=np.float64).ravel()
        if u.size != v.size:
            raise InsemError("u and v differ in size")
        check(_lib.host().insem_solver_set_state(self._h, MEDIA[medium], _dptr(u), _dptr(v), u.size))

    def _all_traces(self):
        ns = check_count(_lib.host().insem_solver_n_samples(self._h))
        total = sum(nc for _, _, nc in self.image.traces) * ns
        buf = np.zeros(max(total, 1))
        check(_lib.host().insem_solver_traces(self._h, _dptr(buf)))
        out, o = {}, 0
        for name, q, nc in self.image.traces:
            out[(name, q)] = buf[o:o + nc * ns].reshape(ns, nc).copy()
            o += nc * ns
        return out

    def trace(self, receiver: str, quantity: str) -> np.ndarray:
        t = self._all_traces()
        if (receiver, quantity) not in t:
            raise InsemError(f"no trace '{quantity}' for receiver '{receiver}'")
        return t[(receiver, quantity)]

    @property
    def traces(self) -> Dict[Tuple[str, str], np.ndarray]:
        return self._all_traces()

    @property
    def times(self) -> np.ndarray:
        ns = check_count(_lib.host().insem_solver_n_samples(self._h))
        return np.arange(ns) * self.image.time.record_every * self.dt

    @property
    def energy(self) -> np.ndarray:
        n = check_count(_lib.host().insem_solver_n_energy(self._h))
        out = np.zeros((max(n, 1), 3))
        check(_lib.host().insem_solver_energy(self._h, _dptr(out)))
        return out[:n]

    def close(self):
        h = getattr(self, "_handle", None)
        if h:
            _lib.host().insem_solver_free(h)
        self._handle = None

    def __del__(self):
        self.close()


# ------------------------------------------------------------------ batch


def reserve_block(n_bytes: int) -> int:
    """A device block of at least n_bytes on the current device for a
    staged context (insem_hip_reserve); its address."""
    p = ctypes.c_void_p()
    _lib.check_hip(_lib.hip().insem_hip_reserve(int(n_bytes), ctypes.byref(p)))
    return int(p.value)


def unreserve_block(block: int) -> None:
    """Gives back a reserved block that no context took (insem_hip_unreserve)."""
    _lib.check_hip(_lib.hip().insem_hip_unreserve(ctypes.c_void_p(block)))


class StagedPower:
    """The host half of a PowerJob's start: the model's context laid out for
    a block from reserve_block(), and its start vectors.  Makes no device
    call, so it may run on any thread; PowerJob.from_staged() finishes it on
    the thread that makes the device calls.  If it fails, the block stays
    reserved."""

    def __init__(self, image: "SetupImage", block: int):
        self._h = check_ptr(_lib.host().insem_power_stage(image._h, ctypes.c_void_p(block)))

    def close(self):
        """Frees it; a context never started gives its block back (a device
        call: on the device-calling thread)."""
        h = getattr(self, "_h", None)
        if h:
            _lib.host().insem_power_staged_free(h)
        self._h = None

    def __del__(self):
        self.close()


class PowerJob:
    """A model's power iteration (lambda_max of M^-1 K, the one its Solver
    runs for the stable time step) queued on a short-lived device context of
    the model alone, without waiting (no counterpart in the reference, which
    estimates the step synchronously in Solver::estimate_stable_dt,
    solver.cpp:583-667).  ready() does not block; lambda_max() waits."""

    def __init__(self, image: "SetupImage"):
        self._h = check_ptr(_lib.host().insem_power_start(image._h))

    @classmethod
    def from_staged(cls, staged: StagedPower) -> "PowerJob":
        """Starts the iteration of a staged context (copies its arrays to
        the device): the job PowerJob(image) would give, bitwise."""
        job = cls.__new__(cls)
        job._h = check_ptr(_lib.host().insem_power_start_staged(staged._h))
        staged.close()
        return job

    def ready(self) -> bool:
        flag = ctypes.c_int32()
        check(_lib.host().insem_power_ready(self._h, ctypes.byref(flag)))
        return bool(flag.value)

    def lambda_max(self) -> float:
        lam = ctypes.c_double()
        check(_lib.host().insem_power_lambda(self._h, ctypes.byref(lam)))
        return lam.value

    def into_solver(self, image: "SetupImage") -> "Solver":
        """The model's Solver on this job's own context (the image the job
        was started from, which the solver takes over): the state Solver(image)
        reaches, without building the context a second time.  The job keeps
        no context afterwards."""
        lib = _lib.host()
        image._owned = False
        s = Solver.__new__(Solver)
        s._handle = check_ptr(lib.insem_power_solver(self._h, image._h))
        image._h = None
        s.image = SetupImage(lib.insem_solver_image(s._handle), owned=False)
        return s

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.host().insem_power_free(h)
        self._h = None

    def __del__(self):
        self.close()


class BatchSolver:
    """Many independent models stepped as one union on the device (no
    counterpart in the reference, which runs one model per Solver).  Each
    model keeps the time step and step count its own Solver would have (its
    own power iteration), and its traces are bitwise those of running it
    alone.  Models must share the polynomial order, coordinate system,
    azimuthal order, media and record cadence."""

    def __init__(self, models, lambdas: Optional[Sequence[float]] = None):
        """lambdas: each model's lambda_max from a PowerJob of it (then the
        batch makes no power iteration of its own), or None."""
        lib = _lib.host()
        models = list(models)
        sims = [m for m in models if isinstance(m, Simulation)]
        built = {}
        if sims:
            # host setups of independent models, several at once (the native
            # builds release the interpreter lock)
            from concurrent.futures import ThreadPoolExecutor
            with ThreadPoolExecutor(max_workers=min(8, len(sims))) as pool:
                for sim, img in zip(sims, pool.map(lambda s: s.build_setup(), sims)):
                    built[id(sim)] = img
        images = [built[id(m)] if isinstance(m, Simulation) else m for m in models]
        if not images:
            raise InsemError("a batch needs at least one model")
        handles = (ctypes.c_void_p * len(images))(*[im._h for im in images])
        for im in images:
            im._owned = False  # the batch takes the images over
        if lambdas is None:
            self._handle = check_ptr(lib.insem_batch_new(handles, len(images)))
        else:
            lam = np.ascontiguousarray(lambdas, dtype=np.float64)
            if lam.shape != (len(images),):
                raise InsemError(f"a batch of {len(images)} models needs as many lambda_max values, not {lam.size}")
            self._handle = check_ptr(lib.insem_batch_new_lambdas(handles, _dptr(lam), len(images)))
        for im in images:
            im._h = None
        self.models = [SetupImage(lib.insem_batch_model_image(self._handle, m), owned=False)
                       for m in range(len(images))]

    @property
    def _h(self):
        h = getattr(self, "_handle", None)
        if not h:
            raise InsemError("the batch has been closed")
        return h

    def _info(self, m):
        d = np.zeros(2)
        i = np.zeros(3, dtype=np.int32)
        check(_lib.host().insem_batch_info(self._h, m, _dptr(d), _iptr(i)))
        return d, i

    @property
    def n_models(self) -> int:
        return len(self.models)

    @property
    def max_steps(self) -> int:
        return int(self._info(-1)[1][1])

    @property
    def step(self) -> int:
        return int(self._info(-1)[1][2])

    def dt(self, m: int) -> float:
        return float(self._info(m)[0][0])

    def n_steps(self, m: int) -> int:
        return int(self._info(m)[1][0])

    @property
    def context(self):
        return _lib.host().insem_batch_context(self._h)

    @property
    def footprint(self) -> int:
        """Device bytes held by the batch."""
        return int(_lib.host().insem_batch_footprint(self._h))

    def advance(self, n_steps: int) -> None:
        check(_lib.host().insem_batch_advance(self._h, int(n_steps)))

    def run(self) -> None:
        self.advance(self.max_steps - self.step)

    def idle(self) -> bool:
        """True when every queued step has finished (does not block)."""
        flag = ctypes.c_int32()
        check(_lib.hip().insem_hip_query(ctypes.c_void_p(self.context), ctypes.byref(flag)))
        return bool(flag.value)

    def device_step(self) -> int:
        """Steps the device has completed (does not block)."""
        n = ctypes.c_int32()
        check(_lib.host().insem_batch_device_step(self._h, ctypes.byref(n)))
        return n.value

    def traces_done(self, m: int) -> Dict[Tuple[str, str], np.ndarray]:
        """Model m's traces (as traces(m)) read without waiting for the steps
        still queued; its last step must be done on the device
        (device_step() >= n_steps(m))."""
        ns = int(self._info(m)[1][1])
        img = self.models[m]
        buf = np.zeros(max(sum(nc for _, _, nc in img.traces) * ns, 1))
        check(_lib.host().insem_batch_traces_done(self._h, m, _dptr(buf)))
        out, o = {}, 0
        for name, q, nc in img.traces:
            out[(name, q)] = buf[o:o + nc * ns].reshape(ns, nc).copy()
            o += nc * ns
        return out

    def traces(self, m: int) -> Dict[Tuple[str, str], np.ndarray]:
        """Model m's traces, keyed and shaped as Solver.traces of that model."""
        ns = int(self._info(m)[1][1])
        img = self.models[m]
        buf = np.zeros(max(sum(nc for _, _, nc in img.traces) * ns, 1))
        check(_lib.host().insem_batch_traces(self._h, m, _dptr(buf)))
        out, o = {}, 0
        for name, q, nc in img.traces:
            out[(name, q)] = buf[o:o + nc * ns].reshape(ns, nc).copy()
            o += nc * ns
        return out

    def close(self):
        h = getattr(self, "_handle", None)
        if h:
            _lib.host().insem_batch_free(h)
        self._handle = None

    def __del__(self):
        self.close()


# -------------------------------------------------------------- simulation


class Simulation:
    """Problem description plus derived discretisation (reference:
    insem::Simulation, simulation.hpp:56-124).  `solver` builds the
    time-stepping state on the device on first access; changing the
    description discards it."""

    def __init__(self, mesh: MacroMesh, discretization: Optional[Discretization] = None):
        disc = discretization or Discretization()
        self.mesh = mesh
        self.discretization = disc
        regions = np.ascontiguousarray(np.asarray(disc.regions, dtype=np.float64).reshape(-1, 5))
        self._h = check_ptr(_lib.host().insem_model_new(mesh._h, disc.order, disc.uniform_level, disc.f_max,
                                                        disc.points_per_wavelength, disc.max_level, _dptr(regions),
                                                        regions.shape[0], disc.azimuthal_order))
        self.materials: Dict[int, Material] = {}
        self.boundaries: Dict[int, str] = {}
        self.sources: List[Source] = []
        self.receivers: List[Receiver] = []
        self.time = TimeConfig()
        self._solver: Optional[Solver] = None
        self.is_setup = False

    def _invalidate(self):
        # A changed description needs a new solver; the old one stays valid
        # for whoever still holds it and frees its device memory when dropped.
        self._solver = None

    def set_material(self, region: int, material: Material) -> None:
        check(_lib.host().insem_model_set_material(self._h, int(region), MEDIA[material.medium], material.vp,
                                                   material.vs, material.rho))
        self.materials[int(region)] = material
        if self.discretization.f_max > 0 or self.discretization.regions:
            self.is_setup = False
        self._invalidate()

    def set_boundary(self, tag: int, type: str) -> None:
        if type not in BOUNDARY_TYPES:
            raise InsemError(f"unknown boundary type '{type}' (expected 'free', 'rigid' or 'absorbing')")
        check(_lib.host().insem_model_set_boundary(self._h, int(tag), BOUNDARY_TYPES[type]))
        self.boundaries[int(tag)] = type
        self._invalidate()

    def add_source(self, source: Source) -> None:
        stf = source.stf
        samples = np.ascontiguousarray(np.asarray(stf.samples or (0.0,), dtype=np.float64))
        check(_lib.host().insem_model_add_source(
            self._h, _SOURCE_TYPES[source.kind], source.position[0], source.position[1], *source.components,
            _STF_TYPES[stf.kind], stf.f0, stf.t0, stf.amplitude, _dptr(samples), len(stf.samples), stf.sample_dt))
        self.sources.append(source)
        self._invalidate()

    def clear_sources(self) -> None:
        check(_lib.host().insem_model_clear_sources(self._h))
        self.sources.clear()
        self._invalidate()

    def add_receiver(self, name: str, x: float, z: float, quantities: Sequence[str], medium: str = "") -> None:
        check(_lib.host().insem_model_add_receiver(self._h, name.encode(), float(x), float(z),
                                                   ",".join(quantities).encode(), medium.encode()))
        self.receivers.append(Receiver(name, (float(x), float(z)), tuple(quantities), medium))
        self._invalidate()

    def clear_receivers(self) -> None:
        check(_lib.host().insem_model_clear_receivers(self._h))
        self.receivers.clear()
        self._invalidate()

    def set_time(self, n_steps: int = 0, duration: float = 0.0, dt: float = 0.0, safety: float = 0.8,
                 record_every: int = 1, energy_every: int = 0) -> None:
        check(_lib.host().insem_model_set_time(self._h, dt, safety, int(n_steps), duration, int(record_every),
                                               int(energy_every)))
        self.time = TimeConfig(dt, safety, int(n_steps), duration, int(record_every), int(energy_every))
        self._invalidate()

    def setup(self) -> None:
        check(_lib.host().insem_model_setup(self._h))
        self.is_setup = True

    def _info(self):
        self.setup()
        out = np.zeros(5, dtype=np.int32)
        check(_lib.host().insem_model_info(self._h, _iptr(out)))
        return out

    @property
    def n_leaves(self) -> int:
        return int(self._info()[0])

    @property
    def max_level(self) -> int:
        return int(self._info()[1])

    def geometry_kind_counts(self) -> Dict[str, int]:
        i = self._info()
        return {"rect": int(i[2]), "affine": int(i[3]), "curved": int(i[4])}

    def leaf_points(self) -> Tuple[np.ndarray, np.ndarray]:
        """(x, z) of every leaf point, [leaf][k] (the geometry of the forest)."""
        n = self.n_leaves * (self.discretization.order + 1) ** 2
        x, z = np.zeros(n), np.zeros(n)
        check(_lib.host().insem_model_leaf_points(self._h, _dptr(x), _dptr(z)))
        return x, z

    def build_setup(self) -> SetupImage:
        """The time-loop state as plain arrays (what the C-ABI consumes)."""
        self.setup()
        return SetupImage(check_ptr(_lib.host().insem_model_build(self._h)))

    @property
    def solver(self) -> Solver:
        if self._solver is None:
            self._solver = Solver(self.build_setup())
        return self._solver

    def run(self) -> None:
        self.solver.run()

    def __del__(self):
        self._invalidate()
        h = getattr(self, "_h", None)
        if h:
            _lib.host().insem_model_free(h)
            self._h = None


# ------------------------------------------------------------------ basis


def gauss_lobatto_jacobi(n_points: int, alpha: int = 0, beta: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """(points, weights) of the Lobatto rule for (1 - x)^alpha (1 + x)^beta
    (reference: basis::gauss_lobatto_jacobi, quadrature.cpp:79-132)."""
    p, w = np.zeros(n_points), np.zeros(n_points)
    check(_lib.host().insem_basis_rule(int(n_points), int(alpha), int(beta), _dptr(p), _dptr(w)))
    return p, w


def gll(n_points: int) -> Tuple[np.ndarray, np.ndarray]:
    return gauss_lobatto_jacobi(n_points, 0, 0)


def glj01(n_points: int) -> Tuple[np.ndarray, np.ndarray]:
    return gauss_lobatto_jacobi(n_points, 0, 1)


def derivative_matrix(nodes) -> np.ndarray:
    """D[i, j] = l_j'(x_i) of the Lagrange basis on `nodes` (lagrange.cpp:77-82)."""
    x = np.ascontiguousarray(nodes, dtype=np.float64)
    out = np.zeros((x.size, x.size))
    check(_lib.host().insem_basis_derivative_matrix(x.size, _dptr(x), _dptr(out)))
    return out

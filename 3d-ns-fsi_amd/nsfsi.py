"""Python host layer over libnsfsi.so (the C-ABI in include/ns_fsi.h).

Mirrors the reference's host surface (Frederik-Kris/3D-NS-FSI): ``ConfigSettings`` (src/ConfigSettings.h),
``Solver`` with ``initialize`` / ``marchTimeStep`` / ``dt`` / ``getConvergenceHistory`` and a ``mesh``
view (src/Solver.h:17-31, src/Mesh.h:79-89), and ``Simulation.run`` (src/Simulation.h:17-32).

The device library is the only compute path: if libnsfsi.so is missing, or no gfx950 GPU is present,
the calls raise; nothing here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnsfsi.so")

NS_MAX_FILTER = 16
NS_MAX_BODIES = 8

STATUS = {
    0: "NS_OK", 1: "NS_ERR_CONFIG", 2: "NS_ERR_GEOMETRY", 3: "NS_ERR_NONFINITE", 4: "NS_ERR_HIP",
    5: "NS_ERR_ARGUMENT", 6: "NS_ERR_STATE", 7: "NS_ERR_UNSUPPORTED",
}
BC = {"inlet": 0, "outlet": 1, "periodic": 2, "symmetry": 3}
BODY = {"none": 0, "cylinder": 1, "sphere": 2}
IC = {"stagnation": 0, "uniform_oblique": 1, "freestream_gaussian": 2, "freestream_uniform": 3}
FIELDS = ["rho", "rho_u", "rho_v", "rho_w", "rho_E", "u", "v", "w", "p", "T", "mu", "kappa"]
NODE_TYPES = ["FluidInterior", "FluidEdge", "FluidGhost", "SolidInactive", "SolidGhost"]


class NSBody(C.Structure):
    """ctypes image of ``ns_body`` (include/ns_fsi.h)."""

    _fields_ = [("type", C.c_int32), ("axis", C.c_int32), ("centre", C.c_double * 3), ("radius", C.c_double)]

    def _key(self):
        return self.type, self.axis, tuple(self.centre), self.radius

    def __eq__(self, other):
        return isinstance(other, NSBody) and self._key() == other._key()

    def __hash__(self):
        return hash(self._key())


class NSConfig(C.Structure):
    """ctypes image of ``ns_config`` (include/ns_fsi.h)."""

    _fields_ = [
        ("NI", C.c_int32), ("NJ", C.c_int32), ("NK", C.c_int32),
        ("stop_criterion", C.c_int32),
        ("stop_time_level", C.c_uint64),
        ("t_end", C.c_double),
        ("conv_stability_limit", C.c_double),
        ("visc_stability_limit", C.c_double),
        ("n_filter_intervals", C.c_int32),
        ("filter_intervals", C.c_uint32 * NS_MAX_FILTER),
        ("n_filter_change_times", C.c_int32),
        ("filter_change_times", C.c_double * NS_MAX_FILTER),
        ("status_report_interval", C.c_double),
        ("save_density", C.c_int32), ("save_momentum", C.c_int32), ("save_energy", C.c_int32),
        ("save_velocity", C.c_int32), ("save_pressure", C.c_int32), ("save_temperature", C.c_int32),
        ("save_viscosity", C.c_int32), ("save_thermal_cond", C.c_int32),
        ("save_ic", C.c_int32), ("save_final", C.c_int32), ("save_intervals", C.c_int32),
        ("save_period", C.c_double), ("save_intervals_start_time", C.c_double), ("save_intervals_end_time", C.c_double),
        ("conv_history", C.c_int32),
        ("gamma", C.c_double), ("Pr", C.c_double), ("R", C.c_double), ("Re", C.c_double),
        ("sutherlands_C2", C.c_double), ("M_0", C.c_double), ("T_0", C.c_double),
        ("L_x", C.c_double), ("L_y", C.c_double), ("L_z", C.c_double),
        ("machine_precision_buffer", C.c_double),
        ("bc", C.c_int32 * 6),
        ("body_type", C.c_int32),
        ("body_axis", C.c_int32),
        ("body_centre", C.c_double * 3),
        ("body_radius", C.c_double),
        ("ic_type", C.c_int32),
        ("ic_seed", C.c_uint64),
        ("ic_amplitude", C.c_double),
        ("inlet_ramp_time", C.c_double),
        ("precision", C.c_int32),
        ("kahan", C.c_int32),
        ("n_extra_bodies", C.c_int32),
        ("pad_bodies", C.c_int32),
        ("extra_bodies", NSBody * (NS_MAX_BODIES - 1)),
    ]


class NSError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


_lib = None


def lib():
    """Loads libnsfsi.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    L.ns_config_default.argtypes = [C.POINTER(NSConfig)]
    L.ns_config_default.restype = None
    L.ns_config_parse.argtypes = [C.c_char_p, C.POINTER(NSConfig), C.c_char_p, C.c_size_t]
    L.ns_create.argtypes = [C.POINTER(NSConfig), C.POINTER(H)]
    L.ns_init.argtypes = [H]
    L.ns_set_initial_field.argtypes = [H, C.POINTER(C.c_double)]
    L.ns_step.argtypes = [H, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    L.ns_run.argtypes = [H, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    L.ns_get_time.argtypes = [H, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    L.ns_get_field.argtypes = [H, C.c_int32, C.POINTER(C.c_double)]
    L.ns_get_node_types.argtypes = [H, C.POINTER(C.c_uint8)]
    L.ns_get_counts.argtypes = [H, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.ns_get_ghosts.argtypes = [H, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                C.POINTER(C.c_int32)]
    L.ns_get_filter_nodes.argtypes = [H, C.POINTER(C.c_int32)]
    L.ns_norm_history.argtypes = [H, C.c_int32, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64)]
    L.ns_totals.argtypes = [H, C.POINTER(C.c_double)]
    L.ns_force_history.argtypes = [H, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64)]
    L.ns_set_dt_schedule.argtypes = [H, C.POINTER(C.c_double), C.c_int64]
    L.ns_set_inflow_profile.argtypes = [H, C.POINTER(C.c_double), C.c_int64]
    L.ns_adjoint_drag.argtypes = [H, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
    L.ns_get_geometry.argtypes = [H, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.ns_profile.argtypes = [H, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
    L.ns_set_option.argtypes = [H, C.c_char_p, C.c_int64]
    L.ns_get_tile.argtypes = [H, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.ns_save.argtypes = [H, C.c_char_p]
    L.ns_restore.argtypes = [H, C.c_char_p]
    L.ns_last_error.argtypes = [H]
    L.ns_last_error.restype = C.c_char_p
    L.ns_destroy.argtypes = [H]
    L.ns_destroy.restype = None
    L.ns_abi_version.restype = C.c_int32
    L.ns_totals_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                    C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
    L.ns_totals_noalloc.restype = C.c_int
    L.ns_noalloc_last_error.argtypes = []
    L.ns_noalloc_last_error.restype = C.c_char_p
    B = C.c_void_p
    L.ns_batch_create.argtypes = [C.POINTER(NSConfig), C.c_int32, C.POINTER(B)]
    L.ns_batch_init.argtypes = [B]
    L.ns_batch_step.argtypes = [B, C.c_int32]
    L.ns_batch_get_time.argtypes = [B, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    L.ns_batch_get_field.argtypes = [B, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    L.ns_batch_norm_history.argtypes = [B, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.c_int64,
                                        C.POINTER(C.c_int64)]
    L.ns_batch_set_option.argtypes = [B, C.c_char_p, C.c_int64]
    L.ns_batch_get_tile.argtypes = [B, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.ns_batch_last_error.argtypes = [B]
    L.ns_batch_last_error.restype = C.c_char_p
    L.ns_batch_destroy.argtypes = [B]
    L.ns_batch_destroy.restype = None
    for name in ("ns_batch_create", "ns_batch_init", "ns_batch_step", "ns_batch_get_time", "ns_batch_get_field",
                 "ns_batch_norm_history", "ns_batch_set_option", "ns_batch_get_tile"):
        getattr(L, name).restype = C.c_int
    for name in ("ns_config_parse", "ns_create", "ns_init", "ns_set_initial_field", "ns_step", "ns_run", "ns_get_time",
                 "ns_get_field", "ns_get_node_types", "ns_get_counts", "ns_get_ghosts", "ns_get_filter_nodes",
                 "ns_norm_history", "ns_force_history", "ns_set_dt_schedule", "ns_set_inflow_profile",
                 "ns_adjoint_drag", "ns_totals", "ns_get_geometry", "ns_profile", "ns_set_option", "ns_get_tile",
                 "ns_save", "ns_restore"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_config() -> NSConfig:
    c = NSConfig()
    lib().ns_config_default(C.byref(c))
    return c


def make_config(**kw) -> NSConfig:
    """Default configuration (the reference's shipped ConfigFile) with keyword overrides.

    Names follow ns_config; ``bc`` takes 6 names (x-min .. z-max), ``body_type``/``ic_type`` take names,
    ``filter_intervals``/``filter_change_times`` take lists, ``body_centre`` a 3-sequence,
    ``precision`` "fp64"/"fp32", ``extra_bodies`` a list of dicts (type, axis, centre, radius) for the bodies
    after the first.
    """
    c = default_config()
    for k, v in kw.items():
        if k == "bc":
            for q, name in enumerate(v):
                c.bc[q] = BC[name] if isinstance(name, str) else int(name)
        elif k == "body_type":
            c.body_type = BODY[v] if isinstance(v, str) else int(v)
        elif k == "body_axis":
            c.body_axis = "xyz".index(v) if isinstance(v, str) else int(v)
        elif k == "ic_type":
            c.ic_type = IC[v] if isinstance(v, str) else int(v)
        elif k == "body_centre":
            for q in range(3):
                c.body_centre[q] = float(v[q])
        elif k == "filter_intervals":
            c.n_filter_intervals = len(v)
            for q, x in enumerate(v):
                c.filter_intervals[q] = int(x)
        elif k == "filter_change_times":
            c.n_filter_change_times = len(v)
            for q, x in enumerate(v):
                c.filter_change_times[q] = float(x)
        elif k == "extra_bodies":
            if len(v) > NS_MAX_BODIES - 1:
                raise ValueError(f"at most {NS_MAX_BODIES - 1} extra bodies")
            c.n_extra_bodies = len(v)
            for q, d in enumerate(v):
                b = c.extra_bodies[q]
                b.type = BODY[d["type"]] if isinstance(d["type"], str) else int(d["type"])
                ax = d.get("axis", 0)
                b.axis = "xyz".index(ax) if isinstance(ax, str) else int(ax)
                for a in range(3):
                    b.centre[a] = float(d["centre"][a])
                b.radius = float(d["radius"])
        elif k == "precision":
            c.precision = {"fp64": 0, "fp32": 1}[v] if isinstance(v, str) else int(v)
        elif k == "stop_criterion":
            c.stop_criterion = {"timesteps": 0, "end_time": 1}[v] if isinstance(v, str) else int(v)
        elif k == "conv_history":
            c.conv_history = {"none": 0, "density": 1, "all": 2}[v] if isinstance(v, str) else int(v)
        else:
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, v)
    c.machine_precision_buffer = np.finfo(np.float64).eps * (c.L_x + c.L_y + c.L_z)
    return c


def config_text(c: NSConfig) -> str:
    """A ConfigFile (libconfig syntax, the reference's keys plus the extension keys) that ns_config_parse reads
    back into the same configuration: integers without a decimal point, floats with one (libconfig's strict
    typing), floats written with 17 significant digits."""
    def f(x):
        r = repr(float(x))
        return r if ("." in r or "e" in r or "n" in r) else r + "."

    inv_bc = {v: k for k, v in BC.items()}
    inv_body = {v: k for k, v in BODY.items()}
    inv_ic = {v: k for k, v in IC.items()}
    b = lambda x: "true" if x else "false"  # noqa: E731
    lines = [
        f"NI = {c.NI}", f"NJ = {c.NJ}", f"NK = {c.NK}",
        f'stopCriterion = "{"timesteps" if c.stop_criterion == 0 else "end_time"}"',
        f"stopTimeLevel = {c.stop_time_level}", f"t_end = {f(c.t_end)}",
        f"convStabilityLimit = {f(c.conv_stability_limit)}", f"viscStabilityLimit = {f(c.visc_stability_limit)}",
        "filterIntervals = [" + ", ".join(str(c.filter_intervals[q]) for q in range(c.n_filter_intervals)) + "]",
        "filterIntervalChangeTimes = [" + ", ".join(f(c.filter_change_times[q]) for q in range(c.n_filter_change_times))
        + "]",
        f"statusReportInterval = {f(c.status_report_interval)}",
        f"saveDensity = {b(c.save_density)}", f"saveMomentum = {b(c.save_momentum)}",
        f"saveEnergy = {b(c.save_energy)}", f"saveVelocity = {b(c.save_velocity)}",
        f"savePressure = {b(c.save_pressure)}", f"saveTemperature = {b(c.save_temperature)}",
        f"saveViscosity = {b(c.save_viscosity)}", f"saveThermalCond = {b(c.save_thermal_cond)}",
        f"saveIC = {b(c.save_ic)}", f"saveFinal = {b(c.save_final)}", f"saveIntervals = {b(c.save_intervals)}",
        f"savePeriod = {f(c.save_period)}", f"saveIntervalsStartTime = {f(c.save_intervals_start_time)}",
        f"saveIntervalsEndTime = {f(c.save_intervals_end_time)}",
        f'saveConvergenceHistory = "{["none", "density", "all"][c.conv_history]}"',
        f"Gamma = {f(c.gamma)}", f"Pr = {f(c.Pr)}", f"R = {f(c.R)}", f"Re = {f(c.Re)}",
        f"sutherlands_C2 = {f(c.sutherlands_C2)}", f"M_0 = {f(c.M_0)}", f"T_0 = {f(c.T_0)}",
        f"L_x = {f(c.L_x)}", f"L_y = {f(c.L_y)}", f"L_z = {f(c.L_z)}",
        "boundaryConditions = [" + ", ".join(f'"{inv_bc[c.bc[q]]}"' for q in range(6)) + "]",
        f'bodyType = "{inv_body[c.body_type]}"', f'bodyAxis = "{"xyz"[c.body_axis]}"',
        "bodyCentre = [" + ", ".join(f(c.body_centre[q]) for q in range(3)) + "]",
        f"bodyRadius = {f(c.body_radius)}", f'initialCondition = "{inv_ic[c.ic_type]}"',
        f"perturbationSeed = {c.ic_seed}", f"perturbationAmplitude = {f(c.ic_amplitude)}",
        f"inletRampTime = {f(c.inlet_ramp_time)}", f'precision = "{"fp64" if c.precision == 0 else "fp32"}"',
        f"kahan = {b(c.kahan)}",
    ]
    if c.n_extra_bodies:
        ent = []
        for q in range(c.n_extra_bodies):
            e = c.extra_bodies[q]
            ent.append(f'{{ type = "{inv_body[e.type]}"; axis = "{"xyz"[e.axis]}"; centre = ['
                       + ", ".join(f(e.centre[a]) for a in range(3)) + f"]; radius = {f(e.radius)}; }}")
        lines.append("extraBodies = (" + ",\n    ".join(ent) + ")")
    return "\n".join(lines) + "\n"


def parse_config(text: str) -> NSConfig:
    """ConfigSettings::loadSettings on a ConfigFile text (src/ConfigSettings.cpp:18-28)."""
    c = NSConfig()
    err = C.create_string_buffer(512)
    s = lib().ns_config_parse(text.encode(), C.byref(c), err, 512)
    if s != 0:
        raise NSError(s, err.value.decode())
    return c


class ConfigSettings:
    """ConfigSettings(filename) of the reference: parses a ConfigFile; ``errorOccurred`` instead of raising."""

    def __init__(self, filename: str):
        self.errorOccurred = False
        self.message = ""
        try:
            with open(filename) as f:
                self.config = parse_config(f.read())
        except (OSError, NSError) as e:
            self.errorOccurred = True
            self.message = str(e)
            self.config = None


class Mesh:
    """Read-only view of the device mesh (Mesh public members, src/Mesh.h:79-89)."""

    def __init__(self, solver: "Solver"):
        self._s = solver
        c = solver.config
        self.NI, self.NJ, self.NK = c.NI, c.NJ, c.NK
        self.nNodesTotal = self.NI * self.NJ * self.NK

    def _geometry(self):
        sp = np.zeros(3)
        of = np.zeros(3)
        self._s._check(lib().ns_get_geometry(self._s._h, _dptr(sp), _dptr(of)))
        return sp, of

    @property
    def dx(self):
        return self._geometry()[0][0]

    @property
    def dy(self):
        return self._geometry()[0][1]

    @property
    def dz(self):
        return self._geometry()[0][2]

    @property
    def positionOffset(self):
        return self._geometry()[1]

    def field(self, name: str) -> np.ndarray:
        """One field as an array indexed [k, j, i] (x fastest)."""
        which = FIELDS.index(name)
        out = np.empty(self.nNodesTotal)
        self._s._check(lib().ns_get_field(self._s._h, which, _dptr(out)))
        return out.reshape(self.NK, self.NJ, self.NI)

    def conserved(self) -> np.ndarray:
        return np.stack([self.field(n) for n in FIELDS[:5]])

    @property
    def nodeType(self) -> np.ndarray:
        out = np.empty(self.nNodesTotal, dtype=np.uint8)
        self._s._check(lib().ns_get_node_types(self._s._h, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out.reshape(self.NK, self.NJ, self.NI)

    def counts(self):
        g, f, n = C.c_int64(), C.c_int64(), C.c_int64()
        self._s._check(lib().ns_get_counts(self._s._h, C.byref(g), C.byref(f), C.byref(n)))
        return g.value, f.value, n.value

    def ghosts(self):
        """(ijk, body-intercept, image point, box lower corner) of every ghost, sorted by reference index."""
        n = self.counts()[0]
        ijk = np.zeros((n, 3), dtype=np.int32)
        bi = np.zeros((n, 3))
        ip = np.zeros((n, 3))
        box = np.zeros((n, 3), dtype=np.int32)
        self._s._check(lib().ns_get_ghosts(self._s._h, ijk.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(bi), _dptr(ip),
                                           box.ctypes.data_as(C.POINTER(C.c_int32))))
        return ijk, bi, ip, box

    def filterNodes(self):
        n = self.counts()[1]
        ijk = np.zeros((n, 3), dtype=np.int32)
        if n:
            self._s._check(lib().ns_get_filter_nodes(self._s._h, ijk.ctypes.data_as(C.POINTER(C.c_int32))))
        return ijk

    def totals(self):
        out = np.zeros(5)
        self._s._check(lib().ns_totals(self._s._h, _dptr(out)))
        return out


class Solver:
    """Solver of the reference (src/Solver.h:17-31) on one MI355X."""

    def __init__(self, params):
        self.config = params.config if isinstance(params, ConfigSettings) else params
        self._h = C.c_void_p()
        s = lib().ns_create(C.byref(self.config), C.byref(self._h))
        if s != 0:
            msg = lib().ns_last_error(self._h).decode() if self._h else ""
            lib().ns_destroy(self._h)
            self._h = C.c_void_p()
            raise NSError(s, msg)
        self.mesh = Mesh(self)

    def _check(self, s):
        if s != 0:
            raise NSError(s, lib().ns_last_error(self._h).decode())

    def initialize(self):
        self._check(lib().ns_init(self._h))

    def setInitialField(self, U5: np.ndarray):
        a = np.ascontiguousarray(U5, dtype=np.float64)
        assert a.size == 5 * self.mesh.nNodesTotal
        self._check(lib().ns_set_initial_field(self._h, _dptr(a)))

    def marchTimeStep(self, nsteps: int = 1):
        """Advances nsteps RK4 steps; returns (t, timeLevel)."""
        t, lv, dt = C.c_double(), C.c_uint64(), C.c_double()
        self._check(lib().ns_step(self._h, int(nsteps), C.byref(t), C.byref(lv), C.byref(dt)))
        return t.value, lv.value

    def time(self):
        t, lv, dt = C.c_double(), C.c_uint64(), C.c_double()
        self._check(lib().ns_get_time(self._h, C.byref(t), C.byref(lv), C.byref(dt)))
        return t.value, lv.value, dt.value

    @property
    def dt(self):
        return self.time()[2]

    def setOption(self, name: str, value: int):
        """Run-time option, e.g. setOption("stage_kernel", 0) for the unfused stage."""
        self._check(lib().ns_set_option(self._h, name.encode(), int(value)))

    def tile(self):
        """(rows per block, z-planes per block) of the fused stage."""
        ty, kc = C.c_int32(), C.c_int32()
        self._check(lib().ns_get_tile(self._h, C.byref(ty), C.byref(kc)))
        return ty.value, kc.value

    def autotune(self, reps: int = 3):
        """Times every fused-stage tile shape on this case and keeps the fastest; returns it."""
        self.setOption("autotune", reps)
        return self.tile()

    def profile(self, nsteps: int):
        """Mean ms per launch of (stage kernel, primitive pass, boundary kernels, reductions) and launches/step."""
        ms = np.zeros(4)
        ln = np.zeros(4, dtype=np.int32)
        self._check(lib().ns_profile(self._h, int(nsteps), _dptr(ms), ln.ctypes.data_as(C.POINTER(C.c_int32))))
        return ms, ln

    def getConvergenceHistory(self):
        out = {}
        for q, name in enumerate(FIELDS[:5]):
            cnt = C.c_int64()
            self._check(lib().ns_norm_history(self._h, q, None, 0, C.byref(cnt)))
            a = np.zeros(cnt.value)
            if cnt.value:
                self._check(lib().ns_norm_history(self._h, q, _dptr(a), cnt.value, C.byref(cnt)))
            out[name] = a
        return out

    def getForceHistory(self) -> np.ndarray:
        """(n, 4) array of (timeLevel, Fx, Fy, Fz): force on the body after each step taken with the
        "body_force" option on (discrete momentum balance over the fluid/ghost faces, FP64)."""
        cnt = C.c_int64()
        self._check(lib().ns_force_history(self._h, None, 0, C.byref(cnt)))
        a = np.zeros((cnt.value, 4))
        if cnt.value:
            self._check(lib().ns_force_history(self._h, _dptr(a), cnt.value, C.byref(cnt)))
        return a

    def setDtSchedule(self, dts):
        """Prescribed step sizes by time level (None: back to the stability limit)."""
        if dts is None:
            self._check(lib().ns_set_dt_schedule(self._h, None, 0))
            return
        a = np.ascontiguousarray(dts, dtype=np.float64)
        self._check(lib().ns_set_dt_schedule(self._h, _dptr(a), a.size))

    def inletPlaneShape(self):
        """(n_b, n_a) shape of the inlet plane's profile array (x inlet: (NK, NJ)); (0, 0) without an inlet."""
        c = self.config
        faces = [c.bc[q] for q in range(6)]
        if BC["inlet"] not in faces:
            return (0, 0)
        axis = faces.index(BC["inlet"]) // 2
        return {0: (c.NK, c.NJ), 1: (c.NK, c.NI), 2: (c.NJ, c.NI)}[axis]

    def setInflowProfile(self, q):
        """Inlet velocity factors per inlet-plane node (None: uniform inlet)."""
        if q is None:
            self._check(lib().ns_set_inflow_profile(self._h, None, 0))
            return
        a = np.ascontiguousarray(q, dtype=np.float64)
        self._check(lib().ns_set_inflow_profile(self._h, _dptr(a), a.size))

    def adjointDrag(self, nsteps: int, n0: int, n1: int, checkpoint: int = 20):
        """J = mean C_D over time levels n0..n1 of an nsteps run from the current initial state, and its
        gradients (ns_adjoint_drag): dict with J, grad_u0 (5, NK, NJ, NI), grad_inflow (inlet plane),
        grad_re and the run's step sizes dts (nsteps + 1)."""
        N = self.mesh.nNodesTotal
        gu = np.zeros(5 * N)
        shp = self.inletPlaneShape()
        gq = np.zeros(int(np.prod(shp)))
        J, gre = C.c_double(), C.c_double()
        dts = np.zeros(nsteps + 1)
        self._check(lib().ns_adjoint_drag(self._h, int(nsteps), int(n0), int(n1), int(checkpoint), C.byref(J),
                                          _dptr(gu), _dptr(gq) if gq.size else None, C.byref(gre), _dptr(dts)))
        c = self.config
        return {"J": J.value, "grad_u0": gu.reshape(5, c.NK, c.NJ, c.NI), "grad_inflow": gq.reshape(shp),
                "grad_re": gre.value, "dts": dts}

    def save(self, path):
        """Checkpoint of the run (ns_save): a restored run continues bit for bit."""
        self._check(lib().ns_save(self._h, os.fsencode(path)))

    def restore(self, path):
        """Loads a checkpoint written by a case of the same grid, boundaries, body and physics (ns_restore);
        call after initialize()."""
        self._check(lib().ns_restore(self._h, os.fsencode(path)))

    def referenceArea(self) -> float:
        """Frontal area for force coefficients: pi r^2 (sphere), 2 r times the span of FluidInterior planes
        along the axis (cylinder)."""
        c = self.config
        if c.body_type == BODY["sphere"]:
            return float(np.pi * c.body_radius ** 2)
        if c.body_type == BODY["cylinder"]:
            n = (c.NI, c.NJ, c.NK)[c.body_axis]
            h = (self.mesh.dx, self.mesh.dy, self.mesh.dz)[c.body_axis]
            return float(2 * c.body_radius * (n - 2) * h)
        raise ValueError("no body")

    def dragCoefficient(self) -> np.ndarray:
        """C_D = Fx / (0.5 rho_0 M_0^2 A_ref) per recorded step (velocities are in units of c_0)."""
        f = self.getForceHistory()
        return f[:, 1] / (0.5 * self.config.M_0 ** 2 * self.referenceArea())

    def close(self):
        if self._h:
            lib().ns_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent FP32 cases advanced together (ns_batch_*): each kernel of the step runs once for
    all cases, each case at its own dt. ``configs`` is a sequence of NSConfig."""

    def __init__(self, configs):
        self.configs = list(configs)
        arr = (NSConfig * len(self.configs))(*self.configs)
        self._h = C.c_void_p()
        s = lib().ns_batch_create(arr, len(self.configs), C.byref(self._h))
        if s != 0:
            msg = lib().ns_batch_last_error(self._h).decode() if self._h else ""
            lib().ns_batch_destroy(self._h)
            self._h = C.c_void_p()
            raise NSError(s, msg)

    def _check(self, s):
        if s != 0:
            raise NSError(s, lib().ns_batch_last_error(self._h).decode())

    def __len__(self):
        return len(self.configs)

    def initialize(self):
        self._check(lib().ns_batch_init(self._h))

    def step(self, nsteps: int = 1):
        self._check(lib().ns_batch_step(self._h, int(nsteps)))

    def time(self, c: int):
        t, lv, dt = C.c_double(), C.c_uint64(), C.c_double()
        self._check(lib().ns_batch_get_time(self._h, int(c), C.byref(t), C.byref(lv), C.byref(dt)))
        return t.value, lv.value, dt.value

    def field(self, c: int, name) -> np.ndarray:
        k = self.configs[c]
        which = FIELDS.index(name) if isinstance(name, str) else int(name)
        out = np.empty(k.NI * k.NJ * k.NK)
        self._check(lib().ns_batch_get_field(self._h, int(c), which, _dptr(out)))
        return out.reshape(k.NK, k.NJ, k.NI)

    def conserved(self, c: int) -> np.ndarray:
        return np.stack([self.field(c, n) for n in FIELDS[:5]])

    def getConvergenceHistory(self, c: int):
        out = {}
        for q, name in enumerate(FIELDS[:5]):
            cnt = C.c_int64()
            self._check(lib().ns_batch_norm_history(self._h, int(c), q, None, 0, C.byref(cnt)))
            a = np.zeros(cnt.value)
            if cnt.value:
                self._check(lib().ns_batch_norm_history(self._h, int(c), q, _dptr(a), cnt.value, C.byref(cnt)))
            out[name] = a
        return out

    def setOption(self, name: str, value: int):
        self._check(lib().ns_batch_set_option(self._h, name.encode(), int(value)))

    def tile(self):
        ty, kc = C.c_int32(), C.c_int32()
        self._check(lib().ns_batch_get_tile(self._h, C.byref(ty), C.byref(kc)))
        return ty.value, kc.value

    def autotune(self, reps: int = 3):
        self.setOption("autotune", reps)
        return self.tile()

    def close(self):
        if self._h:
            lib().ns_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """Simulation of the reference (src/Simulation.h:17-32): initialize, march until the stop criterion."""

    def __init__(self, params):
        self.solver = Solver(params)
        self.t = 0.0
        self.timeLevel = 0

    def run(self):
        self.solver.initialize()
        t, lv = C.c_double(), C.c_uint64()
        self.solver._check(lib().ns_run(self.solver._h, C.byref(t), C.byref(lv)))
        self.t, self.timeLevel = t.value, lv.value
        return self.t, self.timeLevel

"""Host-side mirror of ASL's LBGK interface over the C ABI.

The names follow ASL's driver-facing API (``src/num/aslLBGK.h``,
``src/num/aslLBGKBC.h``, ``src/num/aslNumMethod.h``): ``LBGK`` with
``init()`` / ``execute()`` / ``getRho()`` / ``getVelocity()`` / ``getF()``,
``LBGKUtilities.initF``, and the BC factories ``generateBCNoSlip``,
``generateBCConstantPressure``, ``generateBCConstantVelocity``,
``generateBCConstantPressureVelocity``, ``generateBCVelocity`` with
``initAll`` / ``executeAll``.  Host views use ASL's ghost-inclusive,
z-fastest layout, as numpy arrays of shape (nx+2, ny+2, nz+2).

``Lattice`` is the thin handle wrapper underneath; ``LBGK.execute()`` asks
the library to keep the step pending (``ASLHIP_STEP_DEFER``), so the step
that precedes a host read is the only one that writes rho and u (what ASL's
per-step writes leave visible is exactly that step's output).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (ASLError, BC_CONSTANT_PRESSURE, BC_CONSTANT_PRESSURE_VELOCITY, BC_CONSTANT_VELOCITY,
                   BC_NOSLIP, BC_NOSLIP_MAP, BC_PRESSURE_VELOCITY_MAP, BC_VELOCITY_MAP, BF16, F16, F32, F64,
                   STEP_DEFER, STEP_GRAPH, STEP_WRITE_MACRO, X0, XE, Y0, YE, Z0, ZE, call)

_DT = {"f32": F32, "f64": F64, "bf16": BF16, "f16": F16, F32: F32, F64: F64, BF16: BF16, F16: F16}


def _d3(u):
    a = np.ascontiguousarray(np.asarray((0.0, 0.0, 0.0) if u is None else u, dtype=np.float64).reshape(3))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Lattice:
    """One lattice handle (``aslhip_lattice``): internal block ``size``, template ``q``."""

    def __init__(self, size, q=19, dtype="f32", periodic=False, nu=None):
        _lib.load_library()
        self.size = tuple(int(s) for s in size)
        self.S = tuple(s + 2 for s in self.size)
        self.len = int(np.prod(self.S))
        self.q = int(q)
        self.dtype = _DT[dtype]
        self.host_dtype = np.float64 if self.dtype == F64 else np.float32
        h = ctypes.c_void_p()
        call("aslhip_lattice_create", ctypes.byref(h), *self.size, self.q, self.dtype, 7 if periodic else 0)
        self.h = h
        self._keep = []
        if nu is not None:
            self.set_viscosity(nu)

    def close(self):
        if getattr(self, "h", None):
            _lib.load_library().aslhip_lattice_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters ---------------------------------------------------------
    def set_viscosity(self, nu):
        call("aslhip_set_viscosity", self.h, float(nu))

    def set_compressible(self, flag=True):
        call("aslhip_set_compressible", self.h, 1 if flag else 0)

    def set_closure(self, params, mu=None, sc=None, out_scale=1.0):
        """Learned closure: network parameters (1345 floats, BF16-rounded inside),
        input shift mu / scale sc (7 each) and output scale; None: plain LBGK."""
        if params is None:
            call("aslhip_closure_set", self.h, None, None, None, 0.0)
            return
        fp = ctypes.POINTER(ctypes.c_float)
        arrs = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
                for a in (params, np.zeros(7) if mu is None else mu, np.ones(7) if sc is None else sc)]
        if arrs[0].size != _lib.CLOSURE_NPARAM or arrs[1].size != 7 or arrs[2].size != 7:
            raise ValueError("closure: 1345 parameters and 7-vectors mu, sc expected")
        call("aslhip_closure_set", self.h, *[a.ctypes.data_as(fp) for a in arrs], float(out_scale))

    def set_closure_mode(self, mode):
        """CLOSURE_BF16_MFMA (default, matrix cores) or CLOSURE_FP32 (vector ALUs, reproducible)."""
        call("aslhip_closure_set_mode", self.h, int(mode))

    def closure_features(self):
        """Network inputs (float32, shape size + (7,)) of every interior cell for the next step."""
        out = np.empty(tuple(self.size) + (7,), dtype=np.float32)
        call("aslhip_closure_features", self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return out

    def set_turbulence(self, on=True):
        """LBGKTurbulence's relaxation limiter (with its equilibrium defined; see include/aslhip.h)."""
        call("aslhip_set_turbulence", self.h, 1 if on else 0)

    def set_omega(self, omega):
        """LBGK::setOmega: angular velocity (3 components) of the rotating frame; None turns it off."""
        if omega is None:
            call("aslhip_set_omega", self.h, None)
        else:
            call("aslhip_set_omega", self.h, (ctypes.c_double * 3)(*[float(x) for x in omega]))

    def set_geometry(self, dx=1.0, position=(0.0, 0.0, 0.0)):
        """The asl::Block the lattice was built on: dx and the internal block's
        position, where rotating velocity-map BCs evaluate omega x (r - centre)."""
        call("aslhip_set_geometry", self.h, float(dx), (ctypes.c_double * 3)(*[float(x) for x in position]))

    def set_vmap_skip_outer(self, on=True):
        """Extension, an opt-in defined value: velocity-map density averages
        leave out neighbours in the block's outer ghost layer (otherwise a
        lattice that reads them is refused)."""
        call("aslhip_set_vmap_skip_outer", self.h, int(bool(on)))

    def set_compute_macro(self, rho=True, vel=True):
        call("aslhip_set_compute_macro", self.h, int(bool(rho)), int(bool(vel)))

    # initialisation -----------------------------------------------------
    def init_equilibrium(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        if np.ndim(rho) == 0 and np.ndim(u) == 1:
            _, up = _d3(u)
            call("aslhip_init_equilibrium", self.h, float(rho), up)
            return
        uf = np.ascontiguousarray(np.asarray(u, dtype=self.host_dtype).reshape(3 * self.len))
        rf = None
        if np.ndim(rho) != 0 or float(rho) != 1.0:
            rf = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, dtype=self.host_dtype), self.S).reshape(self.len))
        call("aslhip_init_equilibrium_field", self.h, None if rf is None else rf.ctypes.data, uf.ctypes.data)

    # boundary conditions --------------------------------------------------
    def add_slice_bc(self, kind, faces, p=0.0, u=None):
        _, up = _d3(u)
        bid = ctypes.c_int(-1)
        call("aslhip_bc_add_slice", self.h, int(kind), int(faces), float(p), up, ctypes.byref(bid))
        return bid.value

    def add_map_bc(self, kind, mapv, cdomain=None, p=0.0, u=None, rot=None):
        m = np.ascontiguousarray(np.asarray(mapv, dtype=np.float32).reshape(self.len))
        cd = None if cdomain is None else np.ascontiguousarray(np.asarray(cdomain, dtype=np.float32).reshape(self.len))
        _, up = _d3(u)
        rr = None if rot is None else np.ascontiguousarray(np.asarray(rot, dtype=np.float64).reshape(6))
        bid = ctypes.c_int(-1)
        fp = ctypes.POINTER(ctypes.c_float)
        call("aslhip_bc_add_map", self.h, int(kind), m.ctypes.data_as(fp), None if cd is None else cd.ctypes.data_as(fp),
             float(p), up, None if rr is None else rr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             ctypes.byref(bid))
        return bid.value

    def execute_bc(self, bc_id):
        call("aslhip_bc_execute", self.h, int(bc_id))

    def finalize(self):
        call("aslhip_finalize", self.h)

    # stepping -------------------------------------------------------------
    def step(self, nsteps=1, write_macro=True, defer=False, graph=False):
        flags = ((STEP_WRITE_MACRO if write_macro else 0) | (STEP_DEFER if defer else 0)
                 | (STEP_GRAPH if graph else 0))
        call("aslhip_step", self.h, int(nsteps), flags)

    def synchronize(self):
        call("aslhip_synchronize", self.h)

    def set_tuning(self, knob, value):
        call("aslhip_set_tuning", self.h, int(knob), int(value))

    def tuning(self):
        """The current launch settings: (workgroup size, XCD run length)."""
        b, r = ctypes.c_int(0), ctypes.c_int(0)
        call("aslhip_get_tuning", self.h, ctypes.byref(b), ctypes.byref(r))
        return b.value, r.value

    def tuning_value(self, knob):
        """The value a tuning knob has now (TUNE_CPL: the cells per lane the step kernel runs with)."""
        v = ctypes.c_int(0)
        call("aslhip_tuning_value", self.h, int(knob), ctypes.byref(v))
        return v.value

    def autotune(self, nsteps=20):
        """Time every launch setting over nsteps real steps and keep the fastest: (block, remap)."""
        b, r = ctypes.c_int(0), ctypes.c_int(0)
        call("aslhip_autotune", self.h, int(nsteps), ctypes.byref(b), ctypes.byref(r))
        return b.value, r.value

    def time_steps(self, nsteps, write_macro=False, graph=False, each=False):
        """(ms of the whole call, summed ms of the step kernels; 0 with graph=True)
        [, per-launch ms as a numpy array if each]."""
        tot = ctypes.c_double(0)
        ker = ctypes.c_double(0)
        per = np.zeros(int(nsteps)) if each else None
        flags = (STEP_WRITE_MACRO if write_macro else 0) | (STEP_GRAPH if graph else 0)
        call("aslhip_time_steps", self.h, int(nsteps), flags, ctypes.byref(tot), ctypes.byref(ker),
             None if per is None else per.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return (tot.value, ker.value, per) if each else (tot.value, ker.value)

    # host views -----------------------------------------------------------
    def rho(self):
        out = np.empty(self.len, dtype=self.host_dtype)
        call("aslhip_read_rho", self.h, out.ctypes.data)
        return out.reshape(self.S)

    def velocity(self):
        out = np.empty(3 * self.len, dtype=self.host_dtype)
        call("aslhip_read_velocity", self.h, out.ctypes.data)
        return out.reshape((3,) + self.S)

    def macro_device(self, stream=None):
        """Device addresses (ints) of rho and u, holding what rho() and
        velocity() would return.  stream: the consumer's hipStream_t (an int
        handle): its later work sees the values, and the lattice's next step
        waits for what it queued before that step; None: the host blocks, and
        the values hold until the next step."""
        r, u = ctypes.c_void_p(), ctypes.c_void_p()
        call("aslhip_macro_device", self.h, None if stream is None else ctypes.c_void_p(stream), ctypes.byref(r),
             ctypes.byref(u))
        return r.value, u.value

    def f(self):
        out = np.empty(self.q * self.len, dtype=self.host_dtype)
        call("aslhip_read_f", self.h, out.ctypes.data)
        return out.reshape((self.q,) + self.S)

    def add_vis(self, kind, mapv, u=None, rot=None):
        """Register a visualisation fill of rho / u at the map's boundary nodes
        (VIS_NOSLIP_VEL, VIS_NOSLIP_RHO, or VIS_VELOCITY with a velocity u or
        rot = (axis, centre)); returns its id for execute_vis."""
        m = np.ascontiguousarray(np.asarray(mapv, dtype=np.float32).reshape(self.len))
        _, up = _d3(u if u is not None else (0.0, 0.0, 0.0))
        rr = None
        if rot is not None:
            rr = np.ascontiguousarray(np.asarray(rot, dtype=np.float64).reshape(6))
        vid = ctypes.c_int(-1)
        call("aslhip_vis_add", self.h, int(kind), m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
             up if u is not None else None,
             None if rr is None else rr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(vid))
        return vid.value

    def execute_vis(self, vis_id):
        call("aslhip_vis_execute", self.h, int(vis_id))

    def surface_force(self, mapv):
        """ComputeSurfaceForceMap: momentum exchange at the surface of mapv, shape (3,) + S."""
        m = np.ascontiguousarray(np.asarray(mapv, dtype=np.float32).reshape(self.len))
        out = np.empty(3 * self.len, dtype=self.host_dtype)
        call("aslhip_surface_force", self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.ctypes.data)
        return out.reshape((3,) + self.S)

    def surface_flux(self, mapv):
        """ComputeSurfaceFluxMap: mass exchange at the surface of mapv, shape S."""
        m = np.ascontiguousarray(np.asarray(mapv, dtype=np.float32).reshape(self.len))
        out = np.empty(self.len, dtype=self.host_dtype)
        call("aslhip_surface_flux", self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.ctypes.data)
        return out.reshape(self.S)

    def save_abd(self, path, dx=1.0):
        """Population checkpoint in ABD encoding (``asl_amd.abd``): name "f", the
        ghost-inclusive block, q components (the host view of ``f()``)."""
        from .abd import write_abd
        write_abd(path, "f", self.f(), dx=dx)

    def load_abd(self, path):
        """Restore populations saved by ``save_abd`` (through ``write_f``)."""
        from .abd import read_abd
        name, _, size, _, comp = read_abd(path)
        if name != "f" or tuple(size) != tuple(self.S) or comp.shape[0] != self.q:
            raise ValueError("%s holds %r over %s x %d, not this lattice's populations" % (path, name, size,
                                                                                         comp.shape[0]))
        self.write_f(comp.astype(self.host_dtype))

    def write_f(self, f):
        """Set the populations (shape (q,) + S, ASL's weight-normalised f_i); rho and u follow."""
        a = np.ascontiguousarray(np.asarray(f, dtype=self.host_dtype).reshape(self.q * self.len))
        call("aslhip_write_f", self.h, a.ctypes.data)

    def cell_types(self):
        out = np.empty(self.len, dtype=np.uint8)
        call("aslhip_read_cell_types", self.h, out.ctypes.data)
        return out.reshape(self.S)

    def conservation(self):
        out = np.empty(5)
        call("aslhip_reduce_conservation", self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def stats(self):
        out = (ctypes.c_int64 * 4)()
        call("aslhip_stats", self.h, out)
        return {"counted": out[0], "computed": out[1], "boundary": out[2], "patterns": out[3]}


class Mlp:
    """Trainer of the closure network on the GPU (hand-written HIP kernels)."""

    def __init__(self, params):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float32).reshape(-1))
        if p.size != _lib.CLOSURE_NPARAM:
            raise ValueError("1345 parameters expected")
        h = ctypes.c_void_p()
        call("aslhip_mlp_create", ctypes.byref(h), p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        self.h = h
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None) and _lib._LIB is not None:
            _lib._LIB.aslhip_mlp_destroy(self.h)
            self.h = None

    def set_data(self, x, t):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1, 7))
        t = np.ascontiguousarray(np.asarray(t, dtype=np.float32).reshape(-1))
        if x.shape[0] != t.size:
            raise ValueError("one target per sample expected")
        fp = ctypes.POINTER(ctypes.c_float)
        call("aslhip_mlp_set_data", self.h, x.ctypes.data_as(fp), t.ctypes.data_as(fp), int(t.size))
        self.n = int(t.size)

    def gradient(self, first=0, count=None):
        count = self.n - first if count is None else count
        g = np.empty(_lib.CLOSURE_NPARAM, dtype=np.float32)
        loss = ctypes.c_double(0)
        call("aslhip_mlp_gradient", self.h, int(first), int(count), g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
             ctypes.byref(loss))
        return loss.value, g

    def train(self, steps, batch, lr, seed=1):
        losses = np.empty(max(int(steps), 1))
        call("aslhip_mlp_train", self.h, int(steps), int(batch), float(lr), int(seed),
             losses.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return losses[:steps]

    def params(self):
        p = np.empty(_lib.CLOSURE_NPARAM, dtype=np.float32)
        call("aslhip_mlp_get_params", self.h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return p


class Batch:
    """Independent lattices advanced together: one step-kernel launch per step for all of them."""

    def __init__(self, lattices):
        self.lattices = list(lattices)
        arr = (ctypes.c_void_p * len(self.lattices))(*[l.h for l in self.lattices])
        h = ctypes.c_void_p()
        call("aslhip_batch_create", ctypes.byref(h), arr, len(self.lattices))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib.load_library().aslhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1, write_macro=True, graph=False):
        flags = (STEP_WRITE_MACRO if write_macro else 0) | (STEP_GRAPH if graph else 0)
        call("aslhip_batch_step", self.h, int(nsteps), flags)

    def time_steps(self, nsteps, graph=True, write_macro=False):
        ms = ctypes.c_double(0)
        flags = (STEP_WRITE_MACRO if write_macro else 0) | (STEP_GRAPH if graph else 0)
        call("aslhip_batch_time_steps", self.h, int(nsteps), flags, ctypes.byref(ms))
        return ms.value

    def set_tuning(self, knob, value):
        call("aslhip_batch_set_tuning", self.h, int(knob), int(value))

    def autotune(self, nsteps=10):
        """Time the workgroup sizes over nsteps real steps each and keep the fastest."""
        b = ctypes.c_int(0)
        call("aslhip_batch_autotune", self.h, int(nsteps), ctypes.byref(b))
        return b.value

    def autotune_groups(self, chunks=2):
        """Time the lattice-group sizes (MiB of populations stepped a whole graph
        chunk at a time) over `chunks` real 100-step chunks each; keep the fastest."""
        m = ctypes.c_int(0)
        call("aslhip_batch_autotune_groups", self.h, int(chunks), ctypes.byref(m))
        return m.value


# ---------------------------------------------------------------------------
# ASL-named layer (src/num/aslLBGK.h, src/num/aslLBGKBC.h)
# ---------------------------------------------------------------------------
class VectorTemplate:
    """``asl::VectorTemplate`` for d3q15() / d3q19() / d3q27(): the library's
    direction vectors c (q x 3), weights w and inverse directions, in ASL's order."""

    def __init__(self, q):
        self.q = q
        c = (ctypes.c_int * (3 * q))()
        w = (ctypes.c_double * q)()
        inv = (ctypes.c_int * q)()
        call("aslhip_template", int(q), c, w, inv)
        self.c = np.array(c[:], dtype=np.int64).reshape(q, 3)
        self.w = np.array(w[:])
        self.inv = np.array(inv[:], dtype=np.int64)


def d3q15():
    return VectorTemplate(15)


def d3q19():
    return VectorTemplate(19)


def d3q27():
    return VectorTemplate(27)


class Block:
    """``asl::Block`` (src/data/aslBlocks.h:56-88): size, dx, position."""

    def __init__(self, size, dx=1.0, position=(0.0, 0.0, 0.0)):
        self.size = tuple(int(s) for s in size)
        self.dx = float(dx)
        self.position = tuple(float(p) for p in position)


class NumMethod:
    """``asl::NumMethod`` (src/num/aslNumMethod.h:34-42)."""

    def init(self):
        raise NotImplementedError

    def execute(self):
        raise NotImplementedError


def initAll(v):
    for m in v:
        m.init()


def executeAll(v):
    for m in v:
        m.execute()


class LBGK(NumMethod):
    """``asl::LBGK(Block, nu, const VectorTemplate*, compVel, compRho)``."""

    def __init__(self, block, nu, vt=None, compVel=True, compRho=True, dtype="f32", periodic=False):
        vt = vt or d3q19()
        self.block = block
        self.vectorTemplate = vt
        self.lattice = Lattice(block.size, vt.q, dtype=dtype, periodic=periodic, nu=nu)
        self.lattice.set_compute_macro(compRho, compVel)
        self.lattice.set_geometry(block.dx, block.position)

    def init(self):
        pass

    def setCompressible(self, flag=True):
        self.lattice.set_compressible(flag)

    def setOmega(self, w):
        """Rotating frame (src/num/aslLBGK.cxx:207-208); None switches it off."""
        self.lattice.set_omega(w)

    def execute(self):
        self.lattice.step(1, defer=True)

    def getRho(self):
        return self.lattice.rho()

    def getVelocity(self):
        return self.lattice.velocity()

    def getF(self):
        return self.lattice.f()

    def getVectorTemplate(self):
        return self.vectorTemplate


class LBGKTurbulence(LBGK):
    """``asl::LBGKTurbulence`` (src/num/aslLBGK.h:163-171): LBGK with the per-cell
    limiter of the relaxation rate, fEq being the cell's own equilibrium."""

    def __init__(self, block, nu, vt=None, compVel=True, compRho=True, dtype="f32", periodic=False):
        super().__init__(block, nu, vt, compVel, compRho, dtype, periodic)
        self.lattice.set_turbulence(True)


class LBGKUtilities:
    """``asl::LBGKUtilities(SPLBGK)::initF`` (src/num/aslLBGK.cxx:262-280)."""

    def __init__(self, lbgk):
        self.num = lbgk

    def initF(self, *args):
        if len(args) == 1:
            rho, vel = 1.0, args[0]
        else:
            rho, vel = args
        self.num.lattice.init_equilibrium(rho, vel)


class _BC(NumMethod):
    def __init__(self, lbgk, bc_id):
        self.num = lbgk
        self.id = bc_id

    def init(self):
        pass

    def execute(self):
        self.num.lattice.execute_bc(self.id)


_SLICES = {"X0": X0, "XE": XE, "Y0": Y0, "YE": YE, "Z0": Z0, "ZE": ZE}


def _faces(sl):
    m = 0
    for s in sl:
        m |= _SLICES[s] if isinstance(s, str) else int(s)
    return m


def generateBCNoSlip(lbgk, sl_or_map):
    if isinstance(sl_or_map, np.ndarray):
        return _BC(lbgk, lbgk.lattice.add_map_bc(BC_NOSLIP_MAP, sl_or_map))
    return _BC(lbgk, lbgk.lattice.add_slice_bc(BC_NOSLIP, _faces(sl_or_map)))


def generateBCConstantVelocity(lbgk, v, sl):
    return _BC(lbgk, lbgk.lattice.add_slice_bc(BC_CONSTANT_VELOCITY, _faces(sl), u=v))


def generateBCConstantPressure(lbgk, p, sl_or_map):
    if isinstance(sl_or_map, np.ndarray):
        return _BC(lbgk, lbgk.lattice.add_map_bc(BC_PRESSURE_VELOCITY_MAP, sl_or_map, p=p))
    return _BC(lbgk, lbgk.lattice.add_slice_bc(BC_CONSTANT_PRESSURE, _faces(sl_or_map), p=p))


def generateBCConstantPressureVelocity(lbgk, p, v, sl_or_map):
    if isinstance(sl_or_map, np.ndarray):
        return _BC(lbgk, lbgk.lattice.add_map_bc(BC_PRESSURE_VELOCITY_MAP, sl_or_map, p=p, u=v))
    return _BC(lbgk, lbgk.lattice.add_slice_bc(BC_CONSTANT_PRESSURE_VELOCITY, _faces(sl_or_map), p=p, u=v))


def generateBCVelocity(lbgk, v, mapv, computationalDomain=None, rotation=None):
    return _BC(lbgk, lbgk.lattice.add_map_bc(BC_VELOCITY_MAP, mapv, cdomain=computationalDomain, u=v, rot=rotation))


class _SurfaceIntegral(NumMethod):
    """ComputeSurfaceForceMap / ComputeSurfaceFluxMap (src/num/aslLBGKBC.h:247-279):
    execute() fills ``field`` ((3,) + S force, or S flux) from the current state."""

    def __init__(self, lbgk, mapv, flux):
        self.num = lbgk
        self.map = np.asarray(mapv, dtype=np.float32)
        self.flux = flux
        self.field = None

    def init(self):
        pass

    def execute(self):
        L = self.num.lattice
        self.field = L.surface_flux(self.map) if self.flux else L.surface_force(self.map)


def generateComputeSurfaceForce(lbgk, mapv):
    return _SurfaceIntegral(lbgk, mapv, False)


def generateComputeSurfaceFlux(lbgk, mapv):
    return _SurfaceIntegral(lbgk, mapv, True)


def readSurface(path, block):
    """``asl::readSurface(file, Block b)`` for an STL file: the map on the nodes of
    b (its ghost-inclusive grid), ``(d - 1.2 dx) / dx / 1.8`` with d capped at 3.6 dx."""
    from . import geometry
    return geometry.read_surface(path, block.size, block.dx, block.position)

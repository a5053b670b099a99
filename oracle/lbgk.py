"""ctypes wrapper of the C++ oracle (test infrastructure only).

Mirrors the driver-facing calls of ASL's ``asl::LBGK`` /
``asl::LBGKUtilities`` / LBGK BC factories (``src/num/aslLBGK.h``,
``src/num/aslLBGKBC.h``) so tests can replay a driver's call sequence on the
CPU restatement.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def load_library():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liblbgk_oracle.so")
    if not os.path.exists(path):
        raise RuntimeError("oracle library not built: run `make -C oracle` (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [ctypes.c_int] * 5 + [ctypes.c_uint]
    lib.orc_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_set_threads.restype = ctypes.c_int
    lib.orc_set_viscosity.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.orc_set_compressible.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_set_omega.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_set_turbulence.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_df_map.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                               ctypes.c_int, ctypes.c_void_p]
    lib.orc_triangle_d2.restype = ctypes.c_double
    lib.orc_triangle_d2.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_surface_distance_map.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_void_p]
    lib.orc_set_closure.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float]
    lib.orc_set_closure.restype = ctypes.c_int
    lib.orc_closure_features.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_set_closure_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_surface.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    lib.orc_vis_fill.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.orc_vis_fill.restype = ctypes.c_int
    lib.orc_closure_features.restype = ctypes.c_int
    lib.orc_set_compute_macro.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.orc_set_geometry.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
    lib.orc_set_vmap_skip_outer.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_init_equilibrium.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
    lib.orc_init_equilibrium_field.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.orc_bc_add_slice.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_double, _dp]
    lib.orc_bc_add_slice.restype = ctypes.c_int
    lib.orc_bc_add_map.argtypes = [ctypes.c_void_p, ctypes.c_int, _fp, _fp, ctypes.c_double, _dp, _dp]
    lib.orc_bc_add_map.restype = ctypes.c_int
    lib.orc_bc_execute.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_bc_execute.restype = ctypes.c_int
    lib.orc_lbgk_execute.argtypes = [ctypes.c_void_p]
    lib.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    for name in ("orc_read_rho", "orc_read_velocity", "orc_read_f", "orc_write_f"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, _dp]
    lib.orc_slice_entries.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_slice_entries.restype = ctypes.c_long
    lib.orc_map_masks.argtypes = [ctypes.c_void_p, ctypes.c_int, _u8p]
    lib.orc_conservation.argtypes = [ctypes.c_void_p, _u8p, _dp]
    lib.orc_template.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp, ctypes.POINTER(ctypes.c_int)]
    lib.orc_template.restype = ctypes.c_int
    _LIB = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _vec3(u):
    return np.ascontiguousarray(np.asarray(u if u is not None else (0.0, 0.0, 0.0), dtype=np.float64).reshape(3))


def template(q):
    """(c[Q,3] int, w[Q] float64, inv[Q] int) of the oracle's lattice template."""
    lib = load_library()
    c = (ctypes.c_int * (3 * 27))()
    w = (ctypes.c_double * 27)()
    inv = (ctypes.c_int * 27)()
    n = lib.orc_template(q, c, w, inv)
    if n < 0:
        raise ValueError("unknown template %r" % (q,))
    return (np.array(c[: 3 * n], dtype=np.int64).reshape(n, 3), np.array(w[:n]), np.array(inv[:n], dtype=np.int64))


def df_map(prog, shape, pos, dx, fp64=False):
    """Distance-function map of a postfix program on shape nodes at pos + dx * index."""
    pr = np.ascontiguousarray(np.asarray(prog, dtype=np.float64).reshape(-1))
    sz = np.ascontiguousarray(np.asarray(shape, dtype=np.int32).reshape(3))
    pa = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3))
    out = np.empty(tuple(int(s) for s in sz), dtype=np.float64 if fp64 else np.float32)
    load_library().orc_df_map(pr.ctypes.data, pr.size, sz.ctypes.data, pa.ctypes.data, float(dx), int(bool(fp64)),
                              out.ctypes.data)
    return out


def triangle_d2(p, tri):
    """Squared distance from point p to the triangle tri (3 x 3 vertices)."""
    pa = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(3))
    ta = np.ascontiguousarray(np.asarray(tri, dtype=np.float64).reshape(9))
    return float(load_library().orc_triangle_d2(pa.ctypes.data, ta.ctypes.data))


def surface_distance_map(tri, size, pos, h, cap, shift, div1, div2):
    """readSurface's map of triangles tri ((n, 3, 3), ASL axis order) on a grid of
    size nodes at pos + h * index, float32 of shape size."""
    t = np.ascontiguousarray(np.asarray(tri, dtype=np.float64).reshape(-1))
    sz = np.ascontiguousarray(np.asarray(size, dtype=np.int32).reshape(3))
    pa = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(3))
    ha = np.ascontiguousarray(np.broadcast_to(np.asarray(h, dtype=np.float64), (3,)))
    out = np.empty(tuple(int(s) for s in sz), dtype=np.float32)
    load_library().orc_surface_distance_map(t.ctypes.data, t.size // 9, sz.ctypes.data, pa.ctypes.data, ha.ctypes.data,
                                            float(cap), float(shift), float(div1), float(div2), out.ctypes.data)
    return out


def set_threads(n):
    return load_library().orc_set_threads(int(n))


class OracleLattice:
    """One ASL LBGK lattice on the CPU oracle (internal block nx*ny*nz)."""

    NOSLIP, CONSTANT_PRESSURE, CONSTANT_VELOCITY, CONSTANT_PRESSURE_VELOCITY = 1, 2, 3, 4
    NOSLIP_MAP, PRESSURE_VELOCITY_MAP, VELOCITY_MAP = 5, 6, 7

    def __init__(self, size, q=19, fp64=False, nu=0.1, periodic=False):
        self.lib = load_library()
        self.size = tuple(int(s) for s in size)
        self.S = tuple(s + 2 for s in self.size)
        self.len = int(np.prod(self.S))
        self.q = q
        self.fp64 = fp64
        self.h = self.lib.orc_create(*self.size, q, 1 if fp64 else 0, 7 if periodic else 0)
        if not self.h:
            raise ValueError("oracle: bad lattice arguments")
        self.lib.orc_set_viscosity(self.h, float(nu))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.orc_destroy(h)
            self.h = None

    def set_viscosity(self, nu):
        self.lib.orc_set_viscosity(self.h, float(nu))

    def set_compressible(self, flag=True):
        self.lib.orc_set_compressible(self.h, 1 if flag else 0)

    def set_closure(self, params, mu=None, sc=None, out_scale=1.0):
        """Learned closure: 1345 network parameters (BF16-rounded inside), input
        shift / scale (7 each) and output scale; params None turns it off."""
        if params is None:
            rc = self.lib.orc_set_closure(self.h, None, None, None, 0.0)
        else:
            self._cl = [np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1)) for a in
                        (params, np.zeros(7) if mu is None else mu, np.ones(7) if sc is None else sc)]
            assert self._cl[0].size == 1345 and self._cl[1].size == 7 and self._cl[2].size == 7
            rc = self.lib.orc_set_closure(self.h, self._cl[0].ctypes.data, self._cl[1].ctypes.data,
                                          self._cl[2].ctypes.data, float(out_scale))
        if rc:
            raise ValueError("the closure needs an FP32 D3Q19 lattice")

    def surface_force(self, mapv):
        """ComputeSurfaceForceMap on the current state, shape (3,) + S (cells with |map| >= 0.9999: 0)."""
        m = np.ascontiguousarray(np.asarray(mapv, dtype=np.float32).reshape(self.len))
        out = np.empty(3 * self.len)
        self.lib.orc_surface(self.h, m.ctypes.data, 0, out.ctypes.data)
        return out.reshape((3,) + self.S)

    def surface_flux(self, mapv):
        """ComputeSurfaceFluxMap on the current state, shape S."""
        m = np.ascontiguousarray(np.asarray(mapv, dtype=np.float32).reshape(self.len))
        out = np.empty(self.len)
        self.lib.orc_surface(self.h, m.ctypes.data, 1, out.ctypes.data)
        return out.reshape(self.S)

    def vis_fill(self, kind, mapv, u=None, rot=None):
        """Visualisation fill of rho / u at the map's boundary nodes, in place:
        kind 1 no-slip velocity, 2 no-slip density, 3 velocity u or rotation
        rot = (axis, centre) (generateBCNoSlipVel / NoSlipRho / VelocityVel)."""
        m = np.ascontiguousarray(np.asarray(mapv, dtype=np.float32).reshape(self.len))
        rr = None if rot is None else np.ascontiguousarray(np.asarray(rot, dtype=np.float64).reshape(6))
        uu = None if u is None else _vec3(u)
        if self.lib.orc_vis_fill(self.h, int(kind), m.ctypes.data, None if uu is None else _d(uu),
                                 None if rr is None else rr.ctypes.data) != 0:
            raise ValueError("oracle: bad visualisation fill")

    def set_closure_mode(self, mode):
        """0: FP32 evaluation (bit-identical to the GPU's vector path); 1: BF16 (the matrix-core path)."""
        self.lib.orc_set_closure_mode(self.h, int(mode))

    def closure_features(self):
        """Features (7 per interior cell, float32, shape size + (7,)) of a periodic D3Q19 lattice."""
        out = np.empty(tuple(self.size) + (7,), dtype=np.float32)
        if self.lib.orc_closure_features(self.h, out.ctypes.data):
            raise ValueError("closure features need a periodic D3Q19 lattice")
        return out

    def set_turbulence(self, on=True):
        """LBGKTurbulence (src/num/aslLBGK.cxx:298-355), fEq the equilibrium of the cell's moments."""
        self.lib.orc_set_turbulence(self.h, 1 if on else 0)

    def set_omega(self, omega):
        """LBGK::setOmega: angular velocity (3 floats) of the rotating frame, None: off."""
        if omega is None:
            self.lib.orc_set_omega(self.h, None)
        else:
            self._omega = np.ascontiguousarray(np.asarray(omega, dtype=np.float64).reshape(3))
            self.lib.orc_set_omega(self.h, self._omega.ctypes.data)

    def set_compute_macro(self, rho=True, vel=True):
        self.lib.orc_set_compute_macro(self.h, int(rho), int(vel))

    def set_vmap_skip_outer(self, on=True):
        """Extension: velocity-map density averages skip outer-ghost-layer neighbours."""
        self.lib.orc_set_vmap_skip_outer(self.h, 1 if on else 0)

    def set_geometry(self, dx=1.0, position=(0.0, 0.0, 0.0)):
        self.lib.orc_set_geometry(self.h, float(dx), _d(_vec3(position)))

    def init_f(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        """LBGKUtilities::initF with constants or fields (ASL layout arrays)."""
        if np.ndim(u) == 1 and np.ndim(rho) == 0:
            self.lib.orc_init_equilibrium(self.h, float(rho), _d(_vec3(u)))
            return
        uf = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(3 * self.len))
        if np.ndim(rho) == 0:
            rf = np.full(self.len, float(rho))
        else:
            rf = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(self.len))
        self.lib.orc_init_equilibrium_field(self.h, _d(rf), _d(uf))

    def add_slice_bc(self, kind, faces, p=0.0, u=None):
        r = self.lib.orc_bc_add_slice(self.h, kind, int(faces), float(p), _d(_vec3(u)))
        if r < 0:
            raise ValueError("oracle: bad slice BC")
        return r

    def add_map_bc(self, kind, mapv, cdomain=None, p=0.0, u=None, rot=None):
        m = np.ascontiguousarray(np.asarray(mapv, dtype=np.float32).reshape(self.len))
        cd = None if cdomain is None else np.ascontiguousarray(np.asarray(cdomain, dtype=np.float32).reshape(self.len))
        rr = None if rot is None else np.ascontiguousarray(np.asarray(rot, dtype=np.float64).reshape(6))
        r = self.lib.orc_bc_add_map(
            self.h, kind, m.ctypes.data_as(_fp), None if cd is None else cd.ctypes.data_as(_fp), float(p),
            _d(_vec3(u)), None if rr is None else _d(rr))
        if r < 0:
            raise ValueError("oracle: bad map BC")
        return r

    def execute_bc(self, bc_id):
        if self.lib.orc_bc_execute(self.h, int(bc_id)) != 0:
            raise ValueError("oracle: unknown BC id")

    def execute(self):
        """LBGK::execute(): preProcessing + collide."""
        self.lib.orc_lbgk_execute(self.h)

    def run(self, nsteps, order=()):
        """nsteps x (execute(); execute each BC of ``order``)."""
        arr = (ctypes.c_int * max(1, len(order)))(*order)
        self.lib.orc_run(self.h, int(nsteps), arr, len(order))

    def rho(self):
        out = np.empty(self.len)
        self.lib.orc_read_rho(self.h, _d(out))
        return out.reshape(self.S)

    def velocity(self):
        out = np.empty(3 * self.len)
        self.lib.orc_read_velocity(self.h, _d(out))
        return out.reshape((3,) + self.S)

    def f(self):
        out = np.empty(self.q * self.len)
        self.lib.orc_read_f(self.h, _d(out))
        return out.reshape((self.q,) + self.S)

    def write_f(self, f):
        a = np.ascontiguousarray(np.asarray(f, dtype=np.float64).reshape(self.q * self.len))
        self.lib.orc_write_f(self.h, _d(a))

    def slice_entries(self, bc_id):
        return int(self.lib.orc_slice_entries(self.h, int(bc_id)))

    def map_masks(self, bc_id):
        out = np.empty(self.len, dtype=np.uint8)
        self.lib.orc_map_masks(self.h, int(bc_id), out.ctypes.data_as(_u8p))
        return out.reshape(self.S)

    def conservation(self, mask):
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(self.len))
        out = np.empty(5)
        self.lib.orc_conservation(self.h, m.ctypes.data_as(_u8p), _d(out))
        return out

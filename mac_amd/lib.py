"""ctypes binding of libmac_amd.so (declarations in include/mac_amd.h).

The Python names mirror the reference's projection slots:
``Projection.Init_Iteration`` uploads the brick grid (the matrix-free
counterpart of Algorithms_Solve_Pressure::Init_Iteration), and
``Projection.Divergence_Projection`` runs the solve and the velocity update.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmac_amd.so")

MAC_FP64 = 0
MAC_MIXED = 1

AIR, FLUID, SOLID, INFLOW, INTERIOR, ZERO_PRES = 0, 1, 2, 3, 4, 5
CONST_X, CONST_Y, CONST_Z, EXISTS = 0x08, 0x10, 0x20, 0x80


class MacError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("h", ctypes.c_double * 3),
        ("rtol", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("nu", ctypes.c_int),
        ("omega", ctypes.c_double),
        ("check_every", ctypes.c_int),
        ("stall_iters", ctypes.c_int),
        ("smoother", ctypes.c_int),
        ("cheb_lower", ctypes.c_double),
        ("nu_coarse", ctypes.c_int),
        ("cheb_lower_coarse", ctypes.c_double),
        ("coarse_tail", ctypes.c_int),
        ("viscosity", ctypes.c_double),
        ("coarse_inner", ctypes.c_int),
        ("coarse_damp", ctypes.c_double),
        ("fuse_coarse", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int),
        ("restarts", ctypes.c_int),
        ("fallback_used", ctypes.c_int),
        ("levels", ctypes.c_int),
        ("n_unknowns", ctypes.c_longlong),
        ("n_bricks", ctypes.c_longlong),
        ("n_active_bricks", ctypes.c_longlong),
        ("rel_resid", ctypes.c_double),
        ("true_rel_resid", ctypes.c_double),
        ("max_div", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("ms_solve", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); the exported C-ABI, kept in step with include/mac_amd.h
SIGNATURES = {
    "mac_default_params": (None, [ctypes.POINTER(Params)]),
    "mac_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(Params)]),
    "mac_destroy": (None, [_P]),
    "mac_set_params": (ctypes.c_int, [_P, ctypes.POINTER(Params)]),
    "mac_last_error": (ctypes.c_char_p, []),
    "mac_upload_bricks": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _P, _P]),
    "mac_set_velocity": (ctypes.c_int, [_P, _P, _P, _P]),
    "mac_get_velocity": (ctypes.c_int, [_P, _P, _P, _P]),
    "mac_set_real_velocity": (ctypes.c_int, [_P, _P, _P, _P]),
    "mac_get_pressure": (ctypes.c_int, [_P, _P]),
    "mac_set_pressure": (ctypes.c_int, [_P, _P]),
    "mac_save_velocity": (ctypes.c_int, [_P]),
    "mac_restore_velocity": (ctypes.c_int, [_P]),
    "mac_project": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(Stats)]),
    "mac_divergence": (ctypes.c_int, [_P, _P]),
    "mac_apply_operator": (ctypes.c_int, [_P, _P, _P]),
    "mac_gradient_update": (ctypes.c_int, [_P]),
    "mac_max_divergence": (ctypes.c_int, [_P, _dp]),
    "mac_solve_csr": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_double]),
    "mac_solve_csr_noalloc": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_double,
                                             ctypes.POINTER(ctypes.c_int)]),
    "mac_enable_kernel_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "mac_kernel_times": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _dp,
                                         ctypes.POINTER(ctypes.c_longlong), _dp]),
    "mac_reset_kernel_times": (ctypes.c_int, [_P]),
    "mac_timing_overhead": (ctypes.c_int, [_P, ctypes.c_int, _dp]),
    "mac_sync": (ctypes.c_int, [_P]),
}

_lib = None


def load():
    """Load libmac_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MacError(f"{LIB_PATH} is missing: build it with `python -m mac_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc):
    if rc != 0:
        msg = load().mac_last_error().decode(errors="replace")
        raise MacError(f"mac_amd error {rc}: {msg}")


def default_params(**kw):
    p = Params()
    load().mac_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "h":
            if np.isscalar(v):
                v = (v, v, v)
            for i in range(3):
                p.h[i] = float(v[i])
        else:
            setattr(p, k, v)
    return p


class Projection:
    """A projection context on the current GPU (one per host thread)."""

    def __init__(self, **params):
        self.lib = load()
        self.ctx = ctypes.c_void_p()
        self.params = default_params(**params)
        _check(self.lib.mac_create(ctypes.byref(self.ctx), ctypes.byref(self.params)))
        self.nb = 0

    def close(self):
        if self.ctx:
            self.lib.mac_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **kw):
        for k, v in kw.items():
            if k == "h":
                raise MacError("h is fixed once the context exists")
            setattr(self.params, k, v)
        _check(self.lib.mac_set_params(self.ctx, ctypes.byref(self.params)))

    # -- grid / fields -------------------------------------------------------
    def Init_Iteration(self, keys, flags, u=None, v=None, w=None):
        keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 3)
        flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        nb = keys.shape[0]
        if flags.size != nb * 512:
            raise MacError("flags must hold 512 bytes per brick")
        self.nb = nb
        uu = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (u, v, w)]
        _check(self.lib.mac_upload_bricks(self.ctx, nb, _ptr(keys), _ptr(flags), *[_ptr(a) for a in uu]))

    upload = Init_Iteration

    def set_velocity(self, u, v, w):
        a = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (u, v, w)]
        for x in a:
            if x.size != self.nb * 512:
                raise MacError("velocity arrays must hold 512 values per brick")
        _check(self.lib.mac_set_velocity(self.ctx, *[_ptr(x) for x in a]))

    def set_real_velocity(self, u=None, v=None, w=None):
        """Real speed for the viscous air pressures (None: the projected field)."""
        if u is None:
            _check(self.lib.mac_set_real_velocity(self.ctx, None, None, None))
            return
        a = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (u, v, w)]
        for x in a:
            if x.size != self.nb * 512:
                raise MacError("velocity arrays must hold 512 values per brick")
        _check(self.lib.mac_set_real_velocity(self.ctx, *[_ptr(x) for x in a]))

    def get_velocity(self):
        a = [np.empty(self.nb * 512, dtype=np.float64) for _ in range(3)]
        _check(self.lib.mac_get_velocity(self.ctx, *[_ptr(x) for x in a]))
        return a

    def get_pressure(self):
        p = np.empty(self.nb * 512, dtype=np.float64)
        _check(self.lib.mac_get_pressure(self.ctx, _ptr(p)))
        return p

    def set_pressure(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        _check(self.lib.mac_set_pressure(self.ctx, _ptr(p)))

    def save_velocity(self):
        _check(self.lib.mac_save_velocity(self.ctx))

    def restore_velocity(self):
        _check(self.lib.mac_restore_velocity(self.ctx))

    # -- projection -----------------------------------------------------------
    def Divergence_Projection(self, guess=False, raise_on_noconv=True):
        st = Stats()
        rc = self.lib.mac_project(self.ctx, 1 if guess else 0, ctypes.byref(st))
        if rc != 0 and (raise_on_noconv or rc != -3):
            _check(rc)
        return st.as_dict()

    project = Divergence_Projection

    def divergence(self):
        out = np.empty(self.nb * 512, dtype=np.float64)
        _check(self.lib.mac_divergence(self.ctx, _ptr(out)))
        return out

    def apply_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        y = np.zeros(self.nb * 512, dtype=np.float64)
        _check(self.lib.mac_apply_operator(self.ctx, _ptr(x), _ptr(y)))
        return y

    def gradient_update(self):
        _check(self.lib.mac_gradient_update(self.ctx))

    def max_divergence(self):
        out = ctypes.c_double()
        _check(self.lib.mac_max_divergence(self.ctx, ctypes.byref(out)))
        return out.value

    # -- kernel timing --------------------------------------------------------
    def enable_timing(self, on=True):
        _check(self.lib.mac_enable_kernel_timing(self.ctx, 1 if on else 0))

    def reset_timing(self):
        _check(self.lib.mac_reset_kernel_times(self.ctx))

    def timing_overhead_us(self, n=200):
        """Mean event-pair duration around an empty launch (microseconds)."""
        us = ctypes.c_double()
        _check(self.lib.mac_timing_overhead(self.ctx, int(n), ctypes.byref(us)))
        return us.value

    def kernel_times(self):
        cap = 32
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        n = (ctypes.c_longlong * cap)()
        by = (ctypes.c_double * cap)()
        k = self.lib.mac_kernel_times(self.ctx, cap, names, ms, n, by)
        return {names[i].decode(): {"ms": ms[i], "launches": n[i], "bytes": by[i]} for i in range(k)}

    def sync(self):
        _check(self.lib.mac_sync(self.ctx))


class LinearSolver:
    """Slot A: Policy_Solve_Linear_Umfpack's three calls on the GPU
    (Policy_Solve_Linear_Umfpack.h:28-70).  The CSR arrays are the caller's,
    as in the reference; FactorizeMatrice only records them."""

    def __init__(self, rtol=1e-13):
        self.proj = Projection()
        self.rtol = rtol
        self.mat = None

    def Solve_Linear_FactorizeMatrice(self, n, offset, indice, value):
        self.mat = (int(n), np.ascontiguousarray(offset, dtype=np.int32),
                    np.ascontiguousarray(indice, dtype=np.int32), np.ascontiguousarray(value, dtype=np.float64))

    def Solve_Linear(self, n, offset, indice, value, b, res=None):
        n = int(n)
        off = np.ascontiguousarray(offset, dtype=np.int32)
        idx = np.ascontiguousarray(indice, dtype=np.int32)
        val = np.ascontiguousarray(value, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros(n, dtype=np.float64) if res is None else res
        _check(self.proj.lib.mac_solve_csr(self.proj.ctx, n, _ptr(off), _ptr(idx), _ptr(val), _ptr(b), _ptr(out),
                                           self.rtol))
        return out

    def Solve_Linear_Clean(self):
        self.mat = None


# ---------------------------------------------------------------- time step --
class SimParams(ctypes.Structure):
    _fields_ = [
        ("h", ctypes.c_double * 3),
        ("gravity", ctypes.c_double * 3),
        ("viscosity", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("dt_min", ctypes.c_double),
        ("dt_max", ctypes.c_double),
        ("layer_max", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("particle_rk", ctypes.c_int),
        ("particles_per_inflow_cell", ctypes.c_int),
        ("seed", ctypes.c_ulonglong),
        ("box_lo", ctypes.c_double * 3),
        ("box_hi", ctypes.c_double * 3),
        ("max_particles", ctypes.c_int),
        ("max_bricks", ctypes.c_int),
        ("depth_interior", ctypes.c_int),
        ("depth_remove", ctypes.c_int),
        ("floor_axis", ctypes.c_int),
        ("floor_value", ctypes.c_double),
        ("fill_cube_n", ctypes.c_int),
        ("inflow_speed", ctypes.c_double * 3),
        ("inflow_amp", ctypes.c_double),
        ("inflow_period", ctypes.c_int),
        ("surface_bc", ctypes.c_int),
        ("proj", Params),
    ]


class SimStats(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("nbricks", ctypes.c_int),
        ("nparticles", ctypes.c_int),
        ("nfluid", ctypes.c_longlong),
        ("pcg_iters", ctypes.c_int * 2),
        ("max_div", ctypes.c_double),
        ("ms_topology", ctypes.c_double),
        ("ms_forces", ctypes.c_double),
        ("ms_project", ctypes.c_double),
        ("ms_particles", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("true_rel_resid", ctypes.c_double),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["pcg_iters"] = list(self.pcg_iters)
        return d


SIGNATURES.update({
    "mac_sim_default_params": (None, [ctypes.POINTER(SimParams)]),
    "mac_sim_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(SimParams)]),
    "mac_sim_destroy": (None, [_P]),
    "mac_sim_upload": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _P, _P, _P, ctypes.c_int, _P]),
    "mac_sim_counts": (ctypes.c_int, [_P, _ip, _ip]),
    "mac_sim_get_step": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_ulonglong)]),
    "mac_sim_set_step": (ctypes.c_int, [_P, ctypes.c_ulonglong]),
    "mac_sim_download": (ctypes.c_int, [_P] + [_P] * 11),
    "mac_sim_topology": (ctypes.c_int, [_P]),
    "mac_sim_compute_dt": (ctypes.c_int, [_P, _dp]),
    "mac_sim_forces": (ctypes.c_int, [_P, ctypes.c_int]),
    "mac_sim_advance": (ctypes.c_int, [_P, ctypes.c_double, ctypes.POINTER(SimStats)]),
    "mac_sim_move_particles": (ctypes.c_int, [_P, ctypes.c_double]),
    "mac_sim_interpolate": (ctypes.c_int, [_P, ctypes.c_int, _P, _P]),
    "mac_sim_step": (ctypes.c_int, [_P, ctypes.POINTER(SimStats)]),
    "mac_sim_extrapolate": (ctypes.c_int, [_P, ctypes.c_int]),
    "mac_sim_surface_bc": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    "mac_sim_get_depth": (ctypes.c_int, [_P, _P]),
    "mac_sim_set_depth": (ctypes.c_int, [_P, _P]),
})


def sim_params(**kw):
    p = SimParams()
    load().mac_sim_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("h", "gravity", "box_lo", "box_hi", "inflow_speed"):
            if np.isscalar(v):
                v = (v, v, v)
            arr = getattr(p, k)
            for i in range(3):
                arr[i] = float(v[i])
        elif k == "proj":
            for kk, vv in v.items():
                setattr(p.proj, kk, vv)
        else:
            setattr(p, k, v)
    for i in range(3):
        p.proj.h[i] = p.h[i]
    return p


class Sim:
    """Device-resident time stepper (mac_sim_* of include/mac_amd.h)."""

    def __init__(self, **params):
        self.lib = load()
        self.params = sim_params(**params)
        self.ptr = ctypes.c_void_p()
        _check(self.lib.mac_sim_create(ctypes.byref(self.ptr), ctypes.byref(self.params)))

    def close(self):
        if self.ptr:
            self.lib.mac_sim_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, keys, flags, layer=None, u=None, v=None, w=None, particles=None):
        keys = np.ascontiguousarray(keys, dtype=np.int32).reshape(-1, 3)
        nb = keys.shape[0]
        flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        lay = None if layer is None else np.ascontiguousarray(layer, dtype=np.int8).reshape(-1)
        uu = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (u, v, w)]
        parts = np.zeros((0, 3)) if particles is None else np.ascontiguousarray(particles, dtype=np.float64).reshape(-1, 3)
        _check(self.lib.mac_sim_upload(self.ptr, nb, _ptr(keys), _ptr(flags), _ptr(lay), *[_ptr(a) for a in uu],
                                       parts.shape[0], _ptr(parts)))

    def counts(self):
        nb, npart = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.mac_sim_counts(self.ptr, ctypes.byref(nb), ctypes.byref(npart)))
        return nb.value, npart.value

    def get_step(self):
        n = ctypes.c_ulonglong()
        _check(self.lib.mac_sim_get_step(self.ptr, ctypes.byref(n)))
        return n.value

    def set_step(self, n):
        _check(self.lib.mac_sim_set_step(self.ptr, int(n)))

    def download(self):
        nb, npart = self.counts()
        out = {"keys": np.zeros((nb, 3), np.int32), "flags": np.zeros(nb * 512, np.uint8),
               "layer": np.zeros(nb * 512, np.int8)}
        for k in ("u", "v", "w", "ax", "ay", "az", "p"):
            out[k] = np.zeros(nb * 512)
        out["particles"] = np.zeros((npart, 3))
        _check(self.lib.mac_sim_download(self.ptr, *[_ptr(out[k]) for k in (
            "keys", "flags", "layer", "u", "v", "w", "ax", "ay", "az", "p", "particles")]))
        return out

    def get_depth(self):
        """Per-cell depth (int16, -1 = empty) in download order."""
        nb, _ = self.counts()
        d = np.zeros(nb * 512, np.int16)
        _check(self.lib.mac_sim_get_depth(self.ptr, _ptr(d)))
        return d

    def set_depth(self, depth):
        nb, _ = self.counts()
        d = np.ascontiguousarray(depth, dtype=np.int16).reshape(-1)
        if d.size != nb * 512:
            raise ValueError("depth needs one value per cell of the current bricks")
        _check(self.lib.mac_sim_set_depth(self.ptr, _ptr(d)))

    def topology(self):
        _check(self.lib.mac_sim_topology(self.ptr))

    def compute_dt(self):
        d = ctypes.c_double()
        _check(self.lib.mac_sim_compute_dt(self.ptr, ctypes.byref(d)))
        return d.value

    def forces(self, slot=0):
        _check(self.lib.mac_sim_forces(self.ptr, slot))

    def advance(self, dt):
        st = SimStats()
        _check(self.lib.mac_sim_advance(self.ptr, dt, ctypes.byref(st)))
        return st.as_dict()

    def move_particles(self, dt):
        _check(self.lib.mac_sim_move_particles(self.ptr, dt))

    def interpolate(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.zeros_like(pos)
        _check(self.lib.mac_sim_interpolate(self.ptr, pos.shape[0], _ptr(pos), _ptr(vel)))
        return vel

    def step(self):
        st = SimStats()
        _check(self.lib.mac_sim_step(self.ptr, ctypes.byref(st)))
        return st.as_dict()

    def extrapolate(self, slot=0):
        """Algorithms_Extrapolate::Do on speed slot `slot` (with surface_bc set,
        the free-surface velocity condition first)."""
        _check(self.lib.mac_sim_extrapolate(self.ptr, int(slot)))

    def surface_bc(self, slot=0, dim=3):
        """The free-surface velocity condition on speed slot `slot` (3 = the
        acceleration), in 3-D or (dim=2) on the x-y plane."""
        _check(self.lib.mac_sim_surface_bc(self.ptr, int(slot), int(dim)))

    def set_guess(self, mode, net=None, out_scale=1.0):
        """Projection starting iterate: GUESS_ZERO, GUESS_PREVIOUS or GUESS_LEARNED (with net)."""
        self._net = net  # the library does not own the network
        _check(self.lib.mac_sim_set_guess(self.ptr, int(mode), net.ptr if net is not None else None, float(out_scale)))


# ----------------------------------------------------------------- U-Net ----
SIGNATURES.update({
    "mac_unet_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong]),
    "mac_unet_destroy": (None, [_P]),
    "mac_unet_num_params": (ctypes.c_longlong, [_P]),
    "mac_unet_get_params": (ctypes.c_int, [_P, _P]),
    "mac_unet_set_params": (ctypes.c_int, [_P, _P]),
    "mac_unet_forward": (ctypes.c_int, [_P, _P, _P]),
    "mac_unet_backward": (ctypes.c_int, [_P, _P, _P]),
    "mac_unet_train_step": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "mac_unet_conv_timing": (ctypes.c_int, [_P, ctypes.c_int, _dp, _dp]),
    "mac_unet_wgrad_mode": (ctypes.c_int, [_P, ctypes.c_int]),
    "mac_unet_guess": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_float]),
    "mac_unet_sample": (ctypes.c_int, [_P, _P, _P, ctypes.c_float, _P, _P, _P]),
    "mac_sim_set_guess": (ctypes.c_int, [_P, ctypes.c_int, _P, ctypes.c_float]),
    "mac_sim_capture": (ctypes.c_int, [_P, _P, ctypes.c_float, _P, _P, _P]),
})

GUESS_ZERO, GUESS_PREVIOUS, GUESS_LEARNED = 0, 1, 2


class UNet:
    """The learned-guess network (mac_unet_* of include/mac_amd.h)."""

    def __init__(self, D, H, W, seed=0):
        self.lib = load()
        self.ptr = ctypes.c_void_p()
        self.shape = (D, H, W)
        _check(self.lib.mac_unet_create(ctypes.byref(self.ptr), D, H, W, seed))
        self.nparams = int(self.lib.mac_unet_num_params(self.ptr))

    def close(self):
        if self.ptr:
            self.lib.mac_unet_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.zeros(self.nparams, np.float32)
        _check(self.lib.mac_unet_get_params(self.ptr, _ptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        _check(self.lib.mac_unet_set_params(self.ptr, _ptr(p)))

    def forward(self, x5):
        x5 = np.ascontiguousarray(x5, np.float32).reshape(-1, 5)
        out = np.zeros(x5.shape[0], np.float32)
        _check(self.lib.mac_unet_forward(self.ptr, _ptr(x5), _ptr(out)))
        return out

    def backward(self, gout):
        gout = np.ascontiguousarray(gout, np.float32).reshape(-1)
        g = np.zeros(self.nparams, np.float32)
        _check(self.lib.mac_unet_backward(self.ptr, _ptr(gout), _ptr(g)))
        return g

    def train_step(self, x5, target, mask, lr):
        x5 = np.ascontiguousarray(x5, np.float32).reshape(-1, 5)
        t = np.ascontiguousarray(target, np.float32).reshape(-1)
        m = np.ascontiguousarray(mask, np.float32).reshape(-1)
        loss = ctypes.c_float()
        _check(self.lib.mac_unet_train_step(self.ptr, _ptr(x5), _ptr(t), _ptr(m), float(lr), ctypes.byref(loss)))
        return loss.value

    def guess(self, proj, use_prev=False, out_scale=1.0):
        """Write the learned starting iterate into proj's pressure."""
        _check(self.lib.mac_unet_guess(self.ptr, proj.ctx, 1 if use_prev else 0, float(out_scale)))

    def _pair_buffers(self):
        n = int(np.prod(self.shape))
        return np.zeros((n, 5), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32)

    def sample(self, proj, prev=None, out_scale=1.0):
        """(in5, target, mask) of proj's last projection; prev in brick layout or None."""
        x5, t, m = self._pair_buffers()
        pv = None if prev is None else np.ascontiguousarray(prev, np.float64).reshape(-1)
        _check(self.lib.mac_unet_sample(self.ptr, proj.ctx, _ptr(pv), float(out_scale), _ptr(x5), _ptr(t), _ptr(m)))
        return x5, t, m

    def capture(self, sim, out_scale=1.0):
        """(in5, target, mask) of the simulation's last projection."""
        x5, t, m = self._pair_buffers()
        _check(self.lib.mac_sim_capture(sim.ptr, self.ptr, float(out_scale), _ptr(x5), _ptr(t), _ptr(m)))
        return x5, t, m

    def wgrad_mode(self, mode):
        """0: weight gradients on the matrix cores (default); 1: on FP32 FMA."""
        _check(self.lib.mac_unet_wgrad_mode(self.ptr, int(mode)))

    def conv_timing(self, on):
        ms, fl = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.mac_unet_conv_timing(self.ptr, 1 if on else 0, ctypes.byref(ms), ctypes.byref(fl)))
        return ms.value, fl.value

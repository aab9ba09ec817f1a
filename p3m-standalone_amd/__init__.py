"""Python mirror of the reference's method-table interface over libp3m_hip.so.

The C structs below are byte-identical to fweik/p3m-standalone's types.h
(vector_array_t :126-132, forces_t :152-159, system_t :163-183,
parameters_t :202-215, data_t :252-286, method_t :306-316) and the calls go
through the library's exported method_t objects, exactly as the reference's
main.c drives a method (main.c:361 Init, :378 Influence_function,
common.c:305 Kspace_force, main.c:432-433 Error / Error_k).

The library must be built (make -C p3m-standalone_amd); loading fails loudly
otherwise -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libp3m_hip.so")

METHOD_IDS = (0, 1, 2, 3, 6, 7)
METHOD_NAMES = {0: "p3m-ik", 1: "p3m-ik-i", 2: "p3m-ad", 3: "p3m-ad-i", 6: "p3m-ik-r", 7: "p3m-ad-r"}
FP64, MIXED = 0, 1

OK, EINVAL, ERANGE, ENOMEM, EDEVICE, ENODEVICE = 0, 1, 2, 3, 4, 5

_dp = ctypes.POINTER(ctypes.c_double)


class VectorArray(ctypes.Structure):
    _fields_ = [("size", ctypes.c_int), ("x", _dp), ("y", _dp), ("z", _dp),
                ("fields", ctypes.POINTER(_dp))]


class Forces(ctypes.Structure):
    _fields_ = [("f", ctypes.POINTER(VectorArray)), ("f_k", ctypes.POINTER(VectorArray)),
                ("f_r", ctypes.POINTER(VectorArray))]


class System(ctypes.Structure):
    _fields_ = [("length", ctypes.c_double), ("nparticles", ctypes.c_int),
                ("p", ctypes.POINTER(VectorArray)), ("q", _dp), ("q2", ctypes.c_double),
                ("v", ctypes.POINTER(VectorArray)), ("reference", ctypes.POINTER(Forces)),
                ("epsilon", ctypes.c_double), ("energy", ctypes.c_double)]


class Parameters(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_double), ("rcut", ctypes.c_double),
                ("prefactor", ctypes.c_double), ("mesh", ctypes.c_int), ("ip", ctypes.c_int),
                ("cao", ctypes.c_int), ("cao3", ctypes.c_int), ("precision", ctypes.c_double),
                ("tuning", ctypes.c_int)]


class Runtime(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("t_c", ctypes.c_double), ("t_g", ctypes.c_double),
                ("t_f", ctypes.c_double)]


class Data(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_int), ("G_hat", _dp), ("Qmesh", _dp),
                ("Fmesh", ctypes.POINTER(VectorArray)), ("nshift", _dp), ("Dn", _dp),
                ("dQ", _dp * 2), ("ca_ind", ctypes.POINTER(ctypes.c_int) * 2), ("cf", _dp * 2),
                ("inter", ctypes.c_void_p), ("forward_plans", ctypes.c_int),
                ("backward_plans", ctypes.c_int), ("forward_plan", ctypes.c_void_p * 3),
                ("backward_plan", ctypes.c_void_p * 3), ("neighbor_list", ctypes.c_void_p),
                ("self_force_corrections", _dp), ("method_data", ctypes.c_void_p),
                ("runtime", Runtime)]


class Timing(ctypes.Structure):
    """timing_t (types.h:229-235)."""
    _fields_ = [("avg", ctypes.c_double), ("sgm", ctypes.c_double), ("min", ctypes.c_double),
                ("max", ctypes.c_double), ("n", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k in ("avg", "sgm", "min", "max", "n")}


class RuntimeStat(ctypes.Structure):
    """runtime_stat_t (types.h:237-242)."""
    _fields_ = [("t", Timing), ("t_c", Timing), ("t_g", Timing), ("t_f", Timing)]

    def as_dict(self):
        return {k: getattr(self, k).as_dict() for k in ("t", "t_c", "t_g", "t_f")}


class BatchSystem(ctypes.Structure):
    _fields_ = [("nparticles", ctypes.c_int), ("length", ctypes.c_double), ("mesh", ctypes.c_int),
                ("cao", ctypes.c_int), ("alpha", ctypes.c_double), ("rcut", ctypes.c_double)]


_SP, _PP, _DPT, _FP = (ctypes.POINTER(System), ctypes.POINTER(Parameters), ctypes.POINTER(Data),
                       ctypes.POINTER(Forces))
InitFn = ctypes.CFUNCTYPE(_DPT, _SP, _PP)
InfluenceFn = ctypes.CFUNCTYPE(None, _SP, _PP, _DPT)
KspaceFn = ctypes.CFUNCTYPE(None, _SP, _PP, _DPT, _FP)
ErrorFn = ctypes.CFUNCTYPE(ctypes.c_double, _SP, _PP)
TuneTimerFn = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, _SP, _PP,
                               ctypes.POINTER(RuntimeStat))


class Method(ctypes.Structure):
    _fields_ = [("method_id", ctypes.c_int), ("method_name", ctypes.c_char_p),
                ("method_name_short", ctypes.c_char_p), ("flags", ctypes.c_char),
                ("Init", InitFn), ("Influence_function", InfluenceFn), ("Kspace_force", KspaceFn),
                ("Error", ErrorFn), ("Error_k", ErrorFn)]


_lib = None


def lib():
    """Load libp3m_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libp3m_hip.so is not built; run `make -C %s`" % _HERE)
        try:
            # torch ships its own HIP runtime under the same soname
            # (libamdhip64.so.7); whichever is loaded first serves the whole
            # process, and torch fails ("No HIP GPUs") on another one, so
            # torch is loaded first when present
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        i, d, vp, ip = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)
        sigs = {
            "p3m_hip_last_error": (ctypes.c_char_p, []),
            "p3m_hip_method": (ctypes.POINTER(Method), [i]),
            "p3m_hip_init": (i, [i, _SP, _PP, i, ctypes.POINTER(_DPT)]),
            "p3m_hip_influence_function": (i, [_SP, _PP, _DPT]),
            "p3m_hip_kspace_force": (i, [_SP, _PP, _DPT, _FP]),
            "p3m_hip_free": (i, [_DPT]),
            "p3m_hip_set_precision": (i, [_DPT, i]),
            "p3m_hip_kspace_force_device": (i, [_DPT, i, d, vp, vp, vp, vp, vp, vp]),
            "p3m_hip_synchronize": (i, [_DPT]),
            "p3m_hip_realteil": (i, [_SP, _PP, _FP]),
            "p3m_hip_realteil_device": (i, [_DPT, i, d, vp, vp, vp, vp, d, d, vp, vp]),
            "p3m_hip_calculate_forces": (i, [ctypes.POINTER(Method), _SP, _PP, _DPT, _FP]),
            "p3m_hip_calculate_forces_device": (i, [_DPT, i, d, vp, vp, vp, vp, d, d, vp, vp, vp, vp]),
            "p3m_hip_error": (i, [i, _SP, _PP, _dp, _dp]),
            "p3m_hip_set_error_table": (i, [i, i, i, _dp]),
            "p3m_hip_stage_times": (i, [_DPT, _dp]),
            "p3m_hip_kspace_force_vjp": (i, [_SP, _PP, _DPT, ctypes.POINTER(VectorArray),
                                             ctypes.POINTER(VectorArray), _dp]),
            "p3m_hip_kspace_force_vjp_device": (i, [_DPT, i, d, vp, vp, vp, vp, vp, vp, vp]),
            "p3m_hip_realteil_vjp": (i, [_SP, _PP, ctypes.POINTER(VectorArray),
                                         ctypes.POINTER(VectorArray), _dp]),
            "p3m_hip_realteil_vjp_device": (i, [_DPT, i, d, vp, vp, vp, vp, d, d, vp, vp, vp]),
            "p3m_hip_batch_create": (i, [i, i, i, ctypes.POINTER(BatchSystem),
                                         ctypes.POINTER(vp)]),
            "p3m_hip_batch_free": (i, [vp]),
            "p3m_hip_batch_set_alpha": (i, [vp, i, d]),
            "p3m_hip_batch_kspace_force_device": (i, [vp, vp, vp, vp, vp, vp, vp]),
            "p3m_hip_batch_calculate_forces_device": (i, [vp] + [vp] * 8),
            "p3m_hip_batch_capture": (i, [vp] + [vp] * 8),
            "p3m_hip_batch_replay": (i, [vp]),
            "p3m_hip_batch_synchronize": (i, [vp]),
            "p3m_hip_batch_autotune": (i, [vp] + [vp] * 7 + [ip, ip]),
            "p3m_hip_batch_time_pairs": (i, [vp] + [vp] * 7 + [i, _dp]),
            "p3m_hip_tune_alpha": (d, [i, d, d, d, d]),
            "p3m_hip_tune_scan": (i, [i, _SP, _PP, d, i, TuneTimerFn, vp, ctypes.POINTER(RuntimeStat)]),
            "p3m_hip_tune": (RuntimeStat, [ctypes.POINTER(Method), _SP, _PP, d]),
            "p3m_hip_error_inhomo": (i, [_SP, _PP, i, i, i, i, ctypes.c_char_p, _dp]),
            "p3m_hip_error_map": (i, [_SP, _FP, _FP, i, i, _dp]),
            "p3m_hip_system_new": (i, [i, d, ctypes.POINTER(_SP)]),
            "p3m_hip_system_free": (i, [_SP]),
            "p3m_hip_read_system": (i, [ctypes.c_char_p, ctypes.POINTER(_SP)]),
            "p3m_hip_write_system": (i, [_SP, ctypes.c_char_p]),
            "p3m_hip_read_exact_forces": (i, [_SP, ctypes.c_char_p]),
            "p3m_hip_write_exact_forces": (i, [_SP, ctypes.c_char_p]),
            "p3m_hip_format_out_row": (i, [ctypes.c_char_p, ctypes.c_size_t, _dp]),
            "p3m_hip_ewald_sample_device": (i, [i, d, vp, vp, vp, vp, i, vp, d, d, i, d, vp, vp]),
            "p3m_hip_stage_ring_start": (i, [_DPT, i]),
            "p3m_hip_stage_ring_read": (i, [_DPT, _dp, i, ip]),
            "p3m_hip_stage_ring_read_pairs": (i, [_DPT, _dp, i, ip]),
            "p3m_hip_get_influence_function": (i, [_DPT, _dp]),
            "p3m_hip_get_self_force_coefficients": (i, [_DPT, _dp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class P3MError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("p3m_hip error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != OK:
        raise P3MError(rc, lib().p3m_hip_last_error().decode())


def method(method_id):
    """The library's method_t object for a method id (same as the reference's)."""
    m = lib().p3m_hip_method(method_id)
    if not m:
        raise ValueError("unknown method id %d" % method_id)
    return m.contents


def exported_method(name):
    """One of the exported method_t objects by symbol name (method_p3m_ik, ...)."""
    return Method.in_dll(lib(), name)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _Vec:
    """A vector_array_t over three numpy arrays (kept alive with it)."""

    def __init__(self, x, y, z):
        # private copies: the C side writes through these pointers
        self.arrays = [np.array(a, dtype=np.float64, copy=True) for a in (x, y, z)]
        self.ptrs = (_dp * 3)(*[a.ctypes.data_as(_dp) for a in self.arrays])
        self.c = VectorArray(len(self.arrays[0]), self.ptrs[0], self.ptrs[1], self.ptrs[2],
                             ctypes.cast(self.ptrs, ctypes.POINTER(_dp)))

    def numpy(self):
        return np.stack(self.arrays)

    def set(self, a):
        for d in range(3):
            self.arrays[d][:] = a[d]


class HostSystem:
    """system_t over numpy positions/charges (reference layout, FP64)."""

    def __init__(self, x, y, z, q, length):
        self.pos = _Vec(x, y, z)
        self.q = _c(q)
        n = len(self.q)
        self.c = System(float(length), n, ctypes.pointer(self.pos.c), self.q.ctypes.data_as(_dp),
                        float(np.sum(self.q ** 2)), None, None, 1.0, 0.0)
        self.ptr = ctypes.pointer(self.c)

    @classmethod
    def from_dict(cls, s):
        return cls(s["x"], s["y"], s["z"], s["q"], s["L"])

    @property
    def n(self):
        return self.c.nparticles


class HostForces:
    """forces_t with zeroed f, f_k, f_r (reference layout)."""

    def __init__(self, n):
        z = np.zeros(n)
        self.f, self.f_k, self.f_r = _Vec(z, z, z), _Vec(z, z, z), _Vec(z, z, z)
        assert len({a.ctypes.data for v in (self.f, self.f_k, self.f_r) for a in v.arrays}) == 9
        self.c = Forces(ctypes.pointer(self.f.c), ctypes.pointer(self.f_k.c), ctypes.pointer(self.f_r.c))
        self.ptr = ctypes.pointer(self.c)

    def total(self):
        return self.f.numpy()

    def kspace(self):
        return self.f_k.numpy()

    def real(self):
        return self.f_r.numpy()


def parameters(mesh, cao, alpha, rcut=0.0, tuning=0):
    p = Parameters(float(alpha), float(rcut), 1.0, int(mesh), int(cao) - 1, int(cao), int(cao) ** 3,
                   0.0, int(tuning))
    return p


class Solver:
    """One initialised method (a data_t) -- thin object wrapper used by tests
    and the benchmark.  Every call maps to one C-ABI entry point."""

    def __init__(self, method_id, system, params, precision=FP64):
        self.method_id = method_id
        self.system = system
        self.params = params
        self.pptr = ctypes.pointer(params)
        self.d = _DPT()
        check(lib().p3m_hip_init(method_id, system.ptr, self.pptr, precision, ctypes.byref(self.d)))

    def close(self):
        if self.d:
            check(lib().p3m_hip_free(self.d))
            self.d = _DPT()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_alpha(self, alpha):
        self.params.alpha = float(alpha)
        check(lib().p3m_hip_influence_function(self.system.ptr, self.pptr, self.d))

    def set_precision(self, precision):
        check(lib().p3m_hip_set_precision(self.d, precision))

    def kspace_force(self, forces=None):
        forces = forces or HostForces(self.system.n)
        check(lib().p3m_hip_kspace_force(self.system.ptr, self.pptr, self.d, forces.ptr))
        return forces

    def calculate_forces(self, forces=None):
        forces = forces or HostForces(self.system.n)
        m = lib().p3m_hip_method(self.method_id)
        check(lib().p3m_hip_calculate_forces(m, self.system.ptr, self.pptr, self.d, forces.ptr))
        return forces

    def influence_function(self):
        M = self.params.mesh
        G = np.zeros(M ** 3)
        check(lib().p3m_hip_get_influence_function(self.d, G.ctypes.data_as(_dp)))
        return G

    def self_force_coefficients(self):
        b = np.zeros(81)
        check(lib().p3m_hip_get_self_force_coefficients(self.d, b.ctypes.data_as(_dp)))
        return b

    def stage_times(self):
        t = np.zeros(5)
        check(lib().p3m_hip_stage_times(self.d, t.ctypes.data_as(_dp)))
        return t

    def kspace_force_vjp(self, dF):
        """Gradient of sum(dF * F_k) w.r.t. positions (3, n) and charges (n)
        (host arrays; F_k includes the self-force term of methods 2 and 7)."""
        n = self.system.n
        gin = _Vec(*np.asarray(dF, dtype=np.float64).reshape(3, n))
        z = np.zeros(n)
        dpos = _Vec(z, z, z)
        dq = np.zeros(n)
        check(lib().p3m_hip_kspace_force_vjp(self.system.ptr, self.pptr, self.d, ctypes.byref(gin.c),
                                             ctypes.byref(dpos.c), dq.ctypes.data_as(_dp)))
        return dpos.numpy(), dq

    def kspace_force_vjp_device(self, n, length, x, y, z, q, gF, dpos, dq):
        check(lib().p3m_hip_kspace_force_vjp_device(self.d, n, length, x, y, z, q, gF, dpos, dq))

    def realteil_vjp_device(self, n, length, x, y, z, q, rcut, alpha, gF, dpos, dq):
        check(lib().p3m_hip_realteil_vjp_device(self.d, n, length, x, y, z, q, rcut, alpha, gF, dpos,
                                                dq))

    def stage_ring_start(self, nsteps):
        """Record the stage events of the next `nsteps` device force calls."""
        check(lib().p3m_hip_stage_ring_start(self.d, int(nsteps)))

    def stage_ring_read(self, max_steps):
        """(steps, 5) stage times in ms of the recorded calls."""
        out = np.zeros((max_steps, 5))
        k = ctypes.c_int()
        check(lib().p3m_hip_stage_ring_read(self.d, out.ctypes.data_as(_dp), int(max_steps),
                                            ctypes.byref(k)))
        return out[:k.value]

    def stage_ring_read_pairs(self, max_steps):
        """(steps,) times in ms of the real-space pair kernel alone in the
        recorded calls."""
        out = np.zeros(max_steps)
        k = ctypes.c_int()
        check(lib().p3m_hip_stage_ring_read_pairs(self.d, out.ctypes.data_as(_dp), int(max_steps),
                                                  ctypes.byref(k)))
        return out[:k.value]

    # device-resident entry points (pointers to device memory, e.g. torch tensors)
    def kspace_force_device(self, n, length, x, y, z, q, fk, f):
        check(lib().p3m_hip_kspace_force_device(self.d, n, length, x, y, z, q, fk, f))

    def realteil_device(self, n, length, x, y, z, q, rcut, alpha, fr, energy):
        check(lib().p3m_hip_realteil_device(self.d, n, length, x, y, z, q, rcut, alpha, fr, energy))

    def calculate_forces_device(self, n, length, x, y, z, q, rcut, alpha, f, fk, fr, energy):
        check(lib().p3m_hip_calculate_forces_device(self.d, n, length, x, y, z, q, rcut, alpha, f,
                                                    fk, fr, energy))

    def synchronize(self):
        check(lib().p3m_hip_synchronize(self.d))


def realteil(system, rcut, alpha, forces=None):
    """Realteil on the GPU (realpart.c:124-167): returns forces with f_r filled."""
    forces = forces or HostForces(system.n)
    p = parameters(2, 1, alpha, rcut)
    check(lib().p3m_hip_realteil(system.ptr, ctypes.pointer(p), forces.ptr))
    return forces


def realteil_vjp(system, rcut, alpha, dF):
    """Adjoint of the real-space force (host arrays): (dpos (3, n), dq (n))
    of sum dF . f_r."""
    n = system.n
    g = _Vec(*np.asarray(dF, dtype=np.float64).reshape(3, n))
    z = np.zeros(n)
    dpos = _Vec(z, z, z)
    dq = np.zeros(n)
    p = parameters(2, 1, alpha, rcut)
    check(lib().p3m_hip_realteil_vjp(system.ptr, ctypes.pointer(p), ctypes.pointer(g.c),
                                     ctypes.pointer(dpos.c), dq.ctypes.data_as(_dp)))
    return dpos.numpy(), dq


def error_estimate(method_id, system, params):
    """(total, k-space) rms error estimates of the method (Error / Error_k)."""
    t, k = ctypes.c_double(), ctypes.c_double()
    check(lib().p3m_hip_error(method_id, system.ptr, ctypes.pointer(params), ctypes.byref(t),
                              ctypes.byref(k)))
    return t.value, k.value


def error_inhomo(system, alpha, mesh=64, cao=5, mc=0, uniform=False, out_file=None):
    """Generic_error_estimate_inhomo (p3m-common.c:437-719) at alpha with the
    estimate's own mesh and order (the driver's defaults: 64, 5, mc 0)."""
    p = parameters(mesh, cao, alpha)
    v = ctypes.c_double()
    check(lib().p3m_hip_error_inhomo(system.ptr, ctypes.pointer(p), int(bool(uniform)), int(mesh),
                                     int(cao), int(mc),
                                     os.fsencode(out_file) if out_file else None, ctypes.byref(v)))
    return v.value


def error_map(system, fk, fk_ref, mesh=64, cao=1):
    """Error_map (p3m-common.c:402-435): (mesh, mesh, mesh) real array of
    |fk_ref - fk|^2 per particle assigned with order cao."""
    n = system.n
    f, fr = HostForces(n), HostForces(n)
    f.f_k.set(np.asarray(fk).reshape(3, n))
    fr.f_k.set(np.asarray(fk_ref).reshape(3, n))
    out = np.zeros(2 * mesh ** 3)
    check(lib().p3m_hip_error_map(system.ptr, f.ptr, fr.ptr, int(mesh), int(cao),
                                  out.ctypes.data_as(_dp)))
    return out[0::2].reshape(mesh, mesh, mesh)


def set_error_table(kind, mesh, cao, table):
    t = _c(table)
    assert t.shape == (101,)
    check(lib().p3m_hip_set_error_table(kind, mesh, cao, t.ctypes.data_as(_dp)))


class Batch:
    """Many independent systems evaluated together (p3m_hip_batch_*).

    `systems` is a list of dicts with keys n, L, M, cao, alpha, rcut.  Device
    arrays (pointers) are concatenated system by system; force arrays hold
    three blocks (x, y, z) of the total particle count."""

    def __init__(self, method_id, systems, precision=FP64):
        arr = (BatchSystem * len(systems))(*[
            BatchSystem(int(s["n"]), float(s["L"]), int(s["M"]), int(s["cao"]), float(s["alpha"]),
                        float(s.get("rcut", 0.0))) for s in systems])
        self.method_id = method_id
        self.systems = [dict(s) for s in systems]
        self.ntot = sum(int(s["n"]) for s in systems)
        self.offsets = np.concatenate([[0], np.cumsum([int(s["n"]) for s in systems])]).astype(int)
        self.b = ctypes.c_void_p()
        check(lib().p3m_hip_batch_create(method_id, precision, len(systems), arr, ctypes.byref(self.b)))

    def close(self):
        if self.b:
            check(lib().p3m_hip_batch_free(self.b))
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_alpha(self, s, alpha):
        check(lib().p3m_hip_batch_set_alpha(self.b, int(s), float(alpha)))
        self.systems[s]["alpha"] = float(alpha)

    def kspace_force_device(self, x, y, z, q, fk, f):
        check(lib().p3m_hip_batch_kspace_force_device(self.b, x, y, z, q, fk, f))

    def calculate_forces_device(self, x, y, z, q, f, fk, fr, energy=None):
        check(lib().p3m_hip_batch_calculate_forces_device(self.b, x, y, z, q, f, fk, fr, energy))

    def capture(self, x, y, z, q, f, fk, fr, energy=None):
        check(lib().p3m_hip_batch_capture(self.b, x, y, z, q, f, fk, fr, energy))

    def replay(self):
        check(lib().p3m_hip_batch_replay(self.b))

    def synchronize(self):
        check(lib().p3m_hip_batch_synchronize(self.b))

    def autotune(self, x, y, z, q, f, fk, fr):
        t, g = ctypes.c_int(), ctypes.c_int()
        check(lib().p3m_hip_batch_autotune(self.b, x, y, z, q, f, fk, fr, ctypes.byref(t),
                                           ctypes.byref(g)))
        return t.value, g.value

    def time_pairs(self, x, y, z, q, f, fk, fr, reps=5):
        """Mean ms of the batched half-shell real-space kernel alone over
        `reps` eager force evaluations."""
        ms = ctypes.c_double()
        check(lib().p3m_hip_batch_time_pairs(self.b, x, y, z, q, f, fk, fr, int(reps),
                                             ctypes.byref(ms)))
        return ms.value


def tune_alpha(n, length, q2, rcut, precision):
    """alpha of the reference's tuner for a real-space error target (tuning.c:293)."""
    return lib().p3m_hip_tune_alpha(int(n), float(length), float(q2), float(rcut), float(precision))


def tune(method_id, system, params, precision, form=FP64, timer=None):
    """Tune (tuning.c:250-370): the fastest (mesh, cao) meeting `precision`
    at params.rcut (params.cao fixes the order when nonzero), alpha from the
    real-space target.  Returns (parameters, statistics) -- statistics as
    runtime_stat_t in seconds, t.avg = -1 and params unchanged if nothing met
    the precision.  `timer` (optional) replaces the GPU timer: a callable
    (method_id, mesh, cao, alpha) -> seconds, used for every stage statistic
    of t and t_c (t_g = t_f = 0)."""
    p = Parameters()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(params), ctypes.sizeof(Parameters))
    out = RuntimeStat()
    cb = TuneTimerFn()
    if timer is not None:
        def _cb(ctx, mid, sp, pp, res):
            try:
                v = float(timer(mid, pp.contents.mesh, pp.contents.cao, pp.contents.alpha))
            except Exception:
                return EINVAL
            r = res.contents
            for name in ("t", "t_c", "t_g", "t_f"):
                tv = v if name in ("t", "t_c") else 0.0
                setattr(r, name, Timing(tv, 0.0, tv, tv, 50))
            return OK
        cb = TuneTimerFn(_cb)
    check(lib().p3m_hip_tune_scan(method_id, system.ptr, ctypes.pointer(p), float(precision), form,
                                  cb, None, ctypes.byref(out)))
    return p, out.as_dict()


def _columns(v, n):
    return np.stack([np.ctypeslib.as_array(v.fields[d], shape=(n,)).copy() for d in range(3)])


def read_system(path):
    """Read_system (io.c:100-157): {"x", "y", "z", "q", "L"} from a
    reference system file."""
    sp = _SP()
    check(lib().p3m_hip_read_system(os.fsencode(path), ctypes.byref(sp)))
    try:
        s = sp.contents
        n = s.nparticles
        pos = _columns(s.p.contents, n)
        return {"x": pos[0], "y": pos[1], "z": pos[2],
                "q": np.ctypeslib.as_array(s.q, shape=(n,)).copy(), "L": s.length}
    finally:
        lib().p3m_hip_system_free(sp)


def _c_system(sysd, f=None, fk=None):
    n = len(sysd["q"])
    sp = _SP()
    check(lib().p3m_hip_system_new(n, float(sysd["L"]), ctypes.byref(sp)))
    s = sp.contents
    for d, k in enumerate("xyz"):
        np.ctypeslib.as_array(s.p.contents.fields[d], shape=(n,))[:] = sysd[k]
        if f is not None:
            np.ctypeslib.as_array(s.reference.contents.f.contents.fields[d], shape=(n,))[:] = f[d]
        if fk is not None:
            np.ctypeslib.as_array(s.reference.contents.f_k.contents.fields[d], shape=(n,))[:] = fk[d]
    np.ctypeslib.as_array(s.q, shape=(n,))[:] = sysd["q"]
    s.q2 = float(np.sum(np.asarray(sysd["q"]) ** 2))
    return sp


def write_system(path, sysd):
    """Write_system (io.c:199-225)."""
    sp = _c_system(sysd)
    try:
        check(lib().p3m_hip_write_system(sp, os.fsencode(path)))
    finally:
        lib().p3m_hip_system_free(sp)


def read_exact_forces(path, n):
    """Read_exact_forces (io.c:70-98): (f, f_k) as (3, n) arrays; f_k is
    zero unless the file has the seven columns Write_exact_forces writes."""
    sp = _SP()
    check(lib().p3m_hip_system_new(int(n), 1.0, ctypes.byref(sp)))
    try:
        check(lib().p3m_hip_read_exact_forces(sp, os.fsencode(path)))
        ref = sp.contents.reference.contents
        return _columns(ref.f.contents, n), _columns(ref.f_k.contents, n)
    finally:
        lib().p3m_hip_system_free(sp)


def write_exact_forces(path, f, fk):
    """Write_exact_forces (io.c:159-177): total and k-space force, (3, n)."""
    f, fk = np.asarray(f, dtype=np.float64), np.asarray(fk, dtype=np.float64)
    n = f.shape[1]
    sysd = {"x": np.zeros(n), "y": np.zeros(n), "z": np.zeros(n), "q": np.zeros(n), "L": 1.0}
    sp = _c_system(sysd, f, fk)
    try:
        check(lib().p3m_hip_write_exact_forces(sp, os.fsencode(path)))
    finally:
        lib().p3m_hip_system_free(sp)


def format_out_row(alpha, delta_f, estimate, rs_estimate, k_estimate, delta_f_k, ewald_k_estimate,
                   inhomo_estimate=0.0):
    """One row of the driver's out.dat (main.c:452-455)."""
    v = np.array([alpha, delta_f, estimate, rs_estimate, k_estimate, delta_f_k, ewald_k_estimate,
                  inhomo_estimate], dtype=np.float64)
    buf = ctypes.create_string_buffer(512)
    check(lib().p3m_hip_format_out_row(buf, 512, v.ctypes.data_as(_dp)))
    return buf.value.decode()


def ewald_error_estimates(n, length, q2, alpha, rcut, kmax):
    """Real- and k-space rms force error estimates of the Ewald sum
    (Kolafa-Perram eq. 16, Petersen eq. 11; ewald.c:174-199)."""
    import math
    r = q2 * math.sqrt(rcut / (2.0 * length ** 3)) * math.exp(-(alpha * rcut) ** 2) / ((alpha * rcut) ** 2)
    k = (2.0 * q2 * alpha / length * math.sqrt(1.0 / (math.pi * kmax * n)) *
         math.exp(-(math.pi * kmax / (alpha * length)) ** 2))
    return r, k


def ewald_parameters(n, length, q2, precision, rcut):
    """(alpha, kmax) of an Ewald sum with real-space cutoff rcut whose
    estimated rms force error is <= precision: alpha from the real-space
    estimate, then the smallest kmax that meets it in k space."""
    lo, hi = 1e-3, 50.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if ewald_error_estimates(n, length, q2, mid, rcut, 1)[0] > 0.5 * precision:
            lo = mid
        else:
            hi = mid
    alpha = hi
    kmax = 1
    while ewald_error_estimates(n, length, q2, alpha, rcut, kmax)[1] > 0.5 * precision:
        kmax += 1
    return alpha, kmax


def ewald_sample_device(n, length, x, y, z, q, nsample, idx, alpha, rcut, kmax, f_total,
                        alpha_split=0.0, f_kspace=None):
    """Sampled FP64 Ewald forces (device pointers); see p3m_hip_ewald_sample_device."""
    check(lib().p3m_hip_ewald_sample_device(int(n), float(length), x, y, z, q, int(nsample), idx,
                                            float(alpha), float(rcut), int(kmax), float(alpha_split),
                                            f_total, f_kspace))

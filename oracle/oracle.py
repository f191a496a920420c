"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.  It binds:

* ``liboracle64.so`` -- the float64 restatement of the reference step
  (oracle/g1_oracle.c), with the reference host C-ABI (g1_sizes, g1_c_fk,
  g1_c_fd, g1_c_step; src/host_lib.cpp:7-53) plus batched / RNEA / CRBA
  extensions;
* ``liboracle32.so`` -- the same source at float32 (CPU-baseline port);
* ``_ref/libg1host.so`` and ``_ref/libg1ref_f32.so`` -- the reference's own
  FP64 host library and its float physics core, compiled from the reference
  sources by oracle/Makefile where the reference tree exists.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NB, NQ, NV, NU, NC = 30, 36, 35, 29, 8

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def _bind_host_abi(lib):
    lib.g1_sizes.restype = C.c_int
    lib.g1_sizes.argtypes = [C.POINTER(C.c_int)] * 5
    lib.g1_c_fk.argtypes = [_dp, _dp, _dp]
    lib.g1_c_fd.argtypes = [_dp, _dp, _dp, _dp]
    lib.g1_c_step.argtypes = [_dp, _dp, _dp, C.c_int, C.c_double, _dp, _dp]
    return lib


def _bind_oracle(lib):
    _bind_host_abi(lib)
    lib.g1o_step.argtypes = [_dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                             _dp, _dp, _dp, _dp, _dp]
    lib.g1o_step_batch_f32.argtypes = [C.c_int, C.c_double, _fp, _fp, _fp, _fp, _fp,
                                       _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
    lib.g1o_rnea.argtypes = [_dp, _dp, _dp, _dp]
    lib.g1o_crba.argtypes = [_dp, _dp]
    lib.g1o_m6_solve.argtypes = [_dp, _dp, _dp]
    lib.g1o_quat_integrate.argtypes = [_dp, _dp, C.c_double]
    lib.g1o_hash.restype = C.c_uint
    lib.g1o_hash.argtypes = [C.c_uint]
    lib.g1o_reset_env_f32.argtypes = [C.c_int, C.c_uint, C.c_float, C.c_float, _fp, _fp, _fp]
    lib.g1o_bench_f32.restype = C.c_double
    lib.g1o_bench_f32.argtypes = [C.c_int, C.c_int, C.c_double, _fp, _fp, _fp, _fp, C.c_int]
    return lib


_cache = {}


def _load(name, binder):
    if name not in _cache:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run: make -C oracle)")
        _cache[name] = binder(C.CDLL(path))
    return _cache[name]


def lib64():
    return _load("liboracle64.so", _bind_oracle)


def lib32():
    return _load("liboracle32.so", _bind_oracle)


def ref_host_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libg1host.so"))


def ref_host():
    """The reference's own FP64 host library (src/host_lib.cpp)."""
    return _load(os.path.join("_ref", "libg1host.so"), _bind_host_abi)


def ref_bench():
    lib = _load(os.path.join("_ref", "libg1ref_f32.so"), lambda l: l)
    lib.g1ref_bench_f32.restype = C.c_double
    lib.g1ref_bench_f32.argtypes = [C.c_int, C.c_int, C.c_double, _fp, _fp, _fp, _fp, C.c_int]
    return lib


# ------------------------------------------------------------ conveniences

def fk(qpos, lib=None):
    lib = lib or lib64()
    xpos = np.zeros(NB * 3)
    xmat = np.zeros(NB * 9)
    lib.g1_c_fk(dptr(np.ascontiguousarray(qpos, np.float64)), dptr(xpos), dptr(xmat))
    return xpos.reshape(NB, 3), xmat.reshape(NB, 9)


def fd(qpos, qvel, tau, lib=None):
    lib = lib or lib64()
    qacc = np.zeros(NV)
    lib.g1_c_fd(*(dptr(np.ascontiguousarray(a, np.float64)) for a in (qpos, qvel, tau)), dptr(qacc))
    return qacc


def step(qpos, qvel, ctrl, anchor, nsub, dt, lib=None):
    """Reference-ABI step: returns new (qpos, qvel, anchor, fn)."""
    lib = lib or lib64()
    qp = np.array(qpos, np.float64)
    qv = np.array(qvel, np.float64)
    an = np.array(anchor, np.float64)
    fn = np.zeros(NC)
    lib.g1_c_step(dptr(qp), dptr(qv), dptr(np.ascontiguousarray(ctrl, np.float64)), nsub,
                  dt, dptr(fn), dptr(an))
    return qp, qv, an, fn


def step_full(qpos, qvel, ctrl, anchor, nsub, dt, mu=1.0):
    """Oracle step with every output; returns a dict."""
    lib = lib64()
    out = {"qpos": np.array(qpos, np.float64), "qvel": np.array(qvel, np.float64),
           "anchor": np.array(anchor, np.float64), "qacc": np.zeros(NV), "fn": np.zeros(NC),
           "foot_pos": np.zeros(6), "foot_vel": np.zeros(6), "tau": np.zeros(NV)}
    lib.g1o_step(dptr(out["qpos"]), dptr(out["qvel"]), dptr(np.ascontiguousarray(ctrl, np.float64)),
                 dptr(out["anchor"]), nsub, dt, mu, dptr(out["qacc"]), dptr(out["fn"]),
                 dptr(out["foot_pos"]), dptr(out["foot_vel"]), dptr(out["tau"]))
    return out


def step_batch_f32(qpos, qvel, anchor, ctrl, dt, mu=None, nthreads=0, want=("qacc",)):
    """Teacher-forced single step of float32 SoA state ([dim][nenv]) in float64.

    Returns a dict of float64 SoA arrays for the names in ``want`` (qacc,
    qpos, qvel, anchor, fn, foot_pos, foot_vel, tau)."""
    n = qpos.shape[1]
    shapes = {"qacc": NV, "qpos": NQ, "qvel": NV, "anchor": 2 * NC, "fn": NC,
              "foot_pos": 6, "foot_vel": 6, "tau": NV}
    out = {k: np.zeros((shapes[k], n)) for k in want}
    args = [np.ascontiguousarray(a, np.float32) for a in (qpos, qvel, anchor, ctrl)]
    m = None if mu is None else np.ascontiguousarray(mu, np.float32)
    lib64().g1o_step_batch_f32(n, dt, *(fptr(a) for a in args), fptr(m),
                               *(dptr(out.get(k)) for k in ("qacc", "qpos", "qvel", "anchor",
                                                            "fn", "foot_pos", "foot_vel", "tau")),
                               nthreads)
    return out


def rnea(qpos, qvel, qacc):
    tau = np.zeros(NV)
    lib64().g1o_rnea(*(dptr(np.ascontiguousarray(a, np.float64)) for a in (qpos, qvel, qacc)),
                     dptr(tau))
    return tau


def crba(qpos):
    M = np.zeros(NV * NV)
    lib64().g1o_crba(dptr(np.ascontiguousarray(qpos, np.float64)), dptr(M))
    return M.reshape(NV, NV)


def m6_solve(A21, b6):
    x = np.zeros(6)
    lib64().g1o_m6_solve(dptr(np.ascontiguousarray(A21, np.float64)),
                         dptr(np.ascontiguousarray(b6, np.float64)), dptr(x))
    return x


def quat_integrate(q, w, dt):
    q = np.array(q, np.float64)
    lib64().g1o_quat_integrate(dptr(q), dptr(np.ascontiguousarray(w, np.float64)), dt)
    return q


def reset_env_f32(env, seed, noise, drop):
    qp = np.zeros(NQ, np.float32)
    qv = np.zeros(NV, np.float32)
    an = np.zeros(2 * NC, np.float32)
    lib64().g1o_reset_env_f32(env, seed, noise, drop, fptr(qp), fptr(qv), fptr(an))
    return qp, qv, an

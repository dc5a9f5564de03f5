"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product library never uses
it.  Every function returns host numpy arrays and raises OracleError with the
status / device code of the reference's failure.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the same source (the sanitizer builds
# of oracle/Makefile); the default is the optimised build beside this file.
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class OracleError(RuntimeError):
    def __init__(self, message, status, code):
        super().__init__(message)
        self.status = status
        self.code = code


def build(force=False):
    if "ORACLE_LIB" in os.environ:
        return
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB_PATH)
        L.orc_last_error_message.restype = C.c_char_p
        L.orc_sim_create.restype = C.c_void_p
        L.orc_sim_create.argtypes = [C.c_int, C.c_size_t] + [C.c_void_p] * 7 + [C.c_double] * 4
        for name in ("orc_sim_destroy", "orc_sim_initialize", "orc_sim_rebuild"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_sim_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_sim_partial_forces.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_sim_state.argtypes = [C.c_void_p] + [C.c_void_p] * 9
        L.orc_sim_sizes.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_sim_structure.argtypes = [C.c_void_p] + [C.c_void_p] * 10
        _lib = L
    return _lib


def _chk(status):
    if status != 0:
        L = lib()
        raise OracleError(L.orc_last_error_message().decode(), status,
                          L.orc_last_device_code())


def _a(x, dtype=np.float64):
    return np.ascontiguousarray(x, dtype=dtype)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_threads(n):
    lib().orc_set_num_threads(int(n))


def threads():
    return lib().orc_num_threads()


def key_dtype(bits):
    return np.uint32 if bits == 10 else np.uint64


def bbox(x, y, z):
    x, y, z = _a(x), _a(y), _a(z)
    out = np.zeros(4)
    _chk(lib().orc_bbox(_p(x), _p(y), _p(z), C.c_size_t(len(x)), _p(out)))
    return out


def morton_keys(x, y, z, box, bits):
    x, y, z, box = _a(x), _a(y), _a(z), _a(box)
    n = len(x)
    keys = np.zeros(n, dtype=key_dtype(bits))
    idx = np.zeros(n, dtype=np.uint32)
    _chk(lib().orc_morton_keys(_p(x), _p(y), _p(z), C.c_size_t(n), _p(box), bits,
                               _p(keys), _p(idx)))
    return keys, idx


def frame_capture(x, y, z, sample_count=None, types=None):
    """Sampled bounds [6] and quantized positions [S, 3] (uint16) of a frame,
    plus the sampled types when `types` is given (device_capture.cu:27-153)."""
    x, y, z = _a(x), _a(y), _a(z)
    n = len(x)
    s = n if sample_count is None else int(sample_count)
    bounds = np.zeros(6)
    q = np.zeros((max(s, 1), 3), dtype=np.uint16)
    t_in = None if types is None else _a(types, np.uint8)
    t_out = None if types is None else np.zeros(max(s, 1), dtype=np.uint8)
    _chk(lib().orc_frame_capture(_p(x), _p(y), _p(z), _p(t_in), C.c_size_t(n),
                                 C.c_size_t(s), _p(bounds), _p(q), _p(t_out)))
    return bounds, q[:s], (None if t_out is None else t_out[:s])


def frame_sample_indices(n, s):
    """Source particle of each sample: j * n // s (device_capture.cu:50-51)."""
    return (np.arange(s, dtype=np.uint64) * np.uint64(n)) // np.uint64(s)


def leaf_groups(sorted_keys, capacity=None):
    k = np.ascontiguousarray(sorted_keys)
    n = len(k)
    cap = n if capacity is None else capacity
    uniq = np.zeros(max(n, 1), dtype=k.dtype)
    first = np.zeros(max(n, 1), dtype=np.int32)
    count = np.zeros(max(n, 1), dtype=np.int32)
    g = C.c_int(0)
    _chk(lib().orc_leaf_groups(_p(k), k.dtype.itemsize, n, cap, _p(uniq), _p(first),
                               _p(count), C.byref(g)))
    g = g.value
    return uniq[:g], first[:g], count[:g]


def tree_build(keys):
    k = np.ascontiguousarray(keys)
    n = len(k)
    m = max(n - 1, 1)
    left, right, rf, rl, pl = (np.zeros(m, dtype=np.int32) for _ in range(5))
    parent = np.zeros(2 * n - 1 if n > 0 else 1, dtype=np.int32)
    _chk(lib().orc_tree_build(_p(k), k.dtype.itemsize, n, _p(left), _p(right), _p(parent),
                              _p(rf), _p(rl), _p(pl)))
    if n == 1:
        e = np.zeros(0, dtype=np.int32)
        return dict(left=e, right=e, parent=parent, range_first=e, range_last=e,
                    prefix_length=e)
    return dict(left=left, right=right, parent=parent, range_first=rf, range_last=rl,
                prefix_length=pl)


def tree_com(tree, n, idx, m, x, y, z, first=None, count=None):
    idx = _a(idx, np.uint32)
    m, x, y, z = _a(m), _a(x), _a(y), _a(z)
    tot = 2 * n - 1
    mass, cx, cy, cz = (np.zeros(tot) for _ in range(4))
    visit = np.zeros(max(n - 1, 1), dtype=np.int32)
    left = _a(tree["left"], np.int32) if n > 1 else np.zeros(1, np.int32)
    right = _a(tree["right"], np.int32) if n > 1 else np.zeros(1, np.int32)
    f = None if first is None else _a(first, np.int32)
    c = None if count is None else _a(count, np.int32)
    _chk(lib().orc_tree_com(n, _p(left), _p(right), _p(_a(tree["parent"], np.int32)), _p(f),
                            _p(c), _p(idx), _p(m), _p(x), _p(y), _p(z), _p(mass), _p(cx),
                            _p(cy), _p(cz), _p(visit)))
    return dict(mass=mass, com_x=cx, com_y=cy, com_z=cz, visit_count=visit[: max(n - 1, 0)])


def plan(n_leaves, parent, prefix_length, unique, bits):
    R = 2 * n_leaves - 1
    lv, cnt, off = (np.zeros(R, dtype=np.int32) for _ in range(3))
    m = C.c_int(0)
    pl = _a(prefix_length, np.int32) if len(prefix_length) else np.zeros(1, np.int32)
    u = np.ascontiguousarray(unique, dtype=key_dtype(bits))
    _chk(lib().orc_plan(n_leaves, _p(_a(parent, np.int32)), _p(pl), _p(u), bits, _p(lv),
                        _p(cnt), _p(off), C.byref(m)))
    return lv, cnt, off, m.value


def octree(n_nodes, n_leaves, tparent, range_first, unique, gfirst, gcount, plevel,
           pcount, poffset, bits):
    ch = np.zeros(8 * n_nodes, dtype=np.int32)
    par, lev, first, count = (np.zeros(n_nodes, dtype=np.int32) for _ in range(4))
    prefix = np.zeros(n_nodes, dtype=key_dtype(bits))
    rf = _a(range_first, np.int32) if len(range_first) else np.zeros(1, np.int32)
    u = np.ascontiguousarray(unique, dtype=key_dtype(bits))
    _chk(lib().orc_octree(n_nodes, n_leaves, _p(_a(tparent, np.int32)), _p(rf), _p(u),
                          _p(_a(gfirst, np.int32)), _p(_a(gcount, np.int32)),
                          _p(_a(plevel, np.int32)), _p(_a(pcount, np.int32)),
                          _p(_a(poffset, np.int32)), bits, _p(ch), _p(par), _p(lev),
                          _p(prefix), _p(first), _p(count)))
    return dict(children=ch, parent=par, level=lev, prefix=prefix, first=first, count=count)


def geometry(level, prefix, box, bits):
    n = len(level)
    cx, cy, cz, h = (np.zeros(n) for _ in range(4))
    pre = np.ascontiguousarray(prefix, dtype=key_dtype(bits))
    _chk(lib().orc_geometry(n, _p(_a(level, np.int32)), _p(pre), bits, _p(_a(box)), _p(cx),
                            _p(cy), _p(cz), _p(h)))
    return cx, cy, cz, h


def mass_com(children, parent, first, count, n_particles, idx, m, x, y, z):
    n = len(parent)
    mass, cx, cy, cz = (np.zeros(n) for _ in range(4))
    pend, cc = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    _chk(lib().orc_mass_com(n, n_particles, _p(_a(children, np.int32)),
                            _p(_a(parent, np.int32)), _p(_a(first, np.int32)),
                            _p(_a(count, np.int32)), _p(_a(idx, np.uint32)), _p(_a(m)),
                            _p(_a(x)), _p(_a(y)), _p(_a(z)), _p(mass), _p(cx), _p(cy),
                            _p(cz), _p(pend), _p(cc)))
    return dict(mass=mass, com_x=cx, com_y=cy, com_z=cz, pending=pend, child_count=cc)


def bh_accel(tree, idx, m, x, y, z, theta, softening, G, morton=False, n_particles=None,
             t_begin=0, t_end=None, count=False):
    """tree: dict with children, parent, first, count, mass, com_x/y/z,
    center_x/y/z, half (node arrays)."""
    n = len(idx) if n_particles is None else n_particles
    t_end = n if t_end is None else t_end
    ax, ay, az = np.zeros(n), np.zeros(n), np.zeros(n)
    M = len(tree["parent"])
    npp, npn = C.c_uint64(0), C.c_uint64(0)
    args = [_a(tree[k], np.int32) for k in ("children", "parent", "first", "count")]
    args += [_a(tree[k]) for k in ("mass", "com_x", "com_y", "com_z", "center_x", "center_y",
                                    "center_z", "half")]
    _chk(lib().orc_bh_accel(n, M, 0, *[_p(a) for a in args], _p(_a(idx, np.uint32)),
                            _p(_a(m)), _p(_a(x)), _p(_a(y)), _p(_a(z)), _p(ax), _p(ay),
                            _p(az), C.c_double(theta), C.c_double(softening), C.c_double(G),
                            1 if morton else 0, t_begin, t_end,
                            C.byref(npp) if count else None,
                            C.byref(npn) if count else None))
    if count:
        return ax, ay, az, npp.value, npn.value
    return ax, ay, az


def kick_drift(x, y, z, vx, vy, vz, ax, ay, az, dt):
    arrs = [_a(v).copy() for v in (x, y, z, vx, vy, vz)]
    acc = [_a(v) for v in (ax, ay, az)]
    _chk(lib().orc_kick_drift(*[_p(a) for a in arrs], *[_p(a) for a in acc],
                              C.c_size_t(len(arrs[0])), C.c_double(dt)))
    return arrs


def kick(vx, vy, vz, ax, ay, az, dt):
    arrs = [_a(v).copy() for v in (vx, vy, vz)]
    acc = [_a(v) for v in (ax, ay, az)]
    _chk(lib().orc_kick(*[_p(a) for a in arrs], *[_p(a) for a in acc],
                        C.c_size_t(len(arrs[0])), C.c_double(dt)))
    return arrs


def direct(m, x, y, z, softening, G, targets=None):
    m, x, y, z = _a(m), _a(x), _a(y), _a(z)
    n = len(m)
    t = None if targets is None else _a(targets, np.uint32)
    nt = n if t is None else len(t)
    ax, ay, az = np.zeros(nt), np.zeros(nt), np.zeros(nt)
    _chk(lib().orc_direct(_p(m), _p(x), _p(y), _p(z), C.c_size_t(n), _p(t),
                          C.c_size_t(nt), C.c_double(softening), C.c_double(G), _p(ax),
                          _p(ay), _p(az)))
    return ax, ay, az


def accuracy_inputs(clustered, n=256):
    m, x, y, z = (np.zeros(n) for _ in range(4))
    _chk(lib().orc_accuracy_inputs(1 if clustered else 0, C.c_size_t(n), _p(m), _p(x),
                                   _p(y), _p(z)))
    return m, x, y, z


def random_keys(n, mask=0xFFFFFFFF, start=5, stride=11):
    k = np.zeros(n, dtype=np.uint32)
    _chk(lib().orc_random_keys(C.c_size_t(n), C.c_uint32(mask), start, stride, _p(k)))
    return k


class Sim:
    """The reference's Simulation loop on the oracle."""

    def __init__(self, p, dt, theta, softening, G, bits):
        self.n = len(p["m"])
        self.bits = bits
        self._keep = [_a(p[k]) for k in ("m", "x", "y", "z", "vx", "vy", "vz")]
        self.h = lib().orc_sim_create(bits, self.n, *[_p(a) for a in self._keep],
                                      C.c_double(dt), C.c_double(theta),
                                      C.c_double(softening), C.c_double(G))

    def __del__(self):
        try:
            lib().orc_sim_destroy(self.h)
        except Exception:
            pass

    def initialize(self):
        _chk(lib().orc_sim_initialize(self.h))

    def step(self, k=1):
        _chk(lib().orc_sim_step(self.h, k))

    def rebuild(self):
        _chk(lib().orc_sim_rebuild(self.h))

    def partial_forces(self, t_begin, t_end):
        _chk(lib().orc_sim_partial_forces(self.h, t_begin, t_end))

    def state(self):
        arrs = [np.zeros(self.n) for _ in range(9)]
        _chk(lib().orc_sim_state(self.h, *[_p(a) for a in arrs]))
        return dict(zip(("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"), arrs))

    def structure(self):
        g, m = C.c_int(0), C.c_int(0)
        _chk(lib().orc_sim_sizes(self.h, C.byref(g), C.byref(m)))
        g, m = g.value, m.value
        keys = np.zeros(self.n, dtype=key_dtype(self.bits))
        idx = np.zeros(self.n, dtype=np.uint32)
        ch = np.zeros(8 * m, dtype=np.int32)
        par, first, count = (np.zeros(m, dtype=np.int32) for _ in range(3))
        mass, cx, cy, cz = (np.zeros(m) for _ in range(4))
        _chk(lib().orc_sim_structure(self.h, _p(keys), _p(idx), _p(ch), _p(par), _p(first),
                                     _p(count), _p(mass), _p(cx), _p(cy), _p(cz)))
        return dict(keys=keys, indices=idx, children=ch, parent=par, first=first,
                    count=count, mass=mass, com_x=cx, com_y=cy, com_z=cz, n_groups=g,
                    n_nodes=m)

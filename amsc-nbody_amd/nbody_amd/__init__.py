"""Python binding of the MI355X Barnes-Hut pipeline (ctypes over the C ABI).

The native library ``libnbody_amd.so`` (hand-written HIP kernels for gfx950)
is the only compute path: importing this module without it raises, and every
call runs on the GPU.  torch tensors are used only as device buffers.

The classes and functions mirror the reference's C++ API (src/*.hpp,
include/globalbounding.hpp of sytrams/AMSC-Nbody): ``Bbox``, ``MortonKeys``,
``MortonLeafGroups`` with allocate/build, ``Tree`` with ``build_tree``,
``RadixToOctreePlan``, ``Octree`` with ``build_sparse_octree_topology``,
``compute_octree_geometry``, ``compute_octree_mass_and_center_of_mass``,
``compute_barnes_hut_acceleration``, ``leapfrog_kick_drift``/``leapfrog_kick``
and ``Simulation``.  Failures raise the Python analogue of the reference's
exception class (see ``STATUS_EXCEPTIONS``).
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libnbody_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "nbody_amd.h")


class NbodyError(RuntimeError):
    """Base class of every error raised by the library."""

    def __init__(self, message, status=0, device_code=0):
        super().__init__(message)
        self.status = status
        self.device_code = device_code


class InvalidArgument(NbodyError, ValueError):
    """std::invalid_argument in the reference."""


class LogicError(NbodyError):
    """std::logic_error in the reference."""


class DomainError(NbodyError, ValueError):
    """std::domain_error in the reference."""


class OverflowError_(NbodyError, OverflowError):
    """std::overflow_error in the reference."""


class DeviceError(NbodyError):
    """std::runtime_error in the reference (device codes, HIP failures)."""


STATUS_EXCEPTIONS = {
    -1: InvalidArgument,
    -2: LogicError,
    -3: DeviceError,
    -4: DomainError,
    -5: OverflowError_,
    -6: DeviceError,
    -7: DeviceError,
}

PRECISION_FP64 = 0
PRECISION_MIXED = 1
ORDER_ORIGINAL = 0
ORDER_MORTON = 1
DEFAULT_G = 6.67430e-11  # src/particle.hpp:16


def load_library(path: str = LIB_PATH) -> C.CDLL:
    # torch carries its own HIP runtime with the same soname as the system
    # one.  Loading torch first makes this library bind to that runtime, so
    # the process holds exactly one HIP/HSA runtime; the other order loads
    # two and the second one sees no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(path):
        raise ImportError(
            f"native library {path} is missing: build it with "
            "`make -C amsc-nbody_amd` (or __graft_entry__.build())")
    return C.CDLL(path)


lib = load_library()

_p = C.c_void_p
_i = C.c_int
_d = C.c_double
_sz = C.c_size_t
_u64 = C.c_uint64


class TreeView(C.Structure):
    _fields_ = [("n_leaves", _i), ("n_internal_nodes", _i), ("left", _p),
                ("right", _p), ("parent", _p), ("range_first", _p),
                ("range_last", _p), ("prefix_length", _p), ("mass", _p),
                ("com_x", _p), ("com_y", _p), ("com_z", _p), ("visit_count", _p)]


class OctreeView(C.Structure):
    _fields_ = [("n_particles", _i), ("n_leaves", _i), ("n_nodes", _i), ("root", _i),
                ("children", _p), ("parent", _p), ("level", _p), ("prefix", _p),
                ("first_particle", _p), ("particle_count", _p), ("mass", _p),
                ("com_x", _p), ("com_y", _p), ("com_z", _p), ("center_x", _p),
                ("center_y", _p), ("center_z", _p), ("half_size", _p),
                ("pending_children", _p), ("child_count", _p)]


class SimConfig(C.Structure):
    _fields_ = [("time_step", _d), ("theta", _d), ("softening", _d),
                ("gravitational_constant", _d), ("precision", _i),
                ("bits_per_axis", _i)]


class ParticlesView(C.Structure):
    _fields_ = [("count", _sz), ("mass", _p), ("x", _p), ("y", _p), ("z", _p),
                ("vx", _p), ("vy", _p), ("vz", _p)]


class SimStructure(C.Structure):
    _fields_ = [("n_particles", _i), ("n_groups", _i), ("n_octree_nodes", _i),
                ("bits_per_axis", _i), ("sorted_keys", _p), ("sorted_indices", _p),
                ("unique_keys", _p), ("group_first", _p), ("group_count", _p),
                ("tree", TreeView), ("octree", OctreeView), ("radix_level", _p),
                ("edge_node_count", _p), ("edge_node_offset", _p), ("box", _p),
                ("target_order", _p)]


def _sig(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_sig("nbh_last_error_message", C.c_char_p)
_sig("nbh_last_device_code", _i)
_sig("nbh_version", C.c_char_p)
_sig("nbh_context_create", _i, C.POINTER(_p), _p)
_sig("nbh_context_destroy", _i, _p)
_sig("nbh_bbox_compute", _i, _p, _p, _p, _p, _sz, _p, C.POINTER(_d))
_sig("nbh_morton_keys", _i, _p, _p, _p, _p, _sz, _p, _i, _p, _p)
_sig("nbh_radix_sort_pairs", _i, _p, _p, _p, _sz, _i, _i)
_sig("nbh_leaf_groups_build", _i, _p, _p, _i, _i, _i, _p, _p, _p, C.POINTER(_i))
_sig("nbh_leaf_groups_build_noalloc", _i, _p, _p, _i, _i, _i, _p, _p, _p, _p, C.POINTER(_i))
_sig("nbh_tree_build", _i, _p, C.POINTER(TreeView), _p, _i, _i)
_sig("nbh_tree_center_of_mass", _i, _p, C.POINTER(TreeView), _p, _p, _p, _p, _p, _p, _p)
_sig("nbh_radix_to_octree_plan", _i, _p, C.POINTER(TreeView), _p, _i, _p, _p, _p,
     C.POINTER(_i))
_sig("nbh_octree_build_topology", _i, _p, C.POINTER(OctreeView), C.POINTER(TreeView),
     _p, _p, _p, _i, _i, _i, _p, _p, _p)
_sig("nbh_octree_geometry", _i, _p, C.POINTER(OctreeView), _i, C.POINTER(_d))
_sig("nbh_octree_mass_com", _i, _p, C.POINTER(OctreeView), _p, _p, _p, _p, _p)
_sig("nbh_bh_acceleration", _i, _p, C.POINTER(OctreeView), _p, _p, _p, _p, _p, _p, _p,
     _p, _sz, _d, _d, _d, _i)
_sig("nbh_leapfrog_kick_drift", _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _sz, _d)
_sig("nbh_leapfrog_kick", _i, _p, _p, _p, _p, _p, _p, _p, _sz, _d)
_sig("nbh_gather_morton", _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _sz)
_sig("nbh_sim_create", _i, C.POINTER(_p), C.POINTER(SimConfig), _sz, _p, _p, _p, _p,
     _p, _p, _p, _p)
_sig("nbh_sim_destroy", _i, _p)
_sig("nbh_sim_initialize", _i, _p)
_sig("nbh_sim_step", _i, _p)
_sig("nbh_sim_run", _i, _p, _sz)
_sig("nbh_sim_time", _d, _p)
_sig("nbh_sim_step_number", _sz, _p)
_sig("nbh_sim_particles", _i, _p, C.POINTER(ParticlesView))
_sig("nbh_sim_accelerations", _i, _p, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p))
_sig("nbh_sim_structure_view", _i, _p, C.POINTER(SimStructure))
_sig("nbh_phase_count", _i)
_sig("nbh_phase_name", C.c_char_p, _i)
_sig("nbh_sim_phase_times", _i, _p, C.POINTER(_d), _i)
_sig("nbh_sim_set_timing", _i, _p, _i)
_sig("nbh_sim_count_interactions", _i, _p, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64))
_sig("nbh_generate", _i, _i, _sz, _u64, _p, _p, _p, _p, _p, _p, _p)
_sig("nbh_copy_to_host", _i, _p, _p, _sz)
_sig("nbh_direct_sum", _i, _p, _p, _p, _p, _p, _sz, _p, _sz, _d, _d, _p, _p, _p)
_sig("nbh_sim_energy", _i, _p, C.POINTER(_d), C.POINTER(_d))
_sig("nbh_particle_file_info", _i, C.c_char_p, C.POINTER(_sz), C.POINTER(_i))
_sig("nbh_particle_file_read", _i, C.c_char_p, _sz, _p, _p, _p, _p, _p, _p, _p, _p,
     C.POINTER(_sz))
_sig("nbh_sim_potential", _i, _p, _p)
_sig("nbh_sim_force_vjp", _i, _p, _p, _p, _p, _p, _p, _p, _p)
_sig("nbh_adjoint_create", _i, C.POINTER(_p), C.POINTER(SimConfig), _sz, _i, _i, _p)
_sig("nbh_adjoint_destroy", _i, _p)
_sig("nbh_adjoint_forward", _i, _p, *([_p] * 13))
_sig("nbh_adjoint_backward", _i, _p, *([_p] * 13))
_sig("nbh_adjoint_generation", _i, _p, _p)


class FrameHeader(C.Structure):
    _fields_ = [("version", C.c_uint32), ("header_bytes", C.c_uint32),
                ("sequence", _u64), ("simulation_step", _u64), ("simulation_time", _d),
                ("source_particle_count", _u64), ("particle_count", _u64),
                ("scalar_bytes", C.c_uint32), ("components", C.c_uint32),
                ("payload_bytes", _u64), ("total_steps", _u64),
                ("bounds_min", _d * 3), ("bounds_max", _d * 3),
                ("particle_type_bytes", _u64)]


_sig("nbh_frame_capture", _i, _p, _p, _p, _p, _sz, _sz, _p, C.POINTER(_d))
_sig("nbh_frame_sample_types", _i, _p, _p, _sz, _sz, _p)
_sig("nbh_frame_write_quantized", _i, C.c_char_p, C.c_char_p, _u64, _u64, _d, _sz, _sz,
     C.POINTER(_d), _p, _p, C.c_char_p, _sz)
_sig("nbh_frame_write_float", _i, C.c_char_p, C.c_char_p, _u64, _u64, _d, _u64, _sz, _sz,
     _p, _p, C.c_char_p, _sz)
_sig("nbh_frame_mark_complete", _i, C.c_char_p, C.c_char_p, C.c_char_p, _sz)
_sig("nbh_frame_read_header", _i, C.c_char_p, C.POINTER(FrameHeader))
_sig("nbh_frame_read", _i, C.c_char_p, _sz, _p, _p, C.POINTER(_sz))


def check(status: int) -> None:
    if status == 0:
        return
    message = lib.nbh_last_error_message().decode()
    code = lib.nbh_last_device_code()
    raise STATUS_EXCEPTIONS.get(status, NbodyError)(message, status, code)


def version() -> str:
    return lib.nbh_version().decode()


def phase_names() -> list[str]:
    return [lib.nbh_phase_name(k).decode() for k in range(lib.nbh_phase_count())]


# ------------------------------------------------------------ device memory
def _torch():
    import torch
    return torch


def device():
    torch = _torch()
    if not torch.cuda.is_available():
        raise NbodyError("no HIP device is visible")
    return torch.device("cuda", 0)


def to_device(a, dtype=None):
    """Copy a numpy array (or a list) to a new device tensor."""
    torch = _torch()
    arr = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(arr).to(device())


def empty(n, dtype):
    torch = _torch()
    return torch.empty(int(n), dtype=dtype, device=device())


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


class Context:
    """A stream + workspace; every stage call runs on it."""

    def __init__(self, stream=None):
        self.handle = _p()
        check(lib.nbh_context_create(C.byref(self.handle), stream))

    def close(self):
        if self.handle:
            lib.nbh_context_destroy(self.handle)
            self.handle = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        device()
        _default_ctx = Context()
    return _default_ctx


def key_dtype(bits_per_axis: int):
    torch = _torch()
    if bits_per_axis == 10:
        return torch.int32, np.uint32, 4
    if bits_per_axis == 21:
        return torch.int64, np.uint64, 8
    raise InvalidArgument("bits_per_axis must be 10 or 21", -1)


def host_keys(t, bits_per_axis):
    _, npdt, _ = key_dtype(bits_per_axis)
    return t.cpu().numpy().view(npdt)


# --------------------------------------------------------------- stages
@dataclass
class BboxValues:
    center_x: float
    center_y: float
    center_z: float
    side: float


class Bbox:
    """Bbox (include/globalbounding.hpp:15-36): [cx, cy, cz, side] on device."""

    value_count = 4

    def __init__(self, x=None, y=None, z=None, count=None, ctx=None):
        self.ctx = ctx or default_context()
        self.data = to_device(np.array([0.0, 0.0, 0.0, -1.0]))
        if x is not None:
            self.recompute(x, y, z, count)

    def recompute(self, x, y, z, count=None):
        n = x.numel() if count is None else count
        out = (_d * 4)()
        check(lib.nbh_bbox_compute(self.ctx.handle, ptr(x), ptr(y), ptr(z), n,
                                   ptr(self.data), out))

    def values(self) -> BboxValues:
        v = self.data.cpu().numpy()
        return BboxValues(float(v[0]), float(v[1]), float(v[2]), float(v[3]))

    def device_data(self):
        return self.data


class MortonKeys:
    """MortonKeys (src/morton.hpp:8-49): sorted keys and the permutation."""

    def __init__(self, x, y, z, count=None, bbox: Bbox | None = None,
                 bits_per_axis=10, ctx=None):
        self.ctx = ctx or default_context()
        n = x.numel() if count is None else count
        if bbox is None:
            bbox = Bbox(x, y, z, n, ctx=self.ctx)
        kt, _, _ = key_dtype(bits_per_axis)
        self.bits_per_axis = bits_per_axis
        torch = _torch()
        self.keys = torch.empty(max(n, 1), dtype=kt, device=device())
        self.indices = torch.empty(max(n, 1), dtype=torch.int32, device=device())
        check(lib.nbh_morton_keys(self.ctx.handle, ptr(x), ptr(y), ptr(z), n,
                                  ptr(bbox.data), bits_per_axis, ptr(self.keys),
                                  ptr(self.indices)))
        self.count = n

    def size(self):
        return self.count

    def host_keys(self):
        return host_keys(self.keys[: self.count], self.bits_per_axis)

    def host_indices(self):
        return self.indices[: self.count].cpu().numpy().view(np.uint32)


def radix_sort_pairs(keys, values, key_bits, ctx=None):
    ctx = ctx or default_context()
    key_bytes = keys.element_size()
    check(lib.nbh_radix_sort_pairs(ctx.handle, ptr(keys), ptr(values), keys.numel(),
                                   key_bytes, key_bits))


class MortonLeafGroups:
    """MortonLeafGroups (src/morton_leaf_groups.hpp:8-47)."""

    def __init__(self):
        self.n_particles = 0
        self.n_groups = 0
        self.capacity = 0
        self.unique_keys = None
        self.first_particle = None
        self.particle_count = None
        self.key_bytes = 4


def allocate_morton_leaf_groups(groups: MortonLeafGroups, capacity: int, key_bytes=4):
    if capacity <= 0:
        raise InvalidArgument("MortonLeafGroups capacity must be positive", -1)
    if groups.unique_keys is not None:
        raise LogicError("MortonLeafGroups memory is already allocated", -2)
    torch = _torch()
    kt = torch.int32 if key_bytes == 4 else torch.int64
    groups.unique_keys = empty(capacity, kt)
    groups.first_particle = empty(capacity, torch.int32)
    groups.particle_count = empty(capacity, torch.int32)
    groups.capacity = capacity
    groups.key_bytes = key_bytes
    groups.n_particles = 0
    groups.n_groups = 0


def build_morton_leaf_groups(groups: MortonLeafGroups, sorted_keys, n_particles, ctx=None):
    ctx = ctx or default_context()
    groups.n_particles = 0
    groups.n_groups = 0
    ng = _i(0)
    check(lib.nbh_leaf_groups_build(ctx.handle, ptr(sorted_keys), groups.key_bytes,
                                    n_particles, groups.capacity, ptr(groups.unique_keys),
                                    ptr(groups.first_particle), ptr(groups.particle_count),
                                    C.byref(ng)))
    groups.n_particles = n_particles
    groups.n_groups = ng.value


LEAF_GROUPS_NOALLOC_MAX = 1 << 28


def build_morton_leaf_groups_noalloc(groups: MortonLeafGroups, sorted_keys, n_particles,
                                     d_n_groups=None, stream=None, read_count=True):
    """buildMortonLeafGroups without allocation (no reference counterpart):
    the scan's scratch lives in a static device array, for at most
    LEAF_GROUPS_NOALLOC_MAX particles.  d_n_groups: optional device int32
    tensor receiving the group count; read_count=False returns without a
    host synchronisation (n_groups is then left at 0)."""
    groups.n_particles = 0
    groups.n_groups = 0
    ng = _i(0)
    check(lib.nbh_leaf_groups_build_noalloc(stream, ptr(sorted_keys), groups.key_bytes,
                                            n_particles, groups.capacity,
                                            ptr(groups.unique_keys), ptr(groups.first_particle),
                                            ptr(groups.particle_count),
                                            ptr(d_n_groups) if d_n_groups is not None else None,
                                            C.byref(ng) if read_count else None))
    groups.n_particles = n_particles
    groups.n_groups = ng.value


def free_morton_leaf_groups(groups: MortonLeafGroups):
    groups.__init__()


class Tree:
    """Tree (src/tree_types.hpp:51-108)."""

    def __init__(self):
        self.n_leaves = 0
        self.n_internal_nodes = 0
        for f in ("left", "right", "parent", "range_first", "range_last",
                  "prefix_length", "mass", "com_x", "com_y", "com_z", "visit_count"):
            setattr(self, f, None)

    def view(self) -> TreeView:
        v = TreeView(self.n_leaves, self.n_internal_nodes)
        for f in ("left", "right", "parent", "range_first", "range_last",
                  "prefix_length", "mass", "com_x", "com_y", "com_z", "visit_count"):
            setattr(v, f, ptr(getattr(self, f)))
        return v


def allocate_tree_topology(tree: Tree, n: int):
    if n <= 0:
        raise InvalidArgument("N must be positive", -1)
    if tree.parent is not None:
        raise LogicError("Tree memory is already allocated", -2)
    torch = _torch()
    if n > 1:
        for f in ("left", "right", "range_first", "range_last", "prefix_length"):
            setattr(tree, f, empty(n - 1, torch.int32))
    tree.parent = empty(2 * n - 1, torch.int32)
    tree.n_leaves = n
    tree.n_internal_nodes = n - 1


def allocate_tree(tree: Tree, n: int):
    allocate_tree_topology(tree, n)
    torch = _torch()
    if n > 1:
        tree.visit_count = empty(n - 1, torch.int32)
    for f in ("mass", "com_x", "com_y", "com_z"):
        setattr(tree, f, empty(2 * n - 1, torch.float64))


def build_tree(tree: Tree, sorted_keys, n: int, ctx=None):
    ctx = ctx or default_context()
    v = tree.view()
    check(lib.nbh_tree_build(ctx.handle, C.byref(v), ptr(sorted_keys),
                             sorted_keys.element_size(), n))


def build_tree_from_morton_groups(tree: Tree, groups: MortonLeafGroups, ctx=None):
    if groups.n_particles <= 0:
        raise InvalidArgument("MortonLeafGroups has no particles", -1)
    if groups.n_groups <= 0:
        raise InvalidArgument("MortonLeafGroups has no groups", -1)
    if tree.n_leaves != groups.n_groups:
        raise InvalidArgument("Tree leaf count does not match Morton group count", -1)
    build_tree(tree, groups.unique_keys, groups.n_groups, ctx)


def compute_center_of_mass(tree: Tree, sorted_indices, m, x, y, z, n, ctx=None):
    ctx = ctx or default_context()
    if tree.n_leaves != n:
        raise InvalidArgument("computeCentersOfMass N does not match tree size", -1)
    v = tree.view()
    check(lib.nbh_tree_center_of_mass(ctx.handle, C.byref(v), None, None,
                                      ptr(sorted_indices), ptr(m), ptr(x), ptr(y), ptr(z)))


def compute_grouped_center_of_mass(tree: Tree, groups: MortonLeafGroups, sorted_indices,
                                   m, x, y, z, ctx=None):
    ctx = ctx or default_context()
    if tree.n_leaves != groups.n_groups:
        raise InvalidArgument("Tree leaf count does not match Morton group count", -1)
    v = tree.view()
    check(lib.nbh_tree_center_of_mass(ctx.handle, C.byref(v), ptr(groups.first_particle),
                                      ptr(groups.particle_count), ptr(sorted_indices),
                                      ptr(m), ptr(x), ptr(y), ptr(z)))


class RadixToOctreePlan:
    """RadixToOctreePlan (src/radix_to_octree.hpp:34-67)."""

    def __init__(self):
        self.n_radix_nodes = 0
        self.n_octree_nodes = 0
        self.radix_level = None
        self.edge_node_count = None
        self.edge_node_offset = None


def allocate_radix_to_octree_plan(plan: RadixToOctreePlan, n_radix_nodes: int):
    if n_radix_nodes <= 0:
        raise InvalidArgument("nRadixNodes must be positive", -1)
    if plan.radix_level is not None:
        raise LogicError("RadixToOctreePlan memory is already allocated", -2)
    torch = _torch()
    plan.radix_level = empty(n_radix_nodes, torch.int32)
    plan.edge_node_count = empty(n_radix_nodes, torch.int32)
    plan.edge_node_offset = empty(n_radix_nodes, torch.int32)
    plan.n_radix_nodes = n_radix_nodes
    plan.n_octree_nodes = 0


def build_radix_to_octree_plan(plan: RadixToOctreePlan, tree: Tree, groups: MortonLeafGroups,
                               bits_per_axis=10, ctx=None):
    ctx = ctx or default_context()
    if tree.n_leaves <= 0:
        raise InvalidArgument("The radix tree has no leaves", -1)
    if groups.n_groups != tree.n_leaves:
        raise InvalidArgument("Morton group count does not match radix-tree leaf count", -1)
    if plan.n_radix_nodes != 2 * tree.n_leaves - 1:
        raise InvalidArgument("RadixToOctreePlan size does not match the radix tree", -1)
    plan.n_octree_nodes = 0
    v = tree.view()
    m = _i(0)
    check(lib.nbh_radix_to_octree_plan(ctx.handle, C.byref(v), ptr(groups.unique_keys),
                                       bits_per_axis, ptr(plan.radix_level),
                                       ptr(plan.edge_node_count), ptr(plan.edge_node_offset),
                                       C.byref(m)))
    plan.n_octree_nodes = m.value


OCTREE_TOPOLOGY = ("children", "parent", "level", "prefix", "first_particle", "particle_count")
OCTREE_GEOMETRY = ("center_x", "center_y", "center_z", "half_size")
OCTREE_PHYSICS = ("mass", "com_x", "com_y", "com_z", "pending_children", "child_count")


class Octree:
    """Octree (src/octree_types.hpp:69-178)."""

    def __init__(self, bits_per_axis=10):
        self.n_particles = 0
        self.n_leaves = 0
        self.n_nodes = 0
        self.root = -1
        self.bits_per_axis = bits_per_axis
        for f in OCTREE_TOPOLOGY + OCTREE_GEOMETRY + OCTREE_PHYSICS:
            setattr(self, f, None)

    def view(self) -> OctreeView:
        v = OctreeView(self.n_particles, self.n_leaves, self.n_nodes, self.root)
        for f in OCTREE_TOPOLOGY + OCTREE_GEOMETRY + OCTREE_PHYSICS:
            setattr(v, f, ptr(getattr(self, f)))
        return v


def allocate_octree_topology(octree: Octree, n_nodes: int):
    if n_nodes <= 0:
        raise InvalidArgument("Octree node count must be positive", -1)
    if octree.children is not None:
        raise LogicError("Octree memory is already allocated", -2)
    torch = _torch()
    kt, _, _ = key_dtype(octree.bits_per_axis)
    octree.children = empty(8 * n_nodes, torch.int32)
    for f in ("parent", "level", "first_particle", "particle_count"):
        setattr(octree, f, empty(n_nodes, torch.int32))
    octree.prefix = empty(n_nodes, kt)
    octree.n_nodes = n_nodes


def build_sparse_octree_topology(octree: Octree, tree: Tree, groups: MortonLeafGroups,
                                 plan: RadixToOctreePlan, ctx=None):
    ctx = ctx or default_context()
    if plan.n_radix_nodes != 2 * tree.n_leaves - 1:
        raise InvalidArgument("Radix-to-octree plan does not match the radix tree", -1)
    if plan.n_octree_nodes <= 0:
        raise InvalidArgument("Radix-to-octree plan has no octree nodes", -1)
    if octree.n_nodes != plan.n_octree_nodes:
        raise InvalidArgument("Allocated octree size does not match the construction plan", -1)
    v = octree.view()
    t = tree.view()
    check(lib.nbh_octree_build_topology(
        ctx.handle, C.byref(v), C.byref(t), ptr(groups.unique_keys),
        ptr(groups.first_particle), ptr(groups.particle_count), groups.n_groups,
        groups.n_particles, octree.bits_per_axis, ptr(plan.radix_level),
        ptr(plan.edge_node_count), ptr(plan.edge_node_offset)))
    octree.n_particles = groups.n_particles
    octree.n_leaves = groups.n_groups
    octree.root = 0


def allocate_octree_geometry(octree: Octree):
    if octree.n_nodes <= 0:
        raise InvalidArgument("Octree topology must be allocated before geometry", -1)
    if octree.center_x is not None:
        raise LogicError("Octree geometry is already allocated", -2)
    torch = _torch()
    for f in OCTREE_GEOMETRY:
        setattr(octree, f, empty(octree.n_nodes, torch.float64))


def compute_octree_geometry(octree: Octree, bbox: Bbox, ctx=None):
    ctx = ctx or default_context()
    b = bbox.values()
    box = (_d * 4)(b.center_x, b.center_y, b.center_z, b.side)
    v = octree.view()
    check(lib.nbh_octree_geometry(ctx.handle, C.byref(v), octree.bits_per_axis, box))


def allocate_octree_physical_data(octree: Octree):
    if octree.n_nodes <= 0:
        raise InvalidArgument("Octree topology must be allocated first", -1)
    if octree.mass is not None:
        raise LogicError("Octree physical data is already allocated", -2)
    torch = _torch()
    for f in ("mass", "com_x", "com_y", "com_z"):
        setattr(octree, f, empty(octree.n_nodes, torch.float64))
    octree.pending_children = empty(octree.n_nodes, torch.int32)
    octree.child_count = empty(octree.n_nodes, torch.int32)


def compute_octree_mass_and_center_of_mass(octree: Octree, sorted_indices, m, x, y, z,
                                           ctx=None):
    ctx = ctx or default_context()
    v = octree.view()
    check(lib.nbh_octree_mass_com(ctx.handle, C.byref(v), ptr(sorted_indices), ptr(m),
                                  ptr(x), ptr(y), ptr(z)))


@dataclass
class BarnesHutParameters:
    """BarnesHutParameters (src/barnes_hut.hpp:8-13)."""

    theta: float = 0.5
    softening: float = 1.0e-6
    gravitational_constant: float = DEFAULT_G


def compute_barnes_hut_acceleration(octree: Octree, sorted_indices, m, x, y, z, ax, ay, az,
                                    particle_count, params: BarnesHutParameters,
                                    particle_order=ORDER_ORIGINAL, ctx=None):
    ctx = ctx or default_context()
    v = octree.view()
    check(lib.nbh_bh_acceleration(ctx.handle, C.byref(v), ptr(sorted_indices), ptr(m),
                                  ptr(x), ptr(y), ptr(z), ptr(ax), ptr(ay), ptr(az),
                                  particle_count, params.theta, params.softening,
                                  params.gravitational_constant, particle_order))


def leapfrog_kick_drift(x, y, z, vx, vy, vz, ax, ay, az, count, dt, ctx=None):
    ctx = ctx or default_context()
    check(lib.nbh_leapfrog_kick_drift(ctx.handle, ptr(x), ptr(y), ptr(z), ptr(vx), ptr(vy),
                                      ptr(vz), ptr(ax), ptr(ay), ptr(az), count, dt))


def leapfrog_kick(vx, vy, vz, ax, ay, az, count, dt, ctx=None):
    ctx = ctx or default_context()
    check(lib.nbh_leapfrog_kick(ctx.handle, ptr(vx), ptr(vy), ptr(vz), ptr(ax), ptr(ay),
                                ptr(az), count, dt))


def gather_morton_ordered(sorted_indices, m, x, y, z, count, ctx=None):
    ctx = ctx or default_context()
    torch = _torch()
    out = [empty(count, torch.float64) for _ in range(4)]
    check(lib.nbh_gather_morton(ctx.handle, ptr(sorted_indices), ptr(m), ptr(x), ptr(y),
                                ptr(z), *[ptr(o) for o in out], count))
    return out


def direct_sum(m, x, y, z, targets, softening, G=1.0, ctx=None):
    """FP64 softened direct sum on the device for the listed target indices."""
    ctx = ctx or default_context()
    torch = _torch()
    nt = targets.numel()
    out = [empty(nt, torch.float64) for _ in range(3)]
    check(lib.nbh_direct_sum(ctx.handle, ptr(m), ptr(x), ptr(y), ptr(z), m.numel(),
                             ptr(targets), nt, softening, G, *[ptr(o) for o in out]))
    return out


# ----------------------------------------------------------- particle files
def _particle_arrays(particles: dict, keys=("m", "x", "y", "z", "vx", "vy", "vz")):
    """The named arrays of a particle set as contiguous host arrays (float64,
    uint8 for "type"), all one-dimensional and of one length; InvalidArgument
    otherwise, before any pointer reaches the library."""
    out = []
    for k in keys:
        if k not in particles:
            raise InvalidArgument(f"particle set has no '{k}' array")
        a = np.asarray(particles[k])
        if a.ndim != 1:
            raise InvalidArgument(f"particle array '{k}' must be one-dimensional "
                                  f"(shape {a.shape})")
        out.append(np.ascontiguousarray(a, dtype=np.uint8 if k == "type" else np.float64))
    n = len(out[0])
    for k, a in zip(keys, out):
        if len(a) != n:
            raise InvalidArgument(f"particle array '{k}' has {len(a)} elements, "
                                  f"'{keys[0]}' has {n}")
    return out


def load_particles(path) -> dict:
    """Reads a planar binary particle file (the reference's input format) into
    host arrays m, x, y, z, vx, vy, vz (float64) and type (uint8)."""
    enc = os.fsencode(str(path))
    n, has_types = _sz(0), _i(0)
    check(lib.nbh_particle_file_info(enc, C.byref(n), C.byref(has_types)))
    count = n.value
    out = {k: np.empty(count, dtype=np.float64) for k in ("m", "x", "y", "z", "vx", "vy", "vz")}
    out["type"] = np.empty(count, dtype=np.uint8)
    got = _sz(0)
    check(lib.nbh_particle_file_read(enc, count,
                                     *[out[k].ctypes.data for k in ("m", "x", "y", "z", "vx",
                                                                    "vy", "vz", "type")],
                                     C.byref(got)))
    return out


def save_particles(path, particles: dict, wide_header=True, with_types=False):
    """Writes host arrays in the planar binary particle format."""
    keys = ("m", "x", "y", "z", "vx", "vy", "vz") + (("type",) if with_types else ())
    n = len(_particle_arrays(particles, keys)[0])
    with open(path, "wb") as f:
        f.write(np.array([n], dtype=np.uint64 if wide_header else np.uint32).tobytes())
        for k in ("m", "x", "y", "z", "vx", "vy", "vz"):
            f.write(np.ascontiguousarray(particles[k], dtype=np.float64).tobytes())
        if with_types:
            f.write(np.ascontiguousarray(particles["type"], dtype=np.uint8).tobytes())


# ------------------------------------------------------------------ frames
# The .nbsnap frames of graphical streaming (src/streaming/frame.hpp,
# device_capture.hpp): sampled positions quantized to uint16 over the sample's
# bounds, written through a durable spool directory.
FRAME_VERSION = 3
DEFAULT_STREAM_PARTICLES = 100000  # apps/simulation/main.cpp:36


def _addr(t):
    """Device address of a torch tensor, or an int address as given."""
    if t is None:
        return None
    return t if isinstance(t, int) else t.data_ptr()


def capture_frame(x, y, z, sample_count=None, count=None, ctx=None):
    """Samples `sample_count` of `count` particles (sample j reads particle
    j * count // sample_count) and quantizes their positions on the device.
    Returns (uint16 device tensor of shape [S, 3], numpy bounds
    [min x, min y, min z, max x, max y, max z])."""
    ctx = ctx or default_context()
    torch = _torch()
    n = int(count if count is not None else x.numel())
    s = n if sample_count is None else int(sample_count)
    out = torch.empty((max(s, 1), 3), dtype=torch.uint16, device=device())
    bounds = (_d * 6)()
    check(lib.nbh_frame_capture(ctx.handle, _addr(x), _addr(y), _addr(z), n, s,
                                out.data_ptr(), bounds))
    return out[:s], np.array(bounds[:], dtype=np.float64)


def sample_frame_types(types, sample_count=None, count=None, ctx=None):
    """Samples particle types like capture_frame; raises DeviceError for a
    value outside ParticleType.  Returns a uint8 device tensor."""
    ctx = ctx or default_context()
    torch = _torch()
    n = int(count if count is not None else types.numel())
    s = n if sample_count is None else int(sample_count)
    out = torch.empty(max(s, 1), dtype=torch.uint8, device=device())
    check(lib.nbh_frame_sample_types(ctx.handle, _addr(types), n, s, out.data_ptr()))
    return out[:s]


def _path_buffer():
    return C.create_string_buffer(4096)


class FrameSpool:
    """FrameSpool (src/streaming/frame.hpp:77-114): publishes
    run-<session>-frame-<20-digit sequence>.nbsnap files atomically into
    `directory` and refuses to overwrite one.  session_id None ->
    "<microseconds since the epoch>-<pid>"."""

    def __init__(self, directory, session_id=None):
        import re
        import time
        if not str(directory):
            raise InvalidArgument("Frame spool directory cannot be empty", -1, 0)
        if session_id is None:
            session_id = f"{time.time_ns() // 1000}-{os.getpid()}"
        if not session_id or not re.fullmatch(r"[A-Za-z0-9_-]+", session_id):
            raise InvalidArgument(
                "Frame session id may contain only letters, numbers, '-' and '_'", -1, 0)
        self.directory = os.fspath(directory)
        self.session_id = session_id
        os.makedirs(self.directory, exist_ok=True)

    def _args(self):
        return os.fsencode(self.directory), self.session_id.encode()

    def write_quantized_positions(self, sequence, step, time, positions, bounds,
                                  source_count=0, types=None) -> str:
        """Version 3 with `types`, version 2 (120-byte header) without.
        positions: host uint16 array of shape [S, 3]."""
        q = np.ascontiguousarray(positions, dtype=np.uint16).reshape(-1, 3)
        t = None if types is None else np.ascontiguousarray(types, dtype=np.uint8)
        if t is not None and t.size != q.shape[0]:
            raise InvalidArgument("one particle type per sampled particle is required", -1, 0)
        b = (_d * 6)(*[float(v) for v in np.asarray(bounds, dtype=np.float64).ravel()[:6]])
        path = _path_buffer()
        check(lib.nbh_frame_write_quantized(*self._args(), int(sequence), int(step), float(time),
                                            q.shape[0], int(source_count), b, q.ctypes.data,
                                            None if t is None else t.ctypes.data, path,
                                            len(path)))
        return os.fsdecode(path.value)

    def write_typed_positions(self, sequence, step, time, total_steps, positions, types=None,
                              source_count=0) -> str:
        """Float32 positions: version 2 (80-byte header) with `types`,
        version 1 without."""
        p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        t = None if types is None else np.ascontiguousarray(types, dtype=np.uint8)
        if t is not None and t.size != p.shape[0]:
            raise InvalidArgument("one particle type per particle is required", -1, 0)
        path = _path_buffer()
        check(lib.nbh_frame_write_float(*self._args(), int(sequence), int(step), float(time),
                                        int(total_steps), p.shape[0], int(source_count),
                                        p.ctypes.data, None if t is None else t.ctypes.data,
                                        path, len(path)))
        return os.fsdecode(path.value)

    def write_positions(self, sequence, step, time, positions, source_count=0) -> str:
        return self.write_typed_positions(sequence, step, time, 0, positions, None, source_count)

    def mark_complete(self) -> str:
        path = _path_buffer()
        check(lib.nbh_frame_mark_complete(*self._args(), path, len(path)))
        return os.fsdecode(path.value)


class DeviceFrameWriter:
    """DeviceFrameWriter (src/streaming/device_capture.hpp:10-36): samples at
    most `maximum_particles` (0 = all) of the device arrays x, y, z (torch
    tensors or device addresses) with their particle types, validated once
    here, and writes version-3 quantized frames."""

    def __init__(self, x, y, z, types, maximum_particles=DEFAULT_STREAM_PARTICLES, count=None,
                 ctx=None):
        n = int(count if count is not None else x.numel())
        if n == 0 or x is None or y is None or z is None or types is None:
            raise InvalidArgument("Cannot capture an empty or incomplete device particle view",
                                  -1, 0)
        self.ctx = ctx or default_context()
        self._xyz = (x, y, z)
        self.source_particle_count = n
        self.sample_particle_count = n if maximum_particles == 0 else min(n, maximum_particles)
        self.sampled_types = sample_frame_types(types, self.sample_particle_count, n,
                                                self.ctx).cpu().numpy()

    @classmethod
    def from_simulation(cls, sim, types, maximum_particles=DEFAULT_STREAM_PARTICLES):
        v = sim.particles_view()
        return cls(v.x, v.y, v.z, types, maximum_particles, count=v.count)

    def capture(self):
        """(host uint16 [S, 3] positions, bounds) of the current state."""
        q, bounds = capture_frame(*self._xyz, self.sample_particle_count,
                                  self.source_particle_count, self.ctx)
        return q.cpu().numpy(), bounds

    def write(self, spool: FrameSpool, sequence, step, time) -> str:
        q, bounds = self.capture()
        return spool.write_quantized_positions(sequence, step, time, q, bounds,
                                               self.source_particle_count, self.sampled_types)


def read_frame_header(path) -> dict:
    """readFrameHeader (src/streaming/frame.cpp:507-612) as a dict."""
    h = FrameHeader()
    check(lib.nbh_frame_read_header(os.fsencode(str(path)), C.byref(h)))
    out = {name: getattr(h, name) for name, _ in FrameHeader._fields_}
    out["bounds_min"] = list(h.bounds_min)
    out["bounds_max"] = list(h.bounds_max)
    return out


def read_frame(path):
    """(float32 positions [n, 3], uint8 types [n]) of a frame of any
    version; quantized positions decoded with the frame's bounds."""
    h = read_frame_header(path)
    n = int(h["particle_count"])
    pos = np.empty((n, 3), dtype=np.float32)
    types = np.empty(n, dtype=np.uint8)
    got = _sz(0)
    check(lib.nbh_frame_read(os.fsencode(str(path)), n, pos.ctypes.data, types.ctypes.data,
                             C.byref(got)))
    return pos, types


# --------------------------------------------------------------- workloads
WORKLOADS = {"plummer": 0, "uniform": 1, "disk": 2, "clumps": 3, "blob": 4}


def generate(kind: str, count: int, seed: int = 42):
    """Seeded synthetic particle set: dict of host float64 arrays."""
    arrs = {k: np.empty(count, dtype=np.float64) for k in ("m", "x", "y", "z", "vx", "vy", "vz")}
    check(lib.nbh_generate(WORKLOADS[kind], count, seed,
                           *[arrs[k].ctypes.data for k in ("m", "x", "y", "z", "vx", "vy", "vz")]))
    return arrs


# -------------------------------------------------------------- simulation
class Simulation:
    """Simulation (src/simulation.hpp:24-66) with SimulationConfig
    {timeStep, theta, softening} plus G, precision and key width."""

    def __init__(self, particles: dict, time_step: float, theta: float, softening: float,
                 gravitational_constant: float = DEFAULT_G, precision=PRECISION_FP64,
                 bits_per_axis=21, stream=None):
        self.handle = _p()
        arr = _particle_arrays(particles)
        n = len(arr[0])
        device()
        cfg = SimConfig(time_step, theta, softening, gravitational_constant, precision,
                        bits_per_axis)
        self.config = cfg
        check(lib.nbh_sim_create(C.byref(self.handle), C.byref(cfg), n,
                                 *[a.ctypes.data for a in arr], stream))
        self.count = n

    def close(self):
        if getattr(self, "handle", None):
            lib.nbh_sim_destroy(self.handle)
            self.handle = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self):
        check(lib.nbh_sim_initialize(self.handle))

    def step(self):
        check(lib.nbh_sim_step(self.handle))

    def run(self, steps: int):
        check(lib.nbh_sim_run(self.handle, steps))

    def time(self) -> float:
        return lib.nbh_sim_time(self.handle)

    def step_number(self) -> int:
        return lib.nbh_sim_step_number(self.handle)

    def set_timing(self, enabled: bool):
        check(lib.nbh_sim_set_timing(self.handle, 1 if enabled else 0))

    def phase_times(self) -> dict:
        n = lib.nbh_phase_count()
        buf = (_d * n)()
        lib.nbh_sim_phase_times(self.handle, buf, n)
        return {lib.nbh_phase_name(k).decode(): buf[k] for k in range(n)}

    def particles_view(self) -> ParticlesView:
        v = ParticlesView()
        check(lib.nbh_sim_particles(self.handle, C.byref(v)))
        return v

    def state(self) -> dict:
        """Host copy of mass, positions and velocities."""
        v = self.particles_view()
        return {k: _copy_d2h(getattr(v, f), self.count, np.float64)
                for k, f in (("m", "mass"), ("x", "x"), ("y", "y"), ("z", "z"),
                             ("vx", "vx"), ("vy", "vy"), ("vz", "vz"))}

    def accelerations(self):
        a = [_p(), _p(), _p()]
        check(lib.nbh_sim_accelerations(self.handle, *[C.byref(p) for p in a]))
        return [_copy_d2h(p.value, self.count, np.float64) for p in a]

    def structure(self) -> SimStructure:
        s = SimStructure()
        check(lib.nbh_sim_structure_view(self.handle, C.byref(s)))
        return s

    def structure_host(self) -> dict:
        """Host copies of the last rebuild's keys, permutation, groups, radix tree
        and octree, for parity checks."""
        s = self.structure()
        n, g, m = s.n_particles, s.n_groups, s.n_octree_nodes
        kdt = np.uint32 if s.bits_per_axis == 10 else np.uint64
        o = s.octree
        t = s.tree
        out = {
            "keys": _copy_d2h(s.sorted_keys, n, kdt),
            "indices": _copy_d2h(s.sorted_indices, n, np.uint32),
            "unique": _copy_d2h(s.unique_keys, g, kdt),
            "group_first": _copy_d2h(s.group_first, g, np.int32),
            "group_count": _copy_d2h(s.group_count, g, np.int32),
            "parent_radix": _copy_d2h(t.parent, 2 * g - 1, np.int32),
            "n_groups": g,
            "n_nodes": m,
            "box": _copy_d2h(s.box, 4, np.float64),
            "target_order": (_copy_d2h(s.target_order, n, np.uint32)
                             if s.target_order else None),
            "children": _copy_d2h(o.children, 8 * m, np.int32),
            "parent": _copy_d2h(o.parent, m, np.int32),
            "level": _copy_d2h(o.level, m, np.int32),
            "prefix": _copy_d2h(o.prefix, m, kdt),
            "first": _copy_d2h(o.first_particle, m, np.int32),
            "count": _copy_d2h(o.particle_count, m, np.int32),
            "mass": _copy_d2h(o.mass, m, np.float64),
            "com_x": _copy_d2h(o.com_x, m, np.float64),
            "com_y": _copy_d2h(o.com_y, m, np.float64),
            "com_z": _copy_d2h(o.com_z, m, np.float64),
            "center_x": _copy_d2h(o.center_x, m, np.float64),
            "half": _copy_d2h(o.half_size, m, np.float64),
        }
        if g > 1:
            for k, f in (("left", "left"), ("right", "right"), ("range_first", "range_first"),
                         ("range_last", "range_last"), ("prefix_length", "prefix_length")):
                out[k] = _copy_d2h(getattr(t, f), g - 1, np.int32)
        return out

    def count_interactions(self, with_visits=False):
        pp, pn, wv = _u64(0), _u64(0), _u64(0)
        check(lib.nbh_sim_count_interactions(self.handle, C.byref(pp), C.byref(pn),
                                             C.byref(wv)))
        if with_visits:
            return pp.value, pn.value, wv.value
        return pp.value, pn.value

    def energy(self):
        """(kinetic, potential) of the current state; the potential walks the
        current tree in FP64 (FP64 simulations only)."""
        k, u = _d(0.0), _d(0.0)
        check(lib.nbh_sim_energy(self.handle, C.byref(k), C.byref(u)))
        return k.value, u.value

    def potential(self):
        """Per-particle potential (device tensor, original order)."""
        phi = empty(self.count, _torch().float64)
        check(lib.nbh_sim_potential(self.handle, ptr(phi)))
        return phi

    def force_vjp(self, abx, aby, abz, with_mass=False):
        """(da/dx)^T abar at the current tree, and (da/dm)^T abar as a fourth
        output with with_mass; device float64 tensors in, device tensors out.
        A mixed simulation differentiates its mixed forces (FP32 interactions
        along the mixed walk, FP64 node push-down)."""
        torch = _torch()
        ab = [t.to(device=device(), dtype=torch.float64).contiguous() for t in (abx, aby, abz)]
        g = [empty(self.count, torch.float64) for _ in range(4 if with_mass else 3)]
        check(lib.nbh_sim_force_vjp(self.handle, *[ptr(t) for t in ab], *[ptr(t) for t in g[:3]],
                                    ptr(g[3]) if with_mass else None))
        return g


# ----------------------------------------------------------------- adjoint
class Adjoint:
    """Reverse mode of `steps` KDK steps of the Barnes-Hut pipeline with
    checkpoints every `checkpoint_every` steps (device tensors in and out).
    precision PRECISION_FP64 differentiates the FP64 pipeline,
    PRECISION_MIXED the mixed one (mixed forward steps, mixed force VJPs)."""

    def __init__(self, count: int, time_step: float, theta: float, softening: float,
                 steps: int, checkpoint_every: int, gravitational_constant: float = DEFAULT_G,
                 bits_per_axis=21, stream=None, precision=PRECISION_FP64):
        device()
        self.config = SimConfig(time_step, theta, softening, gravitational_constant,
                                precision, bits_per_axis)
        self.count = count
        self.handle = _p()
        check(lib.nbh_adjoint_create(C.byref(self.handle), C.byref(self.config), count,
                                     steps, checkpoint_every, stream))

    def close(self):
        if getattr(self, "handle", None):
            lib.nbh_adjoint_destroy(self.handle)
            self.handle = _p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, m, x, y, z, vx, vy, vz):
        torch = _torch()
        ins = [t.to(device=device(), dtype=torch.float64).contiguous()
               for t in (m, x, y, z, vx, vy, vz)]
        outs = [empty(self.count, torch.float64) for _ in range(6)]
        check(lib.nbh_adjoint_forward(self.handle, *[ptr(t) for t in ins],
                                      *[ptr(t) for t in outs]))
        return outs

    def generation(self) -> int:
        """Number of forward calls so far; a backward replays the most recent."""
        g = _u64(0)
        check(lib.nbh_adjoint_generation(self.handle, C.byref(g)))
        return g.value

    def backward(self, xb, yb, zb, vxb, vyb, vzb, with_mass=False):
        """Cotangents of (x0, y0, z0, vx0, vy0, vz0), and of the masses as a
        seventh output with with_mass."""
        torch = _torch()
        ins = [t.to(device=device(), dtype=torch.float64).contiguous()
               for t in (xb, yb, zb, vxb, vyb, vzb)]
        outs = [empty(self.count, torch.float64) for _ in range(7 if with_mass else 6)]
        check(lib.nbh_adjoint_backward(self.handle, *[ptr(t) for t in ins],
                                       *[ptr(t) for t in outs[:6]],
                                       ptr(outs[6]) if with_mass else None))
        return outs


def leapfrog_function():
    """torch.autograd.Function over Adjoint: (m, pos[3,N], vel[3,N]) ->
    (pos_K, vel_K); gradients reach m, pos and vel.  For checking the
    adjoint against finite differences."""
    torch = _torch()

    class LeapfrogBH(torch.autograd.Function):
        @staticmethod
        def forward(ctx, adjoint, m, pos, vel):
            out = adjoint.forward(m, pos[0], pos[1], pos[2], vel[0], vel[1], vel[2])
            ctx.adjoint = adjoint
            ctx.generation = adjoint.generation()
            return torch.stack(out[:3]), torch.stack(out[3:])

        @staticmethod
        def backward(ctx, gpos, gvel):
            torch_ = _torch()
            # The adjoint replays the checkpoints of its most recent forward:
            # a later forward on the same object has overwritten this one's.
            if ctx.adjoint.generation() != ctx.generation:
                raise RuntimeError(
                    "LeapfrogBH.backward: the Adjoint ran another forward after this "
                    "output was computed; run forward again before backward, or use "
                    "one Adjoint per trajectory")
            if gpos is None:
                gpos = torch_.zeros(3, ctx.adjoint.count, dtype=torch_.float64, device=device())
            if gvel is None:
                gvel = torch_.zeros(3, ctx.adjoint.count, dtype=torch_.float64, device=device())
            with_mass = ctx.needs_input_grad[1]
            g = ctx.adjoint.backward(gpos[0], gpos[1], gpos[2], gvel[0], gvel[1], gvel[2],
                                     with_mass=with_mass)
            return (None, g[6] if with_mass else None, torch_.stack(g[:3]),
                    torch_.stack(g[3:6]))

    return LeapfrogBH


def _copy_d2h(dptr, count, dtype):
    """Synchronous device-to-host copy of `count` elements at a raw pointer."""
    out = np.empty(int(count), dtype=dtype)
    if count == 0:
        return out
    check(lib.nbh_copy_to_host(out.ctypes.data, dptr, out.nbytes))
    return out


def declared_symbols(header: str = HEADER_PATH) -> list[str]:
    """Names of the functions declared in include/nbody_amd.h."""
    import re
    text = open(header).read()
    decl = re.compile(r"^(?:const\s+)?(?:int|double|size_t|char\s*\*|char)\s*\**\s*"
                      r"(nbh_[a-z0-9_]+)\s*\(", re.M)
    return sorted(set(decl.findall(text)))

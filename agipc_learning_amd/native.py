"""ctypes front end of libmasx.so (include/masx.h).

The product path has no CPU fallback: if the HIP library cannot be loaded, or
a call fails, an exception is raised.  Device pointers are taken from torch
tensors (`tensor.data_ptr()`); torch only provides device memory and streams.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from . import build as _build

LIB_PATH = os.path.join(_build.LIB, "libmasx.so")
HEADER = os.path.join(_build.ROOT, "include", "masx.h")

FP32, FP16, BF16 = 0, 1, 2
RULE_REF_RTZ, RULE_TRUE_REL = 0, 1
STATUS_RUNNING, STATUS_DONE = 0, 1
FAIL_OK, FAIL_BAD_RZ, FAIL_BAD_CURVATURE, FAIL_NONFINITE, FAIL_ITERATION_CAP = 0, 1, 2, 3, 4
PRECISIONS = {"fp32": FP32, "fp16": FP16, "bf16": BF16}


class SystemDesc(C.Structure):
    _fields_ = [("block_rows", C.c_int), ("nb", C.c_longlong), ("n_abd", C.c_int), ("n_fem", C.c_int),
                ("use_mas", C.c_int), ("global_diag", C.c_int), ("bank", C.c_int), ("levels", C.c_int)]


class SolveParams(C.Structure):
    _fields_ = [("rule", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int), ("check_every", C.c_int),
                ("chunk", C.c_int), ("use_graph", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("failure", C.c_int), ("restarts", C.c_int),
                ("true_checks", C.c_int), ("rz0", C.c_double), ("rz", C.c_double),
                ("rel_res_recursive", C.c_double), ("rel_res_true", C.c_double), ("rejected", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PlanInfo(C.Structure):
    _fields_ = [("n_rows", C.c_int), ("n_ext", C.c_int), ("n_sys", C.c_int), ("bank", C.c_int),
                ("nnz", C.c_longlong), ("nb_total", C.c_longlong), ("n_banks", C.c_int),
                ("n_fine_banks", C.c_int), ("n_coarse", C.c_int), ("n_bodies", C.c_int),
                ("n_gal", C.c_longlong), ("inverse_bytes", C.c_longlong)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HierArrays(C.Structure):
    _fields_ = [("going_next", C.c_void_p), ("coarse_table", C.c_void_p), ("fine_mask", C.c_void_p),
                ("prefix", C.c_void_p), ("comp_ptr", C.c_void_p), ("comp_mask", C.c_void_p),
                ("comp_parent", C.c_void_p)]


_lib = None


def header_functions():
    """Names of the functions include/masx.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(masx_[a-z0-9_]+)\s*\(", text)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        _build.ensure()
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"masx HIP library missing: {LIB_PATH}")
    L = C.CDLL(LIB_PATH)
    vp, ip, up, dp, fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.POINTER(C.c_double), \
        C.POINTER(C.c_float)
    sig = {
        "masx_last_error": ([], C.c_char_p),
        "masx_device_info": ([C.c_char_p, C.c_int, ip, C.POINTER(C.c_longlong)], C.c_int),
        "masx_default_solve_params": ([C.POINTER(SolveParams)], None),
        "masx_create": ([C.POINTER(vp)], C.c_int),
        "masx_destroy": ([vp], C.c_int),
        "masx_add_system": ([vp, C.POINTER(SystemDesc), ip, ip, ip], C.c_int),
        "masx_set_structure": ([vp, C.c_int, C.c_longlong, ip, ip], C.c_int),
        "masx_set_topology": ([vp, C.c_int, up, up, up, C.c_int, ip, C.c_int, ip], C.c_int),
        "masx_set_topology_morton": ([vp, C.c_int, up, up, up, C.c_int, dp], C.c_int),
        "masx_set_contacts": ([vp, C.c_int, ip, C.c_int, C.c_int], C.c_int),
        "masx_morton_partition": ([C.c_int, dp, C.c_int, ip, C.c_int, ip], C.c_int),
        "masx_morton_create": ([C.POINTER(vp)], C.c_int),
        "masx_morton_destroy": ([vp], C.c_int),
        "masx_morton_partition_device": ([vp, C.c_int, vp, C.c_int, vp, C.c_longlong, vp, vp], C.c_int),
        "masx_morton_partition_device_noalloc": ([C.c_int, vp, C.c_int, vp, C.c_longlong, vp, vp], C.c_int),
        "masx_stream_rate": ([C.c_longlong, C.c_int, C.c_int, C.POINTER(C.c_double)], C.c_int),
        "masx_set_mas_waves": ([vp, C.c_int], C.c_int),
        "masx_set_spmv_loads": ([vp, C.c_int], C.c_int),
        "masx_set_factor_waves": ([vp, C.c_int], C.c_int),
        "masx_spmv_loads_used": ([vp, ip], C.c_int),
        "masx_set_device_structure": ([vp, C.c_int], C.c_int),
        "masx_device_structure_used": ([vp, ip], C.c_int),
        "masx_plan_array": ([vp, C.c_int, ip, C.c_longlong, C.POINTER(C.c_longlong)], C.c_int),
        "masx_autotune_mas": ([vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)], C.c_int),
        "masx_spmv_layout_check": ([C.c_int, C.c_longlong, ip, ip, C.c_int, C.c_uint, dp, C.POINTER(C.c_longlong),
                                    C.POINTER(C.c_longlong)], C.c_int),
        "masx_spmv_layout_check_bodies": ([C.c_int, C.c_int, C.c_longlong, ip, ip, C.c_int, C.c_uint, dp,
                                           C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)], C.c_int),
        "masx_analyze": ([vp], C.c_int),
        "masx_setup": ([vp], C.c_int),
        "masx_set_values": ([vp, C.c_int, vp, vp], C.c_int),
        "masx_set_values_host": ([vp, C.c_int, dp, vp], C.c_int),
        "masx_set_values_many": ([vp, C.c_int, ip, C.POINTER(vp), vp], C.c_int),
        "masx_set_structure_device": ([vp, C.c_int, C.c_longlong, vp, vp, vp], C.c_int),
        "masx_build": ([vp, C.c_int, vp], C.c_int),
        "masx_solve": ([vp, vp, vp, C.POINTER(SolveParams), C.POINTER(SolveStats), vp], C.c_int),
        "masx_spmv": ([vp, C.c_double, vp, C.c_double, vp, vp], C.c_int),
        "masx_apply_preconditioner": ([vp, vp, vp, vp], C.c_int),
        "masx_set_spmv_tiling": ([vp, C.c_int], C.c_int),
        "masx_converter_create": ([C.POINTER(vp)], C.c_int),
        "masx_converter_destroy": ([vp], C.c_int),
        "masx_convert_triplets": ([vp, C.c_int, C.c_int, vp, vp, vp, C.c_int, vp, vp, vp, ip, vp], C.c_int),
        "masx_autotune": ([vp, C.c_int, ip, dp], C.c_int),
        "masx_time_kernel": ([vp, C.c_int, C.c_int, dp], C.c_int),
        "masx_plan_info_get": ([vp, C.POINTER(PlanInfo)], C.c_int),
        "masx_get_hierarchy": ([vp, C.c_int, ip, ip, ip, ip, up, ip, ip, ip], C.c_int),
        "masx_get_inverse": ([vp, fp, C.c_longlong], C.c_int),
        "masx_bank_of_slot": ([vp, C.c_int, C.c_int, ip], C.c_int),
        "masx_factor_status": ([vp, C.POINTER(C.c_longlong)], C.c_int),
        "masx_hierarchy_host": ([C.c_int, C.c_int, C.c_int, C.c_int, up, up, up, ip, ip, ip, C.c_int, C.c_int,
                                 ip, ip, ip, ip, up, ip, ip, C.c_longlong, ip], C.c_int),
        "masx_fem_create": ([C.POINTER(vp), C.c_int, ip, C.c_longlong, ip, ip], C.c_int),
        "masx_fem_destroy": ([vp], C.c_int),
        "masx_fem_assemble": ([vp, vp, vp, vp, vp, vp, vp, vp, vp], C.c_int),
        "masx_fem_param_grad": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp], C.c_int),
        "masx_hier_create": ([C.POINTER(vp)], C.c_int),
        "masx_hier_destroy": ([vp], C.c_int),
        "masx_hier_build": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_longlong, vp, vp, vp, C.c_int,
                             C.c_int, ip, ip, ip, vp], C.c_int),
        "masx_hier_download": ([vp, ip, ip, up, ip, ip, ip, ip, up, ip, vp], C.c_int),
        "masx_hier_device": ([vp, C.POINTER(HierArrays)], C.c_int),
        "masx_mas_create": ([C.POINTER(vp), C.c_int], C.c_int),
        "masx_mas_destroy": ([vp], C.c_int),
        "masx_mas_init_neighbor": ([vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "masx_mas_upload": ([vp, up, up, up, ip, ip], C.c_int),
        "masx_mas_init_matrix": ([vp, C.c_int], C.c_int),
        "masx_mas_set_preconditioner_bcoo": ([vp, vp, vp, vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp],
                                             C.c_int),
        "masx_mas_preconditioning": ([vp, vp, vp, vp], C.c_int),
        "masx_mas_solver": ([vp], vp),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError("masx: " + load().masx_last_error().decode())


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def torch_float64():
    import torch

    return torch.float64


def _dev(t):
    """Device pointer of a torch tensor (must live on the GPU)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _dev_f64(t, n, name):
    """Device pointer of a contiguous CUDA float64 tensor of exactly n
    values: the library reads or writes that many through the raw pointer."""
    import torch

    if t is None or not t.is_cuda or t.dtype != torch.float64:
        raise ValueError(f"{name}: expected a CUDA float64 tensor")
    if t.numel() != n:
        raise ValueError(f"{name}: expected {n} values, got {t.numel()}")
    return _dev(t)


def morton_partition(positions, bank=32):
    """Vertex partition by Morton order (the one masx_set_topology_morton uses):
    (part_to_real padded to the bank with -1, real_to_part)."""
    L = load()
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    n = len(pos)
    p2r = np.zeros((n + bank - 1) // bank * bank, np.int32)
    r2p = np.zeros(n, np.int32)
    check(L.masx_morton_partition(n, _ptr(pos, C.c_double), bank, _ptr(p2r, C.c_int), len(p2r), _ptr(r2p, C.c_int)))
    return p2r, r2p


class DeviceMorton:
    """Morton vertex partition on the device (masx_morton_*): the same arrays
    as morton_partition, from a CUDA float64 tensor of positions (n, 3)."""

    def __init__(self):
        self.L = load()
        h = C.c_void_p()
        check(self.L.masx_morton_create(C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.masx_morton_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def partition(self, positions, bank=32, stream=None):
        import torch

        pos = positions.contiguous()
        if pos.dtype != torch.float64 or not pos.is_cuda or pos.dim() != 2 or pos.shape[1] != 3:
            raise ValueError("positions must be a CUDA float64 tensor of shape (n, 3)")
        n = pos.shape[0]
        padded = (n + bank - 1) // bank * bank
        p2r = torch.empty(padded, dtype=torch.int32, device=pos.device)
        r2p = torch.empty(n, dtype=torch.int32, device=pos.device)
        check(self.L.masx_morton_partition_device(self.h, n, _dev(pos), bank, _dev(p2r), padded, _dev(r2p),
                                                  stream))
        return p2r, r2p


def stream_rate(nbytes, kind="read", reps=10):
    """Practical streaming rate of the GPU in GB/s (masx_stream_rate): a plain
    read-only stream ("read", bytes read) or copy ("copy", bytes read +
    written) over nbytes of device memory, fastest of reps launches."""
    out = C.c_double()
    check(load().masx_stream_rate(int(nbytes), {"read": 0, "copy": 1}[kind], reps, C.byref(out)))
    return out.value


MORTON_NOALLOC_MAX = 524288  # MASX_MORTON_NOALLOC_MAX


def morton_partition_noalloc(positions, part_to_real, real_to_part, bank=32, stream=None):
    """masx_morton_partition_device_noalloc: the device Morton partition into
    the caller's int32 CUDA tensors part_to_real (>= n rounded up to the bank)
    and real_to_part (n); the library allocates nothing (static scratch, at
    most MORTON_NOALLOC_MAX vertices)."""
    import torch

    pos = positions
    if pos.dtype != torch.float64 or not pos.is_cuda or pos.dim() != 2 or pos.shape[1] != 3 \
            or not pos.is_contiguous():
        raise ValueError("positions must be a contiguous CUDA float64 tensor of shape (n, 3)")
    n = pos.shape[0]
    for t, need, name in ((part_to_real, (n + bank - 1) // bank * bank, "part_to_real"), (real_to_part, n, "real_to_part")):
        if t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous() or t.numel() < need:
            raise ValueError(f"{name}: expected a contiguous CUDA int32 tensor of at least {need} values")
    check(load().masx_morton_partition_device_noalloc(n, _dev(pos), bank, _dev(part_to_real), part_to_real.numel(),
                                                      _dev(real_to_part), stream))


def spmv_layout_check(n_rows, rows, cols, entries_per_thread, seed=0, n_bodies=0):
    """Host check of the FP32 SpMV's tile layout for one system's upper block
    structure (n_bodies affine bodies in rows [0, 4 n_bodies), whose couplings
    to bodies become 12 x 12 blocks): (max |y_layout - y_full| / max |y_full|
    in FP64 on seeded random blocks and x, layout entries, pairs stored once)."""
    L = load()
    r, c = _i32(rows), _i32(cols)
    err, ne, nt = C.c_double(), C.c_longlong(), C.c_longlong()
    check(L.masx_spmv_layout_check_bodies(n_rows, n_bodies, len(r), _ptr(r, C.c_int), _ptr(c, C.c_int),
                                          entries_per_thread, seed, C.byref(err), C.byref(ne), C.byref(nt)))
    return err.value, ne.value, nt.value


def hierarchy_host(n_fem, n_slots, nbr_start, nbr_num, nbr_list, part_to_real, real_to_part, *, pairs=None,
                   node_offset=0, bank=32, levels=0):
    """Library hierarchy computed on the host (no device needed)."""
    L = load()
    s, n, l = _u32(nbr_start), _u32(nbr_num), _u32(nbr_list if len(nbr_list) else np.zeros(1))
    p2r, r2p = _i32(part_to_real), _i32(real_to_part)
    pr = _i32(np.zeros((0, 4)) if pairs is None else pairs).reshape(-1)
    npairs = len(pr) // 4
    pr_p = _ptr(pr, C.c_int) if npairs else None
    lv, tot = C.c_int(), C.c_int()
    check(L.masx_hierarchy_host(n_fem, n_slots, bank, levels, _ptr(s, C.c_uint32), _ptr(n, C.c_uint32),
                                _ptr(l, C.c_uint32), _ptr(p2r, C.c_int), _ptr(r2p, C.c_int), pr_p, npairs,
                                node_offset, C.byref(lv), C.byref(tot), None, None, None, None, None, 0, None))
    Lv, T = lv.value, tot.value
    lc = np.zeros(Lv + 1, np.int32)
    lb = np.zeros(Lv + 1, np.int32)
    fm = np.zeros(n_fem, np.uint32)
    px = np.zeros(n_slots // bank, np.int32)
    gn = np.zeros(max(T, n_fem), np.int32)
    ct = np.zeros((n_fem, 6), np.int32)
    check(L.masx_hierarchy_host(n_fem, n_slots, bank, levels, _ptr(s, C.c_uint32), _ptr(n, C.c_uint32),
                                _ptr(l, C.c_uint32), _ptr(p2r, C.c_int), _ptr(r2p, C.c_int), pr_p, npairs,
                                node_offset, C.byref(lv), C.byref(tot), _ptr(lc, C.c_int), _ptr(lb, C.c_int),
                                _ptr(fm, C.c_uint32), _ptr(px, C.c_int), _ptr(gn, C.c_int), len(gn),
                                _ptr(ct, C.c_int)))
    return dict(levels=Lv, total=T, level_count=lc, level_begin=lb, fine_mask=fm, prefix=px, going_next=gn,
                coarse_table=ct)


class Solver:
    """A batch of independent Newton systems solved by MAS-PCG on the GPU."""

    def __init__(self):
        self.L = load()
        h = C.c_void_p()
        check(self.L.masx_create(C.byref(h)))
        self.h = h
        self.systems = []

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.masx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_system(self, *, block_rows, rows, cols, n_abd=0, n_fem=None, use_mas=True, global_diag=False,
                   bank=32, levels=0):
        rows, cols = _i32(rows), _i32(cols)
        if n_fem is None:
            n_fem = block_rows - 4 * n_abd
        d = SystemDesc(block_rows, len(rows), n_abd, n_fem, int(use_mas), int(global_diag), bank, levels)
        idx = C.c_int()
        check(self.L.masx_add_system(self.h, C.byref(d), _ptr(rows, C.c_int), _ptr(cols, C.c_int), C.byref(idx)))
        self.systems.append(dict(block_rows=block_rows, nb=len(rows), n_abd=n_abd, n_fem=n_fem, bank=bank))
        return idx.value

    def set_structure(self, sys, rows, cols):
        """Replace system `sys`'s upper block structure (re-analysed at the next setup)."""
        r, c = _i32(rows), _i32(cols)
        if len(r) != len(c):
            raise ValueError("rows and cols differ in length")
        check(self.L.masx_set_structure(self.h, sys, len(r), _ptr(r, C.c_int) if len(r) else None,
                                        _ptr(c, C.c_int) if len(c) else None))
        self.systems[sys]["nb"] = len(r)

    def set_structure_device(self, sys, rows, cols, stream=None):
        """As set_structure, from CUDA int32 tensors (copied on the device)."""
        import torch

        for t, name in ((rows, "rows"), (cols, "cols")):
            if not t.is_cuda or t.dtype != torch.int32 or not t.is_contiguous():
                raise ValueError(f"{name}: expected a contiguous CUDA int32 tensor")
        if rows.numel() != cols.numel():
            raise ValueError("rows and cols differ in length")
        n = rows.numel()
        check(self.L.masx_set_structure_device(self.h, sys, n, _dev(rows) if n else None, _dev(cols) if n else None,
                                               stream))
        self.systems[sys]["nb"] = n

    def set_topology(self, sys, nbr_start, nbr_num, nbr_list, part_to_real, real_to_part):
        s, n, l = _u32(nbr_start), _u32(nbr_num), _u32(nbr_list if len(nbr_list) else np.zeros(1))
        p2r, r2p = _i32(part_to_real), _i32(real_to_part)
        check(self.L.masx_set_topology(self.h, sys, _ptr(s, C.c_uint32), _ptr(n, C.c_uint32), _ptr(l, C.c_uint32),
                                       len(nbr_list), _ptr(p2r, C.c_int), len(p2r), _ptr(r2p, C.c_int)))

    def set_topology_morton(self, sys, nbr_start, nbr_num, nbr_list, positions):
        s, n, l = _u32(nbr_start), _u32(nbr_num), _u32(nbr_list if len(nbr_list) else np.zeros(1))
        pos = np.ascontiguousarray(positions, np.float64)
        check(self.L.masx_set_topology_morton(self.h, sys, _ptr(s, C.c_uint32), _ptr(n, C.c_uint32),
                                              _ptr(l, C.c_uint32), len(nbr_list), _ptr(pos, C.c_double)))

    def set_contacts(self, sys, pairs, node_offset=0):
        p = _i32(pairs).reshape(-1)
        check(self.L.masx_set_contacts(self.h, sys, _ptr(p, C.c_int) if len(p) else None, len(p) // 4,
                                       node_offset))

    def analyze(self):
        check(self.L.masx_analyze(self.h))

    def setup(self):
        check(self.L.masx_setup(self.h))

    def _n_ext(self):
        return 3 * sum(d["block_rows"] for d in self.systems)

    def set_values(self, sys, values, stream=None):
        """values: device float64 tensor (nb, 9) or host numpy array."""
        n = 9 * self.systems[sys]["nb"]
        if isinstance(values, np.ndarray):
            v = np.ascontiguousarray(values, np.float64)
            if v.size != n:
                raise ValueError(f"values: expected {n} values, got {v.size}")
            check(self.L.masx_set_values_host(self.h, sys, _ptr(v, C.c_double), stream))
            self._keep = v
            import torch  # the copy is asynchronous on the library stream
            torch.cuda.synchronize()
        else:
            check(self.L.masx_set_values(self.h, sys, _dev_f64(values, n, "values"), stream))

    def set_values_many(self, values, stream=None):
        """values: {system: device float64 tensor (nb, 9)}; one library call."""
        items = sorted(values.items())
        sys_ids = np.array([k for k, _ in items], np.int32)
        ptrs = (C.c_void_p * len(items))(*[_dev_f64(v, 9 * self.systems[k]["nb"], "values").value for k, v in items])
        check(self.L.masx_set_values_many(self.h, len(items), _ptr(sys_ids, C.c_int), ptrs, stream))

    def build(self, precision="fp32", stream=None):
        check(self.L.masx_build(self.h, PRECISIONS.get(precision, precision), stream))

    def solve(self, b, x, *, rule=RULE_TRUE_REL, tol=1e-4, max_iter=0, check_every=50, chunk=8, use_graph=True,
              stream=None):
        p = SolveParams(rule, tol, max_iter, check_every, chunk, int(use_graph))
        st = (SolveStats * len(self.systems))()
        n = self._n_ext()
        check(self.L.masx_solve(self.h, _dev_f64(b, n, "b"), _dev_f64(x, n, "x"), C.byref(p), st, stream))
        return [s.as_dict() for s in st]

    def spmv(self, x, y, a=1.0, b=0.0, stream=None):
        n = self._n_ext()
        check(self.L.masx_spmv(self.h, a, _dev_f64(x, n, "x"), b, _dev_f64(y, n, "y"), stream))

    def apply_preconditioner(self, r, z, stream=None):
        n = self._n_ext()
        check(self.L.masx_apply_preconditioner(self.h, _dev_f64(r, n, "r"), _dev_f64(z, n, "z"), stream))

    def set_spmv_tiling(self, entries_per_thread):
        check(self.L.masx_set_spmv_tiling(self.h, entries_per_thread))

    def autotune(self, reps=50):
        """Times every supported SpMV tiling and keeps the fastest, then the
        other matrix load policy on it."""
        e, ms = C.c_int(), C.c_double()
        check(self.L.masx_autotune(self.h, reps, C.byref(e), C.byref(ms)))
        return dict(entries_per_thread=e.value, ms=ms.value, nontemporal=self.spmv_nontemporal())

    def set_mas_waves(self, waves):
        check(self.L.masx_set_mas_waves(self.h, waves))

    def set_factor_waves(self, waves):
        """MAS factor: 1 or 2 waves per bank, 0 automatic (next build)."""
        check(self.L.masx_set_factor_waves(self.h, waves))

    def set_spmv_loads(self, mode):
        """SpMV matrix loads: 1 nontemporal, 2 ordinary, 0 automatic."""
        check(self.L.masx_set_spmv_loads(self.h, mode))

    def spmv_nontemporal(self):
        v = C.c_int()
        check(self.L.masx_spmv_loads_used(self.h, C.byref(v)))
        return bool(v.value)

    def set_device_structure(self, on):
        """1: a new structure of one MAS system without bodies is analyzed on
        the device at the next setup (the default); 0: on the host."""
        check(self.L.masx_set_device_structure(self.h, 1 if on else 0))

    def device_structure_used(self):
        v = C.c_int()
        check(self.L.masx_device_structure_used(self.h, C.byref(v)))
        return bool(v.value)

    PLAN_ARRAYS = ("row_ptr", "col", "entry_src", "diag_src", "gal_ptr", "gal_src", "bank_ent_ptr", "bank_ent_pos",
                   "bank_ent_src", "spmv_row_ptr", "spmv_col", "spmv_src", "spmv_tptr", "spmv_tiles",
                   "spmv_tile_entries")

    def plan_array(self, name):
        """Host copy of an internal structure array (int32)."""
        which = self.PLAN_ARRAYS.index(name)
        n = C.c_longlong()
        check(self.L.masx_plan_array(self.h, which, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), np.int32)
        check(self.L.masx_plan_array(self.h, which, _ptr(out, C.c_int), out.size, C.byref(n)))
        return out[:n.value]

    def autotune_mas(self, reps=50):
        """Times every supported wave count of the MAS local kernel (built
        precision) and keeps the fastest."""
        w, ms = C.c_int(), C.c_double()
        check(self.L.masx_autotune_mas(self.h, reps, C.byref(w), C.byref(ms)))
        return dict(waves=w.value, ms=ms.value)

    def time_kernel(self, which, reps=20):
        ms = C.c_double()
        check(self.L.masx_time_kernel(self.h, which, reps, C.byref(ms)))
        return ms.value

    def info(self):
        i = PlanInfo()
        check(self.L.masx_plan_info_get(self.h, C.byref(i)))
        return i.as_dict()

    def hierarchy(self, sys=0):
        lv, tot = C.c_int(), C.c_int()
        lb = np.zeros(8, np.int32)
        check(self.L.masx_get_hierarchy(self.h, sys, C.byref(lv), C.byref(tot), None, _ptr(lb, C.c_int), None, None,
                                        None, None))
        d = self.systems[sys]
        Lv, T, n_slots = lv.value, tot.value, int(lb[1])
        lc = np.zeros(Lv + 1, np.int32)
        fm = np.zeros(d["n_fem"], np.uint32)
        px = np.zeros(max(n_slots // d["bank"], 1), np.int32)
        gn = np.zeros(T, np.int32)
        ct = np.zeros((d["n_fem"], 6), np.int32)
        check(self.L.masx_get_hierarchy(self.h, sys, C.byref(lv), C.byref(tot), _ptr(lc, C.c_int), _ptr(lb, C.c_int),
                                        _ptr(fm, C.c_uint32), _ptr(px, C.c_int), _ptr(gn, C.c_int), _ptr(ct, C.c_int)))
        return dict(levels=Lv, total=T, level_count=lc, level_begin=lb[:Lv + 1].copy(), fine_mask=fm, prefix=px,
                    going_next=gn, coarse_table=ct)

    def inverse(self):
        i = self.info()
        nblk = i["bank"] * (i["bank"] + 1) // 2
        out = np.zeros((i["n_banks"], nblk, 9), np.float32)
        check(self.L.masx_get_inverse(self.h, _ptr(out, C.c_float), out.size))
        return out

    def bank_of_slot(self, sys, slot):
        b = C.c_int()
        check(self.L.masx_bank_of_slot(self.h, sys, slot, C.byref(b)))
        return b.value

    def factor_status(self):
        c = C.c_longlong()
        check(self.L.masx_factor_status(self.h, C.byref(c)))
        return c.value


class Converter:
    """Triplets -> unique BCOO sorted by (row, col) on the GPU
    (Converter::convert, linear_system/utils/converter.cu:137-284)."""

    def __init__(self):
        self.L = load()
        h = C.c_void_p()
        check(self.L.masx_converter_create(C.byref(h)))
        self.h = h

    def __del__(self):
        try:
            if self.h is not None and self.h.value:
                self.L.masx_converter_destroy(self.h)
        except Exception:
            pass

    def convert(self, rows, cols, vals, n_block_rows, drop_lower=False, stream=None):
        """rows, cols: int32 CUDA tensors; vals: (n, 9) float64 CUDA tensor.
        Returns (rows, cols, vals) CUDA tensors of the unique blocks."""
        import torch

        n = int(rows.numel())
        out_r = torch.empty(max(n, 1), dtype=torch.int32, device=rows.device)
        out_c = torch.empty_like(out_r)
        out_v = torch.empty((max(n, 1), 9), dtype=torch.float64, device=rows.device)
        m = C.c_int()
        check(self.L.masx_convert_triplets(self.h, n, n_block_rows, _dev(rows), _dev(cols), _dev(vals),
                                           int(drop_lower), _dev(out_r), _dev(out_c), _dev(out_v), C.byref(m),
                                           stream))
        k = m.value
        return out_r[:k], out_c[:k], out_v[:k]


class FemHessian:
    """Stable Neo-Hookean element Hessians assembled on the GPU and the
    parameter gradient of an adjoint solve (femEnergy.cu:432-459, 1641-1855)."""

    def __init__(self, tets, rows, cols):
        self.L = load()
        t = _i32(tets).reshape(-1)
        r, c = _i32(rows), _i32(cols)
        h = C.c_void_p()
        check(self.L.masx_fem_create(C.byref(h), len(t) // 4, _ptr(t, C.c_int), len(r), _ptr(r, C.c_int),
                                     _ptr(c, C.c_int)))
        self.h = h
        self.n_tets, self.nb = len(t) // 4, len(r)
        self.n_rows = int(c.max()) + 1 if len(c) else 0

    def _check_elements(self, dminv, vol, F, mu, lam):
        n = self.n_tets
        for name, a, per in (("dminv", dminv, 9), ("vol", vol, 1), ("F", F, 9), ("mu", mu, 1), ("lam", lam, 1)):
            if a.numel() != n * per or a.dtype != torch_float64():
                raise ValueError(f"{name}: expected {n * per} float64 values")

    def __del__(self):
        try:
            if self.h is not None and self.h.value:
                self.L.masx_fem_destroy(self.h)
        except Exception:
            pass

    def assemble(self, dminv, vol, F, mu, lam, base=None, out=None, stream=None):
        """CUDA float64 tensors per tet (dminv (n, 9), vol (n,), F (n, 9), mu,
        lam (n,)); returns the (nb, 9) block values base + sum of K_e."""
        import torch

        self._check_elements(dminv, vol, F, mu, lam)
        if out is None:
            out = torch.empty((self.nb, 9), dtype=torch.float64, device=mu.device)
        for name, a in (("base", base), ("out", out)):
            if a is not None and (a.numel() != 9 * self.nb or a.dtype != torch.float64):
                raise ValueError(f"{name}: expected {9 * self.nb} float64 values")
        check(self.L.masx_fem_assemble(self.h, _dev(dminv), _dev(vol), _dev(F), _dev(mu), _dev(lam),
                                       _dev(base) if base is not None else None, _dev(out), stream))
        return out

    def param_grad(self, dminv, vol, F, mu, lam, x, y, stream=None):
        """(-y_e^T dK_e/dmu x_e, -y_e^T dK_e/dlambda x_e) per tet; x, y CUDA
        float64 block vectors."""
        import torch

        self._check_elements(dminv, vol, F, mu, lam)
        for name, a in (("x", x), ("y", y)):
            if a.numel() < 3 * self.n_rows or a.dtype != torch.float64:
                raise ValueError(f"{name}: expected at least {3 * self.n_rows} float64 values")
        gm = torch.empty(self.n_tets, dtype=torch.float64, device=mu.device)
        gl = torch.empty_like(gm)
        check(self.L.masx_fem_param_grad(self.h, _dev(dminv), _dev(vol), _dev(F), _dev(mu), _dev(lam), _dev(x),
                                         _dev(y), _dev(gm), _dev(gl), stream))
        return gm, gl


class DeviceHierarchy:
    """MAS hierarchy rebuilt on the GPU (ReorderRealtime,
    TraditionalMAS32Preconditioner.cu:2189-2230); same arrays as
    hierarchy_host, plus the per-bank component lists."""

    def __init__(self):
        self.L = load()
        h = C.c_void_p()
        check(self.L.masx_hier_create(C.byref(h)))
        self.h = h
        self.n_fem = self.n_slots = self.bank = 0
        self.levels = self.total = self.n_components = 0

    def __del__(self):
        try:
            if self.h is not None and self.h.value:
                self.L.masx_hier_destroy(self.h)
        except Exception:
            pass

    def build(self, n_fem, n_slots, nbr_start, nbr_num, nbr_list, part_to_real, real_to_part, *, pairs=None,
              node_offset=0, bank=32, levels=0, stream=None):
        """Topology, partition and pairs are CUDA tensors (uint32 / int32,
        pairs (n, 4) int32 or None).  Returns (levels, total, components)."""
        lv, tot, nc = C.c_int(), C.c_int(), C.c_int()
        n_pairs = 0 if pairs is None else int(pairs.numel()) // 4
        check(self.L.masx_hier_build(self.h, n_fem, n_slots, bank, levels, _dev(nbr_start), _dev(nbr_num),
                                     _dev(nbr_list) if nbr_list.numel() else None, int(nbr_list.numel()),
                                     _dev(part_to_real), _dev(real_to_part), _dev(pairs) if n_pairs else None,
                                     n_pairs, node_offset, C.byref(lv), C.byref(tot), C.byref(nc), stream))
        self.n_fem, self.n_slots, self.bank = n_fem, n_slots, bank
        self.levels, self.total, self.n_components = lv.value, tot.value, nc.value
        return self.levels, self.total, self.n_components

    def download(self, stream=None):
        """Host copies, the keys of hierarchy_host plus comp_ptr, comp_mask, comp_parent."""
        Lv, T = self.levels, self.total
        lc = np.zeros(Lv + 1, np.int32)
        lb = np.zeros(Lv + 1, np.int32)
        fm = np.zeros(self.n_fem, np.uint32)
        px = np.zeros(self.n_slots // self.bank, np.int32)
        gn = np.zeros(max(T, self.n_fem), np.int32)
        ct = np.zeros((self.n_fem, 6), np.int32)
        cp = np.zeros(T // self.bank + 1, np.int32)
        cm = np.zeros(max(self.n_components, 1), np.uint32)
        cq = np.zeros(max(self.n_components, 1), np.int32)
        check(self.L.masx_hier_download(self.h, _ptr(lc, C.c_int), _ptr(lb, C.c_int), _ptr(fm, C.c_uint32),
                                        _ptr(px, C.c_int), _ptr(gn, C.c_int), _ptr(ct, C.c_int), _ptr(cp, C.c_int),
                                        _ptr(cm, C.c_uint32), _ptr(cq, C.c_int), stream))
        k = self.n_components
        return dict(levels=Lv, total=T, level_count=lc, level_begin=lb, fine_mask=fm, prefix=px, going_next=gn,
                    coarse_table=ct, comp_ptr=cp, comp_mask=cm[:k], comp_parent=cq[:k])

    def device_arrays(self):
        """Raw device addresses of the result arrays (valid until the next build)."""
        a = HierArrays()
        check(self.L.masx_hier_device(self.h, C.byref(a)))
        return {name: getattr(a, name) for name, _ in HierArrays._fields_}


class MasPreconditioner:
    """TraditionalMAS32Preconditioner-shaped object over the C ABI
    (TraditionalMAS32Preconditioner.cuh:78-116)."""

    def __init__(self, bank=32):
        self.L = load()
        h = C.c_void_p()
        check(self.L.masx_mas_create(C.byref(h), bank))
        self.h = h

    def __del__(self):
        try:
            if self.h is not None and self.h.value:
                self.L.masx_mas_destroy(self.h)
        except Exception:
            pass

    def init_neighbor(self, vert_num, collision_node_offset, total_neighbor_num, part_map_size):
        check(self.L.masx_mas_init_neighbor(self.h, vert_num, collision_node_offset, total_neighbor_num,
                                            part_map_size))

    def upload(self, neighbor_list, neighbor_start, neighbor_num, part_to_real, real_to_part):
        self._keep = [_u32(neighbor_list if len(neighbor_list) else np.zeros(1)), _u32(neighbor_start),
                      _u32(neighbor_num), _i32(part_to_real), _i32(real_to_part)]
        a, b, c, d, e = self._keep
        check(self.L.masx_mas_upload(self.h, _ptr(a, C.c_uint32), _ptr(b, C.c_uint32), _ptr(c, C.c_uint32),
                                     _ptr(d, C.c_int), _ptr(e, C.c_int)))

    def init_matrix(self, levels=0):
        check(self.L.masx_mas_init_matrix(self.h, levels))

    def set_preconditioner_bcoo(self, values, rows, cols, indices, offset, triplet_num, pairs=None, cp_num=0,
                                precision="fp32", stream=None):
        check(self.L.masx_mas_set_preconditioner_bcoo(self.h, _dev(values), _dev(rows), _dev(cols), _dev(indices),
                                                      offset, triplet_num, _dev(pairs), cp_num,
                                                      PRECISIONS.get(precision, precision), stream))

    def preconditioning(self, R, Z, stream=None):
        check(self.L.masx_mas_preconditioning(self.h, _dev(R), _dev(Z), stream))

"""ctypes binding of the C ABI declared in include/mesh_bem_amd.h.

The library is the in-tree ``libmesh_bem_amd.so`` built by ``_build.build()``.
There is no fallback: if the library is missing or a GPU is absent, calls
raise.  Tensors are passed as raw device pointers together with torch's
current HIP stream.
"""

import atexit
import ctypes
import os

import torch

from ._build import LIB_PATH

MBA_F32, MBA_F64 = 0, 1
MBA_COMPLEX, MBA_REAL = 0, 1


class NsMesh(ctypes.Structure):
    """mba_ns_mesh of include/mesh_bem_amd.h."""

    _fields_ = [("n_faces", ctypes.c_int64), ("face_off", ctypes.c_int64), ("wavenumber", ctypes.c_double),
                ("near_ptr", ctypes.c_void_p), ("near_idx", ctypes.c_void_p), ("zf", ctypes.c_void_p),
                ("ldzf", ctypes.c_int64), ("n_near", ctypes.c_int64)]


class LuMatrix(ctypes.Structure):
    """mba_lu_matrix of include/mesh_bem_amd.h."""

    _fields_ = [("a", ctypes.c_void_p), ("n", ctypes.c_int64), ("lda", ctypes.c_int64), ("ipiv", ctypes.c_void_p),
                ("info", ctypes.c_void_p), ("workspace", ctypes.c_void_p), ("workspace_bytes", ctypes.c_size_t)]


class GemmJob(ctypes.Structure):
    """mba_gemm_job of include/mesh_bem_amd.h."""

    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p), ("m", ctypes.c_int64),
                ("n", ctypes.c_int64), ("k", ctypes.c_int64), ("lda", ctypes.c_int64), ("ldb", ctypes.c_int64),
                ("ldc", ctypes.c_int64), ("alpha", ctypes.c_double), ("beta", ctypes.c_double)]


class LuPrepareJob(ctypes.Structure):
    """mba_lu_prepare_job of include/mesh_bem_amd.h."""

    _fields_ = [("lu", ctypes.c_void_p), ("n", ctypes.c_int64), ("lda", ctypes.c_int64), ("ipiv", ctypes.c_void_p),
                ("aux", ctypes.c_void_p), ("aux_bytes", ctypes.c_size_t)]


class LuSolveJob(ctypes.Structure):
    """mba_lu_solve_job of include/mesh_bem_amd.h."""

    _fields_ = [("lu", ctypes.c_void_p), ("n", ctypes.c_int64), ("lda", ctypes.c_int64), ("aux", ctypes.c_void_p),
                ("b", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("nrhs", ctypes.c_int64), ("work", ctypes.c_void_p)]


class NsRhsMesh(ctypes.Structure):
    """mba_ns_rhs_mesh of include/mesh_bem_amd.h."""

    _fields_ = [("n_faces", ctypes.c_int64), ("face_off", ctypes.c_int64), ("wavenumber", ctypes.c_double),
                ("dirs", ctypes.c_void_p), ("pols", ctypes.c_void_p), ("ndir", ctypes.c_int64),
                ("rwg_fe", ctypes.c_void_p), ("n_edges", ctypes.c_int64), ("v", ctypes.c_void_p)]


class GatherJob(ctypes.Structure):
    """mba_gather_job of include/mesh_bem_amd.h."""

    _fields_ = [("zf", ctypes.c_void_p), ("ldzf", ctypes.c_int64), ("rwg_fe", ctypes.c_void_p),
                ("n_edges", ctypes.c_int64), ("z", ctypes.c_void_p), ("ldz", ctypes.c_int64)]


class ConvertJob(ctypes.Structure):
    """mba_convert_job of include/mesh_bem_amd.h."""

    _fields_ = [("src", ctypes.c_void_p), ("dst", ctypes.c_void_p), ("count", ctypes.c_int64)]


class RefineJob(ctypes.Structure):
    """mba_refine_job of include/mesh_bem_amd.h."""

    _fields_ = [("j", ctypes.c_void_p), ("ldj", ctypes.c_int64), ("d", ctypes.c_void_p), ("ldd", ctypes.c_int64),
                ("n", ctypes.c_int64), ("nrhs", ctypes.c_int64)]


class MgnVMat(ctypes.Structure):
    """mba_mgn_vmat of include/mesh_bem_amd.h."""

    _fields_ = [("p", ctypes.c_void_p * 3), ("idx", ctypes.c_void_p * 3), ("ld", ctypes.c_int64 * 3),
                ("w", ctypes.c_int32 * 3), ("nseg", ctypes.c_int32), ("ones", ctypes.c_int32)]

class MeshPart(ctypes.Structure):
    """mba_mesh_part of include/mesh_bem_amd.h."""

    _fields_ = [("n_verts", ctypes.c_int64), ("n_faces", ctypes.c_int64),
                ("verts", ctypes.POINTER(ctypes.c_double)), ("faces", ctypes.POINTER(ctypes.c_int32))]


_c = ctypes
_vp, _i32p, _i64p, _dp = _c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_void_p
_i32, _i64, _dbl, _sz = _c.c_int32, _c.c_int64, _c.c_double, _c.c_size_t

# name -> (restype, argtypes) for every symbol of include/mesh_bem_amd.h
SIGNATURES = {
    "mba_status_string": (_c.c_char_p, [_c.c_int]),
    "mba_version": (_c.c_int, []),
    "mba_mom_zmat_faceedge_ref": (_c.c_int, [_c.c_int, _vp, _i32p, _vp, _vp, _vp, _vp, _i32, _vp, _vp,
                                             _i32, _dbl, _i64, _i64, _vp, _i64, _vp]),
    "mba_mom_zmat_faceedge_ref_noalloc": (_c.c_int, [_c.c_int, _vp, _i32p, _vp, _vp, _vp, _vp, _i32, _vp, _vp,
                                                     _i32, _dbl, _i64, _i64, _vp, _i64, _vp]),
    "mba_mom_rhs_faceedge_ref": (_c.c_int, [_c.c_int, _vp, _i32p, _vp, _vp, _vp, _vp, _i32, _dbl, _i64,
                                            _i64, _vp, _vp]),
    "mba_mom_interp_faceedge": (_c.c_int, [_c.c_int, _vp, _i32p, _i32p, _vp, _vp, _i64, _i64, _i64,
                                           _i64, _vp, _vp]),
    "mba_mom_vertex_sum": (_c.c_int, [_c.c_int, _vp, _i64p, _i32p, _i64, _i64, _i64, _vp, _vp]),
    "mba_mom_workspace_bytes": (_sz, [_c.c_int, _i64]),
    "mba_mom_zmat_faceedge_ns": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _i64, _dp, _i32, _dbl, _i64p,
                                            _i32p, _i64, _i64, _vp, _i64, _vp, _sz, _vp]),
    "mba_mom_rhs_faceedge_ns": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _i64, _dp, _i32, _dbl, _dp, _dp,
                                           _i64, _vp, _vp]),
    "mba_mom_gather_rwg": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _i64, _i32p, _i64, _vp, _i64, _vp]),
    "mba_mom_gather_rhs": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _i64, _vp, _vp]),
    "mba_mom_batched_workspace_bytes": (_sz, [_c.c_int, _i64, _i32]),
    "mba_mom_zmat_faceedge_ns_batched": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _dp, _i32, _vp, _i32, _vp, _sz,
                                                    _vp]),
    "mba_lu_workspace_bytes": (_sz, [_c.c_int, _i64]),
    "mba_lu_factor": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _i32p, _i32p, _vp, _sz, _vp]),
    "mba_lu_factor_async": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _i32p, _i32p, _vp, _sz, _i32, _i32p, _vp]),
    "mba_gemm_grouped_scratch_bytes": (_sz, [_i32]),
    "mba_gemm_grouped": (_c.c_int, [_c.c_int, _vp, _i32, _vp, _sz, _vp]),
    "mba_lu_solve_prepare_batched_scratch_bytes": (_sz, [_i32]),
    "mba_lu_solve_prepare_batched": (_c.c_int, [_c.c_int, _vp, _i32, _vp, _sz, _vp]),
    "mba_lu_solve_batched_scratch_bytes": (_sz, [_i32]),
    "mba_lu_solve_batched": (_c.c_int, [_c.c_int, _vp, _i32, _vp, _sz, _vp]),
    "mba_lu_solve_batched_persistent_scratch_bytes": (_sz, [_vp, _i32]),
    "mba_lu_solve_batched_persistent": (_c.c_int, [_c.c_int, _vp, _i32, _vp, _sz, _vp, _vp]),
    "mba_lu_batched_scratch_bytes": (_sz, [_c.c_int, _i32, _i64]),
    "mba_lu_factor_batched": (_c.c_int, [_c.c_int, _vp, _i32, _vp, _sz, _i32p, _vp]),
    "mba_lu_batched_persistent_scratch_bytes": (_sz, [_c.c_int, _i32, _i64]),
    "mba_lu_factor_batched_persistent": (_c.c_int, [_c.c_int, _vp, _i32, _vp, _sz, _i32p, _vp]),
    "mba_measure_valu_rates": (_c.c_int, [_vp, _c.c_int, _vp]),
    "mba_lu_factor_batched_persistent_opts": (_c.c_int, [_c.c_int, _vp, _i32, _vp, _vp, _sz, _i32p, _vp]),
    "mba_lu_factor_batched_persistent_profiled": (_c.c_int, [_vp, _i32, _vp, _vp, _sz, _i32p, _vp, _vp, _vp]),
    "mba_lu_factor_profiled": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _i32p, _i32p, _vp, _sz, _vp, _vp,
                                          _vp, _vp]),
    "mba_lu_pivots_to_perm": (_c.c_int, [_i32p, _i64, _i32p, _vp]),
    "mba_lu_solve_workspace_bytes": (_sz, [_c.c_int, _i64, _i64]),
    "mba_lu_solve_aux_bytes": (_sz, [_c.c_int, _i64]),
    "mba_lu_solve_prepare": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _i32p, _vp, _sz, _vp]),
    "mba_lu_solve": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _sz, _vp]),
    "mba_gemm": (_c.c_int, [_c.c_int, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64,
                            _vp]),
    "mba_gemm_variant": (_c.c_int, [_c.c_int, _c.c_int, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl,
                                    _vp, _i64, _vp]),
    "mba_gemm_set_variant": (_c.c_int, [_c.c_int, _c.c_int]),
    "mba_gemm_split_f16_workspace_bytes": (_sz, [_i64, _i64, _i64]),
    "mba_gemm_split_f16": (_c.c_int, [_i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64, _vp, _sz,
                                      _vp]),
    "mba_lu_set_update": (_c.c_int, [_c.c_int]),
    "mba_lu_set_solve": (_c.c_int, [_c.c_int]),
    "mba_lu_set_panel": (_c.c_int, [_c.c_int]),
    "mba_gemm_split_f16_set_tile": (_c.c_int, [_c.c_int]),
    "mba_residual_c128": (_c.c_int, [_vp, _i64, _i64, _vp, _vp, _vp, _i64, _vp]),
    "mba_gemv": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _vp, _vp, _vp]),
    "mba_convert_c128_c64": (_c.c_int, [_vp, _i64, _i64, _i64, _vp, _i64, _vp]),
    "mba_mom_rhs_ns_batched_workspace_bytes": (_sz, [_c.c_int, _i64, _vp, _i32]),
    "mba_mom_rhs_ns_batched": (_c.c_int, [_c.c_int, _vp, _vp, _i64, _vp, _i32, _vp, _i32, _vp, _sz, _vp]),
    "mba_mom_gather_rwg_batched_scratch_bytes": (_sz, [_i32]),
    "mba_mom_gather_rwg_batched": (_c.c_int, [_c.c_int, _vp, _i32, _vp, _sz, _vp]),
    "mba_convert_c128_c64_batched_scratch_bytes": (_sz, [_i32]),
    "mba_convert_c128_c64_batched": (_c.c_int, [_vp, _i32, _vp, _sz, _vp]),
    "mba_refine_update_batched_scratch_bytes": (_sz, [_vp, _i32]),
    "mba_refine_update_batched": (_c.c_int, [_vp, _i32, _vp, _vp, _sz, _vp]),
    "mba_mv32": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _i64, _c.c_int, _vp, _vp, _vp, _vp]),
    "mba_gmres_workspace_bytes": (_sz, [_i64, _i32]),
    "mba_gmres_dots": (_c.c_int, [_c.c_int, _vp, _i64, _i32, _c.c_int, _vp, _i64, _vp, _vp, _sz, _vp]),
    "mba_gmres_update": (_c.c_int, [_c.c_int, _vp, _i64, _i32, _vp, _dbl, _c.c_int, _vp, _i64, _vp]),
    "mba_gmres_scale": (_c.c_int, [_c.c_int, _c.c_int, _vp, _i64, _vp, _vp, _vp]),
    "mba_near_workspace_bytes": (_sz, [_i64]),
    "mba_near_pairs_count": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _i64, _dbl, _i64p, _vp, _sz, _vp]),
    "mba_near_pairs_fill": (_c.c_int, [_i64, _dbl, _i64p, _i32p, _vp, _sz, _vp]),
    "mba_add_diag": (_c.c_int, [_c.c_int, _vp, _i64, _i64, _dbl, _vp, _vp]),
    "mba_helmholtz_decomp_workspace_bytes": (_sz, [_c.c_int, _i64, _i64, _i32, _i64]),
    "mba_helmholtz_decomp": (_c.c_int, [_c.c_int, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i32, _vp, _vp,
                                        _sz, _vp]),
    "mba_radiate_h_workspace_bytes": (_sz, [_c.c_int, _i64, _i64]),
    "mba_radiate_h": (_c.c_int, [_c.c_int, _vp, _i64, _vp, _vp, _vp, _i64, _dbl, _vp, _vp, _sz, _vp]),
    "mba_mgn_gemm": (_c.c_int, [_vp, _c.c_int, _vp, _i64, _c.c_int, _i64, _i64, _i64, _vp, _i64, _vp, _vp,
                                _vp, _c.c_int, _c.c_int, _i32, _vp, _vp]),
    "mba_mgn_segsum": (_c.c_int, [_i64p, _i32p, _i64, _vp, _i64, _i32, _i32, _vp, _i64, _c.c_int, _vp]),
    "mba_mgn_loss": (_c.c_int, [_vp, _vp, _i64p, _i32, _vp, _vp, _vp]),
    "mba_mgn_adam": (_c.c_int, [_vp, _vp, _vp, _vp, _i64, _c.c_float, _c.c_float, _c.c_float, _c.c_float, _i64,
                                _vp]),
    "mba_laplace_layer_ref": (_c.c_int, [_c.c_int, _vp, _i32p, _vp, _vp, _i64, _i64, _vp, _i64p, _i32,
                                         _i32, _vp, _vp]),
    "mba_laplace_pack": (_c.c_int, [_c.c_int, _vp, _i64, _i32, _i64p, _i32p, _i64, _vp, _vp]),
    "mba_laplace_dirichlet_rhs": (_c.c_int, [_c.c_int, _vp, _i64, _vp, _vp, _vp, _vp]),
    "mba_bem2d_layer": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _vp, _i32p, _i64, _i64, _vp, _vp, _i32,
                                   _dbl, _vp, _vp]),
    "mba_bem2d_pack": (_c.c_int, [_c.c_int, _c.c_int, _vp, _i64, _i32, _i64p, _i32p, _i64, _vp, _vp]),
    "mba_bem2d_dirichlet_rhs": (_c.c_int, [_c.c_int, _c.c_int, _vp, _i64, _vp, _vp, _dbl, _dbl, _vp, _vp]),
    "mba_pressure_bc": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _i64, _i64p, _i32p, _vp, _vp, _vp, _i64, _dbl, _dbl,
                                   _dbl, _dbl, _vp, _vp, _vp, _vp, _vp]),
    "mba_pressure_update": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _i64, _i64p, _i32p, _vp, _vp, _vp, _vp, _i64, _dbl,
                                       _dbl, _vp, _vp]),
    "mba_edge_unique_workspace_bytes": (_sz, [_i64, _i64]),
    "mba_edge_unique": (_c.c_int, [_i32p, _i64, _i64, _i32p, _i32p, _i32p, _i64p, _vp, _sz, _vp]),
    "mba_mesh_read": (_c.c_int, [_c.c_char_p, _vp, _vp]),
    "mba_mesh_free": (None, [_vp, _i32]),
    "mba_mesh_write_beast_in": (_c.c_int, [_c.c_char_p, _vp, _i64, _vp, _i64]),
    "mba_loop_subdivide": (_c.c_int, [_vp, _i64, _vp, _i64, _vp, _vp, _vp]),
    "mba_decimate_quadric": (_c.c_int, [_vp, _i64, _vp, _i64, _i64, _vp, _vp, _vp, _vp]),
    "mba_nearest_transfer_workspace_bytes": (_sz, [_c.c_int, _i64, _i64]),
    "mba_nearest_transfer": (_c.c_int, [_c.c_int, _vp, _i64, _vp, _i64, _vp, _c.c_int32, _vp, _vp, _vp, _sz, _vp]),
    "mba_laplace_p0_workspace_bytes": (_sz, [_c.c_int, _i64]),
    "mba_laplace_p1_assemble": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _i64, _dp, _i32, _i64p, _i32p, _vp, _vp,
                                           _vp, _sz, _vp]),
    "mba_laplace_mass_apply": (_c.c_int, [_c.c_int, _vp, _i32p, _i64p, _i32p, _i64, _vp, _dbl, _dbl, _vp, _vp]),
    "mba_laplace_p0_assemble": (_c.c_int, [_c.c_int, _vp, _i32p, _i64, _i64, _dp, _i32, _i64p, _i32p,
                                           _vp, _vp, _i64, _vp, _sz, _vp]),
}

_LIB = None


class BatchedLuOptions(ctypes.Structure):
    """mba_batched_lu_options: 32-column panels per super panel (0: 8) and the
    trailing product (0: FP32 3M, 1: split FP16)."""
    _fields_ = [("super_panels", ctypes.c_int32), ("trailing", ctypes.c_int32)]


class MBAError(RuntimeError):
    pass


def load():
    """Load the HIP library; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise MBAError(
                f"{LIB_PATH} not found: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`"
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue  # exported-symbol completeness is checked by the tests
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
        atexit.register(_quiesce)
    return _LIB


def _quiesce():
    """At interpreter exit: wait for all queued device work, hand the memory
    cached by the stream-ordered allocator (the library's hipMallocAsync
    records) back to the device, then unload the library.  Every HIP shared
    library registers, at load, an exit handler that unregisters its code
    objects from the HIP runtime; left to the C exit handlers it runs after a
    profiler's own exit-time finalisation (rocprofv3's "tool finalization"),
    into a runtime that profiler has already taken apart (the SIGSEGV inside
    exit() of profiled runs, absent from runs that never load this library).
    Unloading here runs that handler while the runtime is whole and takes it
    off the exit list."""
    global _LIB
    try:
        if torch.cuda.is_available() and torch.cuda.is_initialized():
            torch.cuda.synchronize()
            hip = ctypes.CDLL("libamdhip64.so")
            dev = ctypes.c_int(0)
            if hip.hipGetDevice(ctypes.byref(dev)) == 0:
                pool = ctypes.c_void_p()
                if hip.hipDeviceGetDefaultMemPool(ctypes.byref(pool), dev) == 0:
                    hip.hipMemPoolTrimTo(pool, ctypes.c_size_t(0))
                hip.hipDeviceSynchronize()
        if _LIB is not None:
            import _ctypes

            handle, _LIB = _LIB._handle, None
            _ctypes.dlclose(handle)
    except Exception:
        pass


def check(status: int, what: str = ""):
    if status != 0:
        msg = load().mba_status_string(status).decode()
        raise MBAError(f"{what}: {msg} (status {status})")


def ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def stream():
    if not torch.cuda.is_available():
        raise MBAError("no GPU: the mesh-BEM kernels run on a HIP device only")
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def prec_of(dtype: torch.dtype) -> int:
    if dtype in (torch.float32, torch.complex64):
        return MBA_F32
    if dtype in (torch.float64, torch.complex128):
        return MBA_F64
    raise MBAError(f"unsupported dtype {dtype}")


def require_device(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise MBAError("mesh-BEM kernels take device tensors")
        if t is not None and not t.is_contiguous():
            raise MBAError("mesh-BEM kernels take contiguous tensors")

"""ctypes binding of libpiqmc_amd.so (declarations in include/piqmc_amd.h).

The library is built in-tree by pathintegral-qmc_amd/build_lib.py.  There is no
CPU fallback: if the library is missing, importing piqmc fails loudly.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpiqmc_amd.so")

PIQMC_OK, PIQMC_EINVAL, PIQMC_EINDEX, PIQMC_EZERODIV, PIQMC_EHIP, PIQMC_ENOMEM = range(6)

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "piqmc: %s is not built; run `python pathintegral-qmc_amd/build_lib.py` "
        "(or __graft_entry__.build())" % LIB_PATH)

# One HIP runtime per process: torch's, where torch is installed.  Loaded
# first, torch's runtime is the one the library binds to; loaded after the
# library, torch would find the system runtime the library pulled in and see
# no GPU ("No HIP GPUs are available").
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = C.CDLL(LIB_PATH)

_vp = C.c_void_p
_i, _i64, _u32, _u64, _f, _d = C.c_int, C.c_int64, C.c_uint32, C.c_uint64, C.c_float, C.c_double

_SIGS = {
    "piqmc_last_error": ([], C.c_char_p),
    "piqmc_version": ([], _i),
    "piqmc_last_kernel_ms": ([_vp], _i),
    "piqmc_problem_create": ([_i, _i, _vp, C.POINTER(_vp)], _i),
    "piqmc_problem_create_dense": ([_i, _vp, _i64, _i64, C.POINTER(_vp)], _i),
    "piqmc_problem_destroy": ([_vp], _i),
    "piqmc_problem_info": ([_vp, _vp, _vp, _vp, _vp, _vp], _i),
    "piqmc_problem_exact_rows": ([_vp, _vp], _i),
    "piqmc_problem_set_coloring": ([_vp, _i, _vp], _i),
    "piqmc_problem_get_coloring": ([_vp, _vp], _i),
    "piqmc_problem_set_sweep_path": ([_vp, _i], _i),
    "piqmc_problem_set_dense_variant": ([_vp, _i], _i),
    "piqmc_rand_seed": ([_u32, _vp], _i),
    "piqmc_rand_advance": ([_vp, _u64], _i),
    "piqmc_perm_streams": ([_vp, _i, _i, _i, _vp, _i], _i),
    "piqmc_qa_replay": ([_vp, _i, _i, _i, _i, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "piqmc_sa_replay": ([_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "piqmc_quantum_anneal": ([_vp, _i, _i, _i, _f, _i, _vp, _i64, _i64, _vp, _i, _vp, _vp, _vp,
                              _vp], _i),
    "piqmc_anneal": ([_vp, _i, _i, _vp, _i64, _i, _vp, _i, _vp, _vp, _vp, _vp], _i),
    "piqmc_quantum_anneal_dense": ([_vp, _i, _i, _i, _f, _i, _vp, _i64, _i64, _vp, _i64, _i64,
                                    _vp, _vp, _vp, _vp], _i),
    "piqmc_anneal_dense": ([_vp, _i, _i, _vp, _i64, _i, _vp, _i64, _i64, _vp, _vp, _vp, _vp], _i),
    "piqmc_sa_multispin": ([_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "piqmc_sa_colored": ([_vp, _i, _i, _i, _vp, _vp, _u64, _u32, _u32, _i, _i, _vp], _i),
    "piqmc_qa_colored": ([_vp, _i, _i, _i, _i, _f, _vp, _vp, _u64, _u32, _u32, _vp, _i, _i, _vp],
                         _i),
    "piqmc_qa_colored_noalloc": ([_vp, _i, _i, _i, _i, _f, _vp, _vp, _u64, _u32, _u32, _vp, _i, _i,
                                  _vp], _i),
    "piqmc_slice_energies": ([_vp, _i, _vp, _vp, _vp], _i),
    "piqmc_classical_energy": ([_i, _i64, _vp, _vp, _vp, _i, _vp, _vp], _i),
    "piqmc_generate_neighbors": ([_i, _i, _i64, _vp, _vp, _vp, _i, _vp], _i),
    "piqmc_made_create": ([_i, _i, _i, _u64, C.POINTER(_vp)], _i),
    "piqmc_made_destroy": ([_vp], _i),
    "piqmc_made_get_params": ([_vp] * 7, _i),
    "piqmc_made_get_grads": ([_vp] * 7, _i),
    "piqmc_made_logprob": ([_vp, _i, _vp, _vp, _vp], _i),
    "piqmc_made_train_step": ([_vp, _i, _vp, _f, _vp, _vp], _i),
    "piqmc_made_train_epoch": ([_vp, _i, _vp, _vp, _i, _f, _vp, _vp], _i),
    "piqmc_made_sample": ([_vp, _i, _u64, _u32, _vp, _vp, _vp], _i),
    "piqmc_made_set_graph": ([_vp, _i], _i),
    "piqmc_made_mh": ([_vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _f, _f, _u64, _u32, _vp, _vp], _i),
    "piqmc_made_ar": ([_vp, _i, _i, _vp, _f, _u64, _u32, _vp, _vp, _vp], _i),
    "piqmc_made_worldline_mh": ([_vp, _vp, _i, _i, _vp, _f, _f, _u64, _u32, _vp, _vp], _i),
    "piqmc_made_slice_mh": ([_vp, _vp, _i, _i, _vp, _vp, _f, _f, _f, _u64, _u32, _vp, _vp], _i),
    "piqmc_gemm_bf16_nt": ([_vp, _vp, _i, _i, _i, _vp, _vp], _i),
    "piqmc_batch_create": ([C.POINTER(_vp)], _i),
    "piqmc_batch_destroy": ([_vp], _i),
    "piqmc_batch_add": ([_vp, _vp, _i, _i, _i, _i, _f, _vp, _vp, _u64, _u32, _u32, _vp], _i),
    "piqmc_batch_run": ([_vp, _i, _i, _vp], _i),
    "piqmc_batch_prepare": ([_vp, _i, _i, _vp], _i),
    "piqmc_batch_launch": ([_vp, _vp], _i),
    "piqmc_batch_info": ([_vp, _vp, _vp, _vp], _i),
}
for _name, (_args, _res) in _SIGS.items():
    _fn = getattr(lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res

EXPORTED = tuple(_SIGS)


def check(rc):
    """Map a status code to the exception the reference raises for it."""
    if rc == PIQMC_OK:
        return
    msg = (lib.piqmc_last_error() or b"").decode(errors="replace")
    if rc == PIQMC_EINVAL:
        raise ValueError(msg)
    if rc == PIQMC_EINDEX:
        raise IndexError(msg)
    if rc == PIQMC_EZERODIV:
        raise ZeroDivisionError(msg)
    if rc == PIQMC_ENOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def call(name, *args):
    check(getattr(lib, name)(*args))

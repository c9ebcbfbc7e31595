"""Runtime knobs (``stablebear/system.py``), mapped onto this build.

Every pairwise computation runs on the GPU through the HIP library, so the
knobs that steer the reference's CPU/GPU choice are kept for source
compatibility: ``force_cpu(True)`` is refused (there is no CPU compute path),
``set_cuda_threshold`` / ``limit_cpus`` / ``set_parallel_eval_threshold`` /
``set_min_block_side`` are recorded but do not change where work runs;
``set_block_size`` / ``set_tile_cap`` set the pair-tile shape of the HIP
kernels, and ``set_output_budget`` the device memory an output may take at
once (larger outputs run in bands of rows).
"""

from __future__ import annotations

_state = {
    "force_cpu": False,
    "cuda_threshold": 500,
    "device_verbose": False,
    "block": (0, 0),
    "cap": 0,
    "min_block_side": 0,
    "cpus": 0,
    "gpus": 0,
    "parallel_eval_threshold": 500,
    "precision": "fp64",
}


def force_cpu(on: bool):
    if on:
        raise RuntimeError("stablebear_amd computes pairwise integrals on the GPU only; "
                           "force_cpu(True) is not supported")
    _state["force_cpu"] = False


def _apply_shape():
    from . import runtime

    x, y = _state["block"]
    runtime.set_job_tile_shape(x, y, _state["cap"])


def set_block_size(x: int, y: int):
    """Pair-tile shape of the jobs behind pdist / cdist / l2_kernel (rows x
    cols pairs per workgroup, rows + cols <= 64)."""
    if x < 1 or y < 1 or x + y > 64:
        raise ValueError("block size must satisfy 1 <= x, y and x + y <= 64")
    _state["block"] = (int(x), int(y))
    _apply_shape()


def set_tile_cap(cap: int):
    """Breakpoint records per PCF staged in LDS per phase (0 = default)."""
    if cap != 0 and not 2 <= cap <= 4096:
        raise ValueError("tile cap must be 0 or in [2, 4096]")
    _state["cap"] = int(cap)
    _apply_shape()


def set_output_budget(nbytes: int):
    """Device bytes a pdist / cdist / l2_kernel may give its output at once
    (0 = half the free device memory). Larger outputs are computed in bands
    of rows, each downloaded while the next one computes."""
    from . import runtime

    if nbytes < 0:
        raise ValueError("output budget must be >= 0")
    runtime.set_output_budget(int(nbytes))


def limit_cpus(n: int):
    _state["cpus"] = int(n)


def limit_gpus(n: int):
    if n < 1:
        raise ValueError("need at least one GPU")
    _state["gpus"] = int(n)


def set_cuda_threshold(n: int):
    _state["cuda_threshold"] = int(n)


def set_parallel_eval_threshold(n: int):
    _state["parallel_eval_threshold"] = int(n)


def get_parallel_eval_threshold() -> int:
    return _state["parallel_eval_threshold"]


def set_device_verbose(on: bool):
    _state["device_verbose"] = bool(on)


def set_min_block_side(n: int):
    _state["min_block_side"] = int(n)


def set_precision(precision: str):
    """Default arithmetic form for float64 PCFs: "fp64" (default) or "fp32c"."""
    if precision not in ("fp64", "fp32c"):
        raise ValueError("precision must be 'fp64' or 'fp32c'")
    _state["precision"] = precision


def get_precision() -> str:
    return _state["precision"]


def build_type() -> str:
    return "HIP"


def _verbose(msg: str):
    if _state["device_verbose"]:
        print(msg)

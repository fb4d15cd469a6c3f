"""Save and load of fitted exact models in gprx's on-disk format
(src/persist/mod.rs, FORMAT_VERSION = 1 at src/persist/mod.rs:156).

A model directory holds `config.json` (the ModelConfig of
src/persist/config.rs:19-55, serde's externally tagged enums in snake case)
and `model.safetensors` (src/persist/tensors.rs): the training `x` exactly as
the model takes it (column-major, n points by d features, under shape
[n, d]), `y` [n], and with `save_with_factor` the Cholesky factor `l`
[n, n] (column-major, zeros above the diagonal) and `alpha` [n].

The device model keeps the inverse factor and K^-1 rather than L, so
`save_with_factor` forms L on the device (gprx_hip_exact_export_factor) and
a load always factors again on the device at the saved parameters, jitter
policy and precision: a stored `l` / `alpha` is checked (dtype, shape,
finite values) and not uploaded.  At n = 16384 the factor takes about as long
as a 2 GiB upload of L would.

What this backend's model can hold is what a load accepts: kernels built
from constant, rbf, rbf_ard and periodic leaves with sum and product; identity
transforms; `double` or `mixed` precision (mixed writes its factor as f32,
the FP64 factor rounded); accurate `exp`.  Anything else is refused with
PersistFailed(kind "config") before the device is touched.  An `ldlt` file
(an online model) is loaded by factoring again, as the reference does for
an `ldlt` file without a factor (src/persist/mod.rs:996-1009); its point ids
are not kept, and a save writes `llt`.
"""
from __future__ import annotations

import copy
import ctypes as C
import json
import math
import mmap
import os
import struct
import tempfile
from pathlib import Path

import numpy as np

from . import _lib as L
from .errors import GprError, InvalidConfig, raise_for
from .kernels import ConstantKernel, PeriodicKernel, Product, RbfArdKernel, RbfKernel, Sum

FORMAT_VERSION = 1
CONFIG_FILE = "config.json"
TENSOR_FILE = "model.safetensors"
DEFAULT_BOUNDS = (1e-5, 1e5)


class PersistFailed(GprError):
    """GprError::PersistFailed (src/error.rs); kind is "io", "config" or "tensor"."""

    def __init__(self, kind: str, message: str):
        super().__init__(f"persist ({kind}): {message}")
        self.kind = kind


class UnsupportedPersistVersion(GprError):
    """GprError::UnsupportedPersistVersion (src/error.rs:502)."""

    def __init__(self, found, supported: int = FORMAT_VERSION):
        super().__init__(f"config.json format_version {found} is not {supported}")
        self.found, self.supported = found, supported


# ---------------------------------------------------------------- safetensors

# safetensors' Dtype enum order (the Rust crate's declaration order, which its
# derived Ord follows); a file lists its tensors by descending dtype, then name.
_DTYPE_RANK = {"F32": 0, "F64": 1}
_NP = {"F32": np.dtype("<f4"), "F64": np.dtype("<f8")}


def _dtype_name(a: np.ndarray) -> str:
    if a.dtype == np.float64:
        return "F64"
    if a.dtype == np.float32:
        return "F32"
    raise PersistFailed("tensor", f"dtype {a.dtype} is not f32 or f64")


def safetensors_parts(tensors: dict) -> list:
    """The safetensors file of {name: (array, shape)} as a list of buffers:
    an 8-byte little-endian header length and the compact JSON header padded
    with spaces to a multiple of 8, then each tensor's little-endian values
    in the header's order (descending dtype, then name), as
    src/persist/tensors.rs:321-359 lays it out."""
    items = sorted(tensors.items(), key=lambda kv: (-_DTYPE_RANK[_dtype_name(kv[1][0])], kv[0]))
    header, blobs, off = {}, [], 0
    for name, (arr, shape) in items:
        dt = _dtype_name(arr)
        a = np.ascontiguousarray(arr, dtype=_NP[dt]).reshape(-1)
        if a.nbytes != int(np.prod(shape, dtype=np.int64)) * _NP[dt].itemsize:
            raise PersistFailed("tensor", f"{name} holds {a.size} values, not shape {shape}")
        header[name] = {"dtype": dt, "shape": list(shape), "data_offsets": [off, off + a.nbytes]}
        blobs.append(memoryview(a).cast("B"))
        off += a.nbytes
    js = json.dumps(header, separators=(",", ":")).encode()
    js += b" " * (-len(js) % 8)
    return [struct.pack("<Q", len(js)) + js] + blobs


def safetensors_bytes(tensors: dict) -> bytes:
    """safetensors_parts joined into one buffer."""
    return b"".join(bytes(p) for p in safetensors_parts(tensors))


def read_safetensors(path: Path) -> dict:
    """{name: (dtype, shape, buffer)} of a safetensors file, memory-mapped
    (only the pages a load touches are read), with the checks a load needs
    (header inside the file, offsets in bounds and matching the shapes)."""
    try:
        with open(path, "rb") as f:
            size = os.fstat(f.fileno()).st_size
            if size < 8:
                raise PersistFailed("tensor", f"{path} is too short for a safetensors header")
            data = memoryview(mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ))
    except OSError as e:
        raise PersistFailed("io", f"read {path}: {e}") from None
    (hlen,) = struct.unpack("<Q", data[:8])
    if hlen > len(data) - 8:
        raise PersistFailed("tensor", f"{path}: header length {hlen} passes the end of the file")
    try:
        header = json.loads(bytes(data[8:8 + hlen]).decode())
    except (UnicodeDecodeError, json.JSONDecodeError) as e:
        raise PersistFailed("tensor", f"{path}: header is not JSON: {e}") from None
    if not isinstance(header, dict):
        raise PersistFailed("tensor", f"{path}: header is not a JSON object")
    body = data[8 + hlen:]
    out = {}
    for name, info in header.items():
        if name == "__metadata__":
            continue
        try:
            dt, shape, (a, b) = info["dtype"], [int(v) for v in info["shape"]], info["data_offsets"]
        except (KeyError, TypeError, ValueError):
            raise PersistFailed("tensor", f"{path}: tensor {name} has a malformed entry") from None
        if dt not in _NP:
            raise PersistFailed("tensor", f"{path}: tensor {name} has dtype {dt}, not F32 or F64")
        if not (0 <= a <= b <= len(body)) or b - a != int(np.prod(shape)) * _NP[dt].itemsize:
            raise PersistFailed("tensor", f"{path}: tensor {name} offsets {a}..{b} do not fit")
        out[name] = (dt, shape, body[a:b])
    return out


def _tensor(tensors: dict, name: str, shape, dtypes=("F64",)) -> np.ndarray:
    if name not in tensors:
        raise PersistFailed("tensor", f"model.safetensors has no tensor {name}")
    dt, got, raw = tensors[name]
    if dt not in dtypes:
        raise PersistFailed("tensor", f"tensor {name} is {dt}, expected {' or '.join(dtypes)}")
    if got != list(shape):
        raise PersistFailed("tensor", f"tensor {name} has shape {got}, expected {list(shape)}")
    a = np.frombuffer(raw, dtype=_NP[dt])
    step = 1 << 22  # checked in slices: no full-size temporaries for a factor
    if not all(np.isfinite(a[i:i + step]).all() for i in range(0, a.size, step)):
        raise PersistFailed("tensor", f"tensor {name} holds a non-finite value")
    return a


def _write_atomic(path: Path, parts) -> None:
    """Writes the buffers into a temporary file beside `path` and renames it
    over `path` (src/persist/atomic.rs): a reader never sees half a file, and
    the file is never built whole in memory."""
    fd, tmp = tempfile.mkstemp(prefix=f".{path.name}.", dir=path.parent)
    try:
        with os.fdopen(fd, "wb") as f:
            for part in parts:
                f.write(part)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
    except OSError as e:
        try:
            os.unlink(tmp)
        except OSError:
            pass
        raise PersistFailed("io", f"write {path}: {e}") from None


# ---------------------------------------------------------------- config.json

def _bounded(value: float, lo: float, hi: float) -> dict:
    return {"value": float(value), "lo": float(lo), "hi": float(hi)}


def encode_kernel(kernel, bounds) -> dict:
    """KernelJson (src/persist/kernel.rs:48-105) of a spec; `bounds` yields
    the (lo, hi) of each parameter in get_params order.  An n-ary sum or
    product is written as the left-nested pairs `a + b + c` builds."""
    if isinstance(kernel, (Sum, Product)):
        tag = "sum" if isinstance(kernel, Sum) else "product"
        parts = [encode_kernel(c, bounds) for c in kernel.children]
        out = parts[0]
        for p in parts[1:]:
            out = {tag: {"left": out, "right": p}}
        return out
    if isinstance(kernel, ConstantKernel):
        return {"constant": {"constant": _bounded(math.exp(kernel.log_c), *next(bounds))}}
    if isinstance(kernel, RbfKernel):
        return {"rbf": {"lengthscale": _bounded(math.exp(kernel.log_l), *next(bounds))}}
    if isinstance(kernel, RbfArdKernel):
        return {"rbf_ard": {"lengthscales": [_bounded(math.exp(v), *next(bounds))
                                             for v in kernel.log_l]}}
    if isinstance(kernel, PeriodicKernel):
        return {"periodic": {"lengthscale": _bounded(math.exp(kernel.log_l), *next(bounds)),
                             "period": _bounded(math.exp(kernel.log_p), *next(bounds))}}
    raise PersistFailed("config", f"kernel {type(kernel).__name__} has no persisted form")


def _decode_bounded(b, what: str):
    try:
        v, lo, hi = float(b["value"]), float(b["lo"]), float(b["hi"])
    except (KeyError, TypeError, ValueError):
        raise PersistFailed("config", f"{what} is not a {{value, lo, hi}} object") from None
    if not (math.isfinite(v) and v > 0):
        raise PersistFailed("config", f"{what} value {v} is not positive and finite")
    if not (0 < lo < hi and math.isfinite(hi)):
        raise PersistFailed("config", f"{what} bounds ({lo}, {hi}) are not a positive interval")
    return v, (lo, hi)


def _one_key(obj, what: str):
    if not isinstance(obj, dict) or len(obj) != 1:
        raise PersistFailed("config", f"{what} is not a tagged object")
    return next(iter(obj.items()))


def decode_kernel(obj, bounds: list):
    """The spec of a KernelJson; appends each parameter's (lo, hi) to bounds."""
    tag, body = _one_key(obj, "kernel")
    if tag in ("sum", "product"):
        left = decode_kernel(body.get("left") if isinstance(body, dict) else None, bounds)
        right = decode_kernel(body.get("right") if isinstance(body, dict) else None, bounds)
        return left + right if tag == "sum" else left * right
    if not isinstance(body, dict):
        raise PersistFailed("config", f"kernel {tag} has no fields")
    if tag == "constant":
        v, b = _decode_bounded(body.get("constant"), "constant")
        bounds.append(b)
        return ConstantKernel(v)
    if tag == "rbf":
        v, b = _decode_bounded(body.get("lengthscale"), "rbf lengthscale")
        bounds.append(b)
        return RbfKernel(v)
    if tag == "rbf_ard":
        ls = body.get("lengthscales")
        if not isinstance(ls, list) or not ls:
            raise PersistFailed("config", "rbf_ard lengthscales is not a non-empty list")
        vals = []
        for i, e in enumerate(ls):
            v, b = _decode_bounded(e, f"rbf_ard lengthscale {i}")
            vals.append(v)
            bounds.append(b)
        return RbfArdKernel(vals)
    if tag == "periodic":
        vl, bl = _decode_bounded(body.get("lengthscale"), "periodic lengthscale")
        vp, bp = _decode_bounded(body.get("period"), "periodic period")
        bounds += [bl, bp]
        return PeriodicKernel(vl, vp)
    raise PersistFailed("config", f"kernel {tag} is outside this backend's kernels")


def encode_jitter(policy) -> dict:
    """JitterJson (src/persist/config.rs:243-283)."""
    j = policy.initial
    if policy.multiplier == 1.0 and policy.max_jitter == j and policy.max_retries == (1 if j > 0 else 0):
        return {"fixed": {"jitter": j}}
    return {"adaptive": {"initial": policy.initial, "multiplier": policy.multiplier,
                         "max_retries": policy.max_retries, "max_jitter": policy.max_jitter}}


def decode_jitter(obj):
    from .gpr import JitterPolicy
    tag, body = _one_key(obj, "jitter")
    try:
        # the reference builds the policy through its checked constructors
        # (src/persist/config.rs:243-283, src/policy.rs:245-304)
        if tag == "fixed":
            return JitterPolicy.fixed(float(body["jitter"]))
        if tag == "adaptive":
            ini, mul = float(body["initial"]), float(body["multiplier"])
            ret, mx = int(body["max_retries"]), float(body["max_jitter"])
            return JitterPolicy.adaptive(ini, mul, ret, mx)
    except (KeyError, TypeError, ValueError, InvalidConfig):
        raise PersistFailed("config", f"jitter {tag} has invalid fields") from None
    raise PersistFailed("config", f"jitter {tag} is not fixed or adaptive")


def model_config(fitted, has_factor: bool) -> dict:
    """ModelConfig of a fitted model (src/persist/mod.rs:632-656), in the
    struct's field order; fields at their serde defaults are left out as
    the reference leaves them out."""
    gpr = fitted._gpr
    theta = fitted.get_params()
    kernel = copy.deepcopy(gpr.kernel)
    kernel.set_params(theta[:-1])
    bounds = gpr.bounds.reshape(-1, 2) if gpr.bounds is not None else \
        np.tile(DEFAULT_BOUNDS, (theta.size, 1))
    it = iter([tuple(b) for b in bounds])
    cfg = {"format_version": FORMAT_VERSION, "n": fitted.n, "d": fitted.d,
           "has_factor": has_factor, "factor_kind": "llt"}
    if gpr.precision == L.MIXED:
        cfg["precision"] = "mixed"
        # The mixed form's refinement residuals use the kernel evaluated again
        # in f64, the reference's ReevaluateKernel (src/persist/config.rs:171-178).
        cfg["residual"] = "reevaluate_kernel"
    cfg["kernel"] = encode_kernel(kernel, it)
    cfg["likelihood"] = {"noise_variance": _bounded(math.exp(theta[-1]), *next(it))}
    cfg["jitter"] = encode_jitter(gpr.jitter)
    jit = fitted.last_jitter()
    if jit != 0.0:
        cfg["factor_jitter"] = jit
    cfg["distance_cache"] = "always"
    cfg["x_unfitted"] = "identity"
    cfg["y_unfitted"] = "identity"
    cfg["x_transform"] = "identity"
    cfg["y_transform"] = "identity"
    return cfg


# serde_json's default nesting limit, past which the reference refuses a
# config (src/persist/mod.rs:1370-1420)
MAX_JSON_DEPTH = 128


def _json_depth(obj) -> int:
    """Deepest nesting of objects and arrays in a parsed JSON value."""
    deepest, stack = 0, [(obj, 1)]
    while stack:
        v, depth = stack.pop()
        if isinstance(v, dict):
            items = v.values()
        elif isinstance(v, list):
            items = v
        else:
            continue
        deepest = max(deepest, depth)
        stack.extend((c, depth + 1) for c in items)
    return deepest


# The loader of each model kind (src/persist/config.rs:311-336).
_LOADERS = {"exact": "LoadedGpr::load", "sgpr": "LoadedSgpr::load",
            "online_sgpr": "LoadedSgpr::load", "svgp": "LoadedSvgp::load"}


def _expect_model(cfg: dict, expected) -> str:
    """The config's model kind (absent means exact, src/persist/config.rs:311-317),
    or PersistFailed kind "wrong model" naming the loader it needs."""
    kind = cfg.get("model", "exact")
    if kind not in _LOADERS:
        raise PersistFailed("config", f"model {kind!r} is not exact, sgpr, online_sgpr or svgp")
    if kind not in expected:
        raise PersistFailed("wrong model",
                            f"config.json holds a {kind} model; load it with {_LOADERS[kind]}")
    return kind


def _check_config(cfg: dict):
    """Validates what this backend can hold; returns (n, d, has_factor)."""
    if not isinstance(cfg, dict):
        raise PersistFailed("config", "config.json is not a JSON object")
    if cfg.get("format_version") != FORMAT_VERSION:
        raise UnsupportedPersistVersion(cfg.get("format_version"))
    _expect_model(cfg, ("exact",))
    try:
        n, d, has_factor = int(cfg["n"]), int(cfg["d"]), bool(cfg["has_factor"])
        kind = cfg["factor_kind"]
    except (KeyError, TypeError, ValueError):
        raise PersistFailed("config", "config.json lacks n, d, has_factor or factor_kind") from None
    if n < 1 or d < 1:
        raise PersistFailed("config", f"n = {n} and d = {d} must be positive")
    if kind not in ("llt", "ldlt"):
        raise PersistFailed("config", f"factor_kind {kind!r} is not llt or ldlt")
    if cfg.get("precision", "double") not in ("double", "mixed"):
        raise PersistFailed("config", f"precision {cfg.get('precision')!r} is not double or mixed")
    # Either residual loads: the model is factored again in its own form.  The
    # field is read only for mixed precision (src/persist/config.rs:130-140),
    # but must still be one of its two values.
    if cfg.get("residual", "promote_storage") not in ("promote_storage", "reevaluate_kernel"):
        raise PersistFailed("config", f"residual {cfg.get('residual')!r} is not promote_storage "
                                      "or reevaluate_kernel")
    if cfg.get("math", "accurate") != "accurate":
        raise PersistFailed("config", f"math {cfg.get('math')!r} is not accurate")
    if cfg.get("distance") is not None:
        raise PersistFailed("config", "supplied-distance kernels are outside this backend")
    for key in ("x_unfitted", "y_unfitted", "x_transform", "y_transform"):
        if cfg.get(key) != "identity":
            raise PersistFailed("config", f"{key} {cfg.get(key)!r} is not identity")
    return n, d, has_factor


# ---------------------------------------------------------------- save / load

def save_fitted(fitted, directory, with_factor: bool) -> None:
    """save_fitted (src/persist/mod.rs:681-715): the tensors, then config.json."""
    dirp = Path(directory)
    try:
        dirp.mkdir(parents=True, exist_ok=True)
    except OSError as e:
        raise PersistFailed("io", f"create {dirp}: {e}") from None
    n, d = fitted.n, fitted.d
    tensors = {"x": (fitted._x.reshape(-1), (n, d)), "y": (fitted._y, (n,))}
    if with_factor:
        l = np.empty(n * n)
        alpha = np.empty(n)
        jit = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_exact_export_factor(fitted._h, L.dptr(l), n, L.dptr(alpha),
                                                   C.byref(jit), C.byref(err))
        raise_for(rc, err)
        if fitted._gpr.precision == L.MIXED:
            l = l.astype(np.float32)
        tensors["l"] = (l, (n, n))
        tensors["alpha"] = (alpha, (n,))
    cfg = model_config(fitted, with_factor)
    _write_atomic(dirp / TENSOR_FILE, safetensors_parts(tensors))
    _write_atomic(dirp / CONFIG_FILE, [json.dumps(cfg, indent=2).encode()])


def read_model(directory):
    """(gpr, x, y, n, d) of a model directory, every check done on the host."""
    from .gpr import GaussianLikelihood, Gpr
    dirp = Path(directory)
    try:
        raw = (dirp / CONFIG_FILE).read_bytes()
    except OSError as e:
        raise PersistFailed("io", f"read {dirp / CONFIG_FILE}: {e}") from None
    try:
        cfg = json.loads(raw)
    except (UnicodeDecodeError, json.JSONDecodeError, RecursionError) as e:
        raise PersistFailed("config", f"config.json is not JSON: {e}") from None
    if _json_depth(cfg) > MAX_JSON_DEPTH:
        raise PersistFailed("config", f"config.json nests past the recursion limit of "
                                      f"{MAX_JSON_DEPTH}")
    n, d, has_factor = _check_config(cfg)
    bounds: list = []
    kernel = decode_kernel(cfg.get("kernel"), bounds)
    lik = cfg.get("likelihood")
    noise, nb = _decode_bounded(lik.get("noise_variance") if isinstance(lik, dict) else None,
                                "noise_variance")
    bounds.append(nb)
    gpr = Gpr(kernel, GaussianLikelihood(noise)).with_jitter_policy(decode_jitter(cfg.get("jitter")))
    if cfg.get("precision", "double") == "mixed":
        gpr = gpr.with_precision(L.MIXED)
    if any(tuple(b) != DEFAULT_BOUNDS for b in bounds):
        gpr = gpr.with_bounds(np.asarray(bounds, dtype=np.float64).reshape(-1))
    tensors = read_safetensors(dirp / TENSOR_FILE)
    x = _tensor(tensors, "x", (n, d)).copy()
    y = _tensor(tensors, "y", (n,)).copy()
    if has_factor:
        _tensor(tensors, "l", (n, n), ("F64", "F32"))
        _tensor(tensors, "alpha", (n,), ("F64", "F32"))
    return gpr, x, y, n, d


def load(directory):
    """LoadedGpr::load (src/persist/mod.rs:308-310) for an exact model: reads
    and checks the directory on the host, then factors on the device at the
    saved parameters with the saved jitter policy and precision."""
    gpr, x, y, n, d = read_model(directory)
    return gpr.factor(x, n, d, y)


# ---------------------------------------------------------------- sparse models
#
# Sgpr and Svgp directories (src/persist/sparse.rs): config.json is the
# SparseConfig of src/persist/config.rs:534-565 (the exact model's fields with
# `model` and `m`, without the factor); model.safetensors holds x [n, d],
# y [n], z [m, d] and z_train [m, d] (the inducing points in training
# coordinates: z itself under identity transforms), and for Svgp the
# whitened q(u) as q_mean [m] and q_l [m, m] (column-major, zeros above the
# diagonal).  No factor is stored: a load factors again at the saved
# parameters and inducing points.  K_mm is factored with the sparse default
# jitter policy adaptive(1e-8, 10, 5, 1e-3) (src/policy.rs:220-229), the one
# this backend's sparse models use, so that policy is written and the only
# one a load accepts.

def _kmm_jitter():
    from .gpr import JitterPolicy
    return JitterPolicy.adaptive(1e-8, 10.0, 5, 1e-3)


def sparse_config(kind: str, kernel, theta, n: int, m: int, d: int, mixed: bool) -> dict:
    """SparseConfig of a sparse model at log-parameters theta ([kernel | log
    noise]), fields in the struct's order, serde defaults left out."""
    kernel = copy.deepcopy(kernel)
    kernel.set_params(theta[:-1])
    it = iter([DEFAULT_BOUNDS] * len(theta))
    cfg = {"format_version": FORMAT_VERSION, "model": kind, "n": n, "m": m, "d": d}
    if mixed:
        cfg["precision"] = "mixed"
    cfg["kernel"] = encode_kernel(kernel, it)
    cfg["likelihood"] = {"noise_variance": _bounded(math.exp(theta[-1]), *next(it))}
    cfg["jitter"] = encode_jitter(_kmm_jitter())
    cfg["x_unfitted"] = "identity"
    cfg["y_unfitted"] = "identity"
    cfg["x_transform"] = "identity"
    cfg["y_transform"] = "identity"
    return cfg


def save_sparse(kind: str, kernel, theta, x, y, z, n: int, m: int, d: int, mixed: bool, directory,
                q=None) -> None:
    """write_sparse (src/persist/sparse.rs:95-163): the tensors, then
    config.json, each written atomically.  q = (q_mean [m], q_l packed by
    columns, rows i >= j) for Svgp."""
    dirp = Path(directory)
    try:
        dirp.mkdir(parents=True, exist_ok=True)
    except OSError as e:
        raise PersistFailed("io", f"create {dirp}: {e}") from None
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    z = np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
    tensors = {"x": (x, (n, d)), "y": (np.ascontiguousarray(y, dtype=np.float64), (n,)),
               "z": (z, (m, d)), "z_train": (z, (m, d))}
    if q is not None:
        q_mean, packed = q
        q_l = np.zeros((m, m))  # [col, row] here, so the bytes are column-major
        k = 0
        for j in range(m):
            q_l[j, j:] = packed[k:k + m - j]
            k += m - j
        tensors["q_mean"] = (np.ascontiguousarray(q_mean, dtype=np.float64), (m,))
        tensors["q_l"] = (q_l.reshape(-1), (m, m))
    cfg = sparse_config(kind, kernel, np.asarray(theta, dtype=np.float64), n, m, d, mixed)
    _write_atomic(dirp / TENSOR_FILE, safetensors_parts(tensors))
    _write_atomic(dirp / CONFIG_FILE, [json.dumps(cfg, indent=2).encode()])


def read_sparse(directory, kind: str):
    """(kernel, noise, mixed, x, y, z, n, m, d, q) of a sparse directory of
    model `kind`, every check done on the host; q = (q_mean, packed q_l) for
    Svgp, None for Sgpr."""
    dirp = Path(directory)
    try:
        raw = (dirp / CONFIG_FILE).read_bytes()
    except OSError as e:
        raise PersistFailed("io", f"read {dirp / CONFIG_FILE}: {e}") from None
    try:
        cfg = json.loads(raw)
    except (UnicodeDecodeError, json.JSONDecodeError, RecursionError) as e:
        raise PersistFailed("config", f"config.json is not JSON: {e}") from None
    if not isinstance(cfg, dict):
        raise PersistFailed("config", "config.json is not a JSON object")
    if _json_depth(cfg) > MAX_JSON_DEPTH:
        raise PersistFailed("config", f"config.json nests past the recursion limit of "
                                      f"{MAX_JSON_DEPTH}")
    if cfg.get("format_version") != FORMAT_VERSION:
        raise UnsupportedPersistVersion(cfg.get("format_version"))
    found = _expect_model(cfg, ("sgpr", "online_sgpr") if kind == "sgpr" else ("svgp",))
    if found == "online_sgpr":
        raise PersistFailed("config", "online SGPR models are outside this backend")
    try:
        n, m, d = int(cfg["n"]), int(cfg["m"]), int(cfg["d"])
    except (KeyError, TypeError, ValueError):
        raise PersistFailed("config", "config.json lacks n, m or d") from None
    if n < 1 or m < 1 or d < 1:
        raise PersistFailed("config", f"n = {n}, m = {m} and d = {d} must be positive")
    if cfg.get("precision", "double") not in ("double", "mixed"):
        raise PersistFailed("config", f"precision {cfg.get('precision')!r} is not double or mixed")
    if cfg.get("residual", "promote_storage") not in ("promote_storage", "reevaluate_kernel"):
        raise PersistFailed("config", f"residual {cfg.get('residual')!r} is not promote_storage "
                                      "or reevaluate_kernel")
    if cfg.get("math", "accurate") != "accurate":
        raise PersistFailed("config", f"math {cfg.get('math')!r} is not accurate")
    if cfg.get("distance") is not None:
        raise PersistFailed("config", "supplied-distance kernels are outside this backend")
    for key in ("x_unfitted", "y_unfitted", "x_transform", "y_transform"):
        if cfg.get(key) != "identity":
            raise PersistFailed("config", f"{key} {cfg.get(key)!r} is not identity")
    if decode_jitter(cfg.get("jitter")) != _kmm_jitter():
        raise PersistFailed("config", "this backend factors K_mm with the sparse default policy "
                                      "adaptive(1e-8, 10, 5, 1e-3) only")
    bounds: list = []
    kernel = decode_kernel(cfg.get("kernel"), bounds)
    lik = cfg.get("likelihood")
    noise, nb = _decode_bounded(lik.get("noise_variance") if isinstance(lik, dict) else None,
                                "noise_variance")
    bounds.append(nb)
    if any(tuple(b) != DEFAULT_BOUNDS for b in bounds):
        raise PersistFailed("config", "sparse models here take the default (1e-5, 1e5) bounds")
    tensors = read_safetensors(dirp / TENSOR_FILE)
    x = _tensor(tensors, "x", (n, d)).copy()
    y = _tensor(tensors, "y", (n,)).copy()
    z = _tensor(tensors, "z", (m, d)).copy()
    z_train = _tensor(tensors, "z_train", (m, d))
    if not np.array_equal(z_train, z):
        raise PersistFailed("config", "z_train differs from z, which identity transforms make equal")
    q = None
    if kind == "svgp":
        q_mean = _tensor(tensors, "q_mean", (m,)).copy()
        q_l = _tensor(tensors, "q_l", (m, m)).reshape(m, m)  # [col, row]
        if np.any(np.tril(q_l, -1)) or not np.all(np.diagonal(q_l) > 0.0):  # row < col
            raise PersistFailed("tensor", "q_l must be lower triangular with a positive diagonal, "
                                          "and q finite")
        packed = np.concatenate([q_l[j, j:] for j in range(m)])
        q = (q_mean, packed)
    mixed = cfg.get("precision", "double") == "mixed"
    return kernel, noise, mixed, x, y, z, n, m, d, q

"""Spectral helpers and line-search tests (mirror of the hot-path parts of pypty/utils.py).

``fourier_clean`` (utils.py:102-150) band-limits whole object planes with the
plane FFT kernels (``pty_fourier_clean_c64``); ``create_spatial_frequencies``
(utils.py:153-199) builds the q grids and band mask on the host (setup work
done once per run; the kernels evaluate the same grids themselves);
``wolfe_1`` / ``wolfe_2`` (utils.py:326-368) compare host scalars;
``complex_grad_to_phase_grad`` (utils.py:260-276) runs ``pty_phase_grad_c64``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _arrays, _lib


def _planar(array: torch.Tensor):
    """(Ny, Nx, ...) -> contiguous (planes, Ny, Nx) and the inverse mapping."""
    nd = array.dim()
    moved = torch.movedim(array, (0, 1), (nd - 2, nd - 1))
    shape = moved.shape
    flat = moved.reshape(-1, shape[-2], shape[-1]).contiguous()

    def back(t):
        return torch.movedim(t.reshape(shape), (nd - 2, nd - 1), (0, 1))

    return flat, back


def fourier_clean_planar(planes: torch.Tensor, cutoff: float, rolloff: float = 0.0) -> torch.Tensor:
    """In-place band-limit of contiguous (planes, Ny, Nx) complex64 data."""
    _lib.require_device_tensor(planes, "planes")
    if planes.dtype != torch.complex64:
        raise TypeError("pypty_amd.fourier_clean: complex64 required")
    P, ny, nx = planes.shape
    rc = _lib.load().pty_fourier_clean_c64(planes.data_ptr(), P, ny, nx, float(cutoff), float(rolloff),
                                           _lib.stream_ptr(planes.device))
    _lib.check(rc, "fourier_clean")
    return planes


def fourier_clean(array, cutoff=0.66, mask=None, rolloff=0, default_float=None, xp=None):
    """Band-limit over axes (0, 1) (pypty/utils.py:102-150): ifft2_ishift(mask * shift_fft2(array)).

    ``array`` (Ny, Nx[, ...]) is a complex64 torch tensor on a HIP device, a
    numpy array or another library's device array (returned as the same type);
    ``mask`` (DC-centred, as there) replaces the cutoff circle.  Returns a new
    array like the reference.
    """
    if cutoff is None and mask is None:
        return array
    kind = _arrays.kind_of(array)
    if kind != _arrays.TORCH:
        # numpy (copied to the device) or another library's device array (DLPack), returned as such
        out = fourier_clean(_arrays.to_torch(array, dtype=torch.complex64), cutoff, mask, rolloff)
        return _arrays.like(out, kind, array)
    if mask is not None:
        # an explicit mask in the DC-centred layout (utils.py:142-146)
        from . import fft as pfft

        m = _arrays.to_torch(mask, array.device, torch.float32)
        m = m.reshape(m.shape + (1,) * (array.dim() - m.dim()))
        spec = pfft.shift_fft2(array.contiguous(), axes=(0, 1)) * m
        return pfft.ifft2_ishift(spec.contiguous(), axes=(0, 1))
    flat, back = _planar(array)
    if flat.data_ptr() == array.data_ptr():
        flat = flat.clone()
    fourier_clean_planar(flat, float(cutoff), float(rolloff))
    return back(flat).contiguous()


def create_spatial_frequencies(px, py, shape, damping_cutoff_multislice, smooth_rolloff, default_float=np.float32,
                               xp=None):
    """q2, qx, qy, exclude_mask (fftshifted), exclude_mask_ishift as float tensors on the current
    HIP device (pypty/utils.py:153-199)."""
    fd = np.dtype(default_float).type if default_float is not None else np.float32
    qx, qy = np.meshgrid(np.fft.fftfreq(shape, px), np.fft.fftfreq(shape, py), indexing="xy")
    qx, qy = qx.astype(fd), qy.astype(fd)
    q2 = qx ** 2 + qy ** 2
    c = damping_cutoff_multislice
    if c is None:
        c = 2
    if c < 0:
        c = 2 / 3
    max_x = np.max(qx)
    r_max = (max_x * fd(c)) ** 2
    m = (q2 <= r_max).astype(fd)
    if smooth_rolloff != 0:
        r0 = (max_x * (c - smooth_rolloff)) ** 2
        sel = q2 > r0
        m[sel] *= (0.5 * (1 + np.cos(3.141592654 * (q2 - r0) / (r_max - r0))))[sel]
    m = m.astype(fd)
    dev = torch.device("cuda", torch.cuda.current_device())
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)  # noqa: E731
    return t(q2), t(qx), t(qy), t(np.fft.fftshift(m)), t(m)


def get_window(shape, r0, r_max, inverted=True, xp=None):
    """Cosine-tapered square window (pypty/utils.py:656-690), host setup: 1 inside |x|, |y| <= r0
    (radii relative to the half-width), falling to 0 at r_max; inverted (the default) gives 1 - that,
    the region the real-space probe constraint penalises."""
    x = np.arange(shape, dtype=np.float64)
    x = x - np.mean(x)
    r_max = r_max * np.max(x)
    r0 = r0 * np.max(x)
    x, y = np.meshgrid(x, x, indexing="xy")
    mask = np.ones_like(x)
    if r0 < r_max:
        for a in (np.abs(x), np.abs(y)):
            sel = a > r0
            mask[sel] *= (0.5 * (1 + np.cos(3.141592654 * (a - r0) / (r_max - r0))))[sel]
    mask[np.abs(x) > r_max] = 0
    mask[np.abs(y) > r_max] = 0
    return 1 - mask if inverted else mask


def complex_grad_to_phase_grad(grad, array, xp=None):
    """2 Re(-i conj(array) grad) (pypty/utils.py:260-276); numpy or another library's device arrays
    come back as the same type."""
    kind = _arrays.kind_of(grad)
    if kind != _arrays.TORCH:
        out = complex_grad_to_phase_grad(_arrays.to_torch(grad, dtype=torch.complex64).contiguous(),
                                         _arrays.to_torch(array, dtype=torch.complex64).contiguous())
        return _arrays.like(out, kind, grad)
    for t, name in ((grad, "grad"), (array, "array")):
        _lib.require_device_tensor(t, name)
    if grad.shape != array.shape or grad.dtype != torch.complex64 or array.dtype != torch.complex64:
        raise TypeError("pypty_amd.complex_grad_to_phase_grad: matching complex64 tensors required")
    out = torch.empty(grad.shape, dtype=torch.float32, device=grad.device)
    rc = _lib.load().pty_phase_grad_c64(grad.data_ptr(), array.data_ptr(), out.data_ptr(), grad.numel(),
                                        _lib.stream_ptr(grad.device))
    _lib.check(rc, "phase_grad")
    return out


def wolfe_1(value, new_value, d_value, step, wolfe_c1=0.5):
    """Armijo condition (pypty/utils.py:326-348)."""
    return new_value <= value + wolfe_c1 * step * d_value


def wolfe_2(d_value, new_d_value, wolfe_c2=0.9):
    """Curvature condition (pypty/utils.py:350-368)."""
    return -1 * new_d_value <= -1 * d_value * wolfe_c2


def try_to_initialize_beam_current(beam_current, measured_data_shape, default_float=np.float32, xp=None):
    """A beam current of ones, one per pattern, when there is none (pypty/utils.py:1766-1789).  A
    shorter one is padded with ones by the pattern count minus its length on each side, as numpy's
    pad with a scalar width does there, so its values move up by that count.  Returns a device
    float32 tensor (the input's array type is not kept: bfgs_update holds it on the device)."""
    n = int(measured_data_shape[0])
    if beam_current is None:
        return torch.ones(n, dtype=torch.float32, device=torch.device("cuda", torch.cuda.current_device()))
    bc = _arrays.to_torch(beam_current, dtype=torch.float32)
    if bc.device.type != "cuda":
        bc = bc.to(torch.device("cuda", torch.cuda.current_device()))
    bc = bc.reshape(-1)
    if bc.shape[0] < n:
        w = n - bc.shape[0]
        bc = torch.nn.functional.pad(bc, (w, w), value=1.0)
    return bc.contiguous()


_SAFE_FUNCS = {"abs": abs, "min": min, "max": max, "int": int, "float": float, "round": round, "bool": bool}


def string_to_lambda(lambda_string):
    """A "lambda-type" parameter stored as text (pypty/utils.py:2366-2382; params.pkl keeps lambdas as
    source, utils.py:2315-2364) turned back into a callable without eval: the text must be a
    lambda of numeric arguments whose body uses only numbers, its arguments, arithmetic, comparison,
    boolean and conditional expressions, tuples and lists, and abs/min/max/int/float/round/bool.
    Text that is not a lambda is returned unchanged, as the reference returns what eval() rejects; a
    lambda outside that set raises ValueError (pass a callable instead)."""
    import ast
    import operator as op

    if not isinstance(lambda_string, str) or "lambda" not in lambda_string:
        return lambda_string
    try:
        tree = ast.parse(lambda_string.strip(), mode="eval")
    except SyntaxError:
        return lambda_string
    if not isinstance(tree.body, ast.Lambda):
        return lambda_string
    lam = tree.body
    names = [a.arg for a in lam.args.args]
    if lam.args.vararg or lam.args.kwarg or lam.args.kwonlyargs or lam.args.defaults:
        raise ValueError(f"pypty_amd: unsupported lambda signature in {lambda_string!r}")
    bin_ops = {ast.Add: op.add, ast.Sub: op.sub, ast.Mult: op.mul, ast.Div: op.truediv, ast.FloorDiv: op.floordiv,
               ast.Mod: op.mod, ast.Pow: op.pow}
    cmp_ops = {ast.Eq: op.eq, ast.NotEq: op.ne, ast.Lt: op.lt, ast.LtE: op.le, ast.Gt: op.gt, ast.GtE: op.ge}

    def ev(n, env):
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float, bool, type(None))):
            return n.value
        if isinstance(n, ast.Name) and n.id in env:
            return env[n.id]
        if isinstance(n, ast.BinOp) and type(n.op) in bin_ops:
            return bin_ops[type(n.op)](ev(n.left, env), ev(n.right, env))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd, ast.Not)):
            v = ev(n.operand, env)
            return -v if isinstance(n.op, ast.USub) else (+v if isinstance(n.op, ast.UAdd) else not v)
        if isinstance(n, ast.Compare) and all(type(o) in cmp_ops for o in n.ops):
            left = ev(n.left, env)
            for o, c in zip(n.ops, n.comparators):
                right = ev(c, env)
                if not cmp_ops[type(o)](left, right):
                    return False
                left = right
            return True
        if isinstance(n, ast.BoolOp):
            vals = (ev(v, env) for v in n.values)
            return all(vals) if isinstance(n.op, ast.And) else any(vals)
        if isinstance(n, ast.IfExp):
            return ev(n.body, env) if ev(n.test, env) else ev(n.orelse, env)
        if isinstance(n, (ast.Tuple, ast.List)):
            vals = [ev(v, env) for v in n.elts]
            return tuple(vals) if isinstance(n, ast.Tuple) else vals
        if (isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _SAFE_FUNCS
                and not n.keywords):
            return _SAFE_FUNCS[n.func.id](*[ev(a, env) for a in n.args])
        raise ValueError(f"pypty_amd: {lambda_string!r} uses {type(n).__name__}, which is not evaluated; "
                         "pass a callable")

    ev_check = lam.body

    def f(*args):
        if len(args) != len(names):
            raise TypeError(f"{lambda_string!r} takes {len(names)} argument(s)")
        return ev(ev_check, dict(zip(names, args)))

    # reject unsupported constructs now rather than at the first call
    allowed = (ast.BinOp, ast.UnaryOp, ast.Compare, ast.BoolOp, ast.IfExp, ast.Tuple, ast.List, ast.Constant,
               ast.Name, ast.Call, ast.Load, ast.operator, ast.unaryop, ast.cmpop, ast.boolop)
    for node in ast.walk(lam.body):
        bad = not isinstance(node, allowed)
        bad = bad or (isinstance(node, ast.Name) and node.id not in names and node.id not in _SAFE_FUNCS)
        bad = bad or (isinstance(node, ast.Call) and (not isinstance(node.func, ast.Name) or node.keywords
                                                      or node.func.id not in _SAFE_FUNCS))
        bad = bad or (isinstance(node, ast.Constant) and not isinstance(node.value, (int, float, bool, type(None))))
        if bad:
            raise ValueError(f"pypty_amd: {lambda_string!r} uses {type(node).__name__}, which is not evaluated; "
                             "pass a callable")
    f.__doc__ = lambda_string
    return f

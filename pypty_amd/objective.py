"""Loss and gradient over all scan positions (mirror of pypty/objective.py:47-967).

``loss_and_direction`` keeps the reference's 77-argument signature and
10-tuple result.  The far-field multislice path with the lsq_sqrt / ml / lsq /
lsq_sqrt_2 losses, the lsq / gauss / poisson compressed losses (virtual
detectors, objective.py:560-585) and object, probe and position gradients
runs entirely in the gfx950 kernels behind ``pty_loss_and_direction``,
including the per-batch data preprocessing (shift, bin, multiplier, pad;
utils.py:587-653), the static background and its gradient
(objective.py:359-361, 525-526, 588-589) and, with ``load_one_by_one`` and
patterns in host memory, streaming of each batch's patterns through pinned
buffers (objective.py:383-408), the per-position beam current and its
gradient (objective.py:446-450, 688-698), and the per-position aberration
phase plates and their gradient (objective.py:320-326, 459-465, 673-681).
Beam tilts before, inside and after the specimen with their gradients
(objective.py:303-307, 328-350, 428-499, 600-672; multislice_core.py:163-178)
and per-slice distances (multislice_core.py:69-71, 164-168) run in the same
kernels.  The constraint terms (objective.py:765-921) are added by
``pypty_amd.constraints`` on the device.  Arguments that select features
outside that path (near field, phase plates from a file, probe markers,
upsampling, tilts before the specimen together with aberrations) raise
``NotImplementedError`` instead of silently computing something else.

Binning follows the evident intent of utils.py:639-644 (sum of the
data_bin x data_bin blocks); the reference's loop indexes an undefined name
there and fails for data_bin > 1.

``PtychoPass`` is the allocation-owning engine under it: it caches the
device workspace per problem shape and stream so repeated passes
(line-search trials, epochs, benchmark steps) allocate nothing.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _arrays, _lib, constraints
from . import fft as pfft
from . import utils

LOSSES = {"lsq_sqrt": _lib.PTY_LOSS_LSQ_SQRT, "ml": _lib.PTY_LOSS_ML, "lsq": _lib.PTY_LOSS_LSQ,
          "lsq_sqrt_2": _lib.PTY_LOSS_LSQ_SQRT_2, "lsq_compressed": _lib.PTY_LOSS_LSQ_COMPRESSED,
          "gauss_compressed": _lib.PTY_LOSS_GAUSS_COMPRESSED, "poisson_compressed": _lib.PTY_LOSS_POISSON_COMPRESSED}


@dataclass(frozen=True)
class PassConfig:
    """Static description of one loss/gradient evaluation (the pty_problem struct)."""

    n: int
    n_slices: int
    n_obj_modes: int
    n_probe_modes: int
    obj_ny: int
    obj_nx: int
    n_positions: int
    compute_batch: int
    algorithm: str = "lsq_sqrt"
    planar: bool = False
    step_obj: bool = True
    step_probe: bool = True
    step_pos: bool = False
    shift_active: bool = False
    loss_weight: float = 1.0
    slice_distance: float = 1.0
    wavelength: float = 0.0196868
    pixel_size_x: float = 1.0
    pixel_size_y: float = 1.0
    tan_x: float = 0.0
    tan_y: float = 0.0
    damping_cutoff: float = 2.0 / 3.0
    edge_multiplier: float = 1.0
    n_masks: int = 0
    # per-batch preprocessing of raw patterns (pypty/utils.py:587-653)
    raw_n: int = 0
    data_shift_y: int = 0
    data_shift_x: int = 0
    data_bin: int = 1
    data_pad: int = 0
    data_multiplier: float = 1.0
    # keep the wave entering every checkpoint_slices-th slice and recompute the
    # others in the adjoint (0: keep every slice); the results are bit-identical
    checkpoint_slices: int = 0
    # batch i runs on stream i % lanes of up to four (bit-identical results)
    lanes: int = 4
    # also return the gradient of the static background (objective.py:588-589)
    step_static_background: bool = False
    # also return the gradient of the beam current (objective.py:688-698)
    step_beam_current: bool = False
    # aberration rows of n_aberrations coefficients (objective.py:320-326), 0 for none, and
    # whether to return their gradient (objective.py:673-681)
    n_aberrations: int = 0
    n_markers: int = 0
    step_aberrations: bool = False
    # per-slice distances (multislice_core.py:69-71): more than one means one per propagation
    n_distances: int = 0
    # tilt rows (1 or one per position) of the tilt array (objective.py:428-434), the reference's
    # tilt_mode, and whether to return their gradient
    tilt_mode: int = 0
    n_tilts: int = 0
    step_tilts: bool = False

    @property
    def compressed(self) -> bool:
        return "compressed" in self.algorithm

    def to_c(self) -> _lib.PtyProblem:
        if self.algorithm not in LOSSES:
            raise NotImplementedError(f"loss '{self.algorithm}' is not on the accelerated path")
        return _lib.PtyProblem(
            self.n, self.n_slices, self.n_obj_modes, self.n_probe_modes, self.obj_ny, self.obj_nx,
            self.n_positions, self.compute_batch, LOSSES[self.algorithm],
            _lib.PTY_LAYOUT_PLANAR if self.planar else _lib.PTY_LAYOUT_REFERENCE,
            int(bool(self.step_obj)), int(bool(self.step_probe)), int(bool(self.step_pos)),
            int(bool(self.shift_active)), float(self.loss_weight), float(self.slice_distance),
            float(self.wavelength), float(self.pixel_size_x), float(self.pixel_size_y), float(self.tan_x),
            float(self.tan_y), float(self.damping_cutoff), float(self.edge_multiplier), int(self.n_masks),
            int(self.raw_n), int(self.data_shift_y), int(self.data_shift_x), int(self.data_bin), int(self.data_pad),
            float(self.data_multiplier), int(self.checkpoint_slices), int(self.lanes),
            int(bool(self.step_static_background)), int(bool(self.step_beam_current)), int(self.n_aberrations),
            int(self.n_markers), int(bool(self.step_aberrations)), int(self.n_distances), int(self.tilt_mode),
            int(self.n_tilts), int(bool(self.step_tilts)))


class PtychoPass:
    """One forward + loss + adjoint pass over a sequence of positions, on one HIP device.

    Workspaces are cached per (device, size, stream) for the life of the process, so
    repeated passes of the same problem shapes allocate nothing.  Every run takes the
    workspace of the stream it is issued on: passes that share a workspace are issued on
    one stream, whose order keeps them apart, and passes on different streams never share
    one, whichever engine issues them.  ``release_workspaces()`` drops the cache.
    """

    _workspaces: dict = {}

    @classmethod
    def release_workspaces(cls):
        cls._workspaces.clear()

    def __init__(self, cfg: PassConfig, device=None):
        self.cfg = cfg
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.lib = _lib.load()
        self.c_problem = cfg.to_c()
        nbytes = ctypes.c_size_t(0)
        _lib.check(self.lib.pty_workspace_bytes(ctypes.byref(self.c_problem), ctypes.byref(nbytes)), "workspace")
        self.workspace_bytes = int(nbytes.value)
        self.workspace = None  # the workspace of the last run
        self.loss_sse = torch.zeros(2, dtype=torch.float64, device=self.device)

    def run(self, obj, probe_mnn, positions, corrections, measured, sequence, *, obj_grad=None, probe_grad=None,
            pos_grad=None, sim_patterns=None, sync=True, masks=None, measured_by_sequence=False, accumulate=False,
            loss_sse=None, static_background=None, static_background_grad=None, beam_current=None,
            beam_current_grad=None, aberrations=None, aberration_marker=None, aberrations_grad=None, distances=None,
            tilts=None, tilts_grad=None):
        """Launch the pass.  probe_mnn is mode-major (M, n, n).  Returns (loss, sse) as floats
        when ``sync`` else the device tensor [loss, sse].

        measured_by_sequence: row i of ``measured`` holds the pattern of sequence[i] (a streamed batch);
        accumulate: add to the gradients of the previous call on this workspace; loss_sse: a device
        float64 tensor of 2 elements to receive this call's sums (default: the engine's own);
        static_background: (n, n) float32, zero frequency at (n/2, n/2), added squared and band-limited
        to the modelled patterns (objective.py:359-361, 525-526); static_background_grad: (n, n)
        float32 receiving its gradient when the config steps it (objective.py:588-589);
        beam_current: (P,) float32, each position's probe taken |beam_current| times before its shift
        (objective.py:446-450); beam_current_grad: (P,) float32 receiving its gradient on the rows of
        the sequence when the config steps it (objective.py:688-698); aberrations: (n_markers,
        n_aberrations) float32 with aberration_marker (P,) int32, the row of each position, whose
        phase plate multiplies its probe spectrum (objective.py:320-326, 459-465; a row outside
        [0, n_markers) acts as none); aberrations_grad: (n_markers, n_aberrations) float32 receiving
        their gradient when the config steps them (objective.py:673-681); distances: (n_distances,)
        float32 per-slice distances when the config has more than one (multislice_core.py:69-71);
        tilts: (n_tilts, 6) float32, the tilt array plus its correction (objective.py:350, 428-434);
        tilts_grad: (n_tilts, 6) float32 receiving their gradient when the config steps them."""
        c = self.cfg
        for t, name in ((obj, "obj"), (probe_mnn, "probe"), (positions, "positions"), (measured, "measured"),
                        (sequence, "sequence")):
            _lib.require_device_tensor(t, name)
        if obj.dtype != torch.complex64 or probe_mnn.dtype != torch.complex64:
            raise TypeError("pypty_amd: object and probe must be complex64")
        if positions.dtype != torch.int32 or sequence.dtype != torch.int32:
            raise TypeError("pypty_amd: positions and sequence must be int32")
        if measured.dtype != torch.float32:
            raise TypeError("pypty_amd: measured data must be float32")
        if c.step_obj and obj_grad is None:
            raise ValueError("obj_grad buffer required")
        if c.step_probe and probe_grad is None:
            raise ValueError("probe_grad buffer required")
        if (c.step_pos or c.shift_active) and corrections is None:
            raise ValueError("corrections required")
        if c.step_pos and pos_grad is None:
            raise ValueError("pos_grad buffer required")
        if c.compressed:
            _lib.require_device_tensor(masks, "masks")
            if masks.dtype != torch.float32 or tuple(masks.shape) != (c.n_masks, c.n, c.n):
                raise ValueError(f"pypty_amd: masks must be float32 ({c.n_masks}, {c.n}, {c.n})")
            if measured.shape[-1] != c.n_masks:
                raise ValueError("pypty_amd: compressed data need one coefficient per virtual detector")
        key = (self.device, self.workspace_bytes, _lib.stream_ptr(self.device))
        ws = PtychoPass._workspaces.get(key)
        if ws is None:
            ws = torch.empty(self.workspace_bytes, dtype=torch.uint8, device=self.device)
            PtychoPass._workspaces[key] = ws
        self.workspace = ws  # the last run's
        for t, name in ((static_background, "static_background"), (static_background_grad, "static_background_grad")):
            if t is not None:
                _lib.require_device_tensor(t, name)
                if t.dtype != torch.float32 or tuple(t.shape) != (c.n, c.n) or not t.is_contiguous():
                    raise ValueError(f"pypty_amd: {name} must be contiguous float32 ({c.n}, {c.n})")
        if static_background is not None and c.compressed:
            raise NotImplementedError("pypty_amd: a static background with the compressed losses")
        if c.step_static_background and (static_background is None or static_background_grad is None):
            raise ValueError("static_background and static_background_grad required")
        for t, name in ((beam_current, "beam_current"), (beam_current_grad, "beam_current_grad")):
            if t is not None:
                _lib.require_device_tensor(t, name)
                if t.dtype != torch.float32 or tuple(t.shape) != (c.n_positions,) or not t.is_contiguous():
                    raise ValueError(f"pypty_amd: {name} must be contiguous float32 ({c.n_positions},)")
        if c.step_beam_current and (beam_current is None or beam_current_grad is None):
            raise ValueError("beam_current and beam_current_grad required")
        if c.n_aberrations:
            shapes = ((aberrations, "aberrations", torch.float32, (c.n_markers, c.n_aberrations)),
                      (aberration_marker, "aberration_marker", torch.int32, (c.n_positions,)))
            if c.step_aberrations:
                shapes += ((aberrations_grad, "aberrations_grad", torch.float32, (c.n_markers, c.n_aberrations)),)
            for t, name, dt, shape in shapes:
                _lib.require_device_tensor(t, name)
                if t.dtype != dt or tuple(t.shape) != shape or not t.is_contiguous():
                    raise ValueError(f"pypty_amd: {name} must be contiguous {dt} {shape}")
        checks = []
        if c.n_distances > 1:
            checks.append((distances, "distances", (c.n_distances,)))
        if c.n_tilts:
            checks.append((tilts, "tilts", (c.n_tilts, 6)))
            if c.step_tilts:
                checks.append((tilts_grad, "tilts_grad", (c.n_tilts, 6)))
        for t, name, shape in checks:
            _lib.require_device_tensor(t, name)
            if t.dtype != torch.float32 or tuple(t.shape) != shape or not t.is_contiguous():
                raise ValueError(f"pypty_amd: {name} must be contiguous float32 {shape}")
        out = self.loss_sse if loss_sse is None else loss_sse
        ptr = lambda t: None if t is None else t.data_ptr()  # noqa: E731
        bufs = _lib.PtyBuffers(
            ptr(obj), ptr(probe_mnn), ptr(positions), ptr(corrections), ptr(measured), ptr(sequence),
            int(sequence.numel()), ptr(out), ptr(obj_grad), ptr(probe_grad), ptr(pos_grad),
            ptr(sim_patterns), ptr(self.workspace), self.workspace_bytes, ptr(masks) if c.compressed else None,
            _lib.PTY_MEASURED_BY_SEQUENCE if measured_by_sequence else _lib.PTY_MEASURED_BY_POSITION,
            int(bool(accumulate)), ptr(static_background),
            ptr(static_background_grad) if c.step_static_background else None, ptr(beam_current),
            ptr(beam_current_grad) if c.step_beam_current else None,
            ptr(aberrations) if c.n_aberrations else None, ptr(aberration_marker) if c.n_aberrations else None,
            ptr(aberrations_grad) if c.step_aberrations else None,
            ptr(distances) if c.n_distances > 1 else None, ptr(tilts) if c.n_tilts else None,
            ptr(tilts_grad) if c.step_tilts else None)
        rc = self.lib.pty_loss_and_direction(ctypes.byref(self.c_problem), ctypes.byref(bufs),
                                              _lib.stream_ptr(self.device))
        _lib.check(rc, "loss_and_direction")
        if not sync:
            return out
        loss, sse = out.tolist()
        return loss, sse

    def run_streamed(self, obj, probe_mnn, positions, corrections, measured_host, sequence_host, *, obj_grad=None,
                     probe_grad=None, pos_grad=None, masks=None, static_background=None,
                     static_background_grad=None, beam_current=None, beam_current_grad=None, aberrations=None,
                     aberration_marker=None, aberrations_grad=None, distances=None, tilts=None, tilts_grad=None):
        """The same pass with the measured patterns left in host memory (objective.py:383-408,
        load_one_by_one): each batch's rows are gathered into one of two pinned buffers, copied to the
        device on a side stream while the previous batch computes, and the batch runs as an
        accumulating call over its slice of the sequence.  Returns (loss, sse)."""
        c = self.cfg
        seq = np.asarray(sequence_host, dtype=np.int64).ravel()
        src = measured_host if isinstance(measured_host, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(measured_host, dtype=np.float32))
        if src.is_cuda or src.dtype != torch.float32:
            raise TypeError("pypty_amd: streamed patterns must be float32 in host memory")
        B = max(1, int(c.compute_batch))
        row_shape = tuple(src.shape[1:])
        pinned = [torch.empty((B,) + row_shape, dtype=torch.float32).pin_memory() for _ in range(2)]
        staged = [torch.empty((B,) + row_shape, dtype=torch.float32, device=self.device) for _ in range(2)]
        copied = [torch.cuda.Event() for _ in range(2)]  # staged[slot] holds its batch
        free = [torch.cuda.Event() for _ in range(2)]    # the batch that read staged[slot] has run
        main = torch.cuda.current_stream(self.device)
        side = torch.cuda.Stream(self.device)
        starts = list(range(0, len(seq), B))
        sums = torch.zeros((len(starts), 2), dtype=torch.float64, device=self.device)
        seq_dev = torch.as_tensor(seq, dtype=torch.int32).to(self.device)

        def stage(k):
            rows = torch.from_numpy(seq[starts[k]:starts[k] + B])
            slot = k % 2
            copied[slot].synchronize()  # the pinned buffer's previous copy has finished
            torch.index_select(src, 0, rows, out=pinned[slot][:len(rows)])
            with torch.cuda.stream(side):
                side.wait_event(free[slot])
                staged[slot][:len(rows)].copy_(pinned[slot][:len(rows)], non_blocking=True)
                copied[slot].record(side)

        if starts:
            stage(0)
        for k, i0 in enumerate(starts):
            slot = k % 2
            sub = seq_dev[i0:i0 + B]
            main.wait_event(copied[slot])
            self.run(obj, probe_mnn, positions, corrections, staged[slot][:sub.numel()], sub, obj_grad=obj_grad,
                     probe_grad=probe_grad, pos_grad=pos_grad, sync=False, masks=masks, measured_by_sequence=True,
                     accumulate=k > 0, loss_sse=sums[k], static_background=static_background,
                     static_background_grad=static_background_grad, beam_current=beam_current,
                     beam_current_grad=beam_current_grad, aberrations=aberrations,
                     aberration_marker=aberration_marker, aberrations_grad=aberrations_grad, distances=distances,
                     tilts=tilts, tilts_grad=tilts_grad)
            free[slot].record(main)
            if k + 1 < len(starts):
                stage(k + 1)  # the host gathers the next batch while this one computes
        loss, sse = sums.sum(0).tolist()
        return loss, sse


def compute_full_l1_constraint(object, abs_norm_weight, phase_norm_weight, grad_mask=None, return_direction=True,
                               smart_memory=False, *, out_grad=None):
    """Mirror of pypty/objective.py:1286-1320: L1 norms of the phase (phase_norm_weight) and of the
    log-amplitude (abs_norm_weight) of a complex64 object on the device, through
    ``pty_l1_constraint_c64``.  Returns (reg_term, grad): grad is 0.5 (abs_w sign(log |O|) +
    i phase_w sign(arg O)) / conj(O), or None without return_direction.  A weight <= 0 contributes
    nothing, as there.  out_grad (keyword only, not in the reference): add the gradient into this
    tensor instead of a new one.  grad_mask is unused, as in the reference."""
    _lib.require_device_tensor(object, "object")
    if object.dtype != torch.complex64 or not object.is_contiguous():
        raise TypeError("pypty_amd: the object must be contiguous complex64")
    grad = None
    if return_direction:
        grad = out_grad if out_grad is not None else torch.zeros_like(object)
        if grad.shape != object.shape or grad.dtype != torch.complex64 or not grad.is_contiguous():
            raise ValueError("pypty_amd: out_grad must be a contiguous complex64 tensor of the object's shape")
    out = torch.empty(1, dtype=torch.float64, device=object.device)
    partial = torch.empty(512, dtype=torch.float64, device=object.device)
    rc = _lib.load().pty_l1_constraint_c64(object.data_ptr(), object.numel(), float(abs_norm_weight),
                                           float(phase_norm_weight), grad.data_ptr() if grad is not None else None,
                                           out.data_ptr(), partial.data_ptr(), _lib.stream_ptr(object.device))
    _lib.check(rc, "l1_constraint")
    return float(out.item()), grad


def charge_flip(a, delta_phase=0.03, delta_abs=0.14, beta_phase=-0.95, beta_abs=-0.95, fancy_sigma=None, xp=None):
    """Mirror of pypty/objective.py:1014-1110 on the device (pty_charge_flip_c64): per slice and mode,
    the phase and the absorption -log(|a| + 1e-10) below delta times their largest value are
    multiplied by beta, and the object becomes exp(-absorption + i phase).  a: (Ny, Nx, S, K)
    complex64 (torch tensor, numpy or DLPack array); returns a new array of the caller's type.  The
    fancy_sigma deconvolution and reconvolution (objective.py:1058-1108) is not on this path."""
    if fancy_sigma is not None:
        raise NotImplementedError("pypty_amd.charge_flip: fancy_sigma is outside the accelerated path")
    kind, template = _arrays.kind_of(a), a
    t = _arrays.to_torch(a, dtype=torch.complex64)
    if t.device.type != "cuda":
        t = t.to(torch.device("cuda", torch.cuda.current_device()))
    t = t.contiguous().clone()
    planes = int(np.prod(t.shape[2:])) if t.dim() > 2 else 1
    partial = torch.empty(130 * planes, dtype=torch.float32, device=t.device)
    rc = _lib.load().pty_charge_flip_c64(t.data_ptr(), t.numel(), planes, _lib.PTY_LAYOUT_REFERENCE,
                                         float(delta_phase), float(delta_abs), float(beta_phase), float(beta_abs),
                                         partial.data_ptr(), _lib.stream_ptr(t.device))
    _lib.check(rc, "charge_flip")
    return _arrays.like(t, kind, template)


def make_states_orthogonal(probe_states, xp=None):
    """Mirror of pypty/objective.py:1119-1140 on the device (pty_orthogonalize_modes_c64): Gram-Schmidt
    over the probe modes (last axis) in the reference's order, mode i minus its projections
    sum(conj(p_j) p_i) / (1e-10 + sum |p_j|^2) p_j on the modes j < i already made orthogonal, the
    sums in FP64.  Returns a new array of the caller's type."""
    kind, template = _arrays.kind_of(probe_states), probe_states
    t = _arrays.to_torch(probe_states, dtype=torch.complex64)
    if t.device.type != "cuda":
        t = t.to(torch.device("cuda", torch.cuda.current_device()))
    if t.dim() != 3:
        raise NotImplementedError("pypty_amd.make_states_orthogonal: probe markers (4-d probes) are outside the "
                                  "accelerated path")
    t = t.contiguous().clone()
    rc = _lib.load().pty_orthogonalize_modes_c64(t.data_ptr(), t.shape[0] * t.shape[1], t.shape[2],
                                                 _lib.stream_ptr(t.device))
    _lib.check(rc, "orthogonalize_modes")
    return _arrays.like(t, kind, template)


def check_windows_inside(positions: torch.Tensor, sequence: torch.Tensor, n: int, ny: int, nx: int) -> None:
    """Raise if a window [y, y+n) x [x, x+n) of the sequence leaves the object.

    The reference would index out of range (or wrap for negative positions);
    the kernels read the object without bounds checks, so this is enforced here.
    """
    p = positions.index_select(0, sequence.long())
    mn = p.amin(dim=0).tolist() if p.numel() else [0, 0]
    mx = p.amax(dim=0).tolist() if p.numel() else [0, 0]
    if mn[0] < 0 or mn[1] < 0 or mx[0] + n > ny or mx[1] + n > nx:
        raise ValueError(f"pypty_amd: scan windows leave the object: corners in y [{mn[0]}, {mx[0]}], "
                         f"x [{mn[1]}, {mx[1]}], window {n}, object {ny}x{nx}")


def _as_device(x, dtype, device):
    return _arrays.to_torch(x, device, dtype).contiguous()


def _host(x, dtype):
    """A small argument (distances, tilts) as a numpy array, from any array library."""
    if isinstance(x, torch.Tensor) or _arrays.kind_of(x) == _arrays.FOREIGN:
        return _arrays.to_torch(x).cpu().numpy().astype(dtype)
    return np.asarray(x, dtype=dtype)


def _unsupported(name, cond):
    if cond:
        raise NotImplementedError(f"pypty_amd.loss_and_direction: '{name}' is outside the accelerated path")


def loss_and_direction(this_obj, full_probe, this_pos_array, this_pos_correction, this_tilt_array,
                       this_tilts_correction, this_distances, edge_slice_thickness_multiplier, measured_array,
                       algorithm_type, this_wavelength, this_step_probe, this_step_obj, this_step_pos_correction,
                       this_step_tilts, masks, pixel_size_x_A, pixel_size_y_A, recon_type, Cs, defocus_array,
                       alpha_near_field, damping_cutoff_multislice, smooth_rolloff, propmethod,
                       this_chopped_sequence, load_one_by_one, data_multiplier, data_pad, phase_plate_in_h5,
                       this_loss_weight, data_bin, data_shift_vector, upsample_pattern, static_background,
                       this_step_static_background, tilt_mode, aberration_marker, probe_marker, aberrations_array,
                       compute_batch, phase_only_obj, beam_current, this_beam_current_step,
                       this_step_aberrations_array, default_float, default_complex, xp, is_first_epoch, scan_size,
                       scan_step_A, fast_axis_reg_weight_positions, slow_axis_reg_weight_positions,
                       slow_axis_reg_weight_tilts, current_deformation_reg_weight_positions,
                       current_deformation_reg_weight_tilts, fast_axis_reg_weight_tilts, aperture_mask,
                       probe_reg_weight, current_window_weight, current_window, phase_norm_weight, abs_norm_weight,
                       atv_weight, atv_q, atv_p, mixed_variance_weight, mixed_variance_sigma, wedge_mu, beta_wedge,
                       probe_mod_reg_weight, probe_mod_tol, smart_memory, print_flag, data_simulation_flag,
                       data_simulation_exitwave, use_padded_intensity, *, checkpoint_slices=0):
    """Drop-in for pypty.objective.loss_and_direction (pypty/objective.py:47).

    this_obj (Ny, Nx, S, K) and full_probe (N, N, M) complex, positions (P, 2)
    int, corrections (P, 2) and measured_array (P, N, N) float may be torch
    tensors, numpy arrays (as the reference's CPU path passes them; copied to the
    device, the measured patterns once while unchanged) or device arrays of
    another library such as CuPy (viewed through DLPack).  Returns (loss, sse,
    object_grad, probe_grad, pos_grad, tilts_grad, static_background_grad,
    aberrations_array_grad, beam_current_grad, constraint_contributions) like
    the reference, the arrays of this_obj's type; unused gradients are 0 / zero
    arrays as there.

    checkpoint_slices (keyword only, not in the reference): keep the wave
    entering every k-th slice and recompute the others in the adjoint, for
    objects whose full wave stash does not fit; the results are identical.
    """
    _unsupported("recon_type", recon_type != "far_field")
    _unsupported("propmethod", propmethod != "multislice")
    _unsupported("algorithm_type", algorithm_type not in LOSSES)
    _unsupported("tilts before the specimen (tilt_mode 2 and 4) with aberrations",
                 int(tilt_mode) in (2, 4) and aberration_marker is not None and aberrations_array is not None)
    _unsupported("static_background with a compressed loss",
                 static_background is not None and "compressed" in str(algorithm_type))
    _unsupported("probe_marker", probe_marker is not None)
    _unsupported("phase_plate_in_h5", phase_plate_in_h5 is not None)
    _unsupported("upsample_pattern", upsample_pattern != 1)
    _unsupported("data_pad without use_padded_intensity", int(data_pad) != 0 and not use_padded_intensity)
    _unsupported("data_simulation_flag", bool(data_simulation_flag))
    _unsupported("smooth_rolloff", float(smooth_rolloff) != 0.0)
    _unsupported("probe_reg_weight=inf", probe_reg_weight is not None and np.isinf(float(probe_reg_weight)))
    pos_weights = (fast_axis_reg_weight_positions, current_deformation_reg_weight_positions,
                   slow_axis_reg_weight_positions)
    if this_step_pos_correction and any(w != 0 for w in pos_weights) and scan_size is None:
        raise ValueError("pypty_amd.loss_and_direction: the scan-grid constraints on the positions need scan_size")
    dist = _host(this_distances, np.float64).ravel()
    # the tilts in use are the array plus its correction, in the working precision (objective.py:350)
    tt = _host(this_tilt_array, np.float32).reshape(-1, 6) + _host(this_tilts_correction, np.float32).reshape(-1, 6)
    tilt_weights = (current_deformation_reg_weight_tilts, fast_axis_reg_weight_tilts, slow_axis_reg_weight_tilts)
    if this_step_tilts and any(w != 0 for w in tilt_weights) and scan_size is None:
        raise ValueError("pypty_amd.loss_and_direction: the scan-grid constraints on the tilts need scan_size")

    # the caller's array type (numpy, torch, or a device array of another library such as CuPy)
    # is the type of the returned gradients
    kind, template = _arrays.kind_of(this_obj), this_obj
    obj = _arrays.to_torch(this_obj, dtype=torch.complex64).contiguous()
    if obj.device.type != "cuda":
        raise TypeError("pypty_amd: no HIP device for this_obj")
    device = obj.device
    ny, nx, S, K = obj.shape
    full_probe = _arrays.to_torch(full_probe, device, torch.complex64)
    n, _, M = full_probe.shape
    probe_mnn = full_probe.permute(2, 0, 1).contiguous()
    positions = _as_device(this_pos_array, torch.int32, device)
    corrections = _as_device(this_pos_correction, torch.float32, device)
    # load_one_by_one with patterns in host memory streams them per batch (objective.py:383-408)
    stream = bool(load_one_by_one) and not (isinstance(measured_array, torch.Tensor) and measured_array.is_cuda)
    if stream:
        measured = (measured_array.detach().to(torch.float32).contiguous() if isinstance(measured_array, torch.Tensor)
                    else torch.from_numpy(np.ascontiguousarray(measured_array, dtype=np.float32)))
    else:
        # host patterns are copied to the device once and reused while unchanged (they are the same
        # array every epoch)
        measured = _arrays.stage(measured_array, device, torch.float32).contiguous()
    seq = _as_device(this_chopped_sequence, torch.int32, device)
    # static background (objective.py:359-361, 525-526, 588-589): a scalar acts on every pixel
    bg = bg_grad = None
    step_bg = static_background is not None and bool(this_step_static_background)
    if static_background is not None:
        if np.isscalar(static_background):
            bg = torch.full((n, n), float(static_background), dtype=torch.float32, device=device)
        else:
            bg = _as_device(static_background, torch.float32, device)
            _unsupported("static_background not of the window's size", tuple(bg.shape) != (n, n))
        if step_bg:
            bg_grad = torch.empty((n, n), dtype=torch.float32, device=device)
    # beam current (objective.py:446-450, 688-698): each position's probe times |beam_current|
    bc = bc_grad = None
    step_bc = beam_current is not None and bool(this_beam_current_step)
    bc_full = None
    if beam_current is not None:
        # a current longer than the scan (as utils.try_to_initialize_beam_current pads it) keeps its
        # extra entries, which no position indexes and whose gradient is zero
        bc_full = _as_device(beam_current, torch.float32, device).reshape(-1)
        P = positions.shape[0]
        if bc_full.numel() < P:
            raise ValueError("pypty_amd.loss_and_direction: beam_current needs a value per position")
        bc = bc_full[:P].contiguous()
        if step_bc:
            bc_grad = torch.empty_like(bc)
    # aberrations per position (objective.py:320-326, 459-465, 673-681): only with a marker, as there
    ab = ab_marker = ab_grad = None
    step_ab = aberration_marker is not None and aberrations_array is not None and bool(this_step_aberrations_array)
    if aberration_marker is not None and aberrations_array is not None:
        ab = _as_device(aberrations_array, torch.float32, device)
        if ab.dim() != 2 or not 1 <= ab.shape[1] <= _lib.PTY_MAX_ABERRATIONS:
            raise ValueError("pypty_amd.loss_and_direction: aberrations_array must be (rows, 1..64 coefficients)")
        marker_host = _host(aberration_marker, np.int64).ravel()
        if marker_host.shape[0] != positions.shape[0]:
            raise ValueError("pypty_amd.loss_and_direction: aberration_marker needs a row per position")
        if marker_host.size and (marker_host.min() < -ab.shape[0] or marker_host.max() >= ab.shape[0]):
            raise IndexError("pypty_amd.loss_and_direction: aberration_marker indexes a missing row")
        marker_host = np.where(marker_host < 0, marker_host + ab.shape[0], marker_host)  # numpy's negative rows
        ab_marker = torch.as_tensor(marker_host.astype(np.int32)).to(device)
        if step_ab:
            ab_grad = torch.empty_like(ab)
    compressed = "compressed" in algorithm_type
    masks_dev = None
    if compressed:
        if masks is None:
            raise ValueError("pypty_amd.loss_and_direction: compressed losses need the virtual detectors (masks)")
        masks_dev = _as_device(masks, torch.float32, device)
    # per-batch preprocessing (utils.py:587-653; skipped for compressed data as there)
    shift_y, shift_x = (0, 0) if compressed else (int(data_shift_vector[0]), int(data_shift_vector[1]))
    pad = 0 if compressed else int(data_pad) // int(upsample_pattern)
    # per-slice distances (multislice_core.py:69-71): one per propagation
    S_dist = dist.shape[0]
    if S_dist > 1 and S_dist < S - 1:
        raise ValueError(f"pypty_amd.loss_and_direction: {S_dist} slice distances for {S} slices")
    dists_dev = torch.as_tensor(dist.astype(np.float32)).to(device) if S_dist > 1 else None
    # tilts (objective.py:428-499): one row or one per position
    if tt.shape[0] not in (1, positions.shape[0]):
        raise ValueError("pypty_amd.loss_and_direction: the tilt array needs one row or one per position")
    tilts_dev = torch.as_tensor(np.ascontiguousarray(tt)).to(device)
    step_tilts = bool(this_step_tilts)
    tilts_grad = torch.empty_like(tilts_dev) if step_tilts else None
    check_windows_inside(positions, seq, n, ny, nx)
    shift_active = bool(torch.any(corrections != 0).item()) or bool(this_step_pos_correction)
    cutoff = float(damping_cutoff_multislice)
    cfg = PassConfig(
        n=n, n_slices=S, n_obj_modes=K, n_probe_modes=M, obj_ny=ny, obj_nx=nx, n_positions=positions.shape[0],
        compute_batch=int(min(int(compute_batch), seq.numel())), algorithm=algorithm_type, planar=False,
        step_obj=bool(this_step_obj), step_probe=bool(this_step_probe),
        step_pos=bool(this_step_pos_correction), shift_active=shift_active,
        loss_weight=float(this_loss_weight), slice_distance=float(dist[0]), wavelength=float(this_wavelength),
        pixel_size_x=float(pixel_size_x_A), pixel_size_y=float(pixel_size_y_A),
        tan_x=float(tt[0, 3]), tan_y=float(tt[0, 2]), damping_cutoff=cutoff,
        edge_multiplier=float(edge_slice_thickness_multiplier),
        n_masks=int(masks_dev.shape[0]) if compressed else 0, raw_n=0 if compressed else int(measured.shape[-1]),
        data_shift_y=shift_y, data_shift_x=shift_x, data_bin=1 if compressed else int(data_bin), data_pad=pad,
        data_multiplier=1.0 if compressed else float(data_multiplier), checkpoint_slices=int(checkpoint_slices),
        lanes=1 if stream else 4,  # a streamed pass runs one batch per call
        step_static_background=step_bg, step_beam_current=step_bc,
        n_aberrations=0 if ab is None else int(ab.shape[1]), n_markers=0 if ab is None else int(ab.shape[0]),
        step_aberrations=step_ab, n_distances=S_dist if S_dist > 1 else 0, tilt_mode=int(tilt_mode),
        n_tilts=int(tt.shape[0]), step_tilts=step_tilts)
    engine = PtychoPass(cfg, device)
    object_grad = torch.empty_like(obj) if this_step_obj else torch.zeros_like(obj)
    probe_grad_mnn = torch.empty((M, n, n), dtype=torch.complex64, device=device) if this_step_probe else None
    pos_grad = torch.zeros_like(corrections)
    if stream:
        loss, sse = engine.run_streamed(obj, probe_mnn, positions, corrections, measured, seq.cpu().numpy(),
                                        obj_grad=object_grad, probe_grad=probe_grad_mnn, pos_grad=pos_grad,
                                        masks=masks_dev, static_background=bg, static_background_grad=bg_grad,
                                        beam_current=bc, beam_current_grad=bc_grad, aberrations=ab,
                                        aberration_marker=ab_marker, aberrations_grad=ab_grad, distances=dists_dev,
                                        tilts=tilts_dev, tilts_grad=tilts_grad)
    else:
        loss, sse = engine.run(obj, probe_mnn, positions, corrections, measured, seq, obj_grad=object_grad,
                               probe_grad=probe_grad_mnn, pos_grad=pos_grad, masks=masks_dev,
                               static_background=bg, static_background_grad=bg_grad, beam_current=bc,
                               beam_current_grad=bc_grad, aberrations=ab, aberration_marker=ab_marker,
                               aberrations_grad=ab_grad, distances=dists_dev, tilts=tilts_dev, tilts_grad=tilts_grad)
    if loss != loss:
        raise ValueError("A very specific bad thing. Loss is Nan.")
    probe_grad = (probe_grad_mnn.permute(1, 2, 0).contiguous() if this_step_probe
                  else torch.zeros_like(full_probe))
    if tilts_grad is None:
        tilts_grad = torch.zeros_like(tilts_dev)
    # constraints (objective.py:765-921), the reference's fourteen contributions in its order; each
    # adds to the loss (after the loss weight) and to the gradient of what it constrains.  The probe
    # ones are band-limited with the probe gradient (objective.py:927-933); the tilt ones act on the
    # tilts in use (array plus correction) when the tilts are refined.
    cc = np.zeros(14)
    pos_abs = None
    if this_step_pos_correction and any(w != 0 for w in pos_weights):
        pos_abs = positions.to(torch.float64) + corrections.to(torch.float64)
    pos_terms = ((0, constraints.compute_fast_axis_constraint_on_grid, fast_axis_reg_weight_positions),
                 (1, constraints.compute_deformation_constraint_on_grid, current_deformation_reg_weight_positions),
                 (11, constraints.compute_slow_axis_constraint_on_grid, slow_axis_reg_weight_positions))
    tilt_abs = tilts_dev.to(torch.float64) if step_tilts and any(w != 0 for w in tilt_weights) else None
    tilt_terms = {2: (constraints.compute_deformation_constraint_on_grid, current_deformation_reg_weight_tilts),
                  3: (constraints.compute_fast_axis_constraint_on_grid, fast_axis_reg_weight_tilts),
                  12: (constraints.compute_slow_axis_constraint_on_grid, slow_axis_reg_weight_tilts)}
    probe_reg = []
    for k in range(14):
        term, grad = 0.0, None
        if k in tilt_terms:
            fn, w = tilt_terms[k]
            if tilt_abs is not None and w != 0:
                term, grad = fn(tilt_abs, scan_size, w)
                tilts_grad += grad.to(tilts_grad.dtype)
        elif k in (0, 1, 11):
            fn, w = next((f, w) for kk, f, w in pos_terms if kk == k)
            if pos_abs is not None and w != 0:
                term, grad = fn(pos_abs, scan_size, w)
                pos_grad += grad.to(pos_grad.dtype)
        elif k in (4, 5):
            wa, wp = (0.0, phase_norm_weight) if k == 4 else (abs_norm_weight, 0.0)
            if wa != 0 or wp != 0:
                term, _ = compute_full_l1_constraint(obj, wa, wp, None, True, smart_memory, out_grad=object_grad)
        elif k == 6 and this_step_probe and probe_reg_weight:
            term, grad = constraints.compute_probe_constraint(full_probe, aperture_mask, probe_reg_weight, True)
            probe_reg.append(grad)
        elif k == 7 and this_step_probe and current_window_weight:
            term, grad = constraints.compute_window_constraint(full_probe, current_window, current_window_weight)
            probe_reg.append(grad)
        elif k == 8 and atv_weight:
            term, grad = constraints.compute_atv_constraint(obj, atv_weight, atv_q, atv_p, pixel_size_x_A,
                                                            pixel_size_y_A, None, True, smart_memory)
            object_grad += grad
        elif k == 9 and mixed_variance_weight and K > 1:
            term, grad = constraints.compute_mixed_object_variance_constraint(obj, mixed_variance_weight,
                                                                              mixed_variance_sigma, True, smart_memory)
            object_grad += grad
        elif k == 10 and wedge_mu:
            term, grad = constraints.compute_missing_wedge_constraint(obj, pixel_size_x_A, pixel_size_y_A,
                                                                      float(dist[0]), beta_wedge, wedge_mu)
            object_grad += grad
        elif k == 13 and this_step_probe and probe_mod_reg_weight:
            term, grad = constraints.probe_modulation_constraint(full_probe, scan_step_A, pixel_size_x_A,
                                                                 pixel_size_y_A, probe_mod_tol, probe_mod_reg_weight,
                                                                 cutoff)
            probe_reg.append(grad)
        loss += term
        cc[k] = term
    if probe_reg:
        # the probe constraints enter before the probe gradient's band limit, which is linear
        mask_i = utils.create_spatial_frequencies(float(pixel_size_x_A), float(pixel_size_y_A), n, cutoff, 0.0)[4]
        total = probe_reg[0].to(torch.complex64)
        for g in probe_reg[1:]:
            total = total + g.to(torch.complex64)
        spec = pfft.fft2(total.contiguous(), axes=(0, 1)) * mask_i[:, :, None]
        probe_grad += pfft.ifft2(spec.contiguous(), axes=(0, 1))
    constraint_contributions = cc
    # objective.py:377: zeros of the background's shape for an array, 0 for a scalar or none
    if bg_grad is None:
        bg_grad = (torch.zeros_like(bg) if bg is not None and not np.isscalar(static_background) else 0)
    # objective.py:378: zeros of the current's shape when there is one, else 0
    if bc is None:
        bc_grad = 0
    elif bc_grad is None or bc_full.numel() != bc.numel():
        full = torch.zeros_like(bc_full)
        if bc_grad is not None:
            full[:bc.numel()] = bc_grad
        bc_grad = full
    # objective.py:379: zeros of the aberrations' shape when there are some, else 0
    if aberrations_array is None:
        ab_grad = 0
    elif ab_grad is None:
        ab_grad = torch.zeros_like(_as_device(aberrations_array, torch.float32, device))
    out = [_arrays.like(t, kind, template)
           for t in (object_grad, probe_grad, pos_grad, tilts_grad, bg_grad, bc_grad, ab_grad)]
    return (loss, sse, out[0], out[1], out[2], out[3], out[4], out[6], out[5], constraint_contributions)

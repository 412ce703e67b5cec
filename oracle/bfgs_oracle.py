"""numpy restatement of pypty.iterative.bfgs_update for object / probe / positions (test oracle only).

State is a dict {obj (Ny,Nx,S,K), probe (N,N,M), positions (P,2) int, corr (P,2)}
plus a history dict laid out as pypty/iterative.py:1055-1103.  One call is one
epoch of pypty/iterative.py:440-1049 (band-limits :497-514, first direction
:562-572, two-loop :573-681, directional derivative and fallback :683-731,
Wolfe line search :734-836, rejection :838-844, history update :849-932).
"""
from __future__ import annotations

import numpy as np

from . import constraints_oracle as con
from . import ms_oracle as orc


def new_history():
    """pypty/iterative.py:1055-1103 (the variables this restatement supports)."""
    return {"obj_hist_s": [], "probe_hist_s": [], "positions_hist_s": [], "static_background_hist_s": [],
            "beam_current_hist_s": [], "obj_hist_y": [], "probe_hist_y": [], "positions_hist_y": [],
            "static_background_hist_y": [], "beam_current_hist_y": [], "rho_hist": [], "prev_loss": None,
            "prev_sse": None, "obj_grad": None, "probe_grad": None, "pos_grad": None, "static_background_grad": None,
            "beam_current_grad": None, "empty_hist": True, "gamma_obj": 1, "gamma_probe": 1, "gamma_pos": 1,
            "gamma_static": 1, "gamma_beam_current": 1, "aberrations_hist_s": [], "aberrations_hist_y": [],
            "aberrations_grad": None, "gamma_aberrations": 1, "tilts_hist_s": [], "tilts_hist_y": [],
            "tilt_grad": None, "gamma_tilts": 1}


def _d(a, b):
    """2 sum Re(a conj(b)) for complex blocks, sum a b for real ones (iterative.py:597, 652, 683-690);
    a history entry 0 of a variable not refined in its epoch contributes 0."""
    if not isinstance(a, np.ndarray) or not isinstance(b, np.ndarray):
        return 0.0
    if np.iscomplexobj(a):
        return 2.0 * float(np.sum(np.real(a * np.conj(b)), dtype=np.float64))
    return float(np.sum(a.astype(np.float64) * b.astype(np.float64)))


def bfgs_epoch(state, hist, data, *, update_obj=True, update_probe=True, update_pos=False, actual_step=1.0,
               cutoff=2 / 3, update_extra_cut=0.005, wolfe_c1=0.1, wolfe_c2=0.9, max_count=20, reduce_factor=0.1,
               optimism=3, min_step=1e-20, hist_length=10, cdtype=np.complex64, update_bg=False, update_bc=False,
               update_ab=False, update_tilts=False, charge_flip=None, orthogonal=False, phase_only=False):
    """One epoch; mutates state and hist; returns (loss, sse, actual_step, count, d_value, new_d_value, accepted).

    A static background is state["bg"] (N, N) when present; it enters every loss evaluation and,
    with update_bg, is refined as the fourth variable (iterative.py:480, 569, 604-671, 688, 769,
    855-862, 898-903, 1030), a real block like the positions.  A beam current is state["bc"] (P,)
    when present; it enters every loss evaluation and, with update_bc, is refined as the fifth
    (iterative.py:481, 571, 591-673, 690, 771, 853-870, 910-915, 1029).  Aberration rows are
    state["ab"] (rows, coefficients) with data["aberration_marker"]; they enter every loss evaluation
    and, with update_ab, are refined as the sixth (iterative.py:478, 570, 590-672, 689, 770, 854-871,
    904-909, 1028).  Tilts are data["tilts"] (rows, 6) with data["tilt_mode"] plus state["tc"], their
    correction; the loss sees their sum (objective.py:350) and, with update_tilts, the correction is
    refined as the fourth block, after the positions (iterative.py:477, 568, 588-670, 687, 768,
    852-868, 892-897, 1027).  orthogonal makes the probe modes orthogonal before the probe is
    band-limited (iterative.py:498-499); charge_flip = (delta_phase, delta_abs, beta_phase, beta_abs)
    flips the object and restarts the history before the object is band-limited (:502-505).
    phase_only refines the object's phase: its block is the real phase gradient 2 Re(-i conj(O) g)
    of the band-limited gradient (:549-550, 798-799), the update is not band-limited (:674) and a
    step makes |O| exp(i (arg O + t u)), band-limited (:737-746)."""
    fd = np.float32 if np.dtype(cdtype) == np.complex64 else np.float64
    cut = float(fd(cutoff) - fd(update_extra_cut))
    bg0 = state.get("bg", data.get("static_background"))
    update_bg = bool(update_bg) and bg0 is not None
    bc0 = state.get("bc", data.get("beam_current"))
    update_bc = bool(update_bc) and bc0 is not None
    marker = data.get("aberration_marker")
    ab0 = state.get("ab", data.get("aberrations")) if marker is not None else None
    update_ab = bool(update_ab) and ab0 is not None
    tilts = data.get("tilts")
    tc0 = None if tilts is None else state.get("tc", np.zeros_like(tilts, dtype=np.float32))
    update_tilts = bool(update_tilts) and tilts is not None
    on = (update_obj, update_probe, update_pos, update_tilts, update_bg, update_bc, update_ab)
    NB = len(on)
    S_KEYS = ("obj_hist_s", "probe_hist_s", "positions_hist_s", "tilts_hist_s", "static_background_hist_s",
              "beam_current_hist_s", "aberrations_hist_s")
    Y_KEYS = ("obj_hist_y", "probe_hist_y", "positions_hist_y", "tilts_hist_y", "static_background_hist_y",
              "beam_current_hist_y", "aberrations_hist_y")
    G_KEYS = ("gamma_obj", "gamma_probe", "gamma_pos", "gamma_tilts", "gamma_static", "gamma_beam_current",
              "gamma_aberrations")
    for k in S_KEYS + Y_KEYS:
        hist.setdefault(k, [0] * len(hist["rho_hist"]))
    hist.setdefault("gamma_static", 1)
    hist.setdefault("gamma_beam_current", 1)
    hist.setdefault("gamma_aberrations", 1)
    hist.setdefault("gamma_tilts", 1)

    def clean(x):
        return orc.fourier_clean(x, cutoff=cut, fdtype=fd).astype(cdtype)

    def evaluate(v):
        o, p, c, tc, b, cur, ab = v
        tl = None if tilts is None else (np.asarray(tilts, np.float32) + tc).astype(np.float32)
        out = orc.loss_and_direction(
            o, p, state["positions"], c, data["measured"], data["sequence"], dz=data["dz"],
            wavelength=data["wavelength"], pixel_size_x=data["px"], pixel_size_y=data["px"],
            algorithm=data.get("algorithm", "lsq_sqrt"), cutoff=cutoff, step_obj=update_obj,
            step_probe=update_probe, step_pos=update_pos, compute_batch=data.get("compute_batch", 64), cdtype=cdtype,
            static_background=b, step_static_background=update_bg, beam_current=cur, step_beam_current=update_bc,
            aberrations=ab, aberration_marker=None if ab is None else marker, step_aberrations=update_ab,
            distances=data.get("distances"), tilts=tl, tilt_mode=data.get("tilt_mode", 0), step_tilts=update_tilts)
        loss, sse, og, pg, posg = out[:5]
        # L1 constraints of the object (objective.py:811-830), in the loss and object gradient
        for wa, wp in ((0.0, data.get("phase_norm_weight", 0)), (data.get("abs_norm_weight", 0), 0.0)):
            if wa != 0 or wp != 0:
                reg, gl = orc.full_l1_constraint(o, wa, wp)
                loss += reg
                og = (og + gl).astype(og.dtype)
        # further constraint terms (objective.py:765-921): data["constraints"](obj, probe) returns
        # (term, object gradient or None, probe gradient or None); the probe gradient is band-limited
        # with the pass's (objective.py:927-933)
        if data.get("constraints") is not None:
            reg, go, gp = data["constraints"](o, p)
            loss += reg
            if go is not None:
                og = (og + go).astype(og.dtype)
            if gp is not None and update_probe:
                _, _, _, _, m = orc.create_spatial_frequencies(data["px"], data["px"], p.shape[0], cutoff, 0.0,
                                                                 np.float32)
                pg = (pg + orc.ifft2(orc.fft2(gp) * m[:, :, None])).astype(pg.dtype)
        if update_probe:
            pg = clean(pg)
        if update_obj:
            og = clean(og)
            if phase_only:
                og = orc.complex_grad_to_phase_grad(og, o).astype(fd)
        # objective.py:376-378: zeros of the background's (current's) shape when it is not refined
        extra = list(out[5:])
        bgg = extra.pop(0) if b is not None else None
        bcg = extra.pop(0) if cur is not None else None
        abg = extra.pop(0) if ab is not None else None
        tg = extra.pop(0) if tl is not None else None
        return loss, sse, [og, pg, posg, tg if update_tilts else (None if tl is None else np.zeros_like(tc)),
                           bgg if update_bg else (None if b is None else np.zeros_like(b)),
                           bcg if update_bc else (None if cur is None else np.zeros_like(cur)),
                           abg if update_ab else (None if ab is None else np.zeros_like(ab))]

    def ddot(u, g):
        return sum(_d(a, b) for a, b, f in zip(u, g, on) if f)

    if update_probe:
        if orthogonal and state["probe"].shape[-1] > 1:
            state["probe"] = con.make_states_orthogonal(state["probe"])
        state["probe"] = clean(state["probe"])
    if update_obj:
        if charge_flip is not None:
            state["obj"] = con.charge_flip(state["obj"], *charge_flip)
            hist.clear()
            hist.update(new_history())
        state["obj"] = clean(state["obj"])
    x = [state["obj"], state["probe"], state["corr"], tc0, bg0, bc0, ab0]
    if hist["empty_hist"]:
        loss, sse, g = evaluate(x)
        uo = -g[0] if update_obj else None
        dvo = _d(uo, uo) if update_obj else 0.0
        if dvo == 0:
            dvo = 1.0
        u = [uo] + [-g[j] * fd(dvo / _d(g[j], g[j])) if on[j] else None for j in range(1, NB)]
        hist["empty_hist"] = False
    else:
        loss, sse = hist["prev_loss"], hist["prev_sse"]
        g = [hist["obj_grad"], hist["probe_grad"], hist["pos_grad"], hist.get("tilt_grad"),
             hist.get("static_background_grad"), hist.get("beam_current_grad"), hist.get("aberrations_grad")]
        if update_tilts and g[3] is None:
            g[3] = np.zeros_like(tc0)
        if update_bg and g[4] is None:
            g[4] = np.zeros_like(bg0)
        if update_bc and g[5] is None:
            g[5] = np.zeros_like(bc0)
        if update_ab and g[6] is None:
            g[6] = np.zeros_like(ab0)
        if update_probe:
            g[1] = clean(g[1])
        q = [g[j].copy() if on[j] else None for j in range(NB)]
        rhos = hist["rho_hist"]
        alphas = []
        for i in range(len(rhos) - 1, -1, -1):
            a = rhos[i] * ddot([hist[k][i] for k in S_KEYS], q)
            alphas.append(a)
            for j, key in enumerate(Y_KEYS):
                if q[j] is not None:
                    q[j] = q[j] - fd(a) * hist[key][i]
        alphas = alphas[::-1]
        for j, key in enumerate(G_KEYS):
            gm = hist[key]
            if q[j] is not None and gm != 0 and gm == gm:
                q[j] = q[j] * fd(gm)
        for i in range(len(rhos)):
            beta = alphas[i] - rhos[i] * ddot([hist[k][i] for k in Y_KEYS], q)
            for j, key in enumerate(S_KEYS):
                if q[j] is not None:
                    q[j] = q[j] + fd(beta) * hist[key][i]
        u = [None if v is None else -v for v in q]
        if update_obj and not phase_only:
            u[0] = clean(u[0])
    d_value = ddot(u, g)
    if d_value >= 0:
        u = [-g[j] if on[j] else None for j in range(NB)]
        for j, key in enumerate(G_KEYS):
            gm = hist[key]
            if on[j] and gm > 0 and gm == gm:
                u[j] = u[j] * fd(gm)
        d_value = ddot(u, g)
    count = -1
    while True:
        count += 1
        if update_obj and phase_only:
            o_new = clean((np.abs(x[0]) * np.exp(1j * (np.angle(x[0]) + fd(actual_step) * u[0]))).astype(cdtype))
        elif update_obj:
            o_new = clean(x[0] + fd(actual_step) * u[0])
        else:
            o_new = x[0]
        nx = [o_new,
              x[1] + fd(actual_step) * u[1] if update_probe else x[1],
              (x[2] + fd(actual_step) * u[2]).astype(fd) if update_pos else x[2],
              (x[3] + fd(actual_step) * u[3]).astype(fd) if update_tilts else x[3],
              (x[4] + fd(actual_step) * u[4]).astype(fd) if update_bg else x[4],
              (x[5] + fd(actual_step) * u[5]).astype(fd) if update_bc else x[5],
              (x[6] + fd(actual_step) * u[6]).astype(fd) if update_ab else x[6]]
        nloss, nsse, ng = evaluate(nx)
        nd = ddot(u, ng)
        w1 = orc.wolfe_1(loss, nloss, d_value, actual_step, wolfe_c1)
        w2 = orc.wolfe_2(d_value, nd, wolfe_c2)
        if w1 and w2:
            break
        if not w1 and w2:
            actual_step *= reduce_factor
        if not w2 and w1:
            actual_step *= optimism
        if not (w1 or w2):
            actual_step = type(actual_step)(10 ** (np.random.uniform(0, 20) - 20))
        if max_count is not None and count > max_count:
            break
        if actual_step < min_step:
            break
    if not w1:
        hist.clear()
        hist.update(new_history())
        return loss, sse, actual_step, count, d_value, nd, False
    s = [fd(actual_step) * u[j] if on[j] else 0 for j in range(NB)]
    y = [ng[j] - g[j] if on[j] else 0 for j in range(NB)]
    for j in range(NB):
        hist[S_KEYS[j]].append(s[j])
        hist[Y_KEYS[j]].append(y[j])
    hist["prev_loss"], hist["prev_sse"] = nloss, nsse
    (hist["obj_grad"], hist["probe_grad"], hist["pos_grad"], hist["tilt_grad"], hist["static_background_grad"],
     hist["beam_current_grad"], hist["aberrations_grad"]) = ng
    tot = 0.0
    for j, key in enumerate(G_KEYS):
        if on[j]:
            ys = _d(y[j], s[j])
            yy = _d(y[j], y[j]) / (2.0 if np.iscomplexobj(y[j]) else 1.0)
            tot += ys
            hist[key] = ys / yy
    hist["rho_hist"].append(1.0 / tot)
    while len(hist["obj_hist_s"]) > hist_length:
        for key in S_KEYS + Y_KEYS + ("rho_hist",):
            hist[key].pop(0)
    state["obj"], state["probe"] = nx[0], nx[1]
    if update_pos:
        state["corr"] = nx[2]
    if update_tilts:
        state["tc"] = nx[3]
    if update_bg:
        state["bg"] = nx[4]
    if update_bc:
        state["bc"] = nx[5]
    if update_ab:
        state["ab"] = nx[6]
    return loss, sse, actual_step, count, d_value, nd, True

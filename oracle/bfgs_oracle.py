"""numpy restatement of pypty.iterative.bfgs_update for object / probe / positions (test oracle only).

State is a dict {obj (Ny,Nx,S,K), probe (N,N,M), positions (P,2) int, corr (P,2)}
plus a history dict laid out as pypty/iterative.py:1055-1103.  One call is one
epoch of pypty/iterative.py:440-1049 (band-limits :497-514, first direction
:562-572, two-loop :573-681, directional derivative and fallback :683-731,
Wolfe line search :734-836, rejection :838-844, history update :849-932).
"""
from __future__ import annotations

import numpy as np

from . import ms_oracle as orc


def new_history():
    """pypty/iterative.py:1055-1103 (the variables this restatement supports)."""
    return {"obj_hist_s": [], "probe_hist_s": [], "positions_hist_s": [], "obj_hist_y": [], "probe_hist_y": [],
            "positions_hist_y": [], "rho_hist": [], "prev_loss": None, "prev_sse": None, "obj_grad": None,
            "probe_grad": None, "pos_grad": None, "empty_hist": True, "gamma_obj": 1, "gamma_probe": 1,
            "gamma_pos": 1}


def _d(a, b):
    """2 sum Re(a conj(b)) for complex blocks, sum a b for real ones (iterative.py:597, 652, 683-690)."""
    if np.iscomplexobj(a):
        return 2.0 * float(np.sum(np.real(a * np.conj(b)), dtype=np.float64))
    return float(np.sum(a.astype(np.float64) * b.astype(np.float64)))


def bfgs_epoch(state, hist, data, *, update_obj=True, update_probe=True, update_pos=False, actual_step=1.0,
               cutoff=2 / 3, update_extra_cut=0.005, wolfe_c1=0.1, wolfe_c2=0.9, max_count=20, reduce_factor=0.1,
               optimism=3, min_step=1e-20, hist_length=10, cdtype=np.complex64):
    """One epoch; mutates state and hist; returns (loss, sse, actual_step, count, d_value, new_d_value, accepted)."""
    fd = np.float32 if np.dtype(cdtype) == np.complex64 else np.float64
    cut = float(fd(cutoff) - fd(update_extra_cut))

    def clean(x):
        return orc.fourier_clean(x, cutoff=cut, fdtype=fd).astype(cdtype)

    def evaluate(o, p, c):
        # a fixed static background (data["static_background"]) enters the loss only
        loss, sse, og, pg, posg = orc.loss_and_direction(
            o, p, state["positions"], c, data["measured"], data["sequence"], dz=data["dz"],
            wavelength=data["wavelength"], pixel_size_x=data["px"], pixel_size_y=data["px"],
            algorithm=data.get("algorithm", "lsq_sqrt"), cutoff=cutoff, step_obj=update_obj,
            step_probe=update_probe, step_pos=update_pos, compute_batch=data.get("compute_batch", 64), cdtype=cdtype,
            static_background=data.get("static_background"))[:5]
        if update_probe:
            pg = clean(pg)
        if update_obj:
            og = clean(og)
        return loss, sse, og, pg, posg

    def ddot(uo, up, ux, go, gp, gx):
        d = 0.0
        if update_obj:
            d += _d(uo, go)
        if update_probe:
            d += _d(up, gp)
        if update_pos:
            d += _d(ux, gx)
        return d

    if update_probe:
        state["probe"] = clean(state["probe"])
    if update_obj:
        state["obj"] = clean(state["obj"])
    obj, probe, corr = state["obj"], state["probe"], state["corr"]
    if hist["empty_hist"]:
        loss, sse, og, pg, posg = evaluate(obj, probe, corr)
        uo = -og if update_obj else None
        dvo = _d(uo, uo) if update_obj else 0.0
        if dvo == 0:
            dvo = 1.0
        up = -pg * fd(dvo / _d(pg, pg)) if update_probe else None
        ux = -posg * fd(dvo / _d(posg, posg)) if update_pos else None
        hist["empty_hist"] = False
    else:
        loss, sse = hist["prev_loss"], hist["prev_sse"]
        og, pg, posg = hist["obj_grad"], hist["probe_grad"], hist["pos_grad"]
        if update_probe:
            pg = clean(pg)
        q = [og.copy() if update_obj else None, pg.copy() if update_probe else None,
             posg.copy() if update_pos else None]
        rhos = hist["rho_hist"]
        alphas = []
        for i in range(len(rhos) - 1, -1, -1):
            a = rhos[i] * ddot(hist["obj_hist_s"][i], hist["probe_hist_s"][i], hist["positions_hist_s"][i], *q)
            alphas.append(a)
            for j, key in enumerate(("obj_hist_y", "probe_hist_y", "positions_hist_y")):
                if q[j] is not None:
                    q[j] = q[j] - fd(a) * hist[key][i]
        alphas = alphas[::-1]
        for j, key in enumerate(("gamma_obj", "gamma_probe", "gamma_pos")):
            g = hist[key]
            if q[j] is not None and g != 0 and g == g:
                q[j] = q[j] * fd(g)
        for i in range(len(rhos)):
            beta = alphas[i] - rhos[i] * ddot(hist["obj_hist_y"][i], hist["probe_hist_y"][i],
                                              hist["positions_hist_y"][i], *q)
            for j, key in enumerate(("obj_hist_s", "probe_hist_s", "positions_hist_s")):
                if q[j] is not None:
                    q[j] = q[j] + fd(beta) * hist[key][i]
        uo, up, ux = [None if v is None else -v for v in q]
        if update_obj:
            uo = clean(uo)
    d_value = ddot(uo, up, ux, og, pg, posg)
    if d_value >= 0:
        uo = -og if update_obj else None
        up = -pg if update_probe else None
        ux = -posg if update_pos else None
        gs = [hist["gamma_obj"], hist["gamma_probe"], hist["gamma_pos"]]
        uo = uo * fd(gs[0]) if (update_obj and gs[0] > 0 and gs[0] == gs[0]) else uo
        up = up * fd(gs[1]) if (update_probe and gs[1] > 0 and gs[1] == gs[1]) else up
        ux = ux * fd(gs[2]) if (update_pos and gs[2] > 0 and gs[2] == gs[2]) else ux
        d_value = ddot(uo, up, ux, og, pg, posg)
    count = -1
    while True:
        count += 1
        nobj = clean(obj + fd(actual_step) * uo) if update_obj else obj
        nprobe = probe + fd(actual_step) * up if update_probe else probe
        ncorr = (corr + fd(actual_step) * ux).astype(fd) if update_pos else corr
        nloss, nsse, nog, npg, nposg = evaluate(nobj, nprobe, ncorr)
        nd = ddot(uo, up, ux, nog, npg, nposg)
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
    s = [fd(actual_step) * uo if update_obj else 0, fd(actual_step) * up if update_probe else 0,
         fd(actual_step) * ux if update_pos else 0]
    y = [nog - og if update_obj else 0, npg - pg if update_probe else 0, nposg - posg if update_pos else 0]
    for j, (ks, ky) in enumerate((("obj_hist_s", "obj_hist_y"), ("probe_hist_s", "probe_hist_y"),
                                  ("positions_hist_s", "positions_hist_y"))):
        hist[ks].append(s[j])
        hist[ky].append(y[j])
    hist["prev_loss"], hist["prev_sse"] = nloss, nsse
    hist["obj_grad"], hist["probe_grad"], hist["pos_grad"] = nog, npg, nposg
    tot = 0.0
    for j, (key, on) in enumerate((("gamma_obj", update_obj), ("gamma_probe", update_probe),
                                   ("gamma_pos", update_pos))):
        if on:
            ys = _d(y[j], s[j])
            yy = _d(y[j], y[j]) / (2.0 if np.iscomplexobj(y[j]) else 1.0)
            tot += ys
            hist[key] = ys / yy
    hist["rho_hist"].append(1.0 / tot)
    while len(hist["obj_hist_s"]) > hist_length:
        for key in ("obj_hist_s", "probe_hist_s", "positions_hist_s", "obj_hist_y", "probe_hist_y",
                    "positions_hist_y", "rho_hist"):
            hist[key].pop(0)
    state["obj"], state["probe"] = nobj, nprobe
    if update_pos:
        state["corr"] = ncorr
    return loss, sse, actual_step, count, d_value, nd, True

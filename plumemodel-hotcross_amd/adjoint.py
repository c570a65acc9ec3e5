"""torch.autograd interface to the discrete adjoint of the HotCross timestep.

ProbeMisfit wraps one adjoint window of the C-ABI (hc_adj_window): it runs
`nsteps` steps of the timestep on the GPU from a fixed initial state, samples
w and T at the probes every `every` steps, and returns the misfit
    L = sum over samples and probes of 0.5*(wgt_w (w - w_obs)^2 + wgt_t (T - T_obs)^2)
as a 0-d tensor whose gradient flows to three controls:
  qmap        per-cell multiplier of the vent heat flux (npatch values),
  flow_scale  factor on the crossflow series (its amplitude),
  cs          the Smagorinsky constant (cs_h = cs_v).
The gradients come from the hand-written adjoint (checkpoint every `seg`
steps, segments recomputed into an HBM tape), not from torch's autograd of
any operation; the control tensors are plain float64 CPU tensors.

Example (the window starts from the context's state when the Window is
made; every evaluation returns the context to that state):

    win = Window(ctx, probes, every=4, nsteps=400, seg=20, obs=obs)
    q = torch.ones(npatch, dtype=torch.float64, requires_grad=True)
    loss = win.loss(q, torch.tensor(1.0), torch.tensor(0.2))
    loss.backward()          # q.grad = dL/dqmap
"""
import numpy as np
import torch

import hotcross as H


class ProbeMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, qmap, flow_scale, cs, win):
        loss, samples, gq, gs = win.evaluate(qmap.detach().cpu().numpy(), float(flow_scale), float(cs))
        ctx.save_for_backward(torch.from_numpy(gq), torch.tensor(gs[0], dtype=torch.float64),
                              torch.tensor(gs[1] + gs[2], dtype=torch.float64))
        win.last_samples = samples
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        gq, gf, gc = ctx.saved_tensors
        return grad_out * gq, grad_out * gf, grad_out * gc, None


class Window:
    """An adjoint window over a context: a saved initial state, the probes,
    the observations and the window's length and checkpointing (a
    checkpoint every `seg` steps, or with nsnap binomial checkpointing with
    that many snapshot slots)."""

    def __init__(self, ctx, probes, every, nsteps, seg, obs, wgt_w=1.0, wgt_t=1.0, nsnap=None):
        self.ctx = ctx
        self.probes = np.asarray(probes, dtype=np.int32).reshape(-1, 3)
        self.every, self.nsteps, self.seg = int(every), int(nsteps), int(seg)
        self.nsnap = None if nsnap is None else int(nsnap)
        self.obs = np.asarray(obs, dtype=np.float64)
        self.wgt = (float(wgt_w), float(wgt_t))
        self.last_samples = None

    def evaluate(self, qmap, flow_scale, cs):
        """(loss, samples, dL/dqmap, [dL/dflow_scale, dL/dcs_h, dL/dcs_v])
        of the window from the context's current state, which it keeps."""
        self.ctx.set_controls(qmap, flow_scale, cs)
        self.ctx.set_probes(self.probes, self.every, *self.wgt)
        if self.nsnap is not None:
            return self.ctx.adj_window_binomial(self.nsteps, self.nsnap, self.obs, end_state=False)
        return self.ctx.adj_window(self.nsteps, self.seg, self.obs, end_state=False)

    def loss(self, qmap, flow_scale, cs):
        return ProbeMisfit.apply(qmap, flow_scale, cs, self)


def binomial_forward_steps(nsteps, nsnap):
    """Forward steps the binomial window runs for nsnap snapshot slots: the
    advances of the schedule (t l - beta(s+2, t-1) for s = nsnap - 1 free
    slots, t the least with beta(s+1, t) >= l) plus one turn per step."""
    import hotcross as H
    return H.binomial_schedule(nsteps, nsnap)[1]


def plan_window(ctx, nsteps, budget_bytes):
    """The window's checkpointing under an HBM budget: ("fixed", seg, bytes)
    when a fixed-interval window fits (about two forward steps per step: the
    forward sweep and one recompute), else ("binomial", nsnap, bytes) with
    as many snapshot slots as fit (the fewest forward steps for that
    memory); raises if not even one snapshot fits."""
    try:
        seg, need = plan_segment(ctx, nsteps, budget_bytes)
        return "fixed", seg, need
    except ValueError:
        pass
    lo, hi = 1, nsteps
    if ctx.adj_window_binomial_bytes(nsteps, 1) > budget_bytes:
        raise ValueError("not even one snapshot fits the %.1f GB budget" % (budget_bytes / 1e9))
    while lo < hi:  # the largest nsnap that fits
        mid = (lo + hi + 1) // 2
        if ctx.adj_window_binomial_bytes(nsteps, mid) <= budget_bytes:
            lo = mid
        else:
            hi = mid - 1
    return "binomial", lo, ctx.adj_window_binomial_bytes(nsteps, lo)


def plane_probes(p, n=5, spacing=2, height_m=20.0):
    """n x n probes on the plane `height_m` above the vent, centred over the
    source patch (nx/3, ny/2, source80.inc:10-27), `spacing` cells apart; k
    increases downward and the vent is at k = nz."""
    ic, jc = p.nx // 3, p.ny // 2
    kp = p.nz - int(round(height_m / p.dz))
    h = n // 2
    return [(ic + spacing * a, jc + spacing * b, kp) for a in range(-h, n - h) for b in range(-h, n - h)]


def samples_every(p):
    """Steps between 1 Hz samples."""
    return max(1, int(round(1.0 / p.dt)))


def plan_segment(ctx, nsteps, budget_bytes):
    """The checkpoint interval whose checkpoints and tapes take the least
    HBM, and that amount; raises if even that exceeds budget_bytes.  Every
    interval recomputes one forward step per step, so the smallest
    footprint costs nothing in time (with memory for every step's tape the
    recompute could be dropped; at the sizes here it cannot)."""
    best = min(range(1, nsteps + 1), key=lambda s: ctx.adj_window_bytes(nsteps, s))
    need = ctx.adj_window_bytes(nsteps, best)
    if need > budget_bytes:
        raise ValueError("the window needs %.1f GB at best (checkpoint every %d steps), over the %.1f GB budget"
                         % (need / 1e9, best, budget_bytes / 1e9))
    return best, need


"""Trajectory log-likelihood of the pair-energy rate model and its gradient.

A recorded trajectory -- an initial lattice, the events it executed and the
waiting time before each -- has, under serial kinetic Monte Carlo, the
log-likelihood

    log L = sum_k [ log r(event k) - R_k dt_k ],   R_k = sum_c n_c(k) r_c,

where r_c is the rate of class c (a table entry) and n_c(k) the number of
sites in class c before event k.  The engine's adjoint kernel (mskg_adjoint)
replays the trajectory once and returns the class statistics S_c (events
drawn from class c) and A_c = sum_k n_c(k) dt_k, which do not depend on the
rates.  Then

    log L(theta) = sum_c S_c log r_c(theta) - A_c r_c(theta) + const,
    d log L / d theta = sum_c (S_c / r_c - A_c) d r_c / d theta,

and the second line is the backward of TrajectoryLogLikelihood below.  The
rates are the closed forms of models.pair_energy_tables (the pymulskips form
P0*exp(-E/kT), pymulskips/process.py:1082,1146) written in torch, so
autograd carries the gradient to the physical parameters:

    theta = (E_SiSi, E_SiC, E_CC, E2 [eV], flux_Si, flux_C [ML/s],
             nu_Si, nu_C [1/s])

for the two-species model (species 1 Si, 2 C) at a fixed temperature.
The placeholder deposition (1e-8, Get_Prob.f:183) does not depend on theta.
An isolated atom's rate is 1e5 R (Get_Prob.f:305), R the total rate when it
was evaluated, so it does depend on theta through R.  Its terms -- log(1e5 R)
per isolated-atom event and the integral of 1e5 R over the time the atom
stays isolated -- enter log L here as the constants the replay measured
(scalars 1 and 2 of the kernel), not as functions of theta: the gradient
leaves out sum_e [sel_e / R_e - 1e5 dur_e] dR_e/dtheta over the isolated
episodes e.  Each such episode ends after about one event (the atom's rate
is 1e5 times everything else), so the omitted terms average to zero; a
trajectory with isolated-atom events gets a warning and reports their count
(TrajectoryStats.isolated_events).
"""
import warnings

import numpy as np
import torch

from . import _lib
from .models import KB_EV

KEY_ADS, KEY_PH, KEY_ISO, KEY_N, ADJ_NS = _lib.KEY_ADS, _lib.KEY_PH, _lib.KEY_ISO, _lib.KEY_N, _lib.ADJ_NS
PARAMS = ("E_SiSi", "E_SiC", "E_CC", "E2", "flux_Si", "flux_C", "nu_Si", "nu_C")
PH_RATE = float(np.float32(1e-8))


def theta_sic(e_sisi=0.6, e_sic=1.1, e_cc=0.9, e2=0.05, flux_si=2.0, flux_c=1.8, nu_si=1e13, nu_c=1e13):
    return torch.tensor([e_sisi, e_sic, e_cc, e2, flux_si, flux_c, nu_si, nu_c], dtype=torch.float64)


def tables_from_theta(theta, temperature):
    """The engine's rate tables (ptrans_e/ptrans_d/second_factor) at theta:
    the FP64 values the likelihood's r_c(theta) reproduces class by class."""
    from . import models
    t = [float(v) for v in theta]
    pair = [[t[0], t[1]], [t[1], t[2]]]
    kt = KB_EV * temperature
    tab = models.pair_energy_tables(pair, [t[4], t[5]], temperature, 1.0, (14, 6))
    pe = tab["ptrans_e"]
    for s, nu in ((1, t[6]), (2, t[7])):
        idx = np.arange(s - 1, pe.size, 3)
        pe[idx] = pe[idx] * nu
    tab["second_factor"] = np.exp(-np.arange(13) * t[3] / kt)
    return tab


def class_rates(theta, temperature, used):
    """r_c(theta) for the class indices `used` (into the S numbering,
    0..ADJ_NS): evaporation keys, adsorption keys (total deposition rate of
    the site), deposition-by-species indices and the placeholder."""
    kt = KB_EV * temperature
    e = torch.stack([torch.stack([theta[0], theta[1]]), torch.stack([theta[1], theta[2]])])
    flux = theta[4:6]
    nu = theta[6:8]
    out = []
    for c in used:
        c = int(c)
        if c < KEY_ADS:
            s, n1, n2, n3, m = c % 3 + 1, (c // 3) % 4, (c // 12) % 4, (c // 48) % 4, c // 192
            if s > 2 or n3 or not 1 <= n1 + n2 + n3 <= 3:
                raise ValueError(f"class {c} is outside the two-species model")
            energy = n1 * e[s - 1, 0] + n2 * e[s - 1, 1] + m * theta[3]
            out.append(nu[s - 1] * torch.exp(-energy / kt))
        elif c < KEY_PH:
            out.append(flux.sum())  # adsorption site: deposition of either species
        elif c == KEY_PH:
            out.append(torch.tensor(PH_RATE, dtype=torch.float64))
        elif KEY_N <= c < ADJ_NS:
            s = (c - KEY_N) // 3 + 1
            if s > 2:
                raise ValueError("deposition of a third species")
            out.append(flux[s - 1])
        else:
            raise ValueError(f"class {c} has no closed form")
    return torch.stack(out)


class TrajectoryLogLikelihood(torch.autograd.Function):
    """log L = sum_c S_c log r_c - A_c r_c over the classes of a trajectory,
    with the hand-written backward dlogL/dr_c = S_c / r_c - A_c.  S and A
    come from the adjoint kernel (they do not depend on r)."""

    @staticmethod
    def forward(ctx, r, S, A):
        ctx.save_for_backward(r, S, A)
        return (S * torch.log(r) - A * r).sum()

    @staticmethod
    def backward(ctx, g):
        r, S, A = ctx.saved_tensors
        return g * (S / r - A), None, None


class TrajectoryStats:
    """Class statistics of one or more trajectories (summed), as returned by
    the adjoint kernel or the oracle: S (ADJ_NS,), A (KEY_N,), scalars."""

    def __init__(self, S, A, scalars):
        S = np.atleast_2d(S).sum(axis=0)
        A = np.atleast_2d(A).sum(axis=0)
        sc = np.atleast_2d(scalars).sum(axis=0)
        self.isolated_events = int(S[KEY_ISO])
        if S[KEY_ISO] or sc[1]:
            self.iso_const = sc[2] - sc[1]  # isolated atoms: taken as constant in theta
        else:
            self.iso_const = 0.0
        if self.isolated_events:
            warnings.warn(f"{self.isolated_events} isolated-atom events: their log-likelihood terms are taken as "
                          "constant in theta (their rate 1e5 R depends on theta through R)", RuntimeWarning)
        a = np.zeros(ADJ_NS)
        a[:KEY_N] = A
        a[KEY_ISO] = 0.0
        s = S.copy()
        s[KEY_ISO] = 0.0
        self.used = np.nonzero((s != 0) | (a != 0))[0]
        self.S = torch.tensor(s[self.used], dtype=torch.float64)
        self.A = torch.tensor(a[self.used], dtype=torch.float64)
        self.sum_dt = sc[0]


def log_likelihood(theta, stats, temperature):
    """Differentiable log L(theta) of the trajectories summarised in stats."""
    r = class_rates(theta, temperature, stats.used)
    return TrajectoryLogLikelihood.apply(r, stats.S, stats.A) + stats.iso_const


def gradient(theta, stats, temperature):
    th = theta.detach().clone().requires_grad_(True)
    ll = log_likelihood(th, stats, temperature)
    (g,) = torch.autograd.grad(ll, th)
    return ll.detach(), g


def trajectory_stats(engines, logs, dts, every=0):
    """Run the adjoint kernel on engines loaded with the initial states; with
    every > 0 also the class-population checkpoints after every `every`
    events (one int32 array (len // every, KEY_N) per trajectory)."""
    from .engine import adjoint_stats
    return adjoint_stats(engines, logs, dts, every=every)


def waiting_times(log, t0=0.0):
    """dt_k from a log's clock (time after each event)."""
    t = np.asarray(log["time"], np.float64)
    return np.diff(np.concatenate([[t0], t]))


def record_segmented(make_engine, seeds, n_events, segment):
    """Record one serial trajectory per seed (all in one batched launch per
    segment) and checkpoint each trajectory every `segment` events.

    make_engine(seed) returns a loaded engine (its state is the trajectory's
    start); it is also called with seed=None for the checkpoint engines,
    which receive their state by copy.  Returns (logs, dts, segments) where
    segments[i] is the list of (engine at the segment start, log slice,
    waiting-time slice) of trajectory i; the segments replay independently
    (mskg_adjoint) and their statistics add up to the trajectory's."""
    from .engine import run_serial
    gens = []
    for s in seeds:
        g = make_engine(s)
        g.track_classes()
        g.set_event_log(n_events)
        gens.append(g)
    t0 = [g.counters(lists=False)[1] for g in gens]
    starts = [[] for _ in gens]
    done = 0
    while done < n_events:
        k = min(segment, n_events - done)
        for g, st in zip(gens, starts):
            c = make_engine(None)
            c.track_classes()
            c.copy_state_from(g)
            st.append((c, done, done + k))
        run_serial(gens, [k] * len(gens))
        done += k
    logs = [g.event_log() for g in gens]
    for g in gens:
        g.close()
    dts = [waiting_times(lg, t) for lg, t in zip(logs, t0)]
    segments = [[(c, logs[i][a:b], dts[i][a:b]) for c, a, b in starts[i]] for i in range(len(gens))]
    return logs, dts, segments


def segmented_stats(segments, every=0):
    """Adjoint statistics of segmented trajectories in one launch; returns
    per-trajectory (S, A, scalars) summed over each trajectory's segments.
    With every > 0 (a divisor of every segment's length) it also returns,
    per trajectory, the class populations after every `every` events along
    the whole trajectory (the segments' checkpoints in order), as
    (per_trajectory, checkpoints)."""
    flat = [seg for traj in segments for seg in traj]
    if every > 0:
        for _, lg, _ in flat:
            if len(lg) % every:
                raise ValueError(f"a segment of {len(lg)} events is not a multiple of the checkpoint interval {every}")
        S, A, sc, hist = trajectory_stats([c for c, _, _ in flat], [lg for _, lg, _ in flat],
                                          [d for _, _, d in flat], every=every)
    else:
        S, A, sc = trajectory_stats([c for c, _, _ in flat], [lg for _, lg, _ in flat], [d for _, _, d in flat])
    out, ck, k = [], [], 0
    for traj in segments:
        m = len(traj)
        out.append((S[k:k + m].sum(axis=0), A[k:k + m].sum(axis=0), sc[k:k + m].sum(axis=0)))
        if every > 0:
            ck.append(np.concatenate(hist[k:k + m], axis=0))
        k += m
    return (out, ck) if every > 0 else out

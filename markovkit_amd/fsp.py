"""Chemical reaction networks on a box of counts: the host side of finite-state-projection
(FSP) solves on the device.

``ReactionNetwork`` follows include/markovkit/reaction_network.hpp:11-40: a reaction moves x to
x + change at rate(x), and a transition is kept when its rate is positive and no count goes
negative. ``box_generator`` tabulates the generator of the network restricted to the box
0 <= x_s <= bounds[s] (the reference's ``box`` region, fsp.hpp:26-35), with the reference's
boundary rule (generator.hpp:27-44): a transition that leaves the box adds its rate to the
state's leak, and with the absorbing boundary also to its exit rate. States are ordered with
species 0 most significant: index = sum_s x_s * prod_{r > s} (bounds[r] + 1).

Rates are evaluated vectorised over all states (numpy); they must be elementwise functions of
the count columns.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, List, Sequence

import numpy as np


@dataclass
class Reaction:
    change: Sequence[int]
    rate: Callable[[np.ndarray], np.ndarray]  # counts (n_states x species) -> rates (n_states,)


@dataclass
class ReactionNetwork:
    reactions: List[Reaction]

    @property
    def species(self) -> int:
        return len(self.reactions[0].change)


def box_states(bounds: Sequence[int]) -> np.ndarray:
    """All states of the box, species 0 most significant (n x S, int64)."""
    axes = [np.arange(b + 1, dtype=np.int64) for b in bounds]
    grid = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.ravel() for g in grid], axis=1)


def box_generator(network: ReactionNetwork, bounds: Sequence[int], absorbing: bool = True):
    """(first, to, rate, exit, leak): source-major jump lists of the network's generator on
    the box, jumps of a source in reaction order, as mk_generator_create_jumps takes them."""
    bounds = np.asarray(bounds, np.int64)
    X = box_states(bounds)
    n, S = X.shape
    stride = np.ones(S, np.int64)
    for s in range(S - 2, -1, -1):
        stride[s] = stride[s + 1] * (bounds[s + 1] + 1)
    exit_ = np.zeros(n)
    leak = np.zeros(n)
    targets, rates, kept = [], [], []
    for r in network.reactions:
        change = np.asarray(r.change, np.int64)
        rate = np.asarray(r.rate(X), np.float64) * np.ones(n)
        Y = X + change
        valid = (rate > 0.0) & np.all(Y >= 0, axis=1)
        inside = valid & np.all(Y <= bounds, axis=1)
        outside = valid & ~inside
        leak[outside] += rate[outside]
        exit_[inside] += rate[inside]
        targets.append(np.where(inside, Y @ stride, -1))
        rates.append(np.where(inside, rate, 0.0))
        kept.append(inside)
    if absorbing:
        exit_ += leak
    T = np.stack(targets, axis=1)  # n x reactions
    R = np.stack(rates, axis=1)
    K = np.stack(kept, axis=1)
    counts = K.sum(axis=1)
    first = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=first[1:])
    to = T[K]       # row-major over (state, reaction): source-major, reaction order
    rate = R[K]
    return first, to.astype(np.int64), rate, exit_, leak


def toggle_switch(bound: int = 200) -> ReactionNetwork:
    """Genetic toggle switch: U and V each born at 50 / (1 + (other / 10)^2) and dying at
    1.0 per molecule."""
    return ReactionNetwork([
        Reaction([1, 0], lambda x: 50.0 / (1.0 + (x[:, 1] / 10.0) ** 2)),
        Reaction([-1, 0], lambda x: 1.0 * x[:, 0]),
        Reaction([0, 1], lambda x: 50.0 / (1.0 + (x[:, 0] / 10.0) ** 2)),
        Reaction([0, -1], lambda x: 1.0 * x[:, 1]),
    ])


def birth_death(birth: float, death: float) -> ReactionNetwork:
    """One species born at a constant rate and dying at death per molecule."""
    return ReactionNetwork([
        Reaction([1], lambda x: np.full(len(x), birth)),
        Reaction([-1], lambda x: death * x[:, 0]),
    ])


# ---------------------------------------------------------------- finite state projection
# Distributions on a growing support and the two FSP solvers of include/markovkit/fsp.hpp,
# with every exponential on the device (markovkit_amd.expv). Regions are vectorised: a region
# maps an (m x S) array of states to m booleans (the reference's region maps one state).


class Distribution:
    """include/markovkit/distribution.hpp:11-100: states in insertion order, their
    probabilities, and a lookup from state to index."""

    def __init__(self, x=None):
        self._states: List[tuple] = []
        self._p: List[float] = []
        self._index = {}
        if x is not None:
            self.add(x, 1.0)

    def size(self) -> int:
        return len(self._states)

    def __len__(self) -> int:
        return len(self._states)

    def __getitem__(self, i: int) -> tuple:
        return self._states[i]

    def states(self) -> List[tuple]:
        return list(self._states)

    def states_array(self) -> np.ndarray:
        return np.asarray(self._states, np.int64).reshape(len(self._states), -1)

    def probabilities(self) -> np.ndarray:
        return np.asarray(self._p, np.float64)

    def set_probabilities(self, p) -> None:
        p = np.asarray(p, np.float64)
        if p.shape != (self.size(),):
            raise ValueError("one probability per state")
        self._p = [float(v) for v in p]

    def find(self, x):
        return self._index.get(tuple(int(c) for c in x))

    def add(self, x, mass: float = 0.0) -> None:
        """add x with probability `mass`, unless it is already present"""
        key = tuple(int(c) for c in x)
        if key not in self._index:
            self._index[key] = len(self._states)
            self._states.append(key)
            self._p.append(float(mass))

    def remove_if(self, drop) -> None:
        """remove every state i with drop[i] (a boolean array or a predicate of the index),
        keeping the rest in order"""
        if callable(drop):
            drop = [bool(drop(i)) for i in range(self.size())]
        keep = [i for i in range(self.size()) if not drop[i]]
        self._states = [self._states[i] for i in keep]
        self._p = [self._p[i] for i in keep]
        self._index = {x: i for i, x in enumerate(self._states)}

    def normalize(self) -> float:
        """clip negative roundoff and rescale to total probability one; returns the total
        before rescaling"""
        p = np.maximum(self.probabilities(), 0.0)
        total = float(p.sum())
        if total > 0.0:
            p = p / total
        self._p = [float(v) for v in p]
        return total

    def means(self) -> np.ndarray:
        if not self._states:
            return np.zeros(0)
        return self.probabilities() @ self.states_array().astype(np.float64)

    def marginal(self, species: int) -> dict:
        out = {}
        for x, p in zip(self._states, self._p):
            out[x[species]] = out.get(x[species], 0.0) + p
        return out

    def copy(self) -> "Distribution":
        d = Distribution()
        d._states = list(self._states)
        d._p = list(self._p)
        d._index = dict(self._index)
        return d


def everywhere():
    """the region with every state (fsp.hpp:21-24)"""
    return lambda Y: np.ones(len(Y), bool)


def box(bounds):
    """0 <= x_s <= bounds[s], or <= bounds for every species when it is a number
    (fsp.hpp:26-43)"""
    b = np.asarray(bounds, np.int64)
    return lambda Y: np.all((Y >= 0) & (Y <= b), axis=1)


def _transitions(network: ReactionNetwork, X: np.ndarray):
    """targets (n x R x S), rates (n x R) and validity (n x R) of every reaction at every state:
    valid when the rate is positive and no count goes negative (reaction_network.hpp:20-38)"""
    n = len(X)
    Ys, Rs, Vs = [], [], []
    for r in network.reactions:
        rate = np.asarray(r.rate(X), np.float64) * np.ones(n)
        Y = X + np.asarray(r.change, np.int64)
        Ys.append(Y)
        Rs.append(rate)
        Vs.append((rate > 0.0) & np.all(Y >= 0, axis=1))
    return np.stack(Ys, axis=1), np.stack(Rs, axis=1), np.stack(Vs, axis=1)


def expand(p: Distribution, network: ReactionNetwork, inside=None, layers: int = 1) -> None:
    """add `layers` rings of neighbours inside the region; after the first ring only the states
    added by the previous ring can have new neighbours (fsp.hpp:46-60). States are added in
    the reference's order: by source, then by reaction."""
    inside = inside or everywhere()
    begin = 0
    for _ in range(layers):
        end = p.size()
        if end == begin:
            break
        X = p.states_array()[begin:end]
        begin = end
        Y, _, ok = _transitions(network, X)
        S = X.shape[1]
        cand = Y.reshape(-1, S)[ok.reshape(-1)]
        if len(cand):
            cand = cand[inside(cand)]
        for y in cand:
            p.add(y)


def expand_sampled(p: Distribution, network: ReactionNetwork, inside, layers: int,
                   rng: np.random.Generator) -> None:
    """add one neighbour per state and layer, chosen with probability proportional to its rate
    (fsp.hpp:62-84). The random stream is numpy's, not the reference's num::rng."""
    inside = inside or everywhere()
    for _ in range(layers):
        n = p.size()
        X = p.states_array()
        Y, R, ok = _transitions(network, X)
        for i in range(n):
            rates = np.where(ok[i], R[i], 0.0)
            if not ok[i].any():
                continue
            c = np.cumsum(rates)
            k = int(np.searchsorted(c, rng.random() * c[-1], side="right"))
            k = min(k, len(c) - 1)
            y = Y[i, k]
            if inside(y[None, :])[0]:
                p.add(y)


def support_generator(p: Distribution, network: ReactionNetwork, absorbing: bool):
    """(first, to, rate, exit, leak) of the network's generator on the support of p
    (generator.hpp:27-44 with the distribution as the state space): jumps source by source in
    reaction order; a jump leaving the support adds its rate to the leak, and with the
    absorbing boundary also to the exit rate."""
    X = p.states_array()
    n = len(X)
    Y, R, ok = _transitions(network, X)
    Rn = R.shape[1]
    idx = np.full((n, Rn), -1, np.int64)
    for i in range(n):
        for r in range(Rn):
            if ok[i, r]:
                j = p.find(Y[i, r])
                idx[i, r] = -1 if j is None else j
    inside = ok & (idx >= 0)
    outside = ok & (idx < 0)
    exit_ = np.zeros(n)
    leak = np.zeros(n)
    for r in range(Rn):  # per state, in reaction order
        exit_ += np.where(inside[:, r], R[:, r], 0.0)
        leak += np.where(outside[:, r], R[:, r], 0.0)
    if absorbing:
        exit_ = exit_ + leak
    first = np.zeros(n + 1, np.int64)
    np.cumsum(inside.sum(axis=1), out=first[1:])
    return first, idx[inside], R[inside], exit_, leak


@dataclass
class FspSolution:
    """the distribution at each saved time (fsp.hpp:86-107)"""
    times: List[float]
    snapshots: List[Distribution]

    def means(self) -> np.ndarray:
        """means[k, s]: the mean count of species s at times[k]"""
        return np.stack([d.means() for d in self.snapshots])

    def save(self, t: float, p: Distribution) -> None:
        self.times.append(float(t))
        self.snapshots.append(p.copy())


@dataclass
class AdaptiveFspOptions:
    """adaptive_fsp_options (fsp.hpp:109-122), same fields and defaults"""
    inside: Callable = None
    step_scale: float = 1.0
    prune_fraction: float = 0.1
    flux_tolerance: float = 1e-6
    min_probability: float = 0.0
    layers: int = 1
    sampled_expansion: bool = False
    seed: int = 42
    save_every: int = 1000
    max_steps: int = 1 << 28
    krylov_dim: int = 30
    krylov_tol: float = 1e-8


@dataclass
class KrylovFspOptions:
    """krylov_fsp_options (fsp.hpp:190-203), same fields and defaults"""
    inside: Callable = None
    tolerance: float = 1e-6
    initial_step: float = 1.0
    step_growth: float = 2.0
    drop_probability: float = 1e-10
    drop_derivative: float = 1e-16
    layers: int = 1
    save_every: int = 1000
    max_steps: int = 1 << 28
    krylov_dim: int = 30
    krylov_tol: float = 1e-8


def _device_generator(p: Distribution, network: ReactionNetwork, absorbing: bool):
    from .api import Generator
    first, to, rate, exit_, leak = support_generator(p, network, absorbing)
    return Generator.from_jumps(first, to, rate, exit_, leak), exit_, leak


def _round_half_up(x: float) -> int:
    return int(np.floor(x + 0.5))


def solve_adaptive_fsp(network: ReactionNetwork, initial, final_time: float,
                       options: AdaptiveFspOptions = None) -> FspSolution:
    """FSP with flux-controlled steps (fsp.hpp:124-187): expand, propagate with expv on the
    device over the reflecting generator of the support, prune the least likely states whose
    probability flux is negligible, renormalise. The `prune_fraction * n` smallest
    probabilities are the pruning candidates, ties taken in index order."""
    from .api import expv
    o = options or AdaptiveFspOptions()
    p = Distribution(initial)
    sol = FspSolution([], [])
    sol.save(0.0, p)
    rng = np.random.default_rng(o.seed)
    t = 0.0
    step = 1
    while t < final_time and step <= o.max_steps:
        if o.sampled_expansion:
            expand_sampled(p, network, o.inside, o.layers, rng)
        else:
            expand(p, network, o.inside, o.layers)
        Q, exit_, leak = _device_generator(p, network, absorbing=False)
        prob = p.probabilities()
        activity = 0.0
        for i in range(len(prob)):
            activity += prob[i] * exit_[i]
        dt = min(o.step_scale / activity, final_time - t) if activity > 0.0 else final_time - t
        nxt, _ = expv(Q, dt, prob, o.krylov_dim, o.krylov_tol)
        prob = np.maximum(0.0, nxt)
        n = len(prob)
        drop = np.zeros(n, bool)
        if n > 1:
            flux = prob * (exit_ + leak)
            total_flux = float(np.sum(flux))
            candidates = min(_round_half_up(o.prune_fraction * n), n - 1)
            for i in np.argsort(prob, kind="stable")[:candidates]:
                drop[i] = flux[i] < o.flux_tolerance * total_flux
            drop |= prob < o.min_probability
        p.set_probabilities(prob)
        p.remove_if(drop)
        p.normalize()
        t += dt
        if step % o.save_every == 0 or t >= final_time:
            sol.save(t, p)
        step += 1
    if sol.times[-1] < t:
        sol.save(t, p)
    return sol


def solve_krylov_fsp(network: ReactionNetwork, initial, final_time: float,
                     options: KrylovFspOptions = None) -> FspSolution:
    """The Krylov FSP of Vo and Sidje (fsp.hpp:205-254): a step of length tau over the
    absorbing generator of the support is accepted when the absorbed probability stays within
    its share tolerance * (t + tau) / final_time, and halved otherwise; states below
    drop_probability whose probability changes slower than drop_derivative are dropped. Every
    exponential and generator apply runs on the device."""
    from .api import expv
    o = options or KrylovFspOptions()
    p = Distribution(initial)
    sol = FspSolution([], [])
    sol.save(0.0, p)
    t = 0.0
    tau = o.initial_step
    accepted = 0
    attempt = 0
    while t < final_time and attempt < o.max_steps:
        attempt += 1
        tau = min(tau, final_time - t)
        expand(p, network, o.inside, o.layers)
        Q, _, _ = _device_generator(p, network, absorbing=True)
        prob = p.probabilities()
        w, _ = expv(Q, tau, prob, o.krylov_dim, o.krylov_tol)
        mass = float(np.sum(np.maximum(0.0, w)))
        if mass < 1.0 - (o.tolerance * (t + tau) / final_time) and tau > 1e-14:
            tau *= 0.5
            continue
        dp = Q.apply(w)
        prob = np.maximum(0.0, w)
        p.set_probabilities(prob)
        p.remove_if((prob < o.drop_probability) & (np.abs(dp) < o.drop_derivative))
        p.normalize()
        t += tau
        accepted += 1
        if accepted % o.save_every == 0 or t >= final_time:
            sol.save(t, p)
        tau = min(tau * o.step_growth, final_time - t)
    if sol.times[-1] < t:
        sol.save(t, p)
    return sol

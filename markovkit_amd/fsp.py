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

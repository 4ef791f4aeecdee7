"""Host logic of the FSP solvers (markovkit_amd.fsp): the reference's own checks of
distribution, expand and the support generator (tests/test_fsp.cpp:15-66), plus explicit
expected orders of the states that expand adds, and the generator's jump lists against a dense
matrix built state by state from reaction_network::transitions."""
import numpy as np

from markovkit_amd.fsp import (Distribution, Reaction, ReactionNetwork, birth_death, box, expand,
                               support_generator, toggle_switch)


def test_distribution():
    p = Distribution()
    for x, m in [((0,), 0.1), ((1,), 0.2), ((2,), 0.3), ((3,), 0.4), ((3,), 9.0)]:
        p.add(x, m)
    assert p.size() == 4  # add ignores a state already present
    assert abs(p.means()[0] - 2.0) < 1e-15
    p.remove_if(lambda i: i % 2 == 0)
    assert p.states() == [(1,), (3,)]
    assert list(p.probabilities()) == [0.2, 0.4]
    assert p.find((1,)) == 0 and p.find((3,)) == 1 and p.find((2,)) is None
    assert abs(p.normalize() - 0.6) < 1e-15
    assert abs(p.marginal(0)[3] - 2.0 / 3.0) < 1e-15
    p.remove_if(lambda i: True)
    assert p.size() == 0 and p.find((1,)) is None


def dense(first, to, rate, exit_):
    n = len(exit_)
    Q = np.diag(-exit_)
    for j in range(n):
        for q in range(first[j], first[j + 1]):
            Q[to[q], j] += rate[q]
    return Q


def test_expand_and_generator():
    model = birth_death(5.0, 0.1)
    p = Distribution((0,))
    expand(p, model, box(10), 2)
    assert p.size() == 3  # two layers from {0} reach {1} and {2}
    first, to, rate, exit_, leak = support_generator(p, model, absorbing=False)
    assert len(exit_) == 3
    assert exit_[0] == 5.0
    assert abs(leak[2] - 5.0) < 1e-15
    _, _, _, exit_a, _ = support_generator(p, model, absorbing=True)
    assert abs(exit_a[2] - 5.2) < 1e-15
    x = np.array([0.25, 0.5, 0.25])
    assert abs((dense(first, to, rate, exit_) @ x).sum()) < 1e-15  # reflecting conserves


def test_expand_order_and_region():
    net = toggle_switch()
    p = Distribution((0, 0))
    expand(p, net, box([200, 200]), 1)
    # reactions in order: U birth, U death (invalid at 0), V birth, V death (invalid)
    assert p.states() == [(0, 0), (1, 0), (0, 1)]
    expand(p, net, box([1, 200]), 1)
    # every state is a source again on the next call; (2, 0) is outside U <= 1
    assert p.states() == [(0, 0), (1, 0), (0, 1), (1, 1), (0, 2)]


def test_support_generator_against_transitions():
    net = ReactionNetwork([
        Reaction([1, 0], lambda x: 2.0 + 0.5 * x[:, 1]),
        Reaction([-1, 1], lambda x: 1.5 * x[:, 0]),
        Reaction([0, -1], lambda x: np.where(x[:, 1] > 2, 0.0, 0.3 * x[:, 1])),
    ])
    p = Distribution((1, 1))
    expand(p, net, box([3, 3]), 3)
    for absorbing in (False, True):
        first, to, rate, exit_, leak = support_generator(p, net, absorbing)
        Q = dense(first, to, rate, exit_)
        # state by state from the transition rule
        n = p.size()
        Qr = np.zeros((n, n))
        for j, x in enumerate(p.states()):
            out = 0.0
            lk = 0.0
            for r in net.reactions:
                rt = float(r.rate(np.array([x]))[0])
                y = tuple(np.array(x) + np.array(r.change))
                if not rt > 0 or min(y) < 0:
                    continue
                i = p.find(y)
                if i is None:
                    lk += rt
                else:
                    Qr[i, j] += rt
                    out += rt
            Qr[j, j] -= out + (lk if absorbing else 0.0)
            assert leak[j] == lk
        assert np.array_equal(Q, Qr)

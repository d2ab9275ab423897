"""The two exclusion rules of the reference, selected per context (paddi_set_exclusion).

Its CPU path pairs row i with column j iff |i - 1 - j| >= m (src/cpu_kernels/cpu_kernels.cpp:57,
the default); its GPU kernels iff |i - j| >= m - 1 (src/PD3_kernels.cuh:70), which also pairs
i - j = m - 1 and i - j = m.  The check for both is the brute force on explicitly
z-normalised windows under the same rule (oracle.brute_profile).

The planted series shows where the rules differ: window a, which holds a spike, is copied to
window i = a + m0 - 1, sharing one point with it.  Row i's exact copy sits at i - j = m0 - 1,
a pair under the symmetric rule only, and so do the best matches of its neighbours that hold
the second spike; at m0 the CPU rule's top-3 holds one of those rows and the symmetric
rule's does not.
"""
import numpy as np
import pytest

from oracle import oracle as O
from paddi_amd.engine import PaddiError

pytestmark = pytest.mark.gpu
RTOL = 1e-4
M0 = 24
MINL, MAXL, STEP, K = M0 - 2, M0 + 1, 1, 3


def planted(n=3000, a=1200, seed=21):
    rng = np.random.default_rng(seed)
    x = np.cumsum(rng.standard_normal(n))
    x[a + 5] += 6.0 * x[a - 64:a].std()
    x[a + M0 - 1] = x[a]  # the shared point, so the copy leaves window a unchanged
    x[a + M0 - 1:a + 2 * M0 - 1] = x[a:a + M0].copy()
    return x.astype(np.float32), a + M0 - 1


def random_walk(n=4000, seed=22):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.standard_normal(n)).astype(np.float32)


def brute_topk(nn, m, k):
    order = sorted(range(len(nn)), key=lambda i: (-nn[i], i))
    acc = []
    for i in order:
        if len(acc) == k:
            break
        if np.isfinite(nn[i]) and all(abs(i - a) >= m for a in acc):
            acc.append(i)
    return acc, [float(nn[i]) for i in acc]


_REF = {}


def reference(name, m, rule):
    key = (name, m, rule)
    if key not in _REF:
        x = planted()[0] if name == "planted" else random_walk()
        _REF[key] = O.brute_profile(x, m, symmetric=(rule == "symmetric"))[0]
    return _REF[key]


SERIES = {"planted": lambda: planted()[0], "rw": random_walk}


@pytest.fixture
def rule_ctx(ctx):
    yield ctx
    ctx.set_exclusion("cpu")


def test_rule_setting(rule_ctx):
    ctx = rule_ctx
    assert ctx.get_exclusion() == "cpu"
    ctx.set_exclusion("symmetric")
    assert ctx.get_exclusion() == "symmetric"
    with pytest.raises(PaddiError):
        ctx.set_exclusion(2)
    assert ctx.get_exclusion() == "symmetric"
    ctx.set_exclusion("cpu")
    assert ctx.get_exclusion() == "cpu"


def test_the_rules_differ_on_the_planted_copy(rule_ctx):
    """At m0 the copy is a range discord at the CPU rule's median distance and is pruned at
    every r > 0 under the symmetric rule; each rule's flags and distances are its brute
    force's."""
    ctx = rule_ctx
    x, i = planted()
    cpu, sym = reference("planted", M0, "cpu"), reference("planted", M0, "symmetric")
    assert sym[i] < 1e-6 and cpu[i] > np.quantile(cpu, 0.5)
    assert brute_topk(cpu, M0, K)[0] != brute_topk(sym, M0, K)[0]
    for rule, ref in (("cpu", cpu), ("symmetric", sym)):
        ctx.set_exclusion(rule)
        for q in (0.5, 0.99):
            r2 = float(np.quantile(ref, q))
            flag, nn = ctx.range_discords(x, M0, r2)
            near = np.abs(ref - r2) <= 1e-6 * r2
            assert np.array_equal((flag == 1)[~near], (ref > r2)[~near]), (rule, q)
            sel = (flag == 1) & (ref > r2)
            assert (np.abs(nn[sel] - ref[sel]) / ref[sel]).max() <= RTOL, (rule, q)
        assert flag[i] == (1 if rule == "cpu" else 0)


@pytest.mark.parametrize("mode", ["plain", "batched", "learned"])
@pytest.mark.parametrize("name", ["planted", "rw"])
@pytest.mark.parametrize("rule", ["cpu", "symmetric"])
def test_forms_match_the_brute_force_under_each_rule(rule_ctx, rule, name, mode):
    ctx = rule_ctx
    x = SERIES[name]()
    ctx.set_exclusion(rule)
    got = ctx.discords(x, MINL, MAXL, STEP, K, mode=mode)
    for m in range(MINL, MAXL + 1, STEP):
        idx, d = brute_topk(reference(name, m, rule), m, K)
        mine = [g for g in got if g.length == m]
        assert [g.index for g in mine] == idx, (m, [(g.index, g.nn_dist2) for g in mine], idx, d)
        for g, v in zip(mine, d):
            assert abs(g.nn_dist2 - v) <= RTOL * v, (m, g, v)
    if name == "planted":
        cpu = brute_topk(reference(name, M0, "cpu"), M0, K)[0]
        sym = brute_topk(reference(name, M0, "symmetric"), M0, K)[0]
        assert cpu != sym

"""The oracle's BP against an independent statement of the reference's BP (bp_statement.py,
written from hbp_isp.cu), bit for bit after every sweep, and the statement's pieces against
what mathematics fixes.  The chain of evidence: closed forms pin the statement's primitives,
the statement pins the oracle's BP (planes, costs, beliefs and messages after every sweep of
every level, then the map), and the GPU parity tests pin the HIP kernels to the oracle.

The mutation tests plant, one at a time, the mistakes a restatement of hbp_isp.cu could
plausibly make (a swapped sample point, a wrong half-minimum split, a missing max(dc0, .01),
a missing tau_s cap, ...) into a copy of the statement and check that the comparison with the
oracle fails: an oracle carrying the same mistake would fail it the same way."""
import types

import numpy as np
import pytest

import bp_statement as S
import oracle_ffi as O
from hbp_amd import synth


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _setup(seed, W, H, D, ml, alpha=1000.1, tau_s=0.6, ucv=1, right=False, planes=(3, 8), module=S):
    left, right_img, _ = synth.make_pair(seed, W, H, D, planes)
    vo, wo = O.cost_volume(left, right_img, D, right_view=right)
    po = O.Pyramid(right_img if right else left)
    levels = [po.level(l) for l in range(po.levels)]
    bo = O.BP(po, vo, wo, D, alpha=alpha, tau_s=tau_s, min_level=ml, use_cost_vol=ucv)
    st = module.HBP(levels, vo, wo, W, H, D, alpha=alpha, tau_s=tau_s, min_level=ml, use_cost_vol=ucv)
    return po, bo, st


def _first_difference(bo, st, max_sweeps=None):
    """Steps both; returns None when every state and the map agree, else where they first differ."""
    n = len(st.sched) if max_sweeps is None else min(max_sweeps, len(st.sched))
    for i in range(n):
        assert bo.step(1) == 1 and st.step(1) == 1
        so, ss = bo.state(), st.state()
        assert so["level"] == ss["level"]
        for k in ("planes", "cost", "belief", "msg"):
            if not same_bits(so[k], ss[k]):
                return i, k
    if max_sweeps is None:
        do, plo = bo.map()
        ds, pls = st.map()
        if not (same_bits(do, ds) and same_bits(plo, pls)):
            return n, "map"
    return None


CASES = [
    dict(seed=2, W=32, H=24, D=16, ml=1),                       # 5 BP levels, 731 sweeps
    dict(seed=3, W=48, H=36, D=24, ml=2),
    dict(seed=4, W=24, H=20, D=16, ml=1, alpha=-3.5),           # negative lambda
    dict(seed=5, W=36, H=28, D=16, ml=1, ucv=0, alpha=2.0),     # the inlier cost (use_cost_vol = 0)
    dict(seed=6, W=30, H=22, D=12, ml=1, right=True),           # right view
    dict(seed=8, W=20, H=16, D=8, ml=1, tau_s=0.05),            # the tau_s cap binds often
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_oracle_bp_equals_the_statement_after_every_sweep(case):
    po, bo, st = _setup(**case)
    assert len(st.sched) == O.lib.orc_bp_total_sweeps(po.levels, case["ml"], case["D"])
    assert _first_difference(bo, st) is None


def test_the_compared_runs_reach_the_nan_paths():
    """Collinear RANSAC triples give NaN planes (nz = 0, hbp_isp.cu:540-547): the comparison
    covers their costs, messages (CUDA's min(float, float) is fminf, so a NaN phi is tau_s),
    ranks and the map."""
    _, bo, st = _setup(**CASES[0])
    seen = 0
    for _ in range(60):
        bo.step(1)
        st.step(1)
        seen += int(np.isnan(st.state()["planes"]).any())
    assert seen > 10


# ---------------------------------------------------------------- closed forms
def test_lcg_first_values_are_the_bsd_rand_sequence():
    """rand_lcg is the classic BSD / glibc TYPE_0 generator: from seed 1 its outputs are
    1103527590, 377401575, 662824084, 1147902781, 2035015474 (published known answers)."""
    s, out = 1, []
    for _ in range(5):
        s = S.lcg(s)
        out.append(s)
    assert out == [1103527590, 377401575, 662824084, 1147902781, 2035015474]
    # the uniform draw is s / 2^31 * f, in float
    s2, u = S.lcg_uniform(1, 8.0)
    assert s2 == 1103527590 and u == np.float32(np.float32(1103527590) / np.float32(2 ** 31)) * np.float32(8.0)


@pytest.mark.parametrize("abc", [(0.5, -0.25, 7.0), (0.0, 0.0, 3.0), (-1.0, 2.0, 100.0), (0.125, 0.375, -5.5)])
def test_plane_through_three_points_of_a_plane_is_that_plane(abc):
    a, b, c = (np.float32(v) for v in abc)
    pts = [(3, 4), (11, 5), (6, 13)]
    P4 = [np.array([x, y, a * x + b * y + c, 0], np.float32) for x, y in pts]
    got = S.fit_plane(*P4)
    assert np.array_equal(got[:3], np.array([a, b, c], np.float32))


def test_collinear_points_give_nan_not_the_fallback():
    """nz = 0: nz / nz is NaN and |NaN - 1| > FLT_EPSILON is false, so the reference's fallback
    (hbp_isp.cu:542-545) is not taken and the plane is not finite."""
    P4 = [np.array([x, 2 * x, 5.0, 0], np.float32) for x in (1, 2, 4)]
    got = S.fit_plane(*P4)
    assert not np.isfinite(got[:3]).all()


def test_three_diff_ids_are_distinct_and_in_range():
    for tn in range(1, 40):
        for p0 in range(tn):
            for p1 in range(max(tn - 1, 1)):
                for p2 in range(max(tn - 2, 1)):
                    q = S.three_diff(p0, p1, p2, tn)
                    assert all(0 <= v < tn for v in q)
                    if tn >= 3:
                        assert len(set(q)) == 3


def _one_edge_level(plane_s, plane_t, belief_t, c0_color=1.0, dpos=2.0, tau_s=0.6):
    """A two-node level with one edge each way, for the message rule alone."""
    st = types.SimpleNamespace()
    W = 10
    lv = types.SimpleNamespace(V=2, E=2, bd=np.array([[0, 1], [1, 0]]), bdl=np.array([3, 3]),
                               dc0=np.array([c0_color] * 2, np.float32), dc1=np.array([dpos] * 2, np.float32),
                               point=np.array([[0, 1, 10], [12, 13, 22]]), rev=np.array([1, 0]),
                               plane=np.zeros((S.K, 2, 4), np.float32), cost=np.zeros((S.K, 2), np.float32),
                               belief=np.zeros((S.K, 2), np.float32))
    for k in range(S.K):
        lv.plane[k][0] = plane_s[k]
        lv.plane[k][1] = plane_t[k]
        lv.belief[k][1] = belief_t[k]
    h = S.HBP.__new__(S.HBP)
    h.W, h.tau_s = W, np.float32(tau_s)
    h.lv = [lv]
    h.msg = np.zeros((2, S.K), np.float32)
    return h, lv


def test_message_of_identical_planes_is_the_two_half_minima():
    """phi = 0 when every plane is the same (w, which phi ignores, labels the slots): each slot's
    message is min over the first P of (belief_t - msg) plus min over the next P, then the mean
    over the 12 slots is taken off, and the belief of s is its cost plus that message."""
    pl = [np.array([0.25, -0.5, 3.0, k], np.float32) for k in range(S.K)]
    bel = np.array([5, 3, 9, 4, 8, 2, 7, 6, 1, 1, 1, 1], np.float32)
    h, lv = _one_edge_level(pl, pl, bel)
    h.belief_propagation(0)
    raw = np.float32(np.float32(np.float32(0) + bel[:4].min()) + bel[4:8].min())
    mean = np.float32(0)
    for _ in range(S.K):
        mean = np.float32(mean + raw)
    want = np.float32(raw - np.float32(mean / np.float32(S.K)))
    assert raw == np.float32(3 + 2)
    assert np.all(lv.belief[:, 0] == want)


def test_phi_is_capped_at_tau_s_and_c0_floors_dc0():
    """Far-apart planes: every phi is tau_s, so each half-minimum is min(belief) + c0 * tau_s
    with c0 = max(dc0, 0.01) * bdl; dc0 below 0.01 is raised to 0.01."""
    ps = [np.array([0, 0, 0, 0], np.float32)] * S.K
    pt = [np.array([0, 0, 1e6 * (k + 1), 0], np.float32) for k in range(S.K)]
    bel = np.arange(S.K, dtype=np.float32)
    for dc0, c0 in ((0.5, np.float32(0.5) * np.float32(3)), (1e-4, np.float32(0.01) * np.float32(3))):
        h, lv = _one_edge_level(ps, pt, bel, c0_color=dc0, tau_s=0.6)
        h.belief_propagation(0)
        # messages after normalisation are equal across slots (every slot sees the same values)
        m = lv.belief[:, 0]
        assert np.all(m == m[0])
        half = np.float32(bel[0] + c0 * np.float32(0.6))
        half2 = np.float32(bel[4] + c0 * np.float32(0.6))
        raw = np.float32(np.float32(0) + half) + half2
        mean = np.float32(0)
        for _ in range(S.K):
            mean = np.float32(mean + raw)
        assert m[0] == np.float32(raw - np.float32(mean / np.float32(S.K)))


def test_centroid_distance_below_one_counts_as_one():
    """dist_pos = max(dc1, 1) (hbp_isp.cu:330-333): superpixels whose centroids are closer than
    one pixel (which the compared pyramids never produce) smooth as if one pixel apart."""
    ps = [np.array([0, 0, 0.01 * k, 0], np.float32) for k in range(S.K)]
    pt = ps
    bel = np.arange(S.K, dtype=np.float32)

    def belief_of_s(dpos):
        h, lv = _one_edge_level(ps, pt, bel, dpos=dpos, tau_s=100.0)
        h.belief_propagation(0)
        return lv.belief[:, 0].copy()

    assert same_bits(belief_of_s(0.3), belief_of_s(1.0))
    assert same_bits(belief_of_s(0.0), belief_of_s(1.0))
    assert not same_bits(belief_of_s(2.0), belief_of_s(1.0))


def test_rank_ties_keep_the_lower_slots():
    """get_belief_order: if b_j < b_i then rank_i++ else rank_j++, so among equal beliefs the
    lower slot ranks first; select_belief keeps the P best in slot order."""
    pl = [np.array([0.0, 0.0, 1.0, k], np.float32) for k in range(S.K)]
    h, lv = _one_edge_level(pl, pl, np.zeros(S.K, np.float32))
    h.belief_propagation(0)
    assert np.all(lv.belief[:, 0] == lv.belief[0, 0])           # all tied
    assert [lv.plane[k][0][3] for k in range(S.P)] == [0, 1, 2, 3]


def test_select_keeps_the_p_lowest_beliefs_in_slot_order():
    """Distinct beliefs (the costs differ, the messages do not): slots with the P lowest
    beliefs are compacted to the front in their slot order, with their costs."""
    pl = [np.array([0.0, 0.0, 1.0, k], np.float32) for k in range(S.K)]
    h, lv = _one_edge_level(pl, pl, np.zeros(S.K, np.float32))
    cost = np.array([9, 2, 7, 11, 1, 8, 3, 10, 0.5, 6, 5, 4], np.float32)
    lv.cost[:, 0] = cost
    h.belief_propagation(0)
    keep = sorted(np.argsort(cost, kind="stable")[:S.P])
    assert [int(lv.plane[k][0][3]) for k in range(S.P)] == keep
    assert np.array_equal(lv.cost[:S.P, 0], cost[keep])


# ---------------------------------------------------------------- planted mistakes
SRC = open(S.__file__).read()
MUTANTS = [
    ("a sample point of t replaced by one of s", "pts = np.concatenate([B.point[s_], B.point[t_]], axis=1)",
     "pts = np.concatenate([B.point[s_], B.point[s_]], axis=1)"),
    ("x and y of the sample points swapped", "fx = (pts % W).astype(np.float32)", "fx = (pts // W).astype(np.float32)"),
    ("the half-minimum split one slot early", "if p == P:\n                        acc = acc + minimum",
     "if p == P - 1:\n                        acc = acc + minimum"),
    ("one minimum over all 2P slots", "if p == P:\n                        acc = acc + minimum",
     "if False:\n                        acc = acc + minimum"),
    ("no max(dc0, .01)", "dist_color = np.fmax(B.dc0, F(0.01))", "dist_color = B.dc0"),
    ("no tau_s cap", "phi = np.fmin((phi / F(3.0)) / dist_pos, self.tau_s)", "phi = (phi / F(3.0)) / dist_pos"),
    ("NaN phi kept (std::min instead of fminf)", "phi = np.fmin((phi / F(3.0)) / dist_pos, self.tau_s)",
     "phi = np.minimum((phi / F(3.0)) / dist_pos, self.tau_s)"),
    ("division by 3 dropped", "phi = np.fmin((phi / F(3.0)) / dist_pos, self.tau_s)",
     "phi = np.fmin(phi / dist_pos, self.tau_s)"),
    ("boundary length dropped from c0", "c0 = (dist_color * B.bdl.astype(np.float32)).astype(np.float32)",
     "c0 = dist_color.astype(np.float32)"),
    ("passing message without the old message", "msg[:, y] = B.belief[y][t_] - msg[:, y]",
     "msg[:, y] = B.belief[y][t_]"),
    ("no normalisation", "msg_t[:, k] = msg_t[:, k] - mean", "msg_t[:, k] = msg_t[:, k]"),
    ("normalised over 2P slots", "mean = mean / F(K)", "mean = mean / F(2 * P)"),
    ("ping-pong without the reverse edge", "msg[B.rev[e], :] = msg_t[e, :]", "msg[e, :] = msg_t[e, :]"),
    ("rank ties to the higher slot", "lt = bel[j] < bel[i]", "lt = bel[j] <= bel[i]"),
    ("messages compacted by the source's ranks", "if order[k, t] < P:\n                        msg[e, nb]",
     "if order[k, s_[e]] < P:\n                        msg[e, nb]"),
    ("belief sum from zero instead of the cost", "belief = B.cost.copy()                                    # aggregate",
     "belief = np.zeros_like(B.cost)                            # aggregate"),
    ("LCG multiplier off by one", "((1103515245 * s + 12345) & 0xFFFFFFFF) % M31",
     "((1103515246 * s + 12345) & 0xFFFFFFFF) % M31"),
    ("LCG without the mod 2^31", "((1103515245 * s + 12345) & 0xFFFFFFFF) % M31", "((1103515245 * s + 12345) & 0xFFFFFFFF)"),
    ("no distinct-id correction", "p0, p1, p2 = three_diff(p0, p1, p2, n)", "pass"),
    ("plane fit with the normal's sign", "return np.array([-nx, -ny,", "return np.array([nx, ny,"),
    ("sample points of the first triple", "if p == 3 * P - 1:\n                    B.point[v] = ids",
     "if p == 2 * P:\n                    B.point[v] = ids"),
    ("RANSAC on the WTA of the wrong pixel", "self.dwta[i], F(0)]", "self.dwta[(i + 1) % self.N], F(0)]"),
    ("random step without normalize_plane", "a, b, c = F(a / norm), F(b / norm), F(c / norm)", "pass"),
    ("random step's dz not added", "z = F(z + dz)", "z = z"),
    ("child candidate from the child's slot 1", "Hh.plane[p + 2 * P][h] = B.plane[0][c]",
     "Hh.plane[p + 2 * P][h] = B.plane[1][c]"),
    ("neighbour read through the edge's source", "t = int(Hh.bd[nb_pos + umod(s, nb_num), 1])",
     "t = int(Hh.bd[nb_pos + umod(s, nb_num), 0])"),
    ("neighbour seeds from the parent's own CSR", "self.s_neighbor[:V] = B.nb_pos[:V]",
     "self.s_neighbor[:Hh.V] = Hh.nb_pos[:Hh.V]"),
    ("prior copied from the parent's prior slot", "pr = Hh.plane[p][B.clus]", "pr = Hh.plane[p + P][B.clus]"),
    ("interpolation weights swapped", "cost = v * c0 + w * c1", "cost = w * c0 + v * c1"),
    ("dc0 with gamma 10", "GAMMA = 20.0", "GAMMA = 10.0"),
    ("random sweeps in every outer iteration", "rand_ok = not (i <= iter_bp // 4 * 3)", "rand_ok = True"),
]


@pytest.mark.parametrize("name,old,new", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_a_planted_mistake_breaks_the_comparison(name, old, new):
    assert SRC.count(old) == 1, f"mutation site not found once: {old!r}"
    mod = types.ModuleType("bp_mutant")
    exec(compile(SRC.replace(old, new), "bp_mutant", "exec"), mod.__dict__)
    _, bo, st = _setup(seed=4, W=24, H=20, D=16, ml=1, module=mod)
    assert _first_difference(bo, st) is not None, f"the comparison missed: {name}"

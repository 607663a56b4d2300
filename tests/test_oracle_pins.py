"""Pins for the CPU oracle: each check is chosen so that a plausible slip in
the restatement (a dropped term, a wrong sign or index, a transposed operand,
a mis-weighted sample) fails it.  The reference ships no golden vectors or
tests for this path, and cannot be compiled here (Eigen 3.4 and nvcc are
absent), so the pins are closed forms, published known-answer vectors,
independent implementations and exhaustive enumeration."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
import workloads as W

PI32 = np.float64(np.float32(np.pi))


# --------------------------------------------------------------------------
# kernel functions (include/kernel.h:31-59)
# --------------------------------------------------------------------------
def closed_form(kernel, dim, p, n, m, alpha, q):
    d = p - q
    r = np.sqrt((d * d).sum(1))
    tenth = np.float64(np.float32(0.1))
    if kernel == "gravity":
        return -m / (r + tenth)
    if kernel == "helmholtz_x":  # include/kernel.h:68-74
        return m * np.cos(4 * PI32 * r) / (r + tenth)
    if kernel == "helmholtz_y":  # include/kernel.h:77-83
        return m * np.sin(2 * PI32 * r) / (r + tenth)
    if kernel.startswith("gravity_grad_"):  # include/kernel.h:61-66, (q - p) component
        a = "xyz".index(kernel[-1])
        return -m / (r + tenth) ** 3 * (q[:, a] - p[:, a])
    if kernel == "wn":
        dn = (n * d).sum(1)
        if dim == 2:
            return m * dn / (2 * PI32 * (r * r + np.float64(np.float32(1e-5))))
        return m * dn / (4 * PI32 * (r ** 3 + np.float64(np.float32(1e-5))))
    return np.exp(-alpha * r)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", ["gravity", "wn", "smoothdist"])
def test_kernel_closed_forms(kernel, dim):
    rng = np.random.default_rng(1)
    n = 500
    p = rng.normal(size=(n, dim)).astype(np.float32).astype(np.float64)
    nv = rng.normal(size=(n, dim)).astype(np.float32).astype(np.float64)
    m = rng.uniform(0.5, 2, n).astype(np.float32).astype(np.float64)
    q = rng.normal(size=(n, dim)).astype(np.float32).astype(np.float64)
    alpha = 3.0
    o32, o64 = oracle.kernel_pairs(dim, kernel, p, nv, m, alpha, q)
    ref = closed_form(kernel, dim, p, nv, m, alpha, q)
    np.testing.assert_allclose(o64, ref, rtol=1e-13, atol=1e-300)
    den = np.maximum(np.abs(ref), 1e-3) if kernel == "wn" else np.abs(ref)
    assert (np.abs(o32 - ref) / den).max() < 2e-6


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", ["helmholtz_x", "helmholtz_y", "gravity_grad_x",
                                    "gravity_grad_y", "gravity_grad_z"])
def test_unbound_kernel_closed_forms(kernel, dim):
    """The kernels the reference defines but binds to no plugin
    (include/kernel.h:61-83), against float64 closed forms; fp32 error
    relative to the term's magnitude scale (cosine and the (q - p) factor
    have zeros)."""
    if kernel == "gravity_grad_z" and dim == 2:
        pytest.skip("no z component in 2D")
    rng = np.random.default_rng(2)
    n = 500
    p = rng.normal(size=(n, dim)).astype(np.float32).astype(np.float64)
    nv = rng.normal(size=(n, dim)).astype(np.float32).astype(np.float64)
    m = rng.uniform(0.5, 2, n).astype(np.float32).astype(np.float64)
    q = rng.normal(size=(n, dim)).astype(np.float32).astype(np.float64)
    o32, o64 = oracle.kernel_pairs(dim, kernel, p, nv, m, 1.0, q)
    ref = closed_form(kernel, dim, p, nv, m, 1.0, q)
    np.testing.assert_allclose(o64, ref, rtol=1e-12, atol=1e-14)
    r = np.sqrt(((p - q) ** 2).sum(1))
    scale = m / (r + 0.1) if kernel.startswith("helm") else m * r / (r + 0.1) ** 3
    assert (np.abs(o32 - ref) / scale).max() < 1e-5
    # the gradient components are the components of one vector: -m (q - p) / (d + 0.1)^3
    if kernel == "gravity_grad_x":
        comps = [oracle.kernel_pairs(dim, "gravity_grad_" + "xyz"[a], p, nv, m, 1.0, q)[1]
                 for a in range(dim)]
        g = np.stack(comps, 1)
        np.testing.assert_allclose(np.linalg.norm(g, axis=1), m * r / (r + 0.1) ** 3,
                                   rtol=1e-6)
        assert (np.einsum("ij,ij->i", g, q - p) <= 0).all()  # points from q towards p


def test_kernel_signs_and_special_values():
    # gravity is attractive (negative), WN of a dipole pointing away from the
    # query is positive, smooth distance is 1 at zero distance
    p = np.array([[1.0, 0.0, 0.0]])
    q = np.zeros((1, 3))
    n = np.array([[1.0, 0.0, 0.0]])
    g32, _ = oracle.kernel_pairs(3, "gravity", p, n, [2.0], 1.0, q)
    w32, _ = oracle.kernel_pairs(3, "wn", p, n, [1.0], 1.0, q)
    s32, _ = oracle.kernel_pairs(3, "smoothdist", q, n, [1.0], 5.0, q)
    assert np.isclose(g32[0], -2.0 / 1.1, rtol=1e-6)
    assert np.isclose(w32[0], 1.0 / (4 * np.pi * (1 + 1e-5)), rtol=1e-6)
    assert s32[0] == 1.0


# --------------------------------------------------------------------------
# Philox4x32-10
# --------------------------------------------------------------------------
KAT = [  # Random123 known-answer vectors (Salmon et al., SC'11 distribution)
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert oracle.philox(ctr, key).tolist() == expect


TORCH_PHILOX = "/usr/local/lib/python3.10/dist-packages/torch/include/ATen/core/PhiloxRNGEngine.h"


@pytest.mark.skipif(not os.path.exists(TORCH_PHILOX) or shutil.which("g++") is None,
                    reason="PyTorch's Philox header or g++ not available")
def test_philox_matches_pytorch_engine():
    """PyTorch's PhiloxRNGEngine (BSD) is an independent implementation of the
    same generator: its 4-word output for (seed, subsequence, offset) is
    Philox4x32-10 of counter {offset_lo, offset_hi, subseq_lo, subseq_hi}
    under key {seed_lo, seed_hi}."""
    inc = os.path.dirname(os.path.dirname(os.path.dirname(TORCH_PHILOX)))
    src = r'''
#include <ATen/core/PhiloxRNGEngine.h>
#include <cstdio>
int main() {
  unsigned long long seeds[3] = {0ull, 12345ull, 0xfedcba9876543210ull};
  for (auto s : seeds) for (unsigned long long sub = 0; sub < 3; ++sub) {
    at::Philox4_32 eng(s, sub, 0);
    for (int k = 0; k < 4; ++k) printf("%u ", eng());
    printf("\n");
  }
}'''
    with tempfile.TemporaryDirectory() as d:
        cpp, exe = os.path.join(d, "t.cpp"), os.path.join(d, "t")
        open(cpp, "w").write(src)
        r = subprocess.run(["g++", "-std=c++17", "-I", inc, cpp, "-o", exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("PyTorch header does not compile standalone: " + r.stderr[-300:])
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [list(map(int, l.split())) for l in out if l.strip()]
    i = 0
    for s in (0, 12345, 0xfedcba9876543210):
        for sub in range(3):
            ctr = [0, 0, sub & 0xffffffff, sub >> 32]
            key = [s & 0xffffffff, s >> 32]
            assert oracle.philox(ctr, key).tolist() == rows[i]
            i += 1


def test_path_stream_layout():
    seed = 0x123456789
    w = oracle.path_words(seed, group=7, subdomain=3, sample=5, nwords=12)
    for b in range(3):
        blk = oracle.philox([5, 3, 7, b], [seed & 0xffffffff, seed >> 32])
        assert w[4 * b:4 * b + 4].tolist() == blk.tolist()


# --------------------------------------------------------------------------
# tree build (include/tree3d.h:238-371)
# --------------------------------------------------------------------------
def morton_key(cells, k, dim, levels):
    """Integer key of points at dyadic grid cells: at every level the child
    cell is ((z << k) + y) << k + x, concatenated root first."""
    key = np.zeros(len(cells), dtype=object)
    for lev in range(levels):
        shift = (levels - 1 - lev) * k
        digit = np.zeros(len(cells), dtype=np.int64)
        for a in reversed(range(dim)):
            digit = (digit << k) + ((cells[:, a] >> shift) & ((1 << k) - 1))
        key = key * (1 << (dim * k)) + digit
    return key


@pytest.mark.parametrize("dim,k,levels", [(3, 1, 6), (3, 2, 3), (2, 1, 8), (2, 3, 2), (3, 3, 2)])
def test_dyadic_grid_leaf_order_is_morton_order(dim, k, levels):
    """On cell centres of a 2^(k*levels) grid in the unit box every float
    operation of the build is exact, so the leaf order must be the integer
    hierarchical-digit order."""
    g = 1 << (k * levels)
    rng = np.random.default_rng(dim * 10 + k)
    n = min(3000, g ** dim)
    flat = rng.choice(g ** dim, size=n, replace=False)
    cells = np.stack([(flat // g ** a) % g for a in range(dim)], 1).astype(np.int64)
    pts = (cells + 0.5) / g
    t = oracle.Tree(pts, None, np.ones(n), np.zeros(dim), np.ones(dim), split_log=k)
    keys = morton_key(cells, k, dim, levels)
    expect = sorted(range(n), key=lambda i: keys[i])
    assert t.leaf_order().tolist() == expect


def tree_invariants(t, pts):
    nd = t.nodes()
    order = t.leaf_order()
    n = len(pts)
    assert sorted(order.tolist()) == list(range(n))
    assert nd["start"][0] == 0 and nd["end"][0] == n
    p32 = pts.astype(np.float32)[order]
    for v in range(t.num_nodes):
        s, e = nd["start"][v], nd["end"][v]
        assert s < e
        if nd["leaf"][v]:
            assert nd["nchild"][v] == 0
            assert (e - s) <= t.leaf_size or nd["depth"][v] >= t.max_depth
            inside = ((p32[s:e] >= nd["bmin"][v] - 1e-4) & (p32[s:e] <= nd["bmax"][v] + 1e-4))
            assert inside.all()
        else:
            f, c = nd["first_child"][v], nd["nchild"][v]
            assert c >= 1
            kids = range(f, f + c)
            assert nd["start"][f] == s and nd["end"][f + c - 1] == e
            for a, b in zip(kids, list(kids)[1:]):
                assert nd["end"][a] == nd["start"][b]
            assert all(nd["parent"][x] == v for x in kids)
            assert all(nd["depth"][x] == nd["depth"][v] + 1 for x in kids)
    # breadth-first numbering: depth is non-decreasing with the node index
    assert (np.diff(nd["depth"]) >= 0).all()


@pytest.mark.parametrize("k,leaf", [(1, 1), (2, 1), (1, 8), (3, 2)])
def test_tree_invariants_uniform(k, leaf):
    pts, nrm, m, _ = W.uniform_cube(3000, 1, seed=k)
    lo, hi = W.bounds(pts)
    t = oracle.Tree(pts, nrm, m, lo, hi, split_log=k, leaf_size=leaf)
    tree_invariants(t, pts)


def test_tree_invariants_duplicates_hit_max_depth():
    rng = np.random.default_rng(4)
    base = rng.random((50, 3))
    pts = np.concatenate([base, base[:10], base[:3]])
    t = oracle.Tree(pts, None, np.ones(len(pts)), [-0.01] * 3, [1.01] * 3, split_log=1)
    assert t.depth == 100
    tree_invariants(t, pts)


def test_moments_match_float64_sums():
    pts, _, m, _ = W.uniform_cube(2000, 1, seed=3)
    nrm = np.random.default_rng(5).normal(size=pts.shape)
    lo, hi = W.bounds(pts)
    t = oracle.Tree(pts, nrm, m, lo, hi, split_log=2)
    nd = t.nodes()
    order = t.leaf_order()
    p = pts.astype(np.float32).astype(np.float64)[order]
    n = nrm.astype(np.float32).astype(np.float64)[order]
    mm = m.astype(np.float32).astype(np.float64)[order]
    for v in range(0, t.num_nodes, 7):
        s, e = nd["start"][v], nd["end"][v]
        M = mm[s:e].sum()
        C = (mm[s:e, None] * p[s:e]).sum(0) / M
        Nn = (mm[s:e, None] * n[s:e]).sum(0) / M
        assert np.isclose(nd["mass"][v], M, rtol=1e-5)
        np.testing.assert_allclose(nd["centroid"][v], C, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(nd["normal"][v], Nn, rtol=1e-4, atol=1e-5)
        assert np.isclose(nd["mass64"][v], M, rtol=1e-12)
        np.testing.assert_allclose(nd["centroid64"][v], C, rtol=1e-10, atol=1e-12)


# --------------------------------------------------------------------------
# brute force and Barnes-Hut (include/barnes_hut.h:23-78, barnes_hut_gpu.h)
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def small_case():
    pts, _, m, q = W.uniform_cube(3000, 300, seed=7)
    nrm = np.random.default_rng(8).normal(size=pts.shape)
    lo, hi = W.bounds(pts)
    return pts, nrm, m, q, lo, hi


@pytest.mark.parametrize("kernel", ["gravity", "wn", "smoothdist"])
def test_brute_force_against_numpy(small_case, kernel):
    pts, nrm, m, q, lo, hi = small_case
    t = oracle.Tree(pts, nrm, m, lo, hi, alpha=2.0)
    p64 = pts.astype(np.float32).astype(np.float64)
    n64 = nrm.astype(np.float32).astype(np.float64)
    m64 = m.astype(np.float32).astype(np.float64)
    q64 = q.astype(np.float32).astype(np.float64)
    ref = np.array([closed_form(kernel, 3, p64, n64, m64, 2.0, qq[None, :]).sum() for qq in q64])
    got64 = t.brute_force(kernel, q, fp64=True)
    got32 = t.brute_force(kernel, q)
    np.testing.assert_allclose(got64, ref, rtol=1e-11, atol=1e-11)
    absum = np.array([np.abs(closed_form(kernel, 3, p64, n64, m64, 2.0, qq[None, :])).sum()
                      for qq in q64])
    np.testing.assert_allclose(t.abs_sum(kernel, q), absum, rtol=1e-11)
    # fp32 terms each carry ~1 ulp: the error is bounded by the sum's condition
    assert (np.abs(got32 - ref) / absum).max() < 2e-6


def test_barnes_hut_infinite_beta_is_exact(small_case):
    pts, nrm, m, q, lo, hi = small_case
    for k in (1, 2):
        t = oracle.Tree(pts, nrm, m, lo, hi, split_log=k)
        ref = t.brute_force("gravity", q, fp64=True)
        got, cnt = t.barnes_hut("gravity", q, np.inf, counts=True)
        assert (np.abs(got - ref) / np.abs(ref)).max() < 1e-5
        assert (cnt == len(pts)).all()


def test_barnes_hut_tiny_beta_is_root_moment(small_case):
    pts, nrm, m, q, lo, hi = small_case
    t = oracle.Tree(pts, nrm, m, lo, hi)
    nd = t.nodes()
    got, cnt = t.barnes_hut("gravity", q, 1e-30, counts=True)
    c = nd["centroid"][0].astype(np.float64)
    ref = -np.float64(nd["mass"][0]) / (np.linalg.norm(c - q.astype(np.float32), axis=1) + 0.1)
    np.testing.assert_allclose(got, ref, rtol=1e-5)
    assert (cnt == 1).all()


def _bh_recursive(nd, pts32, q, beta, f, group):
    """Depth-first recursion over the exported node arrays: a node is used
    when every query of the group accepts it (ratio >= beta)."""
    tot = np.zeros(len(q))

    def ratio(v):
        if nd["leaf"][v]:
            return np.full(len(q), np.inf)
        return np.linalg.norm(q - nd["centroid"][v], axis=1) / nd["diam"][v]

    def visit(v):
        r = ratio(v) >= beta
        acc = np.all(r.reshape(-1, group), axis=1).repeat(group)
        if acc.all():
            if nd["leaf"][v]:
                for j in range(nd["start"][v], nd["end"][v]):
                    tot[:] += f(pts32[j], nd_mass_pt[j], q)
            else:
                tot[:] += f(nd["centroid"][v], nd["mass"][v], q)
            return
        for c in range(nd["first_child"][v], nd["first_child"][v] + nd["nchild"][v]):
            visit(c)

    nd_mass_pt = None
    return visit, tot, lambda mp: None


def test_barnes_hut_vote_group_semantics(small_case):
    """Vote groups: each group of G consecutive queries uses exactly the nodes
    every member accepts; checked against an independent recursion."""
    pts, nrm, m, q, lo, hi = small_case
    t = oracle.Tree(pts[:400], nrm[:400], m[:400], lo, hi, split_log=1)
    nd = t.nodes()
    order = t.leaf_order()
    p32 = pts[:400].astype(np.float32).astype(np.float64)[order]
    m32 = m[:400].astype(np.float32).astype(np.float64)[order]
    qq = q[:64].astype(np.float32).astype(np.float64)

    def grav(c, mass, qs):
        return -mass / (np.linalg.norm(qs - c, axis=1) + 0.1)

    for group in (1, 32, 64):
        tot = np.zeros(len(qq))
        stack = [0]
        # depth-first, children in order; acceptance decided per group
        live = np.ones(len(qq), bool)

        def visit(v, mask):
            if not mask.any():
                return
            if nd["leaf"][v]:
                r = np.full(len(qq), np.inf)
            else:
                r = np.linalg.norm(qq - nd["centroid"][v].astype(np.float64), axis=1) / nd["diam"][v]
            acc = (r >= 2.0).reshape(-1, group).all(1).repeat(group)
            use = mask & acc
            if use.any():
                if nd["leaf"][v]:
                    for j in range(nd["start"][v], nd["end"][v]):
                        tot[use] += grav(p32[j], m32[j], qq[use])
                else:
                    tot[use] += grav(nd["centroid"][v].astype(np.float64), nd["mass"][v], qq[use])
            rest = mask & ~acc
            if rest.any() and not nd["leaf"][v]:
                for c in range(nd["first_child"][v], nd["first_child"][v] + nd["nchild"][v]):
                    visit(c, rest)

        visit(0, live)
        got = t.barnes_hut("gravity", q[:64], 2.0, vote_group=group)
        np.testing.assert_allclose(got, tot, rtol=2e-5)
    g1 = t.barnes_hut("gravity", q[:64], 2.0, vote_group=1)
    g32 = t.barnes_hut("gravity", q[:64], 2.0, vote_group=32)
    ex = t.brute_force("gravity", q[:64], fp64=True)
    # voting descends at least as far as every member would: never less accurate on average
    assert np.abs(g32 - ex).mean() <= np.abs(g1 - ex).mean() + 1e-7


# --------------------------------------------------------------------------
# stochastic estimator (include/tree_sampler.h:21-113)
# --------------------------------------------------------------------------
def exact_moments_of_estimator(t, pts, m, q, spl):
    """Mean and variance of the estimator for one query, by enumerating every
    sampled point and every Russian-roulette outcome along its path."""
    nd = t.nodes()
    order = t.leaf_order()
    p = pts.astype(np.float32).astype(np.float64)[order]
    mm = m.astype(np.float32).astype(np.float64)[order]
    qq = q.astype(np.float32).astype(np.float64)

    def f_node(v):
        if nd["leaf"][v]:
            s, e = nd["start"][v], nd["end"][v]
            return (-mm[s:e] / (np.linalg.norm(p[s:e] - qq, axis=1) + 0.1)).sum()
        c = nd["centroid"][v].astype(np.float64)
        return -np.float64(nd["mass"][v]) / (np.linalg.norm(c - qq) + 0.1)

    def ratio(v):
        if nd["leaf"][v]:
            return np.inf
        return np.linalg.norm(qq - nd["centroid"][v].astype(np.float64)) / np.float64(nd["diam"][v])

    def child_containing(v, idx):
        for c in range(nd["first_child"][v], nd["first_child"][v] + nd["nchild"][v]):
            if nd["start"][c] <= idx < nd["end"][c]:
                return child_containing(c, idx) if nd["nchild"][c] == 1 else c
        raise AssertionError

    mean_total, var_total = 0.0, 0.0
    r0 = 0
    for si in range(nd["nchild"][r0]):
        S = nd["first_child"][r0] + si
        mean_total += f_node(S)
        if nd["leaf"][S]:
            continue
        nS = nd["end"][S] - nd["start"][S]
        e1 = e2 = 0.0
        for idx in range(nd["start"][S], nd["end"][S]):
            # walk: list of (probability, value) outcomes
            cur, prob, ev, acc = S, 1.0, 1.0, 0.0
            cur_beta = max(1.0, ratio(S))
            while not nd["leaf"][cur]:
                ffr = ratio(cur)
                tp = min(cur_beta / ffr, 1.0)
                # stop here
                e1 += (1.0 / nS) * prob * (1 - tp) * acc
                e2 += (1.0 / nS) * prob * (1 - tp) * acc * acc
                prob *= tp
                ev *= tp
                prev = cur
                cur = child_containing(cur, idx)
                cur_beta = max(cur_beta, ffr)
                kids = range(nd["first_child"][prev], nd["first_child"][prev] + nd["nchild"][prev])
                diff = sum(f_node(c) for c in kids) - f_node(prev)
                acc += nS / (ev * (nd["end"][prev] - nd["start"][prev])) * diff
            e1 += (1.0 / nS) * prob * acc
            e2 += (1.0 / nS) * prob * acc * acc
        mean_total += e1
        var_total += (e2 - e1 * e1) / spl
    return mean_total, var_total


def test_stochastic_enumerated_expectation_is_exact_sum_and_matches_sampling():
    rng = np.random.default_rng(12)
    n = 60
    pts = rng.random((n, 3))
    m = rng.uniform(0.5, 1.5, n)
    lo, hi = np.full(3, -0.01), np.full(3, 1.01)
    t = oracle.Tree(pts, None, m, lo, hi, split_log=1)
    queries = np.array([[0.5, 0.5, 0.5], [1.3, -0.2, 0.4], [0.05, 0.9, 0.5]])
    exact = t.brute_force("gravity", queries, fp64=True)
    spl = 2
    seeds = 4000
    for qi, q in enumerate(queries):
        mu, var = exact_moments_of_estimator(t, pts, m, q, spl)
        # telescoping: the expectation is the exact sum
        assert abs(mu - exact[qi]) <= 1e-5 * abs(exact[qi])
        assert var > 0
        ests = np.array([t.stochastic("gravity", q[None, :], spl, seed=s, group=1)[0]
                         for s in range(seeds)], np.float64)
        se = np.sqrt(var / seeds)
        assert abs(ests.mean() - mu) <= 4 * se
        # sample variance of 4000 draws: relative sd ~ sqrt(2/4000) ~ 2.2%
        assert abs(ests.var(ddof=1) / var - 1.0) <= 0.12


def test_stochastic_group_shares_paths_not_values():
    pts, nrm, m, q, lo, hi = W.uniform_cube(2000, 128, seed=13)[0], None, None, None, None, None
    pts, nrm, m, q = W.uniform_cube(2000, 128, seed=13)
    lo, hi = W.bounds(pts)
    t = oracle.Tree(pts, nrm, m, lo, hi, split_log=2)
    a = t.stochastic("gravity", q, 2, seed=3, group=64)
    # re-evaluating a query alone with the same group id gives the same value
    for i in (0, 5, 64, 127):
        qi = q[i:i + 1]
        # group id of query i is i // 64; reproduce by padding
        pad = np.repeat(qi, (i // 64) * 64 + 1, axis=0)
        b = t.stochastic("gravity", pad, 2, seed=3, group=64)
        assert b[-1] == a[i]


def test_stochastic_root_leaf_is_exact():
    pts = np.random.default_rng(1).random((5, 3))
    t = oracle.Tree(pts, None, np.ones(5), [0, 0, 0], [1, 1, 1], split_log=1, leaf_size=8)
    q = np.random.default_rng(2).random((4, 3))
    np.testing.assert_allclose(t.stochastic("gravity", q, 3), t.brute_force("gravity", q, True),
                               rtol=1e-6)

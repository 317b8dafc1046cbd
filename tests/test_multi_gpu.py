"""GPU parity of the 4-way engine (terminal values, drop-in sweeps, the
batched multi-flop solver), FP64 ICM and bubble factors, and the fused DCFR
step, through the C ABI, against the CPU oracle on the same inputs."""
import ctypes
import os

import numpy as np
import pytest

from conftest import PKG, _load

pytestmark = pytest.mark.gpu

TERM_REL = 2e-5      # FP32 per-opponent masses (FP64 prefix sums) vs the all-FP64 oracle
STATE_REL = 1e-4     # regrets / strategy sums after a few iterations, relative to their max


@pytest.fixture(scope="module")
def M(mtt):
    return _load("mtt_multi", os.path.join(PKG, "multi.py"))


STACKS = {"mixed": [30.0, 45.0, 12.0, 60.0],     # 677 nodes, the 12bb stack called all-in for less
          "config5_seed0": [50.272126549109075, 22.734003532290274, 5.573014295214602,
                            3.739572664639682]}     # 470 nodes, 120 of 221 showdowns with side pots


def _setup(M, seed=0, stacks="mixed"):
    rng = np.random.default_rng(seed)
    game = M.game4([2.5] * 4, STACKS[stacks], M.config5_menu())
    nodes, ni = M.tree4(game)
    flop = np.sort(rng.choice(52, 3, replace=False)).astype(np.uint8)
    ranges = np.zeros((4, 1216), np.float32)
    for p in range(4):
        w = rng.uniform(0, 1, 1176)
        ranges[p, :1176] = w / w.sum()
    bf = rng.uniform(1.0, 1.5, (4, 4))
    return game, nodes, ni, flop, ranges, bf


def _depths(nodes):
    depth = np.zeros(len(nodes), int)
    for n in range(len(nodes)):
        if nodes[n]["type"] == 0:
            for a in range(nodes[n]["num_actions"]):
                depth[nodes[n]["children_offset"] + a] = depth[n] + 1
    dec = [np.nonzero((nodes["type"] == 0) & (depth == d))[0].astype(np.int32) for d in range(depth.max() + 1)]
    return [d for d in dec if len(d)]


@pytest.mark.parametrize("stacks", ["mixed", "config5_seed0"])
def test_4way_terminals_match_oracle(M, mtt, oracle, torch_dev, stacks):
    """Every terminal of the tree (folds, showdowns of 2-4 players, side pots
    where short stacks called all-in for less) against the oracle."""
    import torch
    game, nodes, ni, flop, ranges, bf = _setup(M, 1, stacks)
    print(f"4-way terminals, stacks {stacks}, seed 1, flop", flop.tolist(), flush=True)
    rng = np.random.default_rng(7)
    nn = len(nodes)
    reach = rng.uniform(0, 1e-3, (nn, 4, 1216)).astype(np.float32)
    reach[:, :, 1176:] = 0
    terms = np.nonzero(nodes["type"] != 0)[0].astype(np.int32)
    board = M.board4(flop)
    d_nodes = torch.from_numpy(nodes.view(np.uint8).copy()).to(torch_dev)
    d_reach = torch.from_numpy(reach.ravel()).to(torch_dev)
    d_board = torch.from_numpy(board.view(np.uint8).copy()).to(torch_dev)
    d_bf = torch.from_numpy(bf.astype(np.float32).ravel()).to(torch_dev)
    d_terms = torch.from_numpy(terms).to(torch_dev)
    d_cfv = torch.zeros(nn * 4 * 1216, device=torch_dev)
    mtt.call("mtt_4way_evaluate_terminals_board", d_terms.data_ptr(), len(terms), d_nodes.data_ptr(),
             d_reach.data_ptr(), d_board.data_ptr(), d_bf.data_ptr(), d_cfv.data_ptr(), None)
    torch.cuda.synchronize()
    cfv = d_cfv.cpu().numpy().reshape(nn, 4, 1216)
    # the reference's argument list: combo_rank_pos instead of the board tables
    d_rank = torch.from_numpy(board["rank_pos"][0].astype(np.int16)).to(torch_dev)
    d_cfv2 = torch.zeros_like(d_cfv)
    fl = np.ascontiguousarray(flop, np.uint8)
    mtt.call("mtt_4way_evaluate_terminals", d_terms.data_ptr(), len(terms), d_nodes.data_ptr(), d_reach.data_ptr(),
             d_rank.data_ptr(), d_bf.data_ptr(), d_cfv2.data_ptr(), fl.ctypes.data, None)
    torch.cuda.synchronize()
    assert torch.equal(d_cfv, d_cfv2)
    # a strength order that is not the flop's is refused
    bad = board["rank_pos"][0].astype(np.int16).copy()
    lo = int(np.argmin(board["rank_pos"][0][:1176]))
    hi = int(np.argmax(board["rank_pos"][0][:1176]))
    bad[lo], bad[hi] = bad[hi], bad[lo]
    d_bad = torch.from_numpy(bad).to(torch_dev)
    rc = mtt.lib().mtt_4way_evaluate_terminals(d_terms.data_ptr(), len(terms), d_nodes.data_ptr(), d_reach.data_ptr(),
                                               d_bad.data_ptr(), d_bf.data_ptr(), d_cfv2.data_ptr(), fl.ctypes.data,
                                               None)
    assert rc == mtt.MTT_ERR_ARG and "strength order" in mtt.last_error()
    ob = oracle.board4(flop)
    bf32 = bf.astype(np.float32).astype(np.float64)
    for n in terms:
        ref = oracle.terminal4(ob, nodes[n:n + 1], reach[n], bf32)
        scale = np.abs(ref).max() + 1e-30
        assert np.abs(cfv[n] - ref).max() <= TERM_REL * scale, (n, int(nodes[n]["type"]), int(nodes[n]["active_mask"]))


def test_4way_dropin_sweeps_match_oracle(M, mtt, oracle, torch_dev):
    """A host loop over the reference-style entry points (forward per depth,
    terminals, CFR+ backward per depth) reproduces the oracle's iterations."""
    import torch
    game, nodes, ni, flop, ranges, bf = _setup(M, 2)
    print("4-way drop-in sweeps, seed 2", flush=True)
    nn = len(nodes)
    dev = torch_dev
    levels = [torch.from_numpy(d).to(dev) for d in _depths(nodes)]
    terms = torch.from_numpy(np.nonzero(nodes["type"] != 0)[0].astype(np.int32)).to(dev)
    d_nodes = torch.from_numpy(nodes.view(np.uint8).copy()).to(dev)
    d_board = torch.from_numpy(M.board4(flop).view(np.uint8).copy()).to(dev)
    d_bf = torch.from_numpy(bf.astype(np.float32).ravel()).to(dev)
    reach = torch.zeros(nn * 4 * 1216, device=dev)
    reach[:4 * 1216] = torch.from_numpy(ranges.ravel()).to(dev)
    cfv = torch.zeros_like(reach)
    R = torch.zeros(ni * 3 * 1216, device=dev)
    S = torch.zeros_like(R)
    sig = torch.zeros_like(R)
    h = oracle.FourWay(nodes, ni, flop, ranges, bf.astype(np.float32).astype(np.float64))
    for t in range(1, 4):
        for lv in levels:
            mtt.call("mtt_4way_forward_sweep", lv.data_ptr(), lv.numel(), d_nodes.data_ptr(), R.data_ptr(),
                     sig.data_ptr(), reach.data_ptr(), None)
        mtt.call("mtt_4way_evaluate_terminals_board", terms.data_ptr(), terms.numel(), d_nodes.data_ptr(),
                 reach.data_ptr(), d_board.data_ptr(), d_bf.data_ptr(), cfv.data_ptr(), None)
        for lv in reversed(levels):
            mtt.call("mtt_4way_backward_cfr_plus", lv.data_ptr(), lv.numel(), d_nodes.data_ptr(), sig.data_ptr(),
                     reach.data_ptr(), cfv.data_ptr(), R.data_ptr(), S.data_ptr(), ctypes.c_float(float(t)), None)
        torch.cuda.synchronize()
        h.iterate(1, scheme=0)
        Rc, Sc, root = h.state()
        Rg, Sg = R.cpu().numpy(), S.cpu().numpy()
        assert np.abs(Rg - Rc).max() <= STATE_REL * max(np.abs(Rc).max(), 1e-12), t
        assert np.abs(Sg - Sc).max() <= STATE_REL * max(np.abs(Sc).max(), 1e-12), t
        rg = cfv[:4 * 1216].cpu().numpy()
        assert np.abs(rg - root).max() <= STATE_REL * np.abs(root).max(), t


def test_multi_solver_matches_oracle(M, oracle, torch_dev):
    """Batched DCFR(1.5, 0, 2) over two flops vs the oracle per flop."""
    game, nodes, ni, flop, ranges, bf = _setup(M, 3)
    rng = np.random.default_rng(11)
    flops = np.stack([flop, np.sort(rng.choice([c for c in range(52) if c not in flop], 3, replace=False))]).astype(np.uint8)
    rr = np.stack([ranges, ranges[::-1].copy()])
    print("multi solver, 2 flops, DCFR", flush=True)
    mm = M.Multi(game, flops, rr, bf.astype(np.float32))
    mm.run(6)
    for f in range(2):
        h = oracle.FourWay(nodes, ni, flops[f], rr[f], bf.astype(np.float32).astype(np.float64))
        h.iterate(6, scheme=1, alpha=1.5, beta=0.0, gamma=2.0)
        Rc, Sc, root_c = h.state()
        Rg, Sg, root_g = mm.state(f)
        assert np.abs(Rg - Rc).max() <= 1e-3 * np.abs(Rc).max(), f
        assert np.abs(Sg - Sc).max() <= 1e-3 * np.abs(Sc).max(), f
        assert np.abs(root_g - root_c).max() <= 1e-3 * np.abs(root_c).max(), f
    mm.close()


def test_icm_gpu_matches_oracle(M, oracle, torch_dev):
    rng = np.random.default_rng(5)
    for _ in range(5):
        n = int(rng.integers(2, 10))
        st = rng.uniform(1.0, 80.0, n)
        k = int(rng.integers(1, min(n, 4) + 1))
        pay = np.sort(rng.uniform(0.1, 1.0, k))[::-1].copy()
        g = M.icm(st, pay)
        c = oracle.icm(st, pay)
        assert np.allclose(g, c, rtol=1e-13, atol=1e-15)
    st8 = rng.uniform(5.0, 80.0, 8)
    bg = M.bubble_factors(st8, [0.5, 0.3, 0.2], [0, 1, 2, 3])
    bc = oracle.bubble_factors(st8, [0.5, 0.3, 0.2], [0, 1, 2, 3])
    assert np.allclose(bg, bc, rtol=1e-11)


def test_dcfr_step_gpu_matches_oracle(mtt, oracle, torch_dev):
    import torch
    rng = np.random.default_rng(9)
    na, n, stride = 3, 169, 192
    R = rng.normal(0, 1, (na, stride)).astype(np.float32)
    sg = rng.dirichlet([1.0] * na, stride).T.astype(np.float32).copy()
    S = rng.uniform(0, 1, (na, stride)).astype(np.float32)
    cfv = rng.normal(0, 1, (na, stride)).astype(np.float32)
    reach = rng.uniform(0, 1, n).astype(np.float32)
    d = [torch.from_numpy(x.copy()).to(torch_dev) for x in (R, sg, S, cfv, reach)]
    for _ in range(3):
        mtt.call("mtt_dcfr_step", na, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
                 d[4].data_ptr(), ctypes.c_float(0.7), ctypes.c_float(2.0), ctypes.c_float(0.5), n, stride, None)
        oracle.dcfr_step(na, R, sg, S, cfv, reach, 0.7, 0.5, 2.0)
    torch.cuda.synchronize()
    for g, c in zip(d[:3], (R, sg, S)):
        assert np.allclose(g.cpu().numpy(), c, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("seed,n_nodes", [(0, 470), (1, 677)])
def test_config5_full_size(M, oracle, torch_dev, seed, n_nodes):
    """The 4-way benchmark configuration at full size: 100 seeded flops,
    DCFR(1.5, 0, 2), 200 iterations, bubble factors from FP64 ICM of eight
    stacks.  Seed 0 draws two short stacks (5.6 and 3.7bb behind), whose
    all-in calls leave side pots (470 nodes); seed 1's tree has 677 nodes.
    Flops 0 and 57 match the oracle after 10 iterations; after 200 the root
    strategy sums equal sum_t t^2 * range (every iteration reaches the root
    infoset with the range itself)."""
    game, flops, ranges, bf, st8, pay = M.config5(seed, 100)
    assert np.all(bf >= 1.0)
    mm = M.Multi(game, flops, ranges, bf)
    nodes, ni = mm.nodes, mm.n_info
    print(f"config5 seed {seed}: {len(nodes)} nodes, {ni} infosets, 100 flops, DCFR 200 iterations", flush=True)
    assert len(nodes) == n_nodes
    mm.run(10)
    for f in (0, 57):
        h = oracle.FourWay(nodes, ni, flops[f], ranges[f], bf.astype(np.float64))
        h.iterate(10, scheme=1, alpha=1.5, beta=0.0, gamma=2.0)
        Rc, Sc, root_c = h.state()
        Rg, Sg, root_g = mm.state(f)
        assert np.abs(Rg - Rc).max() <= 1e-3 * np.abs(Rc).max(), f
        assert np.abs(Sg - Sc).max() <= 1e-3 * np.abs(Sc).max(), f
    mm.timing(True)
    mm.run(190)
    kt = mm.kernel_times()
    ms = sum(v[0] for v in kt.values()) / 190
    print(f"config5 seed {seed} iteration: {ms:.3f} ms over 100 flops ({1e3 / ms:.1f} iterations/s); "
          + ", ".join(f"{k} {v[0] / 190:.3f} ms" for k, v in kt.items()), flush=True)
    w = sum(float(t) ** 2 for t in range(1, 201))
    root = nodes[0]
    for f in (0, 99):
        R, S, rootv = mm.state(f)
        assert np.isfinite(R).all() and np.isfinite(rootv).all()
        Ssum = S.reshape(ni, 3, 1216)[root["infoset_offset"]].sum(0)[:1176]
        assert np.allclose(Ssum, w * ranges[f, root["player"], :1176], rtol=1e-4)
    mm.close()

"""GPU parity of the heads-up solver paths (through the C ABI) against the
CPU oracle: strategies and FP64 exploitability on the same seeded inputs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# Tolerances.  Both sides run the same recurrence; FP32 sums differ in order
# (GPU vs CPU), and the differences grow slowly with iterations.
STRAT_LINF_FP32 = 2e-3      # plain FP32 GPU vs FP32 oracle, average strategy
STRAT_LINF_BF16 = 5e-3      # BF16 hi+lo showdown (16-bit reach mantissa) vs FP32 oracle
EXPL_ABS = 2e-3             # % pot, GPU FP64 BR vs oracle FP64 BR of the same strategy
STATE_REL_FP32 = 5e-4       # regrets / strategy sums after 1-3 iterations, relative to their max
STATE_REL_BF16 = 5e-4


def _setup(solver_api, oracle, g, r0, r1, prec):
    nodes, ni, _ = solver_api.tree(g)
    board = list(g.board)[:g.n_board]
    gi, cards = solver_api.alive_index(board)
    return oracle.HU(nodes, ni, board, r0[gi], r1[gi], prec), nodes, ni


def _live_linf(a, b, nodes, n_alive, S_ref):
    """L-inf of the average-strategy difference over (infoset, combo) entries
    that carry strategy mass: sum_a S >= 1e-5 of the infoset's largest.
    Entries below are combos whose reach underflows (Dirichlet(0.5) ranges
    have weights near 0); their average strategy is numerically arbitrary and
    does not enter the exploitability, which is compared separately."""
    a = a.reshape(-1, 4, 1216)
    b = b.reshape(-1, 4, 1216)
    w = S_ref.reshape(-1, 4, 1216).sum(1)
    d = 0.0
    for x in nodes[nodes["type"] == 0]:
        k, na = int(x["infoset_offset"]), int(x["num_actions"])
        m = w[k, :n_alive]
        if m.max() <= 0:
            continue
        keep = m >= 1e-5 * m.max()
        d = max(d, float(np.abs(a[k, :na, :n_alive] - b[k, :na, :n_alive])[:, keep].max()))
    return d


def test_plain_config1_matches_oracle(solver_api, oracle, torch_dev):
    g, r0, r1 = solver_api.config1(0)
    print("config1 plain fp32, seed 0", flush=True)
    s = solver_api.Solver(g, r0, r1, solver_api.MODE_PLAIN_FP32)
    h32, nodes, ni = _setup(solver_api, oracle, g, r0, r1, 1)
    s.run(200)
    h32.iterate(200)
    st = s.strategy()
    n_alive = int(s.info[2])
    assert _live_linf(st, h32.strategy(), nodes, n_alive, h32.state()[1]) < STRAT_LINF_FP32
    e_gpu = s.exploitability()
    e_cpu = h32.exploitability(st, 10.0)
    assert np.allclose(e_gpu[:5], e_cpu[:5], rtol=1e-9, atol=1e-12)
    assert abs(e_gpu[5] - e_cpu[5]) < EXPL_ABS
    s.run(800)
    e = s.exploitability()
    assert e[5] < 0.5 and e[5] < e_gpu[5]
    s.close()


def _multistreet_check(solver_api, oracle, mode, tol_early):
    """Multi-street parity.  CFR+ regret clipping amplifies rounding
    differences at infosets the opponent almost never reaches, so strategies
    are compared after a few iterations and the solve is compared through the
    FP64 exploitability after more."""
    g, r0, r1 = solver_api.config3_small(1)
    s = solver_api.Solver(g, r0, r1, mode)
    h32, nodes, ni = _setup(solver_api, oracle, g, r0, r1, 1)
    for it in range(3):
        s.run(1)
        h32.iterate(1)
        Rg, Sg = s.state()
        Rc, Sc = h32.state()
        assert np.abs(Rg - Rc).max() <= tol_early * np.abs(Rc).max(), it
        assert np.abs(Sg - Sc).max() <= tol_early * np.abs(Sc).max(), it
    s.run(27)
    h32.iterate(27)
    st = s.strategy()
    e_gpu = s.exploitability()
    e_cpu = h32.exploitability(st, 6.0)  # the oracle's FP64 BR of the GPU strategy
    assert np.allclose(e_gpu[:5], e_cpu[:5], rtol=1e-7)
    # constant sum: the two players' chips add to the starting pot times the
    # pair mass (FP32-normalised strategies and FP32 pots in the tree: 1e-6)
    assert abs(e_gpu[2] + e_gpu[3] - 6.0 * e_gpu[4]) <= 1e-6 * 6.0 * e_gpu[4]
    e_ref = h32.exploitability(h32.strategy(), 6.0)
    assert abs(e_gpu[5] - e_ref[5]) < 0.02, (e_gpu[5], e_ref[5])
    s.close()


def test_plain_multistreet_matches_oracle(solver_api, oracle, torch_dev):
    """Turn chance, turn betting, river chance, river betting, turn all-in and
    river showdowns (all-in-only menu, full 49x48 boards)."""
    print("config3-small plain fp32, seed 1", flush=True)
    _multistreet_check(solver_api, oracle, solver_api.MODE_PLAIN_FP32, STATE_REL_FP32)


@pytest.mark.parametrize("stride", [4, 3])
def test_dropin_kernels_manual_iteration(solver_api, oracle, torch_dev, mtt, stride):
    """A host loop over the reference-style kernel entry points (forward per
    depth, per-terminal evaluation, backward per depth) reproduces the
    oracle's first iterations; with per-action buffers of the bet menu's
    four actions, and with the reference's stride of 3
    (honest_multistreet_solver.cu:12) on a menu of at most three actions
    (one bet size and all-in; fold, call, all-in raise)."""
    import torch
    g, r0, r1 = solver_api.config1(2)
    if stride == 3:
        g.menu[2] = solver_api.menu((0.5,), 0.0, 1, 1)
    nodes, ni, _ = solver_api.tree(g)
    assert nodes["num_actions"][nodes["type"] == 0].max() == stride
    board = list(g.board)[:5]
    gi, cards = solver_api.alive_index(board)
    n_alive = len(gi)
    NP, A = 1216, stride
    dev = torch_dev
    combos = np.zeros(NP, dtype=[("c1", "u1"), ("c2", "u1"), ("pad", "V6"), ("mask", "<u8")])
    combos["c1"][:n_alive] = cards[:, 0]
    combos["c2"][:n_alive] = cards[:, 1]
    combos["mask"][:n_alive] = (1 << cards[:, 0].astype(np.uint64)) | (1 << cards[:, 1].astype(np.uint64))
    combos["mask"][n_alive:] = np.uint64(2**64 - 1)
    d_combos = torch.from_numpy(combos.view(np.uint8).copy()).to(dev)
    deck = np.array([c for c in range(52) if c not in board], np.uint8)
    d_deck = torch.from_numpy(deck).to(dev)
    d_nodes = torch.from_numpy(nodes.view(np.uint8).copy()).to(dev)
    nn = len(nodes)
    reach = torch.zeros(nn * 2 * NP, device=dev)
    cfv = torch.zeros(nn * 2 * NP, device=dev)
    R = torch.zeros(ni * A * NP, device=dev)
    S = torch.zeros_like(R)
    sig = torch.zeros_like(R)
    reach[:n_alive] = torch.from_numpy(r0[gi]).to(dev)
    reach[NP:NP + n_alive] = torch.from_numpy(r1[gi]).to(dev)
    # depth lists
    depth = {0: [0]}
    q, dmap = [0], {0: 0}
    terms = []
    while q:
        i = q.pop(0)
        x = nodes[i]
        if x["type"] in (1, 2):
            terms.append(i)
            continue
        depth.setdefault(dmap[i], [])
        if i not in depth[dmap[i]]:
            depth[dmap[i]].append(i)
        for k in range(int(x["num_actions"])):
            c = int(x["children_offset"]) + k
            dmap[c] = dmap[i] + 1
            q.append(c)
    levels = [torch.tensor(depth[d], dtype=torch.int32, device=dev) for d in sorted(depth)]
    d_terms = torch.tensor(terms, dtype=torch.int32, device=dev)
    # the single showdown matrix through the same generator as the solver
    mats = torch.zeros(NP * NP, device=dev)
    cm = np.zeros(NP, dtype=combos.dtype)
    # board with the root's five cards: E[j][h] over root-alive combos
    from itertools import combinations
    sc = oracle.eval7_batch(np.array([[*board, a, b] for a, b in cards], np.uint8)).astype(np.int64)
    m = (1 << cards[:, 0].astype(object)) | (1 << cards[:, 1].astype(object))
    E = np.zeros((NP, NP), np.float32)
    blk = np.array([[bool(m[i] & m[j]) for i in range(n_alive)] for j in range(n_alive)])
    e = (sc[None, :] > sc[:, None]) + 0.5 * (sc[None, :] == sc[:, None])
    e[blk] = 0
    E[:n_alive, :n_alive] = e
    mats.copy_(torch.from_numpy(E.ravel()).to(dev))
    h = oracle.HU(nodes, ni, board, r0[gi], r1[gi], 1)
    L = mtt.lib()
    for t in range(1, 4):
        for lv in levels:
            mtt.check(L.mtt_hu_forward_sweep(lv.data_ptr(), lv.numel(), d_nodes.data_ptr(), d_deck.data_ptr(),
                                             d_combos.data_ptr(), R.data_ptr(), sig.data_ptr(), reach.data_ptr(),
                                             n_alive, A, None), "forward")
        mtt.check(L.mtt_hu_evaluate_terminals(d_terms.data_ptr(), d_terms.numel(), d_nodes.data_ptr(),
                                              reach.data_ptr(), mats.data_ptr(), cfv.data_ptr(), d_deck.data_ptr(),
                                              d_combos.data_ptr(), n_alive, None), "terminals")
        for lv in reversed(levels):
            mtt.check(L.mtt_hu_backward_sweep(lv.data_ptr(), lv.numel(), d_nodes.data_ptr(), d_deck.data_ptr(),
                                              d_combos.data_ptr(), sig.data_ptr(), reach.data_ptr(), cfv.data_ptr(),
                                              R.data_ptr(), S.data_ptr(), ctypes.c_float(float(t)), n_alive, A, None),
                      "backward")
        torch.cuda.synchronize()
        h.iterate(1)
        Rc, Sc = h.state()                      # the oracle's layout: stride 4
        Rc = Rc.reshape(ni, 4, NP)[:, :A].ravel()
        Sc = Sc.reshape(ni, 4, NP)[:, :A].ravel()
        Rg = R.cpu().numpy()
        Sg = S.cpu().numpy()
        scale = max(1.0, np.abs(Rc).max())
        assert np.abs(Rg - Rc).max() < 1e-4 * scale, t
        assert np.abs(Sg - Sc).max() < 1e-4 * max(1.0, np.abs(Sc).max()), t


@pytest.mark.parametrize("river", ["gemm", "sweep", "scan"])
def test_batched_config1_matches_oracle(solver_api, oracle, torch_dev, river, monkeypatch):
    """Batched path on the river subgame, river showdown by the BF16 hi+lo
    GEMM, by the exact strength-order sweep or by prefix scans."""
    monkeypatch.delenv("MTT_RIVER_GEMM", raising=False)
    monkeypatch.setenv("MTT_RIVER_SD", river)
    g, r0, r1 = solver_api.config1(0)
    print(f"config1 batched, river showdown {river}, seed 0", flush=True)
    s = solver_api.Solver(g, r0, r1, solver_api.MODE_BF16_SPLIT)
    h32, nodes, ni = _setup(solver_api, oracle, g, r0, r1, 1)
    s.run(200)
    h32.iterate(200)
    st = s.strategy()
    assert _live_linf(st, h32.strategy(), nodes, int(s.info[2]), h32.state()[1]) < STRAT_LINF_BF16
    e_gpu = s.exploitability()
    e_cpu = h32.exploitability(st, 10.0)
    assert np.allclose(e_gpu[:5], e_cpu[:5], rtol=1e-9, atol=1e-12)
    e_ref = h32.exploitability(h32.strategy(), 10.0)
    assert abs(e_gpu[5] - e_ref[5]) < 0.01
    s.close()


def test_batched_multistreet_matches_oracle(solver_api, oracle, torch_dev):
    """Turn + river street kernels, river GEMM and turn all-in showdowns on
    the all-in-only turn+river game."""
    print("config3-small batched bf16-split, seed 1", flush=True)
    _multistreet_check(solver_api, oracle, solver_api.MODE_BF16_SPLIT, STATE_REL_BF16)


@pytest.mark.parametrize("which", ["config1", "config3_small", "config3"])
def test_generated_sweeps_match_interpreted(solver_api, torch_dev, which, monkeypatch):
    """The per-template sweeps generated at solver creation and the
    interpreted sweeps give the same state after 5 iterations."""
    S = solver_api
    g, r0, r1 = getattr(S, which)(0 if which != "config3_small" else 1)
    print(f"{which}: generated vs interpreted sweeps", flush=True)
    a = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
    assert a.jit_info()[0]
    monkeypatch.setenv("MTT_JIT", "0")
    b = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
    monkeypatch.delenv("MTT_JIT")
    assert not b.jit_info()[0]
    a.run(5)
    b.run(5)
    Ra, Sa = a.state()
    Rb, Sb = b.state()
    dr = np.abs(Ra - Rb).max() / max(np.abs(Rb).max(), 1e-30)
    ds = np.abs(Sa - Sb).max() / max(np.abs(Sb).max(), 1e-30)
    print(f"{which}: max rel diff R {dr:.3g} S {ds:.3g}", flush=True)
    assert dr <= 1e-5 and ds <= 1e-5
    a.close()
    b.close()


@pytest.mark.parametrize("which", ["config3_small", "config3"])
def test_forward_strategy_sum_forms_identical(solver_api, torch_dev, which, monkeypatch):
    """The generated river forward reads the strategy sums either with the
    regrets in its prologue (MTT_JIT_FWD_SMASK=0) or k decisions at a time
    through a bounded buffer descriptor that sends the lanes of zero reach
    out of range (k = 1, 4, 64): the same sums in the same rounding, so the
    states are identical, also after 40 iterations when many strategies are
    pure and whole subtrees carry zero reach."""
    S = solver_api
    g, r0, r1 = getattr(S, which)(0 if which != "config3_small" else 1)
    ref = None
    for k in ("0", "1", "4", "64"):
        monkeypatch.setenv("MTT_JIT_FWD_SMASK", k)
        s = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
        assert s.jit_info()[0]
        s.run(40)
        st = s.state()
        s.close()
        print(f"{which}: strategy-sum form {k} done", flush=True)
        if ref is None:
            ref = st
            assert np.any(ref[1] > 0)
        else:
            assert np.array_equal(st[0], ref[0]) and np.array_equal(st[1], ref[1]), k
    monkeypatch.delenv("MTT_JIT_FWD_SMASK")


@pytest.mark.parametrize("river", ["gemm", "scan"])
def test_batched_config3_matches_oracle(solver_api, oracle, torch_dev, river, monkeypatch):
    """The benchmark's turn+river game (33%/75%/2.5x/all-in per street, nine
    river lines) on the batched path with generated sweeps, against the FP32
    oracle.  Iteration 1: every regret and strategy sum within tolerance.
    Iteration 2: regret matching plays uniform when the positive regrets sum
    to at most 1e-7, and entries whose iteration-1 sum is rounding noise
    around that threshold (value ties) can land on opposite sides in two
    implementations that sum in different orders; such a flip changes that
    combo's regrets by O(pot) there and at its ancestors (and its strategy
    sum there).  So at most 1e-7 of the entries may differ, and each must
    belong to a combo with an iteration-1 positive-regret sum in (0, 1e-6] in
    either implementation.  River showdown by the BF16 hi+lo GEMM or by the
    exact prefix scans."""
    monkeypatch.delenv("MTT_RIVER_GEMM", raising=False)
    monkeypatch.setenv("MTT_RIVER_SD", river)
    g, r0, r1 = solver_api.config3(0)
    print(f"config3 batched ({river} river showdown) vs oracle, 2 iterations", flush=True)
    s = solver_api.Solver(g, r0, r1, solver_api.MODE_BF16_SPLIT)
    assert s.jit_info()[0]
    h32, nodes, ni = _setup(solver_api, oracle, g, r0, r1, 1)
    s.run(1)
    h32.iterate(1)
    Rg, Sg = s.state()
    Rc, Sc = h32.state()
    assert np.abs(Rg - Rc).max() <= STATE_REL_BF16 * np.abs(Rc).max()
    assert np.abs(Sg - Sc).max() <= STATE_REL_BF16 * np.abs(Sc).max()
    # combos with a positive-regret sum near the uniform threshold somewhere
    amb = np.zeros(1216, bool)
    for R in (Rg, Rc):
        pos = np.maximum(R.reshape(ni, 4, 1216), 0).sum(axis=1)
        amb |= ((pos > 0) & (pos <= 1e-6)).any(axis=0)
    del Rg, Sg, Rc, Sc
    s.run(1)
    h32.iterate(1)
    Rg, Sg = s.state()
    Rc, Sc = h32.state()
    tol = STATE_REL_BF16 * np.abs(Rc).max()
    bad = np.argwhere(np.abs(Rg - Rc).reshape(ni, 4, 1216) > tol)
    print(f"iteration 2: {len(bad)} of {Rc.size} regrets beyond tolerance; "
          f"{int(amb.sum())} combos with near-threshold sums", flush=True)
    assert len(bad) <= 1e-7 * Rc.size
    assert amb[bad[:, 2]].all()
    del Rg, Rc
    bad = np.argwhere(np.abs(Sg - Sc).reshape(ni, 4, 1216) > STATE_REL_BF16 * np.abs(Sc).max())
    print(f"iteration 2: {len(bad)} strategy sums beyond tolerance", flush=True)
    assert len(bad) <= 1e-7 * Sc.size
    assert amb[bad[:, 2]].all()
    s.close()


def test_plain_config3_matches_oracle(solver_api, oracle, torch_dev):
    """The plain FP32 path on the benchmark's turn+river game against the
    FP32 oracle after one iteration (every regret and strategy sum), and
    both GPU paths' average strategies after 10 iterations against the FP64
    oracle (strategy L-inf over entries with strategy mass), with their FP64
    exploitabilities and the constant-sum identity of their profiles."""
    g, r0, r1 = solver_api.config3(0)
    print("config3 plain fp32 vs oracle, 1 iteration", flush=True)
    s = solver_api.Solver(g, r0, r1, solver_api.MODE_PLAIN_FP32)
    h32, nodes, ni = _setup(solver_api, oracle, g, r0, r1, 1)
    s.run(1)
    h32.iterate(1)
    Rg, Sg = s.state()
    Rc, Sc = h32.state()
    assert np.abs(Rg - Rc).max() <= STATE_REL_FP32 * np.abs(Rc).max()
    assert np.abs(Sg - Sc).max() <= STATE_REL_FP32 * np.abs(Sc).max()
    del Rg, Sg, Rc, Sc, h32
    b = solver_api.Solver(g, r0, r1, solver_api.MODE_BF16_SPLIT)
    b.run(10)
    s.run(9)
    print("config3: FP64 oracle, 10 iterations", flush=True)
    h64, _, _ = _setup(solver_api, oracle, g, r0, r1, 0)
    h64.iterate(10)
    ref = h64.strategy()
    S64 = h64.state()[1]
    e_ref = h64.exploitability(ref, 6.0)
    n_alive = int(s.info[2])
    w = S64.reshape(ni, 4, 1216).sum(1)[:, :n_alive]
    wm = w.max(1, keepdims=True)
    live = (w >= 1e-5 * wm) & (w > 0)
    heavy = (w >= 0.1 * wm) & (w > 0)
    del S64
    stats = {}
    for name, x in (("plain", s), ("batched", b)):
        st = x.strategy()
        d = np.abs(st.reshape(ni, 4, 1216) - ref.reshape(ni, 4, 1216)).max(1)[:, :n_alive]
        del st
        e = x.exploitability()
        frac = float(np.mean(d[live] > 1e-2))
        stats[name] = (frac, float(d[heavy].max()), e)
        print(f"config3 {name}: strategy L-inf vs FP64 {d[live].max():.3g} over entries with >= 1e-5 of their "
              f"infoset's strategy mass ({frac:.3g} of them differ by > 1e-2), {d[heavy].max():.3g} over entries "
              f"with >= 0.1; exploitability {e[5]:.4f}% (FP64 oracle {e_ref[5]:.4f}%)", flush=True)
        del d
    # FP32 itself moves a few entries at value ties (regret sums landing on
    # either side of the regret-matching threshold), more or fewer with the
    # summation order: the FP32 oracle against the FP64 oracle on this game
    # after 10 iterations gives 5e-8 of the >= 1e-5 entries off by more than
    # 1e-2 and L-inf 0.007 over the >= 0.1 entries; the plain GPU path
    # (different summation order) 3.5e-6 and 0.026.  Exploitability, which
    # these entries barely move, is the tight check.
    for name, (frac, hv, e) in stats.items():
        assert frac <= 1e-4, (name, frac)
        assert hv < 0.1, (name, hv)
        assert abs(e[5] - e_ref[5]) < 0.01, (name, e[5], e_ref[5])
        assert abs(e[2] + e[3] - 6.0 * e[4]) <= 1e-6 * 6.0 * e[4]
    s.close()
    b.close()


@pytest.mark.parametrize("mode", ["sweep", "scan", "gemm"])
@pytest.mark.parametrize("which", ["config1", "config3_small", "config3"])
def test_river_sweep_matches_dense_showdown(solver_api, torch_dev, which, mode, monkeypatch):
    """River showdown values by the strength-order sweep (MTT_RIVER_SD=sweep)
    or the prefix scans (MTT_RIVER_SD=scan) against S r of the same columns
    in FP64, S = sign x blocker of the board's strength order (the
    reference's showdown, equity matrix x reach,
    honest_multistreet_solver.cu:128-196, as 2 (E - 1/2) r).  The sweep sums
    in exact 2^48 fixed point and converts the sum to FP32 from its two 32-bit
    halves (within half an ulp plus 2^-40), so every value agrees to FP32
    rounding (|err| <= 2^-23 |exact| + 2e-12).  The scan sums exactly in
    integer units u = 2^(e - 27) of the column (2^e > the column total), so
    each value is off by at most half a unit per live combo plus the final
    rounding, and on average by far less.  The register-streamed BF16 hi + lo
    GEMM (MTT_RIVER_GEMM=707: every column tile, a last tile of at most 16
    columns by the 16-column tail kernel) agrees to its FP32
    accumulation: within 2e-5 of the column's reach total.  Padded combos
    stay 0."""
    monkeypatch.delenv("MTT_RIVER_GEMM", raising=False)
    if mode == "gemm":
        monkeypatch.delenv("MTT_RIVER_SD", raising=False)
        monkeypatch.setenv("MTT_RIVER_GEMM", "707")
    else:
        monkeypatch.setenv("MTT_RIVER_SD", mode)
    S = solver_api
    g, r0, r1 = getattr(S, which)(1 if which == "config3_small" else 0)
    s = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
    s.run(3)
    nb = 1 if which == "config1" else 49 * 48
    for board in sorted({0, nb // 2, nb - 1}):
        reach, vals, cards, grank, nl = s.river_showdown(board)
        c = cards[:nl].astype(np.int64)
        gr = grank[:nl].astype(np.int64)
        sign = np.sign(gr[:, None] - gr[None, :]).astype(np.float64)
        cm = np.zeros((nl, 52))
        cm[np.arange(nl), c[:, 0]] = 1.0
        cm[np.arange(nl), c[:, 1]] = 1.0
        Sm = np.where(cm @ cm.T > 0, 0.0, sign)
        R = np.rint(reach[:, :nl].astype(np.float64) * 2.0 ** 48)
        exact = (R @ Sm.T) * 2.0 ** -48
        err = np.abs(vals[:, :nl].astype(np.float64) - exact)
        print(f"{which} board {board} ({mode}): {nl} combos, {reach.shape[0]} columns, max |value| "
              f"{np.abs(exact).max():.4g}, max err {err.max():.3g}", flush=True)
        assert np.abs(exact).max() > 0
        if mode == "gemm":
            # the BF16 hi + lo reach the GEMM multiplies is what the export returns (hi + lo), so only the
            # FP32 accumulation of the MFMAs (2 x 68 K steps of 16 products) separates it from the FP64 sum
            tot = reach[:, :nl].astype(np.float64).sum(axis=1)[:, None]
            assert np.all(err <= 2e-5 * tot + 1e-12), (board, float((err / (tot + 1e-30)).max()))
        elif mode == "sweep":
            bound = 2.0 ** -23 * np.abs(exact) + 2e-12
            if not np.all(err <= bound):
                ci, hi = np.unravel_index(np.argmax(err / bound), err.shape)
                same = np.flatnonzero(gr == gr[hi])
                print(f"worst: column {ci}, combo {hi} (cards {c[hi]}, tie group of {len(same)}), exact "
                      f"{exact[ci, hi]:.10g}, value {vals[ci, hi]:.10g}, reach {reach[ci, hi]:.6g}; "
                      f"{int((err > bound).sum())} entries over the bound", flush=True)
            assert np.all(err <= bound), (board, float(err.max()))
        else:
            tot = reach[:, :nl].astype(np.float64).sum(axis=1)
            unit = np.where(tot > 0, 2.0 ** (np.floor(np.log2(np.maximum(tot, 1e-300))) + 1 - 27), 0.0)[:, None]
            assert np.all(err <= nl * unit + 2.0 ** -23 * np.abs(exact) + 1e-12), (board, float(err.max()))
            assert err.mean() <= 4.0 * unit.max() + 1e-12
        assert not np.any(vals[:, nl:])
    s.close()


@pytest.mark.parametrize("which", ["config3_small", "config3"])
def test_river_gemm_forms_agree(solver_api, torch_dev, which, monkeypatch):
    """Every river showdown GEMM form the autotuner chooses from (hero-major,
    with S as bytes or at two bits per pair, the latter with the reach
    fetched one or two chunks ahead;
    column-major with all, half or a third of the column tiles per
    workgroup, prefetching 2 or 3 chunks) accumulates in the same order, so
    the solver states after 3 iterations are identical."""
    S = solver_api
    g, r0, r1 = getattr(S, which)(1 if which == "config3_small" else 0)
    ref = None
    monkeypatch.setenv("MTT_RIVER_GEMM", "102")
    s = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
    ntiles = (s.river_showdown(0)[0].shape[0] + 31) // 32
    s.close()
    # the column-major forms the autotuner tries: all, half or a third of the tiles per workgroup
    colmajor = sorted({f"{-(-ntiles // d)}{pd}" for d in (1, 2, 3) for pd in (2, 3) if -(-ntiles // d) <= 9})
    for shape in ("102", "101", "100", *colmajor):
        monkeypatch.setenv("MTT_RIVER_GEMM", shape)
        s = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
        s.run(3)
        R, Sm = s.state()
        s.close()
        print(f"{which}: river GEMM form {shape} done", flush=True)
        if ref is None:
            ref = (R, Sm)
        else:
            assert np.array_equal(R, ref[0]) and np.array_equal(Sm, ref[1]), shape
    # the register-streamed forms (5, 6, 7 or 12 hero tiles per wave) share one
    # accumulation order of their own (64-villain chunks): identical to each
    # other after 3 iterations, and after one iteration (uniform strategies,
    # so the same reach goes into both products) within FP32 rounding of the
    # forms above
    ref7 = None
    for shape in ("707", "706", "705", "712"):
        monkeypatch.setenv("MTT_RIVER_GEMM", shape)
        s = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
        s.run(3)
        R, Sm = s.state()
        s.close()
        print(f"{which}: river GEMM form {shape} done", flush=True)
        if ref7 is None:
            ref7 = (R, Sm)
        else:
            assert np.array_equal(R, ref7[0]) and np.array_equal(Sm, ref7[1]), shape
    one = {}
    for shape in ("102", "707"):
        monkeypatch.setenv("MTT_RIVER_GEMM", shape)
        s = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
        s.run(1)
        one[shape] = s.state()[0]
        s.close()
    scale = max(float(np.abs(one["102"]).max()), 1e-30)
    dR = float(np.abs(one["707"] - one["102"]).max())
    print(f"{which}: one iteration, form 707 vs 102: max |dR| {dR:.3g} of max |R| {scale:.3g}", flush=True)
    assert dR <= 1e-5 * scale, dR
    monkeypatch.delenv("MTT_RIVER_GEMM")

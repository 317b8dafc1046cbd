"""GPU parity of the batched flop-subgame solver (mtt_batch_*): instances of
different trees (including the widest and deepest menus), DCFR, against the
CPU oracle per instance; graph replay against eager launches; the full-size
512-instance configuration."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE_REL = 1e-3   # BF16 hi+lo reach against exact W - L counts vs the FP32 oracle, after a few iterations


def _pick(S, games):
    """Indices of a few structurally different instances: the most slots, the
    deepest, the smallest, and one in between."""
    info = []
    for k, g in enumerate(games):
        nodes, ni, _ = S.tree(g)
        dec = nodes["type"] == 0
        info.append((int(nodes["num_actions"][dec].sum()), len(nodes), k))
    info.sort()
    return sorted({info[-1][2], info[0][2], info[len(info) // 2][2], info[-9][2]})


def _oracle_state(S, oracle, g, r, iters):
    nodes, ni, _ = S.tree(g)
    board = list(g.board)[:3]
    gi, _ = S.alive_index(board)
    h = oracle.HU(nodes, ni, board, r[0][gi], r[1][gi], 1)
    h.iterate(iters, scheme=1, alpha=1.5, beta=0.0, gamma=2.0)
    return h.state()


def test_batch_matches_oracle(solver_api, oracle, torch_dev):
    S = solver_api
    games, ranges = S.config4(0, 64)
    picks = _pick(S, games)
    print("batch instances", picks, flush=True)
    sub = [games[k] for k in picks]
    b = S.Batch(sub, ranges[picks])
    b.run(5)
    for j, k in enumerate(picks):
        Rg, Sg = b.state(j)
        Rc, Sc = _oracle_state(S, oracle, games[k], ranges[k], 5)
        assert np.abs(Rg - Rc).max() <= STATE_REL * np.abs(Rc).max(), k
        assert np.abs(Sg - Sc).max() <= STATE_REL * np.abs(Sc).max(), k
    ht, ms = b.tuning()
    assert ht in (2, 4, 12, 13, 14)
    b.close()


def test_batch_wide_trees_match_oracle(solver_api, oracle, torch_dev):
    """The widest trees of the 512-instance configuration (68 and 74 action
    slots: regrets held in three 32-register vectors) batched with narrow
    ones, against the oracle after 5 iterations."""
    S = solver_api
    games, ranges = S.config4(0, 512)
    slots = []
    for k, g in enumerate(games):
        nodes, _, _ = S.tree(g)
        slots.append((int(nodes["num_actions"][nodes["type"] == 0].sum()), k))
    slots.sort()
    picks = sorted({slots[-1][1], slots[-4][1], slots[0][1], slots[len(slots) // 2][1]})
    assert slots[-1][0] > 64 and slots[-4][0] > 64
    print("wide batch instances", [(k, dict((b, a) for a, b in slots)[k]) for k in picks], flush=True)
    b = S.Batch([games[k] for k in picks], ranges[picks])
    b.run(5)
    for j, k in enumerate(picks):
        Rg, Sg = b.state(j)
        Rc, Sc = _oracle_state(S, oracle, games[k], ranges[k], 5)
        assert np.abs(Rg - Rc).max() <= STATE_REL * np.abs(Rc).max(), k
        assert np.abs(Sg - Sc).max() <= STATE_REL * np.abs(Sc).max(), k
    b.close()


def test_batch_graph_replay_matches_eager(solver_api, torch_dev):
    """Iterations replayed from the captured graph equal eager launches, over
    several replays with host reads in between."""
    S = solver_api
    games, ranges = S.config4(1, 16)
    a = S.Batch(games, ranges)
    e = S.Batch(games, ranges)
    e.set_graph(False)
    for step in (1, 3, 4):
        a.run(step)
        e.run(step)
        for k in (0, 7, 15):
            Ra, Sa = a.state(k)
            Re, Se = e.state(k)
            assert np.array_equal(Ra, Re) and np.array_equal(Sa, Se), (step, k)
    a.close()
    e.close()


def test_flop_gemm_forms_agree(solver_api, torch_dev, monkeypatch):
    """Every flop showdown GEMM form the autotuner chooses from (W - L staged
    in LDS with 2 or 4 hero tiles per workgroup; register-streamed with 2, 3
    or 4 hero tiles per wave) adds the four BF16 products of a K step in the
    same order per accumulator, so the states after 4 iterations are
    identical for every instance (one or two column tiles; the last hero
    group holds 2 of 2, 3 or 4 tiles, and nothing may land outside an
    instance's own values)."""
    S = solver_api
    games, ranges = S.config4(1, 24)
    ref = None
    for form in ("2", "4", "12", "13", "14"):
        monkeypatch.setenv("MTT_FLOP_GEMM", form)
        b = S.Batch(games, ranges)
        assert b.tuning()[0] == int(form)
        b.run(4)
        st = [b.state(k) for k in range(len(games))]
        b.close()
        print(f"flop GEMM form {form} done", flush=True)
        if ref is None:
            ref = st
        else:
            for (Ra, Sa), (Rb, Sb) in zip(st, ref):
                assert np.array_equal(Ra, Rb) and np.array_equal(Sa, Sb), form
    monkeypatch.delenv("MTT_FLOP_GEMM")


@pytest.mark.skip(reason="the persistent flop iteration is switched off: its first GPU runs did not finish within "
                         "three minutes (twice), while a CPU simulation of its queue order and counters completes; "
                         "cause unknown, not run again")
def test_persistent_iteration_matches_phase_launches(solver_api, torch_dev, monkeypatch):
    """The persistent iteration (one launch per iteration: workgroups take the
    forward halves, GEMM hero groups and backward slices of every instance
    from a queue, waiting on per-instance completion counters) runs the same
    bodies as the three phase launches, so the states are identical: on the
    512-instance configuration's 8 widest and its deepest trees mixed with 56
    others (three register vectors, 8-level stacks), graph-replayed and
    launched one by one, over several runs."""
    S = solver_api
    games, ranges = S.config4(0, 512)
    info = []
    for k, g in enumerate(games):
        nodes, _, _ = S.tree(g)
        info.append((int(nodes["num_actions"][nodes["type"] == 0].sum()), len(nodes), k))
    picks = sorted({k for _, _, k in sorted(info)[-8:]} | {max(info, key=lambda x: x[1])[2]} | set(range(56)))
    sub, rsub = [games[k] for k in picks], ranges[picks]
    a = S.Batch(sub, rsub)
    p = S.Batch(sub, rsub)
    e = S.Batch(sub, rsub)
    e.set_graph(False)
    for step in (1, 2, 5):
        monkeypatch.setenv("MTT_FLOP_PERSISTENT", "0")
        a.run(step)
        monkeypatch.setenv("MTT_FLOP_PERSISTENT", "1")
        p.run(step)
        e.run(step)
        assert a.iter_mode()[0] == 0 and p.iter_mode()[0] == 1 and e.iter_mode()[0] == 1
        for j in range(len(picks)):
            Ra, Sa = a.state(j)
            for o in (p, e):
                Ro, So = o.state(j)
                assert np.array_equal(Ra, Ro) and np.array_equal(Sa, So), (step, picks[j])
        print(f"persistent vs phase launches: {len(picks)} instances equal after {a.iterations} iterations",
              flush=True)
    monkeypatch.delenv("MTT_FLOP_PERSISTENT")
    for b in (a, p, e):
        b.close()


def test_config4_full_size(solver_api, oracle, torch_dev):
    """512 seeded instances, DCFR(1.5, 0, 2), 300 iterations in graph
    replay.  Instances 0, 311 and the last (511) match the oracle after 3
    iterations.  After 300, for instances 0, 255, 511 and the widest tree:
    the FP64 best response of the batch's average strategy equals the
    oracle's FP64 best response of the same strategy (to the FP32 rounding
    of the equity matrix the GPU reads); its exploitability is
    within 0.005% of the pot of the FP64 oracle's own 300-iteration DCFR solve;
    the profile is constant-sum; and the average strategies agree with the
    FP64 oracle's where they carry strategy mass."""
    import time
    S = solver_api
    games, ranges = S.config4(0, 512)
    print("config4: 512 instances, DCFR 300 iterations", flush=True)
    b = S.Batch(games, ranges)
    b.run(3)
    for k in (0, 311, 511):
        Rg, Sg = b.state(k)
        Rc, Sc = _oracle_state(S, oracle, games[k], ranges[k], 3)
        assert np.abs(Rg - Rc).max() <= STATE_REL * np.abs(Rc).max(), k
        assert np.abs(Sg - Sc).max() <= STATE_REL * np.abs(Sc).max(), k
    t = time.time()
    b.run(292)
    dt = time.time() - t
    print(f"config4: {292 / dt:.1f} iterations/s over 512 instances (graph replay); GEMM tuning {b.tuning()}; "
          f"iteration form {b.iter_mode()}", flush=True)
    assert b.iter_mode()[0] == 0
    b.timing(True)
    b.run(5)
    kt = b.kernel_times()
    print("config4 per-kernel ms/iteration:", {k: round(v[0] / v[1], 4) for k, v in kt.items()}, flush=True)
    assert b.iterations == 300
    wide = max(range(512), key=lambda k: b.tree_info(k)[1])
    for k in sorted({0, 255, 511, wide}):
        g = games[k]
        nodes, ni, _ = S.tree(g)
        board = list(g.board)[:3]
        gi, _ = S.alive_index(board)
        pot = float(g.pot)
        st = b.strategy(k)
        e = b.exploitability(k)
        h = oracle.HU(nodes, ni, board, ranges[k][0][gi], ranges[k][1][gi], 0)
        e_same = h.exploitability(st, pot)
        # the GPU best response reads E = (W + T/2) / 990 rounded to FP32; the oracle's is exact
        assert np.allclose(e[:5], e_same[:5], rtol=1e-7, atol=1e-12), (k, e, e_same)
        assert abs(e[2] + e[3] - pot * e[4]) <= 1e-6 * pot * e[4]
        h.iterate(300, scheme=1, alpha=1.5, beta=0.0, gamma=2.0)
        ref = h.strategy()
        e_ref = h.exploitability(ref, pot)
        w = h.state()[1].reshape(ni, 4, 1216).sum(1)[:, :len(gi)]
        live = (w >= 1e-3 * w.max(1, keepdims=True)) & (w > 0)
        d = np.abs(st.reshape(ni, 4, 1216) - ref.reshape(ni, 4, 1216)).max(1)[:, :len(gi)]
        print(f"config4 instance {k} ({len(nodes)} nodes): exploitability {e[5]:.4f}% pot (FP64 oracle solve "
              f"{e_ref[5]:.4f}%), strategy L-inf vs FP64 {d[live].max():.3g} over entries with >= 1e-3 of their "
              f"infoset's strategy mass", flush=True)
        # measured: exploitabilities equal to 1e-4 % pot, L-inf 1e-4
        assert abs(e[5] - e_ref[5]) < 0.005, (k, e[5], e_ref[5])
        assert d[live].max() < 0.01, k
    b.close()

"""Generated per-template sweeps of the batched path: the source emitted for a
game's street templates, and a hipRTC compile of it for gfx950 (no device
needed)."""
import re

import pytest

from conftest import PKG, _load
import os


@pytest.fixture(scope="module")
def S(mtt):
    return _load("mtt_solver", os.path.join(PKG, "solver.py"))


def test_sweep_source_has_one_case_per_template(S, monkeypatch):
    g, _, _ = S.config3(0)
    src = S.sweep_source(g)
    nodes, ni, _ = S.tree(g)
    # the turn template, then one template per river line (one per turn leaf)
    n_river = len(re.findall(r"^case \d+: \{", src.split("mtt_jit_fwd1")[1].split("mtt_jit_bwd1")[0], re.M))
    n_turn = len(re.findall(r"^case \d+: \{", src.split("mtt_jit_fwd0")[1].split("mtt_jit_bwd0")[0], re.M))
    assert n_turn == 1
    assert n_river >= 2
    for k in ("mtt_jit_fwd0", "mtt_jit_bwd0", "mtt_jit_fwd1", "mtt_jit_bwd1"):
        assert f"void {k}(" in src
    # every action slot of a river template is read once for its regret and once for its strategy sum,
    # and the strategy sum is written back by a plain store (one thread owns each combo): no atomics.
    # By default the sums go through the bounded buffer descriptor after the walk (zero-reach lanes out
    # of range); MTT_JIT_FWD_SMASK=0 reads them with the regrets in the prologue.
    # By default each slice's regrets come from the register set the previous slice loaded (Rn*: the
    # first slice's before the slice loop, each later slice's during the walk before it, so every slot
    # has two such loads); MTT_JIT_FWD_XPRE=0 has each slice load its own (Rl).
    fwd1 = src.split("void mtt_jit_fwd1(")[1].split("void mtt_jit_bwd1(")[0]
    n_rv = len(re.findall(r"const float Rv(\d+) = Rn\1;", fwd1))
    assert n_rv > 0
    assert 2 * n_rv == len(re.findall(r"\bRn\d+ = Rp\[", fwd1))
    assert not re.search(r"const float Rv\d+ = Rl\[", fwd1)
    monkeypatch.setenv("MTT_JIT_FWD_XPRE", "0")
    fwd1_own = S.sweep_source(g).split("void mtt_jit_fwd1(")[1].split("void mtt_jit_bwd1(")[0]
    assert n_rv == len(re.findall(r"const float Rv\d+ = Rl\[", fwd1_own)) and "Rn0" not in fwd1_own
    monkeypatch.delenv("MTT_JIT_FWD_XPRE")
    assert n_rv == len(re.findall(r"const float Sv\d+ = __builtin_bit_cast\(float, __builtin_amdgcn_raw_buffer_load_b32\(rsS, so\d+, ", fwd1))
    assert n_rv == len(re.findall(r"__builtin_amdgcn_raw_buffer_store_b32\(__builtin_bit_cast\(uint32_t, Sv\d+ \+ wr\d+ \* ", fwd1))
    assert "Atomic" not in fwd1
    monkeypatch.setenv("MTT_JIT_FWD_SMASK", "0")
    fwd1 = S.sweep_source(g).split("void mtt_jit_fwd1(")[1].split("void mtt_jit_bwd1(")[0]
    assert n_rv == len(re.findall(r"const float Sv\d+ = Sl\[", fwd1))
    assert n_rv == len(re.findall(r"\bSg\[\d+\] = Sv\d+ \+ wr \* ", fwd1))
    assert "Atomic" not in fwd1 and "rsS" not in fwd1
    assert "#pragma once" not in src


def test_sweep_source_deterministic(S):
    g, _, _ = S.config1(0)
    assert S.sweep_source(g) == S.sweep_source(g)


@pytest.mark.parametrize("which", ["config1", "config3"])
def test_sweeps_compile_for_gfx950(S, which):
    g, _, _ = getattr(S, which)(0)
    nbytes, secs = S.sweep_compile(g)
    assert nbytes > 1000

"""Error paths of the solve handles on the GPU: inputs the library must
refuse with MTT_ERR_ARG (NaN, negative or empty ranges; bad bubble factors;
flops with repeated cards), and a device allocation that fails -- the device
memory filled first -- which must come back as MTT_ERR_OOM with the library
usable again once memory is free."""
import os

import numpy as np
import pytest

from conftest import PKG, _load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S(mtt):
    return _load("mtt_solver", os.path.join(PKG, "solver.py"))


@pytest.fixture(scope="module")
def M(mtt, S):
    return _load("mtt_multi", os.path.join(PKG, "multi.py"))


def _status(mtt, fn, *args, **kw):
    with pytest.raises(mtt.MTTError) as e:
        fn(*args, **kw)
    return e.value.status, str(e.value)


def test_solver_refuses_bad_ranges(mtt, S, torch_dev):
    g, r0, r1 = S.config1(0)
    for label, bad in (("NaN", np.where(np.arange(1326) == 700, np.nan, r0)),
                       ("negative", np.where(np.arange(1326) == 5, -0.25, r0)),
                       ("infinite", np.where(np.arange(1326) == 9, np.inf, r0)),
                       ("empty", np.zeros(1326)),
                       ("too large", r0 * 1e6)):
        st, msg = _status(mtt, S.Solver, g, bad.astype(np.float32), r1, S.MODE_BF16_SPLIT)
        print(f"range {label}: {msg}", flush=True)
        assert st == mtt.MTT_ERR_ARG, (label, msg)
    st, msg = _status(mtt, S.Solver, g, r0, r1, 7)
    assert st == mtt.MTT_ERR_ARG, msg
    # the library still works after the refusals
    s = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
    s.run(3)
    assert np.isfinite(s.exploitability()).all()
    s.close()


def test_batch_and_multi_refuse_bad_inputs(mtt, S, M, torch_dev):
    games, ranges = S.config4(0, 4)
    bad = ranges.copy()
    bad[2, 1, 17] = np.nan
    st, msg = _status(mtt, S.Batch, games, bad)
    assert st == mtt.MTT_ERR_ARG, msg
    game, flops, rg, bf, _, _ = M.config5(0, 3)
    f2 = flops.copy()
    f2[1] = (7, 7, 30)
    st, msg = _status(mtt, M.Multi, game, f2, rg, bf)
    assert st == mtt.MTT_ERR_ARG, msg
    r2 = rg.copy()
    r2[0, 2, 11] = -1.0
    st, msg = _status(mtt, M.Multi, game, flops, r2, bf)
    assert st == mtt.MTT_ERR_ARG, msg
    b2 = np.array(bf, np.float32).copy()
    b2.flat[5] = np.nan
    st, msg = _status(mtt, M.Multi, game, flops, rg, b2)
    assert st == mtt.MTT_ERR_ARG, msg


def test_allocation_failure_is_reported_and_recovered(mtt, S, torch_dev):
    """Fill the device's free memory but 2 GiB, then ask for the turn+river
    solver (~18 GB): it must fail with MTT_ERR_OOM (or MTT_ERR_HIP, should
    the runtime's own allocation fail first) and a message, leaving nothing
    behind; after the memory is released the same call succeeds."""
    import torch
    g, r0, r1 = S.config3(0)
    free, total = torch.cuda.mem_get_info()
    print(f"device memory: {free / 2**30:.1f} GiB free of {total / 2**30:.1f}", flush=True)
    hog = torch.empty(int(free - (2 << 30)), dtype=torch.uint8, device=torch_dev)
    try:
        st, msg = _status(mtt, S.Solver, g, r0, r1, S.MODE_BF16_SPLIT)
        print(f"solver with the memory full: {msg}", flush=True)
        assert st in (mtt.MTT_ERR_OOM, mtt.MTT_ERR_HIP), msg
    finally:
        del hog
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    free2, _ = torch.cuda.mem_get_info()
    assert free2 >= free - (1 << 30), "memory of the failed creation not returned"
    s = S.Solver(g, r0, r1, S.MODE_BF16_SPLIT)
    s.run(2)
    assert np.isfinite(s.exploitability()).all()
    s.close()

"""The calibration kernels behind bench.py --ceilings (include/tbj_calib.h):
the scatter-run pattern writes every output slot exactly once (a permutation
of the input), and the random-access kernels run and report a time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bins,run", [(1, 4096), (256, 16), (2048, 2), (8192, 1)])
def test_calibrate_runs_is_a_permutation(tbj, bins, run):
    import torch
    n = 8 * bins * run * 4  # four tiles per XCD
    src = torch.arange(n, dtype=torch.int64, device="cuda:0")
    dst = torch.full((n,), -1, dtype=torch.int64, device="cuda:0")
    ms = tbj.calibrate_runs(src, dst, bins, run)
    out = dst.cpu().numpy()
    assert ms > 0
    assert np.array_equal(np.sort(out), np.arange(n))
    # tile t (XCD t % 8) writes its run of bin b at region (t % 8, b), slot (t // 8) * run
    ntiles, tile = n // (bins * run), bins * run
    t, j = 13 % ntiles, tile - 1
    x, tl, b, off = t % 8, t // 8, j // run, j % run
    cap = ntiles // 8 * run
    assert out[(x * bins + b) * cap + tl * run + off] == t * tile + j


def test_calibrate_rejects_bad_shapes(tbj):
    import torch
    src = torch.zeros(1000, dtype=torch.int64, device="cuda:0")
    with pytest.raises(tbj.TbjError):
        tbj.calibrate_runs(src, src.clone(), 256, 16)  # n not a multiple of 8 tiles


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_calibrate_random_runs(tbj, kind):
    import torch
    table = torch.full((1 << 20,), 0x01010101, dtype=torch.int64, device="cuda:0")
    sink = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    before = table.clone()
    ms = tbj.calibrate_random(table, 8 << 20, 1 << 20, kind, sink)
    assert ms > 0
    if kind == 2:  # 2^20 atomic increments, all inside the table
        diff = (table.view(torch.int32) - before.view(torch.int32)).sum().item()
        assert diff == 1 << 20
    else:
        assert torch.equal(table, before)

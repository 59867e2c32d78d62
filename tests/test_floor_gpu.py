"""fmx_gather_floor, the measurement kernel that reads X rows in a given
order and nothing else: it runs for every supported row width, clamps
column ids past the end of X instead of reading beyond it, leaves X
untouched, and refuses row widths it does not support."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,dt", [(4, "float32"), (16, "float32"), (32, "float32"), (8, "float64"),
                                  (16, "float64"), (256, "float32")])
def test_gather_floor_runs_and_clamps(fmx, k, dt):
    import torch
    rows = 5000
    X = torch.rand((rows, k), dtype=getattr(torch, dt), device="cuda")
    X0 = X.clone()
    g = torch.Generator(device="cuda")
    g.manual_seed(k)
    cols = torch.randint(0, rows + 100, (123457,), device="cuda", generator=g, dtype=torch.int32)
    fmx.gather_floor(X, cols)
    fmx.gather_floor(X, cols[:1])
    torch.cuda.synchronize()
    assert torch.equal(X, X0)


def test_gather_floor_errors(fmx):
    import torch
    X = torch.rand((100, 3), device="cuda")
    with pytest.raises(fmx.FmxError) as e:  # 12-byte rows
        fmx.gather_floor(X, torch.zeros(10, dtype=torch.int32, device="cuda"))
    assert e.value.status == fmx.EINVAL
    X = torch.rand((100, 24), device="cuda")
    with pytest.raises(fmx.FmxError) as e:  # 96-byte rows: not 16 times a power of two
        fmx.gather_floor(X, torch.zeros(10, dtype=torch.int32, device="cuda"))
    assert e.value.status == fmx.EINVAL

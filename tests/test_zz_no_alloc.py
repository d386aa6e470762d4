"""mfa_reduce_noalloc (partial sums in the library's file-scope device array)
against mfa_reduce (an allocated workspace), through the C ABI: the same bits
for every reduction, on two calls in a row, the second on a smaller box, so
that partial sums the first call left in the array cannot leak into the
second; FP64 and reduced storage."""
import numpy as np
import pytest

from muphfasa_amd import abi
from oracle import oracle as O
from tests.gpu_util import random_state

pytestmark = pytest.mark.gpu

KINDS = ["mass", "momentum", "health", "comp_sum"]


def _fields(dev, lat, n, dtype, seed):
    import torch
    L = O.Lattice(lat)
    f0 = np.ascontiguousarray(O.valid(L, random_state(L, n, seed=seed)))
    t64 = torch.from_numpy(f0).to(dev)
    f64 = abi.field(t64, n, 0, L.D)
    if dtype == "f64":
        return L, t64, f64
    tt = {"f32": torch.float32, "f16dev": torch.float16, "bf16dev": torch.bfloat16}[dtype]
    t = torch.empty(t64.shape, dtype=tt, device=dev)
    f = abi.field(t, n, 0, L.D, abi.DTYPE[dtype])
    abi.convert(f64, f, abi.LATTICE[lat])
    return L, t, f


@pytest.mark.parametrize("lat,sizes", [("D3Q19", ([40, 33, 30], [13, 7, 6])),
                                       ("D2Q9", ([300, 170], [17, 5])),
                                       ("D3Q27", ([20, 24, 18], [3, 2, 1]))])
@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16dev"])
def test_noalloc_equals_reduce_on_two_calls(dev, lat, sizes, dtype):
    import torch
    li = abi.LATTICE[lat]
    for call, n in enumerate(sizes):
        print(f"case noalloc lattice={lat} dtype={dtype} call={call} n={n}", flush=True)
        L, t, f = _fields(dev, lat, n, dtype, seed=100 + call)
        for kind in KINDS:
            if kind == "comp_sum" and dtype != "f64":
                continue
            want = torch.full((32,), float("nan"), dtype=torch.float64, device=dev)
            got = torch.full((32,), float("nan"), dtype=torch.float64, device=dev)
            abi.reduce(f, li, abi.RED[kind], want.data_ptr())
            abi.reduce_noalloc(f, li, abi.RED[kind], got.data_ptr())
            torch.cuda.synchronize()
            nv = {"mass": 1, "momentum": L.D, "health": 2, "comp_sum": L.Q}[kind]
            w, g = want.cpu().numpy()[:nv], got.cpu().numpy()[:nv]
            assert np.isfinite(w).all()
            assert np.array_equal(w, g), (kind, w, g)
            assert np.isnan(got.cpu().numpy()[nv:]).all()


def test_noalloc_refuses_more_than_27_components(dev):
    import torch
    t = torch.zeros((28, 1, 4, 4), dtype=torch.float64, device=dev)
    out = torch.zeros(32, dtype=torch.float64, device=dev)
    with pytest.raises(abi.MfaError) as e:
        abi.reduce_noalloc(abi.field(t, [4, 4], 0, 2), abi.LATTICE["D2Q9"], abi.RED["comp_sum"],
                           out.data_ptr())
    assert "27 components" in str(e.value)

"""fsp_net_forward_static_device, the network's forward pass that reserves
no device memory (activations in the library's static device array), against
fsp_net_forward_device, which reserves its activations, on the same inputs:
bit for bit, on two calls in a row, the second on a smaller grid."""
import numpy as np
import pytest

from fluidisim_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def features(shape, seed):
    """Random binary16 network input, 4 channels per cell."""
    n = shape[0] * shape[1] * shape[2]
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(4 * n, generator=g).to(torch.float16)
    return x.view(torch.int16).cuda()


def test_static_forward_equals_the_reserving_forward():
    from fluidisim_amd.learned import WEIGHTS, Network
    net = Network()
    net.load(WEIGHTS)
    # the largest grid the static array holds, then a smaller one
    for shape, seed in (((128, 32, 64), 1), ((64, 16, 32), 2)):
        n = shape[0] * shape[1] * shape[2]
        x = features(shape, seed)
        y_ref = net.forward(x, shape)
        y = torch.full((n,), float("nan"), device="cuda")
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        st = net.forward_static(x, shape, y)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        print(f"[static forward] {shape}: status {st}, free bytes {free0} -> {free1}", flush=True)
        assert st == _lib.FSP_OK
        assert free0 == free1
        a, b = y.cpu().numpy(), y_ref.cpu().numpy()
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"max diff {np.abs(a - b).max()}"
    net.close()


def test_static_forward_refuses_grids_beyond_its_array():
    from fluidisim_amd.learned import WEIGHTS, Network
    net = Network()
    net.load(WEIGHTS)
    shape = (128, 64, 64)  # 524 288 cells, twice the static array's grid
    n = shape[0] * shape[1] * shape[2]
    x = features(shape, 3)
    y = torch.zeros(n, device="cuda")
    torch.cuda.synchronize()
    assert net.forward_static(x, shape, y) == _lib.FSP_SHAPE
    assert float(y.abs().max()) == 0.0
    # an unsupported shape (x not a multiple of 64) is refused as by fsp_net_forward_device
    assert net.forward_static(x, (96, 16, 32), y) == _lib.FSP_SHAPE
    net.close()

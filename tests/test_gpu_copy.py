"""GPU: mgpu_copy, the plain HBM copy kernel behind the bench's copy ceiling."""
import pytest
import torch

import mgpu_hip
from gpu_util import banner

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [0, 1, 15, 16, 17, 4096, 16384 - 16, 16384, 16384 + 48,
                                    (1 << 20) + 3, (1 << 26) + 16 * 1000 + 5])
def test_copy_kernel(nbytes):
    """mgpu_copy (the HBM copy behind the bench's copy ceiling) copies every
    byte, whole 16-byte words, partial tiles and the byte tail alike, and
    touches nothing past the end."""
    banner("copy", nbytes)
    g = torch.Generator(device="cpu").manual_seed(nbytes)
    src = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, generator=g).to("cuda")
    dst = torch.full((nbytes + 64,), 0xA5, dtype=torch.uint8, device="cuda")
    mgpu_hip.copy(dst[:nbytes], src)
    torch.cuda.synchronize()
    assert bool(torch.equal(dst[:nbytes], src))
    assert bool((dst[nbytes:] == 0xA5).all())


def test_copy_rejects_misaligned():
    src = torch.zeros(64, dtype=torch.uint8, device="cuda")
    dst = torch.zeros(64, dtype=torch.uint8, device="cuda")
    with pytest.raises(mgpu_hip.MgpuError):
        mgpu_hip.copy(dst[1:33], src[:32])

"""The L1 + D-SSIM kernels on constant images against the closed form of
SSIM at an interior pixel (the same mixed-second-difference construction as
test_oracle_stages.py::test_ssim_constant_images_closed_form): pins the
kernel's C1, C2 and window taps (ssim.wgsl:29-34) independently of the
oracle. The GPU's f32 partial sums bound the precision at ~1e-5 per image
sum, far below the 1e-2 a swap of C1 and C2 moves the value."""

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _gpu_ssim_sum(gpu, w, h, a, b):
    import splattar_amd as sa
    img = gpu.alloc(np.float32, (h, w, 4))
    dl = gpu.alloc(np.float32, (h, w, 4))
    img.write(np.full((h, w, 4), a, np.float32))
    loss = sa.SsimLoss(gpu, w, h, 0.2, img, dl, dssim_scale=1.0)
    loss.target.write(np.full((h, w, 4), b, np.float32))
    loss.encode()
    l1, dssim = loss.read_losses()
    loss.close()
    img.free()
    dl.free()
    assert abs(l1 - abs(np.float32(a) - np.float32(b))) < 1e-6
    return (1.0 - dssim) * w * h


@pytest.mark.parametrize("a,b", [(0.5, 0.3), (0.9, 0.1), (0.05, 0.7)])
def test_ssim_kernel_constant_images_closed_form(gpu, a, b):
    taps = np.array([0.26601818, 0.21300237, 0.10936069, 0.03599398, 0.00759866, 0.00102838])
    S = (taps[0] + 2 * taps[1:].sum()) ** 2
    C1, C2 = 1e-4, 9e-4
    a32, b32 = float(np.float32(a)), float(np.float32(b))
    want = ((2 * a32 * b32 * S * S + C1) * (2 * a32 * b32 * S * (1 - S) + C2) /
            (((a32 * a32 + b32 * b32) * S * S + C1) * ((a32 * a32 + b32 * b32) * S * (1 - S) + C2)))
    announce(f"ssim closed form a={a} b={b}")
    W, H = 17, 19
    got = (_gpu_ssim_sum(gpu, W + 1, H + 1, a, b) - _gpu_ssim_sum(gpu, W + 1, H, a, b)
           - _gpu_ssim_sum(gpu, W, H + 1, a, b) + _gpu_ssim_sum(gpu, W, H, a, b))
    print(f"  interior SSIM {got:.7f} vs closed form {want:.7f}")
    assert abs(got - want) < 1e-4

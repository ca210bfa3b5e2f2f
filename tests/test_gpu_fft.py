"""The library's 2D inverse FFT on its own (svdi_fft2_inverse) against torch.fft.ifft2 (rocFFT)
on the same complex64 planes, at the grid sizes the imager uses: G = 192 (128^2 images), 1536
(1024^2, the wavefront kernels), 3072 (2048^2, radix-2 steps around them), 6144 (4096^2, the
general Stockham kernels) and odd grids (whole plane kept).  The window is the imager's: the
central N = 2G/3 rows and columns of an even grid.

Bound: max |ours - rocFFT| <= 2e-6 of max |rocFFT| (both FP32; an FFT's rounding grows like
log2(G) ulp of the output's scale).
"""
import pytest
import torch

from conftest import announce
import svdimager_amd as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("G,P", [(192, 4), (1536, 3), (3072, 2), (6144, 1), (15, 3), (255, 2), (1215, 1)])
def test_fft_matches_rocfft(gpu_ctx, G, P):
    announce(f"fft2_inverse G={G} P={P}")
    gen = torch.Generator(device="cuda").manual_seed(G)
    planes = torch.complex(torch.randn(P, G, G, device="cuda", generator=gen),
                           torch.randn(P, G, G, device="cuda", generator=gen)).contiguous()
    if G % 2 == 0:
        N = (2 * G) // 3
        lo = G // 2 - N // 2
        rows = cols = N
    else:
        lo, rows, cols = 0, G, G
    ours = S.fft2_inverse(gpu_ctx, planes, lo, rows, lo, cols)
    ref = torch.fft.ifft2(planes, norm="forward")[:, lo:lo + rows, lo:lo + cols]
    torch.cuda.synchronize()
    scale = float(ref.abs().max())
    err = float((ours - ref).abs().max()) / scale
    print(f"G={G}: max |ours - rocFFT| / max = {err:.2e}", flush=True)
    assert err <= 2e-6

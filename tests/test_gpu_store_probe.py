"""mf_hip_store_probe(_f32), the write-ceiling probe of the order-2 output shape: it writes every value of F, J, H
(and nothing past them) for matrix and Voigt row widths, tail tiles included, so the bench's ceiling covers the same
bytes as an evaluation."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nf", [9, 6])
@pytest.mark.parametrize("f32", [False, True])
def test_probe_writes_exactly_the_three_regions(nf, f32):
    import torch
    import matfun_amd
    lib = matfun_amd.library()
    dt = torch.float32 if f32 else torch.float64
    n = 997                                   # (15 full wavefronts of 64 and one of 37)
    pad = 64
    bufs = [torch.zeros(n * k + pad, dtype=dt, device="cuda") for k in (nf, nf * nf, nf ** 3)]
    fn = lib.mf_hip_store_probe_f32 if f32 else lib.mf_hip_store_probe
    print(f"store probe nf={nf} f32={f32} n={n}", flush=True)
    assert fn(n, nf, *[b.data_ptr() for b in bufs], torch.cuda.current_stream().cuda_stream) == 0
    torch.cuda.synchronize()
    for b, k in zip(bufs, (nf, nf * nf, nf ** 3)):
        body, tail = b[:n * k], b[n * k:]
        assert bool((body >= 1).all()) and bool((body <= 64).all())
        assert bool((tail == 0).all())

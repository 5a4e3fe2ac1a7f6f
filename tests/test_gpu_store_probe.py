"""mf_hip_store_probe(_f32) and mf_hip_store_probe_in(_f32), the ceiling probes of the order-2 output shape (the
second after reading the input rows, as an evaluation does): they write every value of F, J, H (and nothing past
them) for matrix and Voigt row widths, tail tiles included, so the bench's ceilings cover the same bytes as an
evaluation."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nf", [9, 6])
@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("with_input", [False, True])
def test_probe_writes_exactly_the_three_regions(nf, f32, with_input):
    import torch
    import matfun_amd
    lib = matfun_amd.library()
    dt = torch.float32 if f32 else torch.float64
    n = 997                                   # (15 full wavefronts of 64 and one of 37)
    pad = 64
    bufs = [torch.zeros(n * k + pad, dtype=dt, device="cuda") for k in (nf, nf * nf, nf ** 3)]
    st = torch.cuda.current_stream().cuda_stream
    print(f"store probe nf={nf} f32={f32} with_input={with_input} n={n}", flush=True)
    if with_input:
        x = torch.rand(n * nf, dtype=torch.float64, device="cuda")
        fn = lib.mf_hip_store_probe_in_f32 if f32 else lib.mf_hip_store_probe_in
        assert fn(n, x.data_ptr(), nf, nf, *[b.data_ptr() for b in bufs], st) == 0
    else:
        fn = lib.mf_hip_store_probe_f32 if f32 else lib.mf_hip_store_probe
        assert fn(n, nf, *[b.data_ptr() for b in bufs], st) == 0
    torch.cuda.synchronize()
    for b, k in zip(bufs, (nf, nf * nf, nf ** 3)):
        body, tail = b[:n * k], b[n * k:]
        assert bool((body >= 1).all()) and bool((body <= 64).all())
        assert bool((tail == 0).all())

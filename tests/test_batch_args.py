"""The heterogeneous-batch API passes raw device pointers to the C ABI, so the Python layer checks every tensor before
anything is launched: type, contiguity, element count and device of A, alpha, F, J, H (matfun_amd._check_batch_args).
These run on the CPU: every rejection happens before a plan or a launch is needed."""
import pytest
import torch

from matfun_amd import _check_batch_args

N, NJ, NH = 10, 6, 4          # a batch of 10 matrices, 6 of order >= 1, 4 of order 2


def good_out(dt=torch.float64):
    return torch.empty(N * 9, dtype=dt), torch.empty(NJ * 81, dtype=dt), torch.empty(NH * 729, dtype=dt)


def check(**kw):
    args = dict(n=N, nj=NJ, nh=NH, A=torch.empty(N, 3, 3, dtype=torch.float64), alpha=None, out=good_out())
    args.update(kw)
    _check_batch_args(**args)


def test_valid_host_tensors_fail_only_on_the_device():
    with pytest.raises(ValueError, match="on the GPU"):
        check()


@pytest.mark.parametrize("A, what", [
    (torch.empty(N, 3, 3, dtype=torch.float32), "dtype"),                       # FP32 input
    (torch.empty(N - 1, 3, 3, dtype=torch.float64), "elements"),                # short
    (torch.empty(N + 1, 3, 3, dtype=torch.float64), "elements"),                # long: not the plan's batch
    (torch.empty(3, 3, N, dtype=torch.float64).permute(2, 0, 1), "contiguous"),  # strided view
    (None, "torch CUDA tensor"),
    ([[0.0] * 9] * N, "torch CUDA tensor"),
])
def test_bad_input(A, what):
    with pytest.raises(ValueError, match=what):
        check(A=A)


@pytest.mark.parametrize("alpha, what", [
    (torch.empty(N, dtype=torch.float32), "dtype"),
    (torch.empty(N - 1, dtype=torch.float64), "elements"),
    (torch.empty(2 * N, dtype=torch.float64)[::2], "contiguous"),
])
def test_bad_alpha(alpha, what):
    with pytest.raises(ValueError, match=what):
        check(alpha=alpha)


def test_bad_outputs():
    F, J, H = good_out()
    with pytest.raises(ValueError, match="F: .*at least 90"):
        check(out=(F[:-1], J, H))
    with pytest.raises(ValueError, match="J: .*at least 486"):
        check(out=(F, J[:-1], H))
    with pytest.raises(ValueError, match="H: .*at least 2916"):
        check(out=(F, J, H[:-1]))
    with pytest.raises(ValueError, match="H: the batch has"):
        check(out=(F, J, None))
    with pytest.raises(ValueError, match="dtype"):               # FP64 buffers for the FP32 form
        check(out=(F, J, H), precision="mixed")
    Ff, Jf, Hf = good_out(torch.float32)
    with pytest.raises(ValueError, match="dtype"):
        check(out=(Ff, Jf, Hf))
    with pytest.raises(ValueError, match="on the GPU"):          # the FP32 form with FP32 buffers passes the rest
        check(out=(Ff, Jf, Hf), precision="mixed")
    with pytest.raises(ValueError, match="precision"):
        check(precision="fp16")


def test_empty_batch_and_unneeded_rows():
    _check_batch_args(0, 0, 0, None, None, (torch.empty(0), torch.empty(0), torch.empty(0)))
    # no J / H rows in the batch: J and H may be omitted (the input still has to be on the GPU)
    with pytest.raises(ValueError, match="on the GPU"):
        _check_batch_args(N, 0, 0, torch.empty(N * 9, dtype=torch.float64), None,
                          (torch.empty(N * 9, dtype=torch.float64), None, None))


def test_sequence_inputs_are_both_checked():
    A = torch.empty(N * 9, dtype=torch.float64)
    with pytest.raises(ValueError, match="A: .*elements"):
        _check_batch_args(N, NJ, NH, (A, A[:-9]), None, good_out(), names=("A0", "A"))
    with pytest.raises(ValueError, match="A0: .*dtype"):
        _check_batch_args(N, NJ, NH, (A.float(), A), None, good_out(), names=("A0", "A"))

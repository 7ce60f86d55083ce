"""The allocation-free batched banded solve (loma_band5_solve_dev_noalloc)
against the allocating entry point loma_band5_solve_dev, through the C ABI:
two calls in a row, the second with a smaller batch, each bit for bit equal
to the allocating entry's result on the same systems (the same kernel runs
on both), the first also against dense FP64 solves.  A batch beyond the
fixed scratch is refused before any device work."""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_band import _dense_solve, _systems

pytestmark = pytest.mark.gpu


def _solve(fn, gpu, band, rhs, n, nsys, rhs_fp32):
    import torch
    b = torch.from_numpy(band).to(gpu)
    r = torch.from_numpy(rhs.astype(np.complex64 if rhs_fp32 else np.complex128)).to(gpu)
    x = torch.full((nsys, n), float("nan"), dtype=torch.complex128, device=gpu)
    fn(b.data_ptr(), r.data_ptr(), x.data_ptr(), n, nsys, rhs_fp32)
    torch.cuda.synchronize()
    return x.cpu().numpy()


@pytest.mark.parametrize("rhs_fp32", [False, True])
def test_noalloc_matches_allocating_entry(gpu, rhs_fp32):
    from loma_phd_amd import capi
    for call, (n, nsys) in enumerate(((513, 300), (257, 67))):
        print(f"noalloc call {call}: n={n} nsys={nsys} fp32 rhs={rhs_fp32}", flush=True)
        band, rhs = _systems(nsys, n, seed=1000 + 17 * call + n)
        want = _solve(capi.band5_solve_dev, gpu, band, rhs, n, nsys, rhs_fp32)
        got = _solve(capi.band5_solve_dev_noalloc, gpu, band, rhs, n, nsys, rhs_fp32)
        assert np.array_equal(got, want), (call, np.abs(got - want).max())
        if call == 0:
            rhs_used = rhs.astype(np.complex64).astype(np.complex128) if rhs_fp32 else rhs
            assert rel_l2(got[:24], _dense_solve(band[:24], rhs_used[:24])) < 1e-12


def test_noalloc_refuses_batches_beyond_scratch(gpu):
    import torch
    from loma_phd_amd import capi
    n, nsys = 513, 1636  # 513*5*1636 = 4,196,340 > 4,194,304 doubles
    b = torch.zeros((1,), dtype=torch.float64, device=gpu)
    with pytest.raises(capi.LomaError) as e:
        capi.band5_solve_dev_noalloc(b.data_ptr(), b.data_ptr(), b.data_ptr(), n, nsys)
    assert e.value.code == capi.LOMA_ERR_ARG

"""xtb_forces_noalloc (no device allocation, scratch in a fixed device
array) against xtb_forces on the same converged batch: two calls in a row,
the second on a smaller batch, and the size limit."""

import ctypes as ct

import numpy as np
import pytest
import torch

from dxtb_amd import _lib
from dxtb_amd import workloads as W
from helpers import run_native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _noalloc(pipe):
    """Gradient from the allocation-free entry point on the pipeline's
    converged state, into a fresh buffer; the stream is synchronised."""
    lib = _lib.load()
    g = torch.full_like(pipe.grad, float("nan"))
    st = torch.cuda.current_stream()
    rc = lib.xtb_forces_noalloc(ct.byref(pipe.batch), ct.byref(pipe.dt.s), pipe.positions.data_ptr(),
                                pipe.cn.data_ptr(), pipe.gamma.data_ptr(), ct.byref(pipe.state),
                                g.data_ptr(), ct.c_void_p(st.cuda_stream))
    assert rc == _lib.XTB_OK
    torch.cuda.synchronize()
    return g.view(-1, 3).cpu().numpy()


def test_noalloc_forces_equal_forces_two_calls():
    rng = np.random.default_rng(51)
    big = [W.chain(rng, 30), W.water_cluster(rng, 20), W.chain(rng, 12)]
    small = [W.chain(rng, 4)]
    for mols in (big, small):
        print(f"case: forces_noalloc, {sum(len(z) for z, _ in mols)} atoms", flush=True)
        pipe = run_native(mols, forces=True)
        ref = pipe.gradient().cpu().numpy()
        got = _noalloc(pipe)
        # the same kernels; only the order of the atomic additions differs
        assert np.abs(got - ref).max() < 1e-12


def test_noalloc_rejects_batches_over_the_limit():
    pipe = run_native([W.chain(np.random.default_rng(52), 3)], forces=True)
    lib = _lib.load()
    b = _lib.XtbBatch.from_buffer_copy(pipe.batch)
    b.nat_total = _lib.NOALLOC_MAX_ATOMS + 1
    g = torch.empty(3, dtype=torch.float64, device="cuda:0")
    rc = lib.xtb_forces_noalloc(ct.byref(b), ct.byref(pipe.dt.s), pipe.positions.data_ptr(),
                                pipe.cn.data_ptr(), pipe.gamma.data_ptr(), ct.byref(pipe.state),
                                g.data_ptr(), ct.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _lib.XTB_ETOOBIG

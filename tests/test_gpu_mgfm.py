"""GPU parity of the batched MGFM kernel against the oracle: bit-exact
outputs and identical Newton iteration counts (FP64, no FMA contraction)."""
import numpy as np
import pytest

from oracle import pyoracle
from pgdg_p2_amd import workloads


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_mgfm_batch_bit_exact(cuda, variant, mode):
    import torch
    import pgdg_p2_amd as P
    n = 100_003
    inp = workloads.mgfm_faces(n, seed=5 + variant)
    gl, pil, gr, pir = workloads.MGFM_EOS
    print(f"mgfm parity variant={variant} mode={mode} n={n}", flush=True)
    ref, it_ref = pyoracle.mgfm_batch(inp, gl, pil, gr, pir, variant=variant, mode=mode,
                                      rtol=1e-12, max_iter=50 if mode == 0 else 100)
    out, it = P.mgfm_batch(torch.from_numpy(inp).to(cuda), gl, pil, gr, pir, variant=variant,
                           tol_mode=mode, rtol=1e-12, max_iter=50 if mode == 0 else 100)
    torch.cuda.synchronize()
    out = out.cpu().numpy()
    it = it.cpu().numpy()
    np.testing.assert_array_equal(it, it_ref)
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))


@pytest.mark.gpu
def test_mgfm_batch_guess_and_empty(cuda):
    import torch
    import pgdg_p2_amd as P
    inp = workloads.mgfm_faces(1000, seed=9)
    gl, pil, gr, pir = workloads.MGFM_EOS
    guess = np.maximum(inp[2], inp[3]) * 0.5
    ref, it_ref = pyoracle.mgfm_batch(inp, gl, pil, gr, pir, variant=0, mode=1, max_iter=100,
                                      guess=guess)
    out, it = P.mgfm_batch(torch.from_numpy(inp).to(cuda), gl, pil, gr, pir, variant=0,
                           tol_mode=1, max_iter=100, pstar0=torch.from_numpy(guess).to(cuda))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(it.cpu().numpy(), it_ref)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    e = torch.empty((6, 0), dtype=torch.float64, device=cuda)
    o, i = P.mgfm_batch(e, gl, pil, gr, pir)
    assert o.shape == (3, 0)


@pytest.mark.gpu
def test_device_sqrt_div_correctly_rounded(cuda):
    # The whole-step parity relies on IEEE-correct FP64 sqrt and division on
    # the device; compare against numpy on random operands.
    import torch
    rng = np.random.default_rng(1)
    a = np.exp(rng.uniform(-300, 300, 1_000_000))
    b = np.exp(rng.uniform(-300, 300, 1_000_000))
    ta, tb = torch.from_numpy(a).to(cuda), torch.from_numpy(b).to(cuda)
    np.testing.assert_array_equal(torch.sqrt(ta).cpu().numpy(), np.sqrt(a))
    np.testing.assert_array_equal((ta / tb).cpu().numpy(), a / b)

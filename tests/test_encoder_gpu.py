"""GPU LDPC encoder (the benchmark's test-signal source) vs the oracle's
encoder (Gaussian elimination of the core, an independent method): bit-exact
codewords, and H*c = 0."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CASES = [(1, 2), (1, 13), (1, 52), (1, 104), (1, 208), (1, 384), (1, 26), (1, 240),
         (2, 2), (2, 15), (2, 56), (2, 60), (2, 112), (2, 224), (2, 352), (2, 384)]


@pytest.mark.parametrize("bg,Z", CASES)
def test_gpu_encoder_matches_oracle(bg, Z):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from aerial_amd import workloads as W
    print("[case] encoder BG%d Z=%d (lifting set %d)" % (bg, Z, O.ils(Z)), flush=True)
    ns = 22 if bg == 1 else 10
    info = np.stack([O.random_bits(100 + i, ns * Z) for i in range(3)])
    cw = W.encode(bg, Z, torch.from_numpy(info).cuda())
    torch.cuda.synchronize()
    got = cw.cpu().numpy()
    for i in range(3):
        ref = O.encode(bg, Z, info[i])
        assert np.array_equal(got[i], ref)
        assert O.syndrome(bg, Z, 46 if bg == 1 else 42, got[i]) == 0


def test_gpu_encoder_matches_reference_goldens():
    """MATLAB encoder output shipped with the reference (BG1, Zc = 52)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import os
    from aerial_amd import workloads as W
    print("[case] encoder reference goldens BG1 Z=52", flush=True)
    f = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "encoder_goldens.npz"))
    src = np.concatenate([f["src_1_0"], f["src_752_0"]])
    enc = np.concatenate([f["enc_1_0"], f["enc_752_0"]])
    cw = W.encode(1, 52, torch.from_numpy(src).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(cw.cpu().numpy(), enc)

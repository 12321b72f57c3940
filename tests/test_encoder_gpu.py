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

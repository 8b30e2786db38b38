"""qsim_sample_noalloc (no device allocation: the library's fixed per-GPU
scratch array) against qsim_sample on the same states and seeds: identical
indices, on consecutive calls with a smaller input second, for both
precisions; its refusals (too large for the scratch, a pending layout change
with no second buffer) come before any work and leave the state usable."""
import numpy as np
import pytest

from oracle import circuits as C
from oracle import permute as P

pytestmark = pytest.mark.gpu


def _sv(hip, n, dtype):
    return hip.StatevectorF32(n) if dtype == np.complex64 else hip.Statevector(n)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_noalloc_sampling_equals_sampling(hip, dtype):
    for n, shots, seed in [(18, 20000, 7), (11, 3000, 8), (3, 17, 9)]:
        print(f"sample_noalloc n={n} shots={shots} dtype={np.dtype(dtype).name}", flush=True)
        sv = _sv(hip, n, dtype)
        sv.from_numpy(C.random_state(n, seed))
        want = sv.sample(shots, seed=seed)
        got = sv.sample_noalloc(shots, seed=seed)
        assert got.dtype == np.uint64 and got.shape == (shots,)
        assert np.array_equal(got, want)
        assert int(got.max()) < (1 << n)


def test_noalloc_sampling_refusals_and_pending_layout(hip):
    n = 14
    rng = np.random.default_rng(4)
    L0 = list(range(n))
    L1 = [int(x) for x in rng.permutation(n)]
    L2 = [int(x) for x in rng.permutation(n)]
    psi0 = C.random_state(n, 21)
    a, b = hip.StatevectorF32(n), hip.StatevectorF32(n)
    a.from_numpy(psi0)
    b.from_numpy(psi0)
    a.permute(L0, L1)
    b.permute(L0, L1)
    # no second buffer yet: refused, the state untouched
    with pytest.raises(RuntimeError, match="second buffer"):
        a.sample_noalloc(100, seed=1)
    ref = P.permute(psi0, L0, L1).astype(np.complex64)
    assert np.array_equal(a.to_numpy(), ref)  # this read gives it the second buffer
    a.permute(L1, L2)
    b.permute(L1, L2)
    assert np.array_equal(a.sample_noalloc(5000, seed=3), b.sample(5000, seed=3))
    # beyond the scratch: refused before any work
    big = hip.NOALLOC_SCRATCH_BYTES // 16 + 1
    with pytest.raises(RuntimeError, match="sample_noalloc: .* bytes of scratch"):
        a.sample_noalloc(big, seed=5)
    assert np.array_equal(a.sample_noalloc(64, seed=5), a.sample(64, seed=5))

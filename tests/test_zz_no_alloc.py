"""gpr_sort_pairs_device_noalloc (fixed scratch array, no allocation) against
gpr_sort_pairs_device (library-allocated scratch): bit-identical sorted keys
and indices, on two calls in a row with the second one smaller, and the size
limit rejected with an error.  Runs through the C ABI."""
import numpy as np
import pytest

import gpr
import oracle_py as O

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20


def _sort(fn, keys):
    import torch
    dk = torch.from_numpy(keys.view(np.int32)).cuda()
    sk = torch.full_like(dk, -1)
    si = torch.full_like(dk, -1)
    fn(dk.data_ptr(), sk.data_ptr(), si.data_ptr(), len(keys))
    torch.cuda.synchronize()
    return sk.cpu().numpy().view(np.uint32), si.cpu().numpy().view(np.uint32)


def test_noalloc_sort_equals_allocating_sort_twice_second_smaller():
    rng = np.random.default_rng(31)
    # the first call uses the whole array (2^20 keys, 1024 buckets, a long run
    # of equal keys that takes the large-bucket path); the second is smaller,
    # with another digit count, and must not see the first call's contents
    for n in (MAX_KEYS, 77777):
        print(f"[noalloc] n={n}", flush=True)
        keys = rng.integers(0, 1 << 30, n).astype(np.uint32)
        keys[: n // 5] = keys[n // 2]
        ak, ai = _sort(gpr.sort_pairs_device, keys)
        bk, bi = _sort(gpr.sort_pairs_device_noalloc, keys)
        assert np.array_equal(ak, bk) and np.array_equal(ai, bi), n
        if n <= 1 << 17:  # the oracle as a third opinion where it is quick
            ok, oi = O.sort_pairs(keys)
            assert np.array_equal(bk, ok) and np.array_equal(bi, oi), n


def test_noalloc_sort_rejects_more_than_its_maximum():
    import torch
    n = MAX_KEYS + 1
    dk = torch.zeros(n, dtype=torch.int32, device="cuda")
    with pytest.raises(gpr.GprError, match="maximum"):
        gpr.sort_pairs_device_noalloc(dk.data_ptr(), dk.data_ptr(), dk.data_ptr(), n)

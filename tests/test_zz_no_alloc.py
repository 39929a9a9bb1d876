"""marsmd_radix_sort_pairs_static (the sort with the library's own fixed workspace,
for callers that cannot allocate) against marsmd_radix_sort_pairs with a caller's
workspace, through the C ABI: bit-identical keys and values on consecutive calls (the
largest size first, then smaller ones, so later calls run over a workspace the earlier
ones filled), both equal to a stable argsort of the masked keys."""
import numpy as np
import pytest

from marsmd import _capi as api

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 21


def _sorted_both(torch, keys, key_bits):
    n = len(keys)
    vals = np.arange(n, dtype=np.uint32)
    kin = torch.from_numpy(keys.view(np.int32)).cuda()
    vin = torch.from_numpy(vals.view(np.int32)).cuda()
    outs = [torch.empty_like(kin) for _ in range(4)]
    nbytes = api.lib().marsmd_radix_sort_scratch_bytes(n)
    ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device="cuda")
    api.call("marsmd_radix_sort_pairs", kin.data_ptr(), vin.data_ptr(), outs[0].data_ptr(), outs[1].data_ptr(), n,
             key_bits, ws.data_ptr(), nbytes, None)
    api.call("marsmd_radix_sort_pairs_static", kin.data_ptr(), vin.data_ptr(), outs[2].data_ptr(),
             outs[3].data_ptr(), n, key_bits, None)
    torch.cuda.synchronize()
    return vals, [o.cpu().numpy().view(np.uint32) for o in outs]


@pytest.mark.parametrize("sizes", [[(MAX_KEYS, 30), (100_003, 30), (4097, 10)]])
def test_static_workspace_sort_equals_the_allocating_sort(cuda, sizes):
    rs = np.random.default_rng(11)
    for n, bits in sizes:
        print(f"case: static-workspace sort, n={n}, key_bits={bits}", flush=True)
        # heavy duplicates (a quarter of the key range) and bits above key_bits set
        keys = (rs.integers(0, max(1, (1 << bits) // 4), n, dtype=np.uint64) |
                (rs.integers(0, 4, n, dtype=np.uint64) << 30)).astype(np.uint32)
        vals, (ka, va, kb, vb) = _sorted_both(cuda, keys, bits)
        assert np.array_equal(ka, kb), f"keys differ at n={n}"
        assert np.array_equal(va, vb), f"values differ at n={n}"
        order = np.argsort(keys & np.uint32((1 << bits) - 1), kind="stable")
        assert np.array_equal(vb, vals[order]), f"not a stable sort at n={n}"
        assert np.array_equal(kb, keys[order])


def test_static_workspace_sort_rejects_more_keys_than_its_maximum(cuda):
    buf = cuda.empty(4 * (MAX_KEYS + 1), dtype=cuda.int32, device="cuda")
    out = [cuda.empty_like(buf) for _ in range(3)]
    with pytest.raises(api.MarsError, match="maximum"):
        api.call("marsmd_radix_sort_pairs_static", buf.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                 out[2].data_ptr(), MAX_KEYS + 1, 30, None)

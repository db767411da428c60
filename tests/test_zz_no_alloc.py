"""mek_expm_dense_noalloc, the exponential that allocates nothing (its
working memory is a static device array of the library), against the
allocating mek_expm_dense through the C ABI: consecutive calls with falling
sizes, across the in-LDS (k <= 62) and workspace (k > 62) forms, bit for bit;
sizes beyond the static array refused."""
import numpy as np
import pytest
import scipy.linalg

from conftest import mek

pytestmark = pytest.mark.gpu


def _hessenberg(k, seed, scale):
    rng = np.random.default_rng(seed)
    return np.triu(rng.standard_normal((k, k)), -1) * scale


def _noalloc(H):
    k = H.shape[0]
    d_a = mek.DeviceBuffer.from_array(np.asfortranarray(H).ravel(order="F"))
    d_e = mek.DeviceBuffer(8 * k * k)
    mek._check(mek.lib().mek_expm_dense_noalloc(k, d_a.ptr, d_e.ptr, None))
    mek.synchronize()
    return d_e.to_array().reshape(k, k, order="F")


def test_noalloc_matches_allocating_entry(gpu):
    # sizes fall call after call: every call reuses the scratch the previous,
    # larger one left behind
    for k, seed, scale in [(128, 1, 0.4), (100, 2, 1.5), (63, 3, 0.8), (62, 4, 2.0), (30, 5, 1.0), (1, 6, 3.0)]:
        H = _hessenberg(k, seed, scale)
        E_ref = mek.expm_dense(H)
        E = _noalloc(H)
        print(f"[no-alloc] k={k} max|E|={np.abs(E).max():.3e}", flush=True)
        np.testing.assert_array_equal(E, E_ref)
        ref = scipy.linalg.expm(H)
        assert np.linalg.norm(E - ref) <= 1e-12 * np.linalg.norm(ref), k


def test_noalloc_refuses_sizes_beyond_the_scratch(gpu):
    d = mek.DeviceBuffer(8 * 4)
    for k in (0, 129):
        st = mek.lib().mek_expm_dense_noalloc(k, d.ptr, d.ptr, None)
        assert st == -1 and b"k in [1, 128]" in mek.lib().mek_last_error(), k

"""hd_block_rms_noalloc (the block RMS with its scratch in a fixed device
array) against hd_block_rms, bit for bit, through the C ABI: two calls in a
row, the second with a smaller input (the array keeps the first call's
contents), and a shape beyond the array's size refused without a launch."""
import ctypes as C

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = pytest.mark.gpu

f32 = np.float32


@pytest.fixture(scope="module")
def L():
    if not gpu_available():
        pytest.skip("no GPU")
    from heimdall_multibeam_amd import stages
    L = stages.lib()
    sz, vp = C.c_size_t, C.c_void_p
    L.hd_block_rms_noalloc.argtypes = [vp, vp, sz, sz, sz, sz, sz, vp]
    L.hd_block_rms_noalloc.restype = C.c_int
    return L


def _case(L, rng, n_series, beam_stride, beam_count, nbeams, pitch):
    from heimdall_multibeam_amd.hip import DeviceBuffer
    # the leaves are read at positions below (nbeams-1)*beam_stride + beam_count
    assert (nbeams - 1) * beam_stride + beam_count <= pitch
    host = rng.normal(0, 3, (n_series, pitch)).astype(f32)
    d = DeviceBuffer(host.nbytes)
    d.upload(host)
    a, b = DeviceBuffer(n_series * 4), DeviceBuffer(n_series * 4)
    print("case noalloc block rms: %d series, stride %d, count %d, %d beams"
          % (n_series, beam_stride, beam_count, nbeams), flush=True)
    rc = L.hd_block_rms(d.address, a.address, beam_stride, beam_count, nbeams, n_series, pitch,
                        None)
    assert rc == 0
    rc = L.hd_block_rms_noalloc(d.address, b.address, beam_stride, beam_count, nbeams, n_series,
                                pitch, None)
    assert rc == 0
    want = a.download().view(np.uint32)
    got = b.download().view(np.uint32)
    np.testing.assert_array_equal(got, want)
    assert np.all(np.isfinite(got.view(f32))) and np.all(got.view(f32) > 0)
    return got.view(f32)


def test_noalloc_block_rms_matches_allocating_entry_point(L):
    rng = np.random.default_rng(3)
    # the 1024-channel gulp's block-RMS shape (np = 70 partials per series)
    _case(L, rng, 300, 262144, 262144 - 36540 - 4096, 1, 225668)
    # a smaller second call: fewer series, three beams, short rows
    _case(L, rng, 5, 6000, 5000, 3, 18008)


def test_noalloc_block_rms_refuses_beyond_its_array(L):
    from heimdall_multibeam_amd import HD_MEM_ALLOC_FAILED
    from heimdall_multibeam_amd.hip import DeviceBuffer
    d, out = DeviceBuffer(1 << 12), DeviceBuffer(1 << 12)
    # 2000 series x 1000 partials > 2^20 words: refused before any launch
    rc = L.hd_block_rms_noalloc(d.address, out.address, 3125 * 1000, 3125 * 1000, 1, 2000,
                                3125 * 1000, None)
    assert rc == HD_MEM_ALLOC_FAILED
    # an invalid shape is refused as by hd_block_rms
    assert L.hd_block_rms_noalloc(d.address, out.address, 10, 20, 1, 1, 10, None) != 0

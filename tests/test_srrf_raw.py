"""The SRRF.raw file of the reference's caller (multispot5_place_methods.cc:
270-274 write, 284-289 read): raw float32, x fastest, read back bit for bit;
a short file is an error.  The device post-processing (NaN -> 0, negative ->
0, max starting at 1, :260-269) is compared with numpy on the GPU."""
import numpy as np
import pytest

from rfba_amd import srrf


def test_raw_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    vol = rng.normal(size=(5, 6, 7)).astype(np.float32)
    vol[0, 0, 0] = np.nan
    p = str(tmp_path / "SRRF.raw")
    srrf.write_raw(p, vol)
    assert (tmp_path / "SRRF.raw").stat().st_size == vol.size * 4
    back = srrf.read_raw(p, vol.shape)
    assert np.array_equal(back, vol, equal_nan=True)
    # x fastest: the file is the C-order byte image of [z][y][x]
    assert np.array_equal(np.fromfile(p, np.float32).reshape(vol.shape), vol, equal_nan=True)


def test_raw_short_file_is_an_error(tmp_path):
    p = str(tmp_path / "SRRF.raw")
    srrf.write_raw(p, np.zeros(10, np.float32))
    with pytest.raises(srrf.SRRFError):
        srrf.read_raw(p, (11,))


@pytest.mark.gpu
@pytest.mark.parametrize("scale", [0.5, 30.0])
def test_postprocess_device_matches_numpy(gpu, scale):
    rng = np.random.default_rng(1)
    vol = (rng.normal(size=(9, 33, 65)) * scale).astype(np.float32)
    vol.ravel()[::17] = np.nan
    d = srrf.DeviceBuffer.from_array(vol)
    fmax = srrf.postprocess_device(d, vol.size)
    got = d.to_array(vol.shape, np.float32)
    d.close()
    ref = np.where(np.isnan(vol), 0.0, vol)
    ref = np.where(ref < 0, 0.0, ref).astype(np.float32)
    assert np.array_equal(got, ref)
    assert fmax == max(1.0, float(ref.max()))

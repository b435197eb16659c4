"""CPU: the .di3 density container against the reference's own sample file.

tests/golden/par_0_head.di3bin holds the first 33 + 4*64 bytes of the reference's
ZQ_OpticalFlow3D/par3D1/par_0.di3 (2,097,185 bytes: a 128 x 64 x 64 single-channel FP32 volume), the
container ZQ_SmokeGrid3D::ExportDensity writes (ZQ_SmokeGrid3D.cpp:55-81).  The sample's type tag
carries bytes after its NUL ("float\\0" then stale memory), which a reader must ignore."""
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEAD = os.path.join(ROOT, "tests", "golden", "par_0_head.di3bin")
REF_FILE = "/root/reference/ZQ_OpticalFlow3D/par3D1/par_0.di3"
REF_LEN = 2097185


def test_reference_sample_header_and_samples(tmp_path):
    from zq_smokesimulation_amd import scene
    head = open(HEAD, "rb").read()
    assert head[:6] == b"float\0" and head[6:16] != b"\0" * 10  # stale bytes after the tag's NUL
    # the sample padded back to its full length: the header describes exactly the payload
    full = tmp_path / "par_0.di3"
    full.write_bytes(head + b"\0" * (REF_LEN - len(head)))
    dims, _ = scene.read_di3(full, header_only=True)
    assert dims == (128, 64, 64, 1) and 33 + 4 * 128 * 64 * 64 == REF_LEN
    dims, data = scene.read_di3(full)
    assert data.shape == (64, 64, 128)
    want = np.frombuffer(head[33:], "<f4")
    np.testing.assert_array_equal(data.reshape(-1)[:64], want)
    # the truncated sample itself is refused (payload shorter than the header says)
    with pytest.raises(ValueError):
        scene.read_di3(HEAD, header_only=True)


def test_layout_written_by_hand_round_trips(tmp_path):
    from zq_smokesimulation_amd import scene
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, 4, 7)).astype(np.float32)
    p = tmp_path / "x.di3"
    p.write_bytes(b"float" + b"\0" * 11 + struct.pack("<4i", 7, 4, 5, 1) + b"\0" + a.tobytes())
    dims, b = scene.read_di3(p)
    assert dims == (7, 4, 5, 1)
    np.testing.assert_array_equal(a, b)
    bad = tmp_path / "bad.di3"
    bad.write_bytes(b"double" + b"\0" * 10 + struct.pack("<4i", 7, 4, 5, 1) + b"\0" + a.tobytes())
    with pytest.raises(ValueError):
        scene.read_di3(bad)


@pytest.mark.skipif(not os.path.exists(REF_FILE), reason="the reference tree is not present on this machine")
def test_reference_sample_file_itself():
    from zq_smokesimulation_amd import scene
    dims, data = scene.read_di3(REF_FILE)
    assert dims == (128, 64, 64, 1)
    assert np.isfinite(data).all()
    np.testing.assert_array_equal(data.reshape(-1)[:64], np.frombuffer(open(HEAD, "rb").read()[33:], "<f4"))

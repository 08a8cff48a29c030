"""HDR image files (naivetrace_amd/csrc/image_io.cpp): PFM and uncompressed OpenEXR, read back by
readers written from the published file layouts (PFM: text header, little-endian floats, bottom
row first; OpenEXR 2.0: magic 20000630, version 2, attribute header, offset table, scanline
chunks with the channels in alphabetical order)."""
import struct

import numpy as np
import pytest

from naivetrace_amd import abi
from naivetrace_amd.renderer import write_rgb_image


def read_pfm(path):
    data = open(path, "rb").read()
    head = data.split(b"\n", 3)
    assert head[0] == b"PF"
    w, h = (int(x) for x in head[1].split())
    scale = float(head[2])
    assert scale < 0  # little-endian
    px = np.frombuffer(head[3], dtype="<f4", count=w * h * 3).reshape(h, w, 3)
    return px[::-1]


def read_exr(path):
    data = open(path, "rb").read()
    assert struct.unpack_from("<I", data, 0)[0] == 20000630
    assert struct.unpack_from("<I", data, 4)[0] == 2
    pos = 8
    attrs = {}
    while data[pos] != 0:
        name_end = data.index(b"\0", pos)
        name = data[pos:name_end].decode()
        type_end = data.index(b"\0", name_end + 1)
        typ = data[name_end + 1:type_end].decode()
        size = struct.unpack_from("<i", data, type_end + 1)[0]
        attrs[name] = (typ, data[type_end + 5:type_end + 5 + size])
        pos = type_end + 5 + size
    pos += 1
    assert attrs["compression"] == ("compression", b"\0")
    xmin, ymin, xmax, ymax = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = xmax - xmin + 1, ymax - ymin + 1
    chl = attrs["channels"][1]
    names, p = [], 0
    while chl[p] != 0:
        e = chl.index(b"\0", p)
        names.append(chl[p:e].decode())
        assert struct.unpack_from("<i", chl, e + 1)[0] == 2  # FLOAT
        p = e + 1 + 16
    assert names == sorted(names)
    offsets = struct.unpack_from(f"<{h}Q", data, pos)
    img = np.zeros((h, w, len(names)), np.float32)
    for off in offsets:
        y, nbytes = struct.unpack_from("<ii", data, off)
        assert nbytes == w * 4 * len(names)
        rows = np.frombuffer(data, dtype="<f4", count=w * len(names), offset=off + 8).reshape(len(names), w)
        for k in range(len(names)):
            img[y - ymin, :, k] = rows[k]
    return {n: img[..., k] for k, n in enumerate(names)}


@pytest.fixture
def image():
    rng = np.random.default_rng(2)
    img = rng.exponential(1.0, (7, 11, 3)).astype(np.float32)
    img[0, 0] = [1e30, 0.0, -2.5]
    return img


def test_pfm_round_trip(tmp_path, image):
    p = tmp_path / "a.pfm"
    write_rgb_image(image, p, abi.IMAGE_PFM)
    np.testing.assert_array_equal(read_pfm(p), image)


def test_exr_round_trip(tmp_path, image):
    p = tmp_path / "a.exr"
    write_rgb_image(image, p, abi.IMAGE_EXR)
    ch = read_exr(p)
    np.testing.assert_array_equal(np.stack([ch["R"], ch["G"], ch["B"]], axis=-1), image)


def test_bad_arguments(tmp_path, image):
    with pytest.raises(Exception):
        write_rgb_image(image, tmp_path / "a.xyz", 7)
    with pytest.raises(Exception):
        write_rgb_image(image, "/nonexistent/dir/a.exr", abi.IMAGE_EXR)

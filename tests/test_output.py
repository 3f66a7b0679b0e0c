"""The .lbake writer (lightbaker_amd/output.py, output.rs:95-163 with encode.rs:76-150) on the CPU:
its containers against the formats' definitions (KTX2: the identifier, the header's nine words,
the index, the level index, the data format descriptor's sizes; raw deflate; Zstandard frames),
and a write/read round trip of every chunk and header field.
"""
from __future__ import annotations

import struct
from types import SimpleNamespace

import numpy as np
import pytest

from lightbaker_amd import output as O
from lightbaker_amd.scenes import reference_test_scene


def test_dfd_sizes_and_samples():
    # KDF 1.3 basic block: 24 bytes + 16 a sample, after the 4-byte total
    b = O.dfd(O.BC5)
    assert len(b) == 4 + 24 + 2 * 16 and struct.unpack_from("<I", b)[0] == len(b)
    assert struct.unpack_from("<HH", b, 8) == (2, 24 + 32)
    assert b[12:16] == bytes([133, 1, 1, 0]) and b[16:20] == bytes([3, 3, 0, 0]) and b[20] == 16
    assert struct.unpack_from("<HBB", b, 28) == (0, 63, 0)
    assert struct.unpack_from("<HBB", b, 44) == (64, 63, 1)
    r = O.dfd(O.RGB9E5)
    assert len(r) == 4 + 24 + 3 * 16 and r[12] == 1 and r[20] == 4
    assert [struct.unpack_from("<HBB", r, 28 + 16 * i) for i in range(3)] == \
        [(0, 8, 0), (9, 8, 1), (18, 8, 2)]
    assert struct.unpack_from("<II", r, 28 + 8) == (0, 0xFFFFFFFF)


def test_ktx2_layout_and_round_trip():
    rng = np.random.default_rng(1)
    data = rng.integers(0, 4, 64 * 32).astype(np.uint8).tobytes()
    k = O.ktx2(O.BC5, 64, 32, data, level=3)
    assert k[:12] == bytes([0xAB, 0x4B, 0x54, 0x58, 0x20, 0x32, 0x30, 0xBB, 0x0D, 0x0A, 0x1A, 0x0A])
    assert struct.unpack_from("<9I", k, 12) == (141, 1, 64, 32, 0, 0, 1, 1, 2)
    dfd_at, dfd_len, kvd_at, kvd_len = struct.unpack_from("<4I", k, 48)
    assert dfd_at == 80 + 24 and dfd_len == len(O.dfd(O.BC5)) and kvd_at == kvd_len == 0
    assert struct.unpack_from("<QQ", k, 64) == (0, 0)
    at, length, size = struct.unpack_from("<3Q", k, 80)
    assert at % 8 == 0 and at >= dfd_at + dfd_len and at + length == len(k) and size == len(data)
    # a Zstandard frame (magic 0xFD2FB528) holding the data
    assert struct.unpack_from("<I", k, at)[0] == 0xFD2FB528
    r = O.read_ktx2(k)
    assert r["data"] == data and (r["width"], r["height"], r["vk_format"]) == (64, 32, 141)


def test_deflate_is_raw():
    data = bytes(range(256)) * 50
    d = O.deflate(data)
    assert d[:2] != b"\x78\xda"  # no zlib wrapper
    assert O.inflate(d) == data


def fake_bake(scene, rng, aw=64, ah=32, ds=2, volume=True):
    iw, ih = -(-aw // ds), -(-ah // ds)
    info = {"atlas_width": aw, "atlas_height": ah, "indirect_width": iw, "indirect_height": ih,
            "tiles": 4, "tiles_baked": 3, "texels_baked": 1234, "volume_origin": [0.0, 0.5, 0.0],
            "volume_step": 2.0, "volume_size": [3, 2, 3], "volume_scale": 1.75}
    return SimpleNamespace(
        info=info, direct_bc5=rng.integers(0, 256, aw * ah).astype(np.uint8),
        indirect_rgb9e5=rng.integers(0, 2 ** 32, (ih, iw), dtype=np.uint64).astype(np.uint32),
        charts=rng.random((5, 2, 4)).astype(np.float32),
        tri_charts=rng.integers(0, 5, 40).astype(np.uint32),
        validity=rng.integers(0, 6, (ah, aw)).astype(np.uint8),
        volume=np.zeros((2, 3, 3, 4, 4), np.float32) if volume else None,
        volume_bytes=rng.integers(0, 256, (18, 16)).astype(np.uint8) if volume else None)


@pytest.mark.parametrize("volume", [True, False])
def test_write_read_round_trip(volume):
    scene = reference_test_scene(8.0)
    b = fake_bake(scene, np.random.default_rng(2), volume=volume)
    f = O.write(scene, b, scene_hash="ab" * 32, user={"k": 1}, seconds=1.5, level=3)
    assert f[:4] == b"LBAK" and struct.unpack_from("<I", f, 4)[0] == 1
    r = O.read(f)
    h = r["header"]
    assert [c["name"] for c in h["chunks"]] == \
        ["direct", "indirect", "charts", "triangleCharts", "validity"] + (["volume"] if volume else [])
    assert all(c["offset"] % 8 == 0 for c in h["chunks"])
    assert r["direct"]["data"] == b.direct_bc5.tobytes()
    assert r["direct"]["vk_format"] == 141 and r["indirect"]["vk_format"] == 123
    assert (r["indirect"]["width"], r["indirect"]["height"]) == (32, 16)
    assert r["indirect"]["data"] == b.indirect_rgb9e5.astype("<u4").tobytes()
    assert r["charts"] == b.charts.tobytes() and r["triangleCharts"] == b.tri_charts.tobytes()
    assert r["validity"] == b.validity.tobytes()
    if volume:
        assert r["volume"] == b.volume_bytes.tobytes()
        assert h["volume"] == {"origin": [0.0, 0.5, 0.0], "step": 2.0, "size": [3, 2, 3],
                               "scale": 1.75}
    else:
        assert h["volume"] is None and "volume" not in r
    assert h["scene"] == "ab" * 32 and h["user"] == {"k": 1}
    assert h["atlas"]["width"] == 64 and h["atlas"]["indirectWidth"] == 32
    assert h["atlas"]["charts"] == 5
    assert h["stats"]["tilesBaked"] == 3 and h["stats"]["seconds"] == 1.5
    # output.rs:128-129: pi (h / 3 + 2 z / 3) over a face looking up; the test scene's sky
    want = [np.pi * (scene.horizon[c] / 3 + 2 * scene.zenith[c] / 3) for c in range(3)]
    assert np.allclose(h["skyIrradiance"], want, rtol=1e-6)
    assert h["meshes"] == [{"name": m.name, "triangles": len(m.positions) // 3}
                           for m in scene.meshes]


def test_volume_chunks_follow_the_ground_plan():
    scene = reference_test_scene(8.0)
    # volume 0..64 x 0.5..4.5 x 0..64 step 2 over 16 m tiles: probes at x = 0, 2, .., 64 fall in
    # tiles 0..4 along x and z
    assert scene.settings.tile_size == 16.0
    n, baked = O.volume_chunks(scene)
    assert n == 5 * 5 and baked == n
    nx, ny, nz = 33, 3, 33
    dirty = np.zeros(nx * ny * nz, bool)
    dirty[0] = True
    assert O.volume_chunks(scene, dirty) == (25, 1)

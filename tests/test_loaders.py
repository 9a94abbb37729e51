"""Scene loaders (gamesplat_amd.loaders; per-Gaussian decode in the library's
gs_decode_ply / gs_decode_sog), CPU only.

- PLY: activation against numpy (exp scale, sigmoid opacity, normalised
  quaternion), channel-major f_rest repacked to RGB triples exactly, the
  defaults for absent properties, header errors.
- SOG: an archive built here from raw bytes (lossless WebP, exact RGB under
  zero alpha) decodes to a numpy restatement of the format: 16-bit means
  lerped and un-logged, codebook scales and DC, the smallest-three
  quaternion (omitted slot = mode, unit norm), opacity a/255, palette SH
  addressed by (label % 64) * coefficients + k, label / 64; malformed
  archives raise LoaderError.
- The real scene inside the reference's docs/index.data (a SOG ZIP found
  from its end-of-central-directory record; skipped without the reference
  tree): 1,538,841 finite Gaussians with unit quaternions and SH degree 3.
"""
import io
import json
import struct
import zipfile
from pathlib import Path

import numpy as np
import pytest

from gamesplat_amd import loaders

INDEX_DATA = Path("/root/reference/docs/index.data")


def test_ply_activation(tmp_path):
    rng = np.random.default_rng(1)
    n = 500
    f = {"x": rng.normal(0, 3, n), "y": rng.normal(0, 3, n), "z": rng.normal(0, 3, n),
         "f_dc_0": rng.normal(0, .5, n), "f_dc_1": rng.normal(0, .5, n), "f_dc_2": rng.normal(0, .5, n)}
    f.update({f"f_rest_{k}": rng.normal(0, .1, n) for k in range(45)})
    f["opacity"] = rng.normal(0, 2, n)
    f.update({f"scale_{a}": rng.normal(-4.5, .8, n) for a in range(3)})
    f.update({f"rot_{a}": rng.normal(0, 1, n) for a in range(4)})
    f = {k: np.asarray(v, np.float32) for k, v in f.items()}
    loaders.write_ply(tmp_path / "s.ply", f)
    g = loaders.load_ply(tmp_path / "s.ply")
    assert g.shape == (n, 59)
    np.testing.assert_array_equal(g[:, 0], f["x"])
    np.testing.assert_array_equal(g[:, 2], f["z"])
    np.testing.assert_allclose(g[:, 3:6], np.exp(np.stack([f[f"scale_{a}"] for a in range(3)], 1).astype(np.float64)),
                               rtol=2e-7)
    q = np.stack([f[f"rot_{a}"] for a in range(4)], 1).astype(np.float64)
    np.testing.assert_allclose(g[:, 6:10], q / np.linalg.norm(q, axis=1, keepdims=True), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(g[:, 13], 1 / (1 + np.exp(-f["opacity"].astype(np.float64))), rtol=1e-6)
    np.testing.assert_array_equal(g[:, 10], f["f_dc_0"])
    for k in range(15):
        for c in range(3):
            np.testing.assert_array_equal(g[:, 14 + 3 * k + c], f[f"f_rest_{c * 15 + k}"])
    # absent properties take the loader's defaults
    loaders.write_ply(tmp_path / "m.ply", {"x": f["x"], "y": f["y"], "z": f["z"]})
    m = loaders.load_ply(tmp_path / "m.ply")
    assert (m[:, 3:6] == np.float32(0.01)).all() and (m[:, 6] == 1).all() and (m[:, 7:10] == 0).all()
    assert (m[:, 13] == 1).all() and (m[:, 10:13] == 0).all() and (m[:, 14:] == 0).all()


def test_ply_errors(tmp_path):
    p = tmp_path / "bad.ply"
    p.write_bytes(b"ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nproperty float y\n"
                  b"property float z\nend_header\n0 0 0\n")
    with pytest.raises(loaders.LoaderError):
        loaders.load_ply(p)
    p.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty double x\nend_header\n")
    with pytest.raises(loaders.LoaderError):
        loaders.load_ply(p)


def _webp(rgba):
    from PIL import Image

    buf = io.BytesIO()
    Image.fromarray(rgba, "RGBA").save(buf, format="WEBP", lossless=True, exact=True)
    return buf.getvalue()


def _sog(n, w, rng, bands=3, pal=100, mode_byte=None, label_limit=None, version=2):
    h = (n + w - 1) // w
    img = lambda: rng.integers(0, 256, (h, w, 4), dtype=np.uint8)  # noqa: E731
    ml, mu, sc, s0 = img(), img(), img(), img()
    qt = img()
    qt[..., :3] = rng.integers(20, 236, (h, w, 3))  # stored components inside the unit ball
    qt[..., 3] = rng.integers(252, 256, (h, w)) if mode_byte is None else mode_byte
    coefs = {1: 3, 2: 8, 3: 15}[bands]
    cen = rng.integers(0, 256, ((pal + 63) // 64, 64 * coefs, 4), dtype=np.uint8)
    lab = np.zeros((h, w, 4), np.uint8)
    labels = rng.integers(0, pal if label_limit is None else label_limit, (h, w))
    lab[..., 0], lab[..., 1] = labels & 255, labels >> 8
    meta = {"version": version, "count": n,
            "means": {"mins": [-2.0, -1.5, -3.0], "maxs": [2.5, 1.0, 3.5], "files": ["means_l.webp", "means_u.webp"]},
            "scales": {"codebook": list(np.linspace(-7, -1, 256)), "files": ["scales.webp"]},
            "quats": {"files": ["quats.webp"]},
            "sh0": {"codebook": list(np.linspace(-2, 2, 256)), "files": ["sh0.webp"]},
            "shN": {"count": pal, "bands": bands, "codebook": list(np.linspace(-1, 1, 256)),
                    "files": ["shN_centroids.webp", "shN_labels.webp"]}}
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w") as z:
        z.writestr("meta.json", json.dumps(meta))
        for name, a in [("means_l", ml), ("means_u", mu), ("scales", sc), ("quats", qt), ("sh0", s0),
                        ("shN_centroids", cen), ("shN_labels", lab)]:
            z.writestr(f"{name}.webp", _webp(a))
    return buf.getvalue(), meta, dict(ml=ml, mu=mu, sc=sc, qt=qt, s0=s0, cen=cen, lab=lab)


def test_sog_decode_matches_format():
    rng = np.random.default_rng(2)
    n, w = 1000, 48
    data, meta, im = _sog(n, w, rng)
    g = loaders.load_sog(data)
    flat = {k: v.reshape(-1, 4)[:n] for k, v in im.items() if k not in ("cen",)}
    f32 = np.float32
    # means: 16-bit value lerped in float32, then sign(n) (e^|n| - 1)
    q = (flat["mu"][:, :3].astype(np.uint32) << 8) | flat["ml"][:, :3]
    mins, maxs = np.array(meta["means"]["mins"], f32), np.array(meta["means"]["maxs"], f32)
    nv = mins + (maxs - mins) * (q.astype(f32) / f32(65535))
    pos = np.sign(nv) * np.expm1(np.abs(nv).astype(np.float64))
    np.testing.assert_allclose(g[:, 0:3], pos, rtol=1e-6, atol=1e-7)
    scb = np.array(meta["scales"]["codebook"], f32)
    np.testing.assert_allclose(g[:, 3:6], np.exp(scb[flat["sc"][:, :3]].astype(np.float64)), rtol=2e-7)
    s0cb = np.array(meta["sh0"]["codebook"], f32)
    np.testing.assert_array_equal(g[:, 10:13], s0cb[flat["s0"][:, :3]])
    np.testing.assert_array_equal(g[:, 13], flat["s0"][:, 3].astype(f32) / f32(255))
    # smallest three: stored components in order, the omitted one in slot `mode`
    stored = (flat["qt"][:, :3].astype(f32) / f32(255) - f32(0.5)) * f32(2) / np.sqrt(f32(2))
    mode = flat["qt"][:, 3].astype(int) - 252
    for i in range(n):
        others = [k for k in range(4) if k != mode[i]]
        np.testing.assert_array_equal(g[i, 6 + np.array(others)], stored[i])
    np.testing.assert_allclose(np.linalg.norm(g[:, 6:10].astype(np.float64), axis=1), 1.0, atol=1e-6)
    # palette SH: label -> centroid pixel ((label % 64) * 15 + k, label / 64)
    label = flat["lab"][:, 0].astype(int) | (flat["lab"][:, 1].astype(int) << 8)
    cb = np.array(meta["shN"]["codebook"], f32)
    for i in range(0, n, 37):
        for k in range(15):
            px = im["cen"][label[i] // 64, (label[i] % 64) * 15 + k, :3]
            np.testing.assert_array_equal(g[i, 14 + 3 * k:17 + 3 * k], cb[px])


def test_sog_fewer_bands_leave_zeros():
    rng = np.random.default_rng(3)
    data, _, _ = _sog(300, 32, rng, bands=1, pal=70)
    g = loaders.load_sog(data)
    assert (g[:, 14 + 9:] == 0).all() and np.abs(g[:, 14:23]).sum() > 0


def test_sog_errors():
    rng = np.random.default_rng(4)
    with pytest.raises(loaders.LoaderError):
        loaders.load_sog(_sog(100, 16, rng, mode_byte=251)[0])  # invalid quaternion mode
    with pytest.raises(loaders.LoaderError):
        loaders.load_sog(_sog(100, 16, rng, pal=10, label_limit=20)[0])  # label beyond the palette
    with pytest.raises(loaders.LoaderError):
        loaders.load_sog(_sog(100, 16, rng, version=1)[0])
    data = _sog(100, 16, rng)[0]
    src = zipfile.ZipFile(io.BytesIO(data))
    buf = io.BytesIO()
    with zipfile.ZipFile(buf, "w") as z:
        for i in src.infolist():
            if i.filename != "quats.webp":
                z.writestr(i.filename, src.read(i))
    with pytest.raises(loaders.LoaderError):
        loaders.load_sog(buf.getvalue())  # a referenced image is missing


@pytest.mark.skipif(not INDEX_DATA.exists(), reason="reference tree not present")
def test_real_scene_from_reference_bundle():
    raw = INDEX_DATA.read_bytes()
    pos = raw.rfind(b"PK\x05\x06")
    cd_size, cd_off, comment = struct.unpack("<IIH", raw[pos + 12:pos + 22])
    start = pos - cd_size - cd_off
    g = loaders.load_sog(raw[start:pos + 22 + comment])
    assert g.shape == (1_538_841, 59)
    assert np.isfinite(g).all()
    np.testing.assert_allclose(np.linalg.norm(g[:, 6:10].astype(np.float64), axis=1), 1.0, atol=1e-5)
    assert (g[:, 13] >= 0).all() and (g[:, 13] <= 1).all() and (g[:, 3:6] > 0).all()
    assert np.abs(g[:, 14 + 24:]).max() > 0  # third-band coefficients present


def _planes_reference(g):
    """The documented plane layout (include/gs_api.h GS_SCENE_PLANES) built
    with numpy: P0 position + opacity, P1 rotation, P2 scale + DC r, P3 DC g b
    + rest 0..1, P4+j rest 2+4j..5+4j, P14 rest 42..44 and 0."""
    n = g.shape[0]
    rest = g[:, 14:59]
    planes = np.zeros((15, n, 4), np.float32)
    planes[0] = np.c_[g[:, 0:3], g[:, 13]]
    planes[1] = g[:, 6:10]
    planes[2] = np.c_[g[:, 3:6], g[:, 10]]
    planes[3] = np.c_[g[:, 11:13], rest[:, 0:2]]
    for j in range(10):
        planes[4 + j] = rest[:, 2 + 4 * j:6 + 4 * j]
    planes[14, :, 0:3] = rest[:, 42:45]
    return planes


def test_planes_layout_and_ply_straight_to_planes(tmp_path):
    """gs_planes_from_gaussians writes the documented layout, and decoding a
    PLY straight into it (gs_decode_ply_planes) equals decoding to Gaussians
    and repacking, bit for bit."""
    import ctypes as C

    from gamesplat_amd import abi, loaders

    rng = np.random.default_rng(41)
    n = 5003
    f = {"x": rng.normal(0, 3, n), "y": rng.normal(0, 3, n), "z": rng.normal(0, 3, n)}
    f.update({f"f_dc_{c}": rng.normal(0, .5, n) for c in range(3)})
    f.update({f"f_rest_{k}": rng.normal(0, .05, n) for k in range(45)})
    f["opacity"] = rng.normal(0.5, 2, n)
    f.update({f"scale_{a}": rng.normal(-3.5, .6, n) for a in range(3)})
    f.update({f"rot_{a}": rng.normal(0, 1, n) for a in range(4)})
    loaders.write_ply(tmp_path / "s.ply", f)
    g = loaders.load_ply(tmp_path / "s.ply")
    lib = abi.load()
    planes = np.empty((15, n, 4), np.float32)
    abi.check(lib.gs_planes_from_gaussians(g.ctypes.data, n, planes.ctypes.data))
    np.testing.assert_array_equal(planes, _planes_reference(g))
    rows, lay = loaders.read_ply(tmp_path / "s.ply")
    direct = np.empty((15, n, 4), np.float32)
    abi.check(lib.gs_decode_ply_planes(rows.ctypes.data, n, C.byref(lay), direct.ctypes.data))
    np.testing.assert_array_equal(direct.view(np.uint32), planes.view(np.uint32))

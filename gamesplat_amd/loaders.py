"""Scene loaders: binary 3DGS PLY and SOG v2 archives to Gaussians.

The container work happens here (PLY header and rows through numpy, the SOG
ZIP through zipfile, its lossless WebP images through Pillow, meta.json
through json); the per-Gaussian arithmetic is the library's gs_decode_ply /
gs_decode_sog, which apply the reference loaders' float operations
(src/gaussian_loader_ply.cpp:14-228, src/gaussian_loader_sog.cpp:126-551).
Both return a float32 array of shape (N, 59) in GsGaussian layout, ready for
Context.upload.  Errors (unsupported format, missing files or fields,
mismatched images, invalid quaternion modes or palette labels) raise
LoaderError, as the reference loaders fail with a message and no scene.
"""
from __future__ import annotations

import ctypes as C
import io
import json
import zipfile
from pathlib import Path

import numpy as np

from . import _lib

GAUSSIAN_FLOATS = _lib.GAUSSIAN_FLOATS


class LoaderError(RuntimeError):
    pass


GsPlyLayout = _lib.GsPlyLayout
GsSogData = _lib.GsSogData


def _lib_handle():
    return _lib.load()


# ---- PLY -------------------------------------------------------------------

def _ply_header(f):
    props, count, binary_le = [], 0, False
    while True:
        line = f.readline()
        if not line:
            raise LoaderError("PLY header has no end_header")
        text = line.decode("ascii", "replace").strip()
        if text.startswith("format "):
            binary_le = text.split()[1] == "binary_little_endian"
        elif text.startswith("element vertex "):
            count = int(text.split()[2])
        elif text.startswith("property "):
            parts = text.split()
            if parts[1] != "float":
                raise LoaderError(f"unsupported PLY property type: {text}")
            props.append(parts[2])
        elif text == "end_header":
            return props, count, binary_le, f.tell()


def read_ply(path):
    """Binary little-endian 3DGS PLY with float properties: its rows (count,
    properties) float32 and their GsPlyLayout."""
    with open(path, "rb") as f:
        props, count, binary_le, offset = _ply_header(f)
    if not binary_le:
        raise LoaderError("only binary_little_endian PLY is supported")
    if count <= 0 or not props:
        raise LoaderError("PLY has no vertices")
    index = {n: i for i, n in enumerate(props)}
    for n in ("x", "y", "z"):
        if n not in index:
            raise LoaderError(f"PLY is missing property {n}")
    rows = np.memmap(path, dtype="<f4", mode="r", offset=offset, shape=(count, len(props)))
    lay = GsPlyLayout()
    lay.stride = len(props)
    get = lambda n: index.get(n, -1)  # noqa: E731
    lay.x, lay.y, lay.z = get("x"), get("y"), get("z")
    lay.scale[:] = [get(f"scale_{a}") for a in range(3)] if all(f"scale_{a}" in index for a in range(3)) \
        else [-1, -1, -1]
    lay.rot[:] = [get(f"rot_{a}") for a in range(4)] if all(f"rot_{a}" in index for a in range(4)) \
        else [-1, -1, -1, -1]
    lay.opacity = get("opacity")
    lay.dc[:] = [get(f"f_dc_{c}") for c in range(3)] if all(f"f_dc_{c}" in index for c in range(3)) \
        else [-1, -1, -1]
    lay.rest[:] = [get(f"f_rest_{k}") for k in range(45)]
    return np.ascontiguousarray(rows), lay


def load_ply(path) -> np.ndarray:
    """Binary little-endian 3DGS PLY with float properties, decoded to Gaussians."""
    rows, lay = read_ply(path)
    out = np.empty((rows.shape[0], GAUSSIAN_FLOATS), np.float32)
    st = _lib_handle().gs_decode_ply(rows.ctypes.data, rows.shape[0], C.byref(lay), out.ctypes.data)
    if st != 0:
        raise LoaderError(f"PLY decode failed ({st})")
    return out


def upload_ply(ctx, path) -> int:
    """Decodes a PLY straight into the device scene (gs_upload_ply: the rows
    are decoded chunk by chunk into page-locked plane buffers while the
    previous chunk copies).  Returns the Gaussian count."""
    rows, lay = read_ply(path)
    _lib.check(_lib_handle().gs_upload_ply(ctx.handle, rows.ctypes.data, rows.shape[0], C.byref(lay)))
    ctx.n = rows.shape[0]
    return ctx.n


def write_ply(path, fields: dict) -> None:
    """Binary PLY with the given float properties (name -> (N,) array), in order."""
    names = list(fields)
    n = len(next(iter(fields.values())))
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    header += [f"property float {k}" for k in names]
    header.append("end_header")
    data = np.stack([np.asarray(fields[k], np.float32) for k in names], 1)
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode())
        f.write(data.astype("<f4").tobytes())


# ---- SOG -------------------------------------------------------------------

def _rgba(archive: dict, name: str) -> np.ndarray:
    from PIL import Image

    if name not in archive:
        raise LoaderError(f"SOG archive is missing {name}")
    try:
        with Image.open(io.BytesIO(archive[name])) as im:
            if im.format != "WEBP":
                raise LoaderError(f"{name} is not a WebP image")
            return np.ascontiguousarray(np.asarray(im.convert("RGBA"), np.uint8))
    except LoaderError:
        raise
    except Exception as e:  # Pillow raises several types for corrupt data
        raise LoaderError(f"cannot decode {name}: {e}") from e


def _codebook(meta: dict, key: str):
    try:
        cb = [float(x) for x in meta[key]["codebook"]]
        files = meta[key]["files"]
    except (KeyError, TypeError, ValueError) as e:
        raise LoaderError(f"SOG {key} is missing required fields") from e
    if len(cb) < 256:
        raise LoaderError(f"SOG {key} codebook has {len(cb)} entries")
    return cb[:256], files[0]


def read_archive(source) -> dict:
    """Name -> bytes of every file in a SOG ZIP (a path, bytes or a file object)."""
    src = io.BytesIO(source) if isinstance(source, (bytes, bytearray, memoryview)) else source
    with zipfile.ZipFile(src) as z:
        return {i.filename: z.read(i) for i in z.infolist() if not i.is_dir()}


def read_sog(source):
    """SOG v2 archive (a path, bytes or a file object) as GsSogData, with the
    decoded images it points to (keep them alive while it is used)."""
    archive = read_archive(source if not isinstance(source, (str, Path)) else str(source))
    if "meta.json" not in archive:
        raise LoaderError("SOG archive is missing meta.json")
    try:
        meta = json.loads(archive["meta.json"])
    except ValueError as e:
        raise LoaderError(f"meta.json: {e}") from e
    if meta.get("version") != 2:
        raise LoaderError(f"unsupported SOG version {meta.get('version')}")
    count = int(meta.get("count", 0))
    if count <= 0:
        raise LoaderError("SOG count out of range")
    try:
        mins, maxs = meta["means"]["mins"], meta["means"]["maxs"]
        mfiles = meta["means"]["files"]
        qfile = meta["quats"]["files"][0]
    except (KeyError, TypeError, IndexError) as e:
        raise LoaderError("SOG meta is missing means or quats") from e
    scb, sfile = _codebook(meta, "scales")
    s0cb, s0file = _codebook(meta, "sh0")
    ml, mu = _rgba(archive, mfiles[0]), _rgba(archive, mfiles[1])
    sc, qt, s0 = _rgba(archive, sfile), _rgba(archive, qfile), _rgba(archive, s0file)
    shape = ml.shape
    if any(im.shape != shape for im in (mu, sc, qt, s0)):
        raise LoaderError("SOG per-Gaussian image dimensions do not match")
    if count > shape[0] * shape[1]:
        raise LoaderError("SOG Gaussian count exceeds the image capacity")
    d = GsSogData()
    d.count = count
    d.means_mins[:] = [float(x) for x in mins[:3]]
    d.means_maxs[:] = [float(x) for x in maxs[:3]]
    d.scales_codebook[:] = scb
    d.sh0_codebook[:] = s0cb
    d.means_l, d.means_u, d.scales, d.quats, d.sh0 = (im.ctypes.data for im in (ml, mu, sc, qt, s0))
    keep = [ml, mu, sc, qt, s0]
    if "shN" in meta:
        shn = meta["shN"]
        try:
            bands, pal = int(shn["bands"]), int(shn["count"])
            cb = [float(x) for x in shn["codebook"]][:256]
            cfile, lfile = shn["files"][0], shn["files"][1]
        except (KeyError, TypeError, ValueError, IndexError) as e:
            raise LoaderError("SOG shN is missing required fields") from e
        if not 1 <= bands <= 3 or not 0 < pal <= 65536 or len(cb) < 256:
            raise LoaderError("SOG shN bands, count or codebook out of range")
        cen, lab = _rgba(archive, cfile), _rgba(archive, lfile)
        if count > lab.shape[0] * lab.shape[1]:
            raise LoaderError("SOG shN label image is too small for the Gaussian count")
        coefs = {1: 3, 2: 8, 3: 15}[bands]
        if cen.shape[1] < 64 * coefs or cen.shape[0] < (pal + 63) // 64:
            raise LoaderError("SOG shN centroid image is too small")
        d.shN_bands, d.shN_count = bands, pal
        d.shN_codebook[:] = cb
        d.shN_centroids, d.centroids_width, d.centroids_height = cen.ctypes.data, cen.shape[1], cen.shape[0]
        d.shN_labels = lab.ctypes.data
        keep += [cen, lab]
    return d, keep


def load_sog(source) -> np.ndarray:
    """SOG v2 archive (a path, bytes or a file object), decoded to Gaussians."""
    d, keep = read_sog(source)
    out = np.empty((d.count, GAUSSIAN_FLOATS), np.float32)
    st = _lib_handle().gs_decode_sog(C.byref(d), out.ctypes.data)
    del keep
    if st != 0:
        raise LoaderError("SOG decode failed: invalid quaternion mode or palette label")
    return out


def upload_sog(ctx, source) -> int:
    """Decodes a SOG archive straight into the device scene (gs_upload_sog,
    pipelined like upload_ply).  Returns the Gaussian count."""
    d, keep = read_sog(source)
    code = _lib_handle().gs_upload_sog(ctx.handle, C.byref(d))
    del keep
    if code != 0:
        raise LoaderError(f"SOG upload failed: {_lib_handle().gs_last_error().decode()}")
    ctx.n = d.count
    return ctx.n


def load_scene(path) -> np.ndarray:
    """By extension: .ply or .sog."""
    p = Path(path)
    if p.suffix.lower() == ".ply":
        return load_ply(p)
    if p.suffix.lower() == ".sog":
        return load_sog(p)
    raise LoaderError(f"unsupported scene format: {p.suffix}")

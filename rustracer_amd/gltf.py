"""glTF 2.0 scenes loaded into a Scene (rustracer-core/src/loader.rs:265-650).

The reference imports the file with the `gltf` crate (1.x; buffers, images via
the `image` crate) and bakes it into world-space meshes, materials, role
texture arrays and lights. This module does the same over this library's
scene API:

- `.gltf` (JSON, buffers as data URIs or files beside it) and `.glb`;
- the default scene's node hierarchy, depth first in document order, with
  world = parent * local and glam's evaluation order for matrix products and
  point/vector transforms (f32 throughout, no fused multiply-add);
- triangles, strips (alternating winding) and fans; implicit indices; smooth
  normals when none are authored; mirrored transforms flip the winding;
- metallic-roughness materials with the crate's defaults, alphaMode MASK as
  alpha cutout and BLEND as a dielectric; base-colour, metallic-roughness,
  emissive and normal textures compacted into one array per role (sRGB decode
  for base colour and emissive, normal scale baked into x/y), mixed sizes
  dropped;
- an area light per emitting triangle, and the default sky when the file has
  no environment light.

Images are decoded with Pillow. 8-bit PNG decoding is exact, so those textures
equal the crate's; JPEG decoders may differ by a unit in some texels, and
Pillow delivers 16-bit colour PNGs as 8-bit (the crate keeps 16 bits). The node
matrix of a translation/rotation/scale triple and the inverse used for normals
follow the usual formulas in f32 but not necessarily the crate's and glam's
exact operation order: for identity, translation and axis-aligned scale they
are exact, for general rotations they may differ in the last bit.
"""
from __future__ import annotations

import base64
import ctypes as C
import json
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from . import _ffi
from ._ffi import check, lib, ptr

F32 = np.float32
MAX_ROLE_LAYERS = 240  # loader.rs:22
SKY_ZENITH = (0.55, 0.6, 0.75)
SKY_HORIZON = (0.95, 0.93, 0.9)
ROLE_ALBEDO, ROLE_MR, ROLE_EMISSIVE, ROLE_NORMAL = 0, 1, 2, 3
_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


class GltfError(ValueError):
    pass


# ─────────────────────────── texture decoding (loader.rs:28-131) ───────────────────────────

def srgb_to_linear(values: np.ndarray) -> np.ndarray:
    """loader.rs:28-34, evaluated by the library in f32 with the C library's powf."""
    a = np.ascontiguousarray(values, F32).copy()
    check(lib().rt_srgb_to_linear(ptr(a), a.size), "rt_srgb_to_linear")
    return a


@dataclass
class Texture:
    """A linear RGBA f32 texture, (height, width, 4) (rustracer-core texture.rs)."""
    data: np.ndarray

    @property
    def width(self) -> int:
        return self.data.shape[1]

    @property
    def height(self) -> int:
        return self.data.shape[0]

    @staticmethod
    def solid(rgba) -> "Texture":
        return Texture(np.asarray(rgba, F32).reshape(1, 1, 4).copy())


def decode_to_rgba(img) -> np.ndarray | None:
    """loader.rs:36-116 for an image as the `image` crate delivers it: 8-bit grey,
    RGB and RGBA and 16-bit RGB(A) become RGBA f32 (unorm / 255 or / 65535,
    alpha 1 when absent); grey+alpha and 16-bit grey are not supported (None)."""
    mode = img.mode
    if mode == "P":  # palette images are expanded by the decoder
        img = img.convert("RGBA" if "transparency" in img.info else "RGB")
        mode = img.mode
    elif mode == "1":
        img = img.convert("L")
        mode = "L"
    if mode == "L":
        v = np.asarray(img, np.uint8).astype(F32) / F32(255.0)
        return np.stack([v, v, v, np.ones_like(v)], axis=-1)
    if mode in ("RGB", "RGBA"):
        a = np.asarray(img, np.uint8).astype(F32) / F32(255.0)
        if mode == "RGB":
            a = np.concatenate([a, np.ones(a.shape[:2] + (1,), F32)], axis=-1)
        return np.ascontiguousarray(a)
    return None


def build_role(master, used, srgb: bool, scales=None):
    """loader.rs:133-234 (build_role, and build_role_scaled when `scales` is
    given): the used textures (in material order, each once) of the first used
    texture's size, at most MAX_ROLE_LAYERS, sRGB-decoded or with the normal
    scale baked into x and y. Returns the role list and, per master texture,
    its layer or None."""
    mapping = [None] * len(master)
    if not master:
        return [], mapping
    first = next((u for u in used if u is not None), 0)
    ref_w, ref_h = master[first].width, master[first].height
    role = []
    for m in used:
        if m is None or mapping[m] is not None:
            continue
        tex = master[m]
        if (tex.width, tex.height) != (ref_w, ref_h) or len(role) >= MAX_ROLE_LAYERS:
            continue  # the reference warns and drops it
        copy = tex.data.copy()
        if scales is not None:
            scale = F32(scales[m] if m < len(scales) else 1.0)
            if abs(float(scale) - 1.0) > 1e-6:
                copy[..., 0] *= scale
                copy[..., 1] *= scale
        elif srgb:
            copy[..., :3] = srgb_to_linear(copy[..., :3])
        mapping[m] = len(role)
        role.append(Texture(copy))
    return role, mapping


def sample_nearest(tex: Texture, u: float, v: float):
    """loader.rs:253-263: repeat addressing, nearest texel."""
    if tex.width == 0 or tex.height == 0:
        return np.zeros(4, F32)

    def wrap(t):
        t = F32(t)
        fr = F32(t - F32(np.trunc(t)))
        r = F32(np.fmod(fr, F32(1.0)))
        return F32(r + F32(1.0)) if r < 0 else r

    def index(t, n):
        x = F32(t * F32(n))
        i = 0 if not x > 0 else int(x)  # `as usize` saturates; NaN -> 0
        return min(i, n - 1)

    return tex.data[index(wrap(v), tex.height), index(wrap(u), tex.width)]


# ─────────────────────────── f32 vector / matrix arithmetic in glam's order ───────────────────────────

def _dot3(a, b):
    return F32(F32(F32(a[0] * b[0]) + F32(a[1] * b[1])) + F32(a[2] * b[2]))


def _cross(a, b):
    return np.array([a[1] * b[2] - b[1] * a[2], a[2] * b[0] - b[2] * a[0], a[0] * b[1] - b[0] * a[1]], F32)


def _normalize(v):
    return (v * (F32(1.0) / np.sqrt(_dot3(v, v)))).astype(F32)


def _mat_mul(a, b):
    """glam Mat4 * Mat4 (column-major 4x4 arrays, a[col][row]): each result
    column is ((a.x * b.x + a.y * b.y) + a.z * b.z) + a.w * b.w."""
    out = np.zeros((4, 4), F32)
    for j in range(4):
        col = a[0] * b[j][0]
        col = col + a[1] * b[j][1]
        col = col + a[2] * b[j][2]
        col = col + a[3] * b[j][3]
        out[j] = col
    return out


def _transform_points(m, p):
    """glam Mat4::transform_point3 over an (n, 3) array."""
    res = m[0][:3] * p[:, 0:1]
    res = m[1][:3] * p[:, 1:2] + res
    res = m[2][:3] * p[:, 2:3] + res
    res = m[3][:3] + res
    return res.astype(F32)


def _transform_vectors(m, v):
    """glam Mat4::transform_vector3 over an (n, 3) array."""
    res = m[0][:3] * v[:, 0:1]
    res = m[1][:3] * v[:, 1:2] + res
    res = m[2][:3] * v[:, 2:3] + res
    return res.astype(F32)


def _inverse(m):
    """4x4 inverse (column-major in and out), computed in f64 and rounded to f32."""
    a = m.T.astype(np.float64)  # rows
    try:
        inv = np.linalg.inv(a)
    except np.linalg.LinAlgError:
        inv = np.full((4, 4), np.inf)
    return inv.T.astype(F32)


def _determinant(m):
    return float(np.linalg.det(m.astype(np.float64)))


def _node_matrix(node) -> np.ndarray:
    """The node's local transform, column-major (glTF `matrix`, or T * R * S)."""
    if "matrix" in node:
        return np.asarray(node["matrix"], F32).reshape(4, 4)
    t = np.asarray(node.get("translation", (0, 0, 0)), F32)
    x, y, z, w = (F32(c) for c in node.get("rotation", (0, 0, 0, 1)))
    s = np.asarray(node.get("scale", (1, 1, 1)), F32)
    x2, y2, z2 = x + x, y + y, z + z
    xx2, xy2, xz2 = x2 * x, x2 * y, x2 * z
    yy2, yz2, zz2 = y2 * y, y2 * z, z2 * z
    sx2, sy2, sz2 = x2 * w, y2 * w, z2 * w
    one = F32(1.0)
    r = np.array([[one - yy2 - zz2, xy2 + sz2, xz2 - sy2, 0],
                  [xy2 - sz2, one - xx2 - zz2, yz2 + sx2, 0],
                  [xz2 + sy2, yz2 - sx2, one - xx2 - yy2, 0],
                  [0, 0, 0, 1]], F32)
    tm = np.eye(4, dtype=F32)
    tm[3, :3] = t
    sm = np.diag(np.array([s[0], s[1], s[2], 1.0], F32))
    return _mat_mul(_mat_mul(tm, r), sm)


# ─────────────────────────── file access ───────────────────────────

class _File:
    """The JSON document, its binary buffers and image bytes."""

    def __init__(self, path: str):
        self.base = os.path.dirname(os.path.abspath(path))
        raw = open(path, "rb").read()
        glb_bin = None
        if raw[:4] == b"glTF":
            version, length = struct.unpack_from("<II", raw, 4)
            if version != 2:
                raise GltfError(f"unsupported GLB version {version}")
            off, doc = 12, None
            while off + 8 <= min(length, len(raw)):
                clen, ctype = struct.unpack_from("<II", raw, off)
                chunk = raw[off + 8:off + 8 + clen]
                if ctype == 0x4E4F534A:
                    doc = json.loads(chunk.decode("utf-8"))
                elif ctype == 0x004E4942 and glb_bin is None:
                    glb_bin = chunk
                off += 8 + clen + (-clen % 4)
            if doc is None:
                raise GltfError("GLB without a JSON chunk")
        else:
            doc = json.loads(raw.decode("utf-8"))
        self.doc = doc
        asset = doc.get("asset", {})
        if not str(asset.get("version", "2.0")).startswith("2"):
            raise GltfError(f"unsupported glTF version {asset.get('version')}")
        self.buffers = []
        for i, b in enumerate(doc.get("buffers", [])):
            if "uri" in b:
                data = self._uri(b["uri"])
            elif glb_bin is not None and i == 0:
                data = glb_bin
            else:
                raise GltfError(f"buffer {i} has no data")
            if len(data) < b.get("byteLength", 0):
                raise GltfError(f"buffer {i} is shorter than its byteLength")
            self.buffers.append(data)

    def _uri(self, uri: str) -> bytes:
        if uri.startswith("data:"):
            head, _, payload = uri.partition(",")
            if not head.endswith(";base64"):
                raise GltfError("only base64 data URIs are supported")
            return base64.b64decode(payload)
        from urllib.parse import unquote
        return open(os.path.join(self.base, unquote(uri)), "rb").read()

    def view_bytes(self, index: int) -> tuple[bytes, int, int | None]:
        v = self.doc["bufferViews"][index]
        buf = self.buffers[v["buffer"]]
        off = v.get("byteOffset", 0)
        return buf[off:off + v["byteLength"]], v["byteLength"], v.get("byteStride")

    def accessor(self, index: int) -> np.ndarray:
        """An accessor's elements as an (count, width) array of its component type."""
        a = self.doc["accessors"][index]
        dtype = np.dtype(_COMPONENT[a["componentType"]]).newbyteorder("<")
        width = _WIDTH[a["type"]]
        count = a["count"]
        out = np.zeros((count, width), dtype)
        if "bufferView" in a:
            data, _, stride = self.view_bytes(a["bufferView"])
            elem = dtype.itemsize * width
            stride = stride or elem
            off = a.get("byteOffset", 0)
            if count and off + stride * (count - 1) + elem > len(data):
                raise GltfError(f"accessor {index} reads past its buffer view")
            raw = np.frombuffer(data, np.uint8, count=(stride * (count - 1) + elem) if count else 0, offset=off)
            rows = np.lib.stride_tricks.as_strided(raw, shape=(count, elem), strides=(stride, 1)) if count else raw
            out = np.ascontiguousarray(rows).view(dtype).reshape(count, width).copy()
        sparse = a.get("sparse")
        if sparse:
            n = sparse["count"]
            idx_desc, val_desc = sparse["indices"], sparse["values"]
            idata, _, _ = self.view_bytes(idx_desc["bufferView"])
            idx = np.frombuffer(idata, np.dtype(_COMPONENT[idx_desc["componentType"]]).newbyteorder("<"), count=n,
                                offset=idx_desc.get("byteOffset", 0))
            vdata, _, _ = self.view_bytes(val_desc["bufferView"])
            vals = np.frombuffer(vdata, dtype, count=n * width, offset=val_desc.get("byteOffset", 0))
            out[idx.astype(np.int64)] = vals.reshape(n, width)
        return out

    def image(self, index: int):
        from PIL import Image  # the image decoder (Pillow)
        import io

        im = self.doc["images"][index]
        if "uri" in im:
            data = self._uri(im["uri"])
        else:
            data, _, _ = self.view_bytes(im["bufferView"])
        img = Image.open(io.BytesIO(data))
        img.load()
        return img


def _unorm_to_f32(a: np.ndarray) -> np.ndarray:
    """The crate's into_f32 for texture coordinates: u8 / 255, u16 / 65535."""
    if a.dtype == np.uint8:
        return a.astype(F32) / F32(255.0)
    if a.dtype == np.uint16:
        return a.astype(F32) / F32(65535.0)
    return a.astype(F32)


# ─────────────────────────── materials ───────────────────────────

@dataclass
class _Pending:
    material: np.ndarray  # one MATERIAL record
    albedo: int | None = None
    mr: int | None = None
    emissive: int | None = None
    normal: int | None = None
    normal_scale: float = 1.0


def _material_record(kind, albedo, roughness, metallic, ior, emissive, cutout) -> np.ndarray:
    m = np.zeros(1, _ffi.MATERIAL)
    m["albedo"][0] = (albedo[0], albedo[1], albedo[2], 1.0)
    m["emissive"][0] = emissive
    m["roughness"], m["metallic"], m["ior"], m["kind"] = roughness, metallic, ior, kind
    m["flags"] = 1 if cutout else 0
    for t in ("tex_albedo", "tex_mr", "tex_emissive", "tex_normal"):
        m[t] = _ffi.TEX_NONE
    return m


def _materials(doc) -> list:
    """loader.rs:284-347: materials with their raw glTF texture indices."""
    out = []
    for mat in doc.get("materials", []):
        pbr = mat.get("pbrMetallicRoughness", {})
        base = pbr.get("baseColorFactor", (1.0, 1.0, 1.0, 1.0))
        rough = F32(pbr.get("roughnessFactor", 1.0))
        metal = F32(pbr.get("metallicFactor", 1.0))
        rough = F32(1.0) if (np.isnan(rough) or rough < 0) else rough
        metal = F32(0.0) if (np.isnan(metal) or metal < 0) else metal
        mode = mat.get("alphaMode", "OPAQUE")
        kind = _ffi.MAT_DIELECTRIC if mode == "BLEND" else _ffi.MAT_LAMBERTIAN
        rec = _material_record(kind, base[:3], rough, metal, 1.5, mat.get("emissiveFactor", (0.0, 0.0, 0.0)),
                               mode == "MASK")
        tex = lambda info: None if info is None else info["index"]
        normal = mat.get("normalTexture")
        out.append(_Pending(rec, tex(pbr.get("baseColorTexture")), tex(pbr.get("metallicRoughnessTexture")),
                            tex(mat.get("emissiveTexture")), tex(normal),
                            float(normal.get("scale", 1.0)) if normal else 1.0))
    if not out:  # loader.rs:350-359: Material::lambertian(0.8)
        out.append(_Pending(_material_record(_ffi.MAT_LAMBERTIAN, (0.8, 0.8, 0.8), 1.0, 0.0, 1.5, (0, 0, 0), False)))
    for p in out:  # loader.rs:361-367: a textured emissive with a zero factor uses the texture as is
        e = p.material["emissive"][0].astype(F32)
        if p.emissive is not None and np.sqrt(_dot3(e, e)) < F32(1e-6):
            p.material["emissive"][0] = (1.0, 1.0, 1.0)
    return out


# ─────────────────────────── meshes ───────────────────────────

@dataclass
class _Mesh:
    positions: np.ndarray
    normals: np.ndarray
    texcoords: np.ndarray
    indices: np.ndarray
    material_id: int
    extras: dict = field(default_factory=dict)


def _triangulate(mode: int, idx: np.ndarray) -> np.ndarray:
    if mode == 4:
        n = len(idx) // 3
        return idx[:3 * n].reshape(n, 3)
    if mode == 5:  # strips wind alternately
        tris = [(idx[i], idx[i + 1], idx[i + 2]) if i % 2 == 0 else (idx[i], idx[i + 2], idx[i + 1])
                for i in range(max(len(idx) - 2, 0))]
        return np.array(tris, np.uint32).reshape(-1, 3)
    if mode == 6:
        tris = [(idx[0], idx[i], idx[i + 1]) for i in range(1, max(len(idx) - 1, 1))]
        return np.array(tris, np.uint32).reshape(-1, 3)
    return np.zeros((0, 3), np.uint32)


def _bake_primitive(f: _File, prim, world: np.ndarray, n_materials: int) -> _Mesh | None:
    """loader.rs:512-650."""
    attrs = prim.get("attributes", {})
    if "POSITION" not in attrs:
        return None
    positions = f.accessor(attrs["POSITION"]).astype(F32)
    if len(positions) == 0:
        return None
    n_verts = len(positions)
    mode = prim.get("mode", 4)
    if mode not in (4, 5, 6):
        return None  # points and lines: skipped (the reference warns)
    if "indices" in prim:
        idx = f.accessor(prim["indices"]).reshape(-1).astype(np.uint32)
    else:
        idx = np.arange(n_verts, dtype=np.uint32)
    tris = _triangulate(mode, idx)
    if len(tris) == 0:
        return None
    if np.any(tris >= n_verts):
        raise GltfError("primitive index out of range of its positions")
    if "NORMAL" in attrs:
        normals = f.accessor(attrs["NORMAL"]).astype(F32)
    else:
        acc = np.zeros((n_verts, 3), F32)
        for a, b, c in tris:
            n = _cross(positions[b] - positions[a], positions[c] - positions[a])
            acc[a] += n
            acc[b] += n
            acc[c] += n
        normals = np.empty_like(acc)
        for i, n in enumerate(acc):
            normals[i] = _normalize(n) if _dot3(n, n) > F32(1e-12) else (0.0, 1.0, 0.0)
    if len(normals) != n_verts:
        normals = np.tile(np.array([0.0, 1.0, 0.0], F32), (n_verts, 1))
    if "TEXCOORD_0" in attrs:
        texcoords = _unorm_to_f32(f.accessor(attrs["TEXCOORD_0"]))
    else:
        texcoords = np.zeros((n_verts, 2), F32)
    normal_mat = _inverse(world).T.copy()
    mirrored = _determinant(world) < 0.0
    pos_w = _transform_points(world, positions)
    nt = _transform_vectors(normal_mat, normals)
    n_out = normals.copy()
    for i in range(n_verts):
        if _dot3(nt[i], nt[i]) > F32(1e-12):
            n_out[i] = _normalize(nt[i])
    if mirrored:
        tris = tris[:, [0, 2, 1]]
    mat = prim.get("material")
    last = max(n_materials - 1, 0)
    material_id = last if mat is None else min(mat, last)
    return _Mesh(pos_w, n_out, texcoords.astype(F32), np.ascontiguousarray(tris.reshape(-1), np.uint32), material_id)


def _triangle_emits(mesh: _Mesh, t: int, mat: np.ndarray, tex_emissive: list) -> bool:
    """loader.rs:236-251."""
    e = mat["emissive"].astype(F32).reshape(3)
    layer = int(mat["tex_emissive"])
    if layer == _ffi.TEX_NONE:
        return bool(np.sqrt(_dot3(e, e)) > F32(1e-4))
    if layer >= len(tex_emissive):
        return False
    i0, i1, i2 = mesh.indices[3 * t:3 * t + 3]
    uv = lambda i: mesh.texcoords[i] if i < len(mesh.texcoords) else np.zeros(2, F32)
    c_uv = ((uv(i0) + uv(i1)) + uv(i2)) / F32(3.0)
    c = sample_nearest(tex_emissive[layer], c_uv[0], c_uv[1])
    return bool(_dot3(c[:3], e) > F32(0.05))


# ─────────────────────────── the loader ───────────────────────────

def load_gltf(path: str):
    """loader.rs:265-510: a Scene with the file's meshes, materials, role
    texture arrays, area lights and (unless the file has one) the default sky."""
    from . import Scene

    f = _File(path)
    doc = f.doc
    # textures, in glTF texture order (loader.rs:269-281)
    master = []
    for t in doc.get("textures", []):
        src = t.get("source")
        data = None
        if src is not None:
            data = decode_to_rgba(f.image(src))
        master.append(Texture(data) if data is not None else Texture.solid((0.5, 0.5, 0.5, 1.0)))
    pending = _materials(doc)
    used = lambda attr: [getattr(p, attr) for p in pending]
    role_albedo, map_albedo = build_role(master, used("albedo"), True)
    role_mr, map_mr = build_role(master, used("mr"), False)
    role_emissive, map_emissive = build_role(master, used("emissive"), True)
    scales = [1.0] * len(master)  # loader.rs:375-388: the first scale seen for each texture
    for p in pending:
        if p.normal is not None and not (abs(scales[p.normal] - p.normal_scale) > 1e-4 and scales[p.normal] != 1.0):
            scales[p.normal] = p.normal_scale
    role_normal, map_normal = build_role(master, used("normal"), False, scales=scales)

    def remap(mapping, raw):
        if raw is None or raw >= len(mapping) or mapping[raw] is None:
            return _ffi.TEX_NONE
        return mapping[raw]

    materials = []
    for p in pending:
        m = p.material
        m["tex_albedo"] = remap(map_albedo, p.albedo)
        m["tex_mr"] = remap(map_mr, p.mr)
        m["tex_emissive"] = remap(map_emissive, p.emissive)
        m["tex_normal"] = remap(map_normal, p.normal)
        materials.append(m)

    # the default scene's nodes, depth first in document order (loader.rs:403-445)
    meshes = []
    scenes = doc.get("scenes", [])
    root = scenes[doc["scene"]] if "scene" in doc and scenes else (scenes[0] if scenes else None)
    stack = [(n, np.eye(4, dtype=F32)) for n in reversed(root.get("nodes", []))] if root else []
    nodes = doc.get("nodes", [])
    while stack:
        ni, parent = stack.pop()
        node = nodes[ni]
        world = _mat_mul(parent, _node_matrix(node))
        if "mesh" in node:
            for prim in doc["meshes"][node["mesh"]].get("primitives", []):
                m = _bake_primitive(f, prim, world, len(materials))
                if m is not None:
                    meshes.append(m)
        for c in reversed(node.get("children", [])):
            stack.append((c, world))

    scene = Scene()
    for m in materials:
        scene.add_material(m)
    for role, textures in ((ROLE_ALBEDO, role_albedo), (ROLE_MR, role_mr), (ROLE_EMISSIVE, role_emissive),
                           (ROLE_NORMAL, role_normal)):
        for t in textures:
            scene.add_texture(role, t.data)
    for m in meshes:
        scene.add_mesh(m.positions, m.indices, m.material_id, normals=m.normals, uvs=m.texcoords)

    # area lights from emitting triangles (loader.rs:447-469)
    tri_offset = 0
    for m in meshes:
        mat = materials[m.material_id][0] if m.material_id < len(materials) else None
        n_tris = len(m.indices) // 3
        if mat is not None:
            e = mat["emissive"].astype(F32)
            emits = int(mat["tex_emissive"]) != _ffi.TEX_NONE or np.sqrt(_dot3(e, e)) > F32(1e-4)
            if emits:
                for t in range(n_tris):
                    if _triangle_emits(m, t, mat, role_emissive):
                        scene.add_light(_ffi.LIGHT_AREA, color=tuple(float(x) for x in e), intensity=1.0,
                                        triangle_index=tri_offset + t)
        tri_offset += n_tris
    # loader.rs:471-490: the default sky unless an environment light exists; a
    # glTF file brings none, so it is always added (scene.toml may replace it)
    scene.add_light(_ffi.LIGHT_ENVIRONMENT, vec_a=SKY_ZENITH, vec_b=SKY_HORIZON, intensity=1.0)
    return scene

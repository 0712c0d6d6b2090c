"""Scene ingestion: Rendu's .scene files, OBJ meshes and PNG textures into a rendu_amd Scene.

Host-side restatement of what the path-tracer app does before PathTracer::render:
* Codable::decode (src/engine/system/Codable.cpp:185-260): '#' comments, 'key: values' lines,
  nested 'a: b: values' sub-tokens, '* object' roots and '- item' array elements;
* Scene::init (src/engine/scene/Scene.cpp:31-131): objects, materials (matched by name after
  loading), point / directional / spot lights, background (colour, image, cube), camera; objects
  sorted by material type (Scene.cpp:104-107, stable here);
* Codable::decodeTransformation (Codable.cpp:79-104): model = T * R(angle, axis) * S;
* Resources::getMesh (ResourcesManager.cpp:306-340) with Mesh(stream, Indexed) (Mesh.cpp:100-260:
  only the first three vertices of a face, 'vt' v flipped to 1 - v, vertices shared per distinct
  'v/vt/vn' string) and Mesh::computeNormals (Mesh.cpp:327-350) when normals are missing;
* Resources::getTexture (ResourcesManager.cpp:345-530): 'r,g,b,a' strings give 8 x 8 constant
  textures, file names resolve to .png (then .jpg, .jpeg, .bmp, .tga, .exr) anywhere under the
  resource roots, then to the first level of a mipmapped file (name_0), cubemaps to name_px, _nx,
  _ny, _py, _pz, _nz or their first level (name_0_px, ...); 8-bit images become value / 255 with
  4 channels (Image::loadLDR, Image.cpp:291-330), half-float EXR channels are read by name
  (Image::loadHDR); cube faces are flipped vertically;
* the atmosphere background ('sun'): Rendu's precomputed scattering table as its texture.

* Mesh::computeTangentsAndBitangents(force) (Mesh.cpp:356-508), forced as the path-tracer app
  loads its scenes (pathtracer/main.cpp:90): MikkTSpace corner frames and Rendu's vertex split, in
  the library (rdu_mesh_tangent_frames).

Differences: PNG (8-bit, non-interlaced) and uncompressed EXR are decoded here; JPEG goes through Pillow's decoder (libjpeg), whose inverse DCT
and chroma upsampling can differ from stb_image's (Rendu's) by a level or two in some texels.
Animations are not supported (scenes load in their initial state).
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import scene as S
from .camera import Camera
from .output import read_exr, read_png
from .workloads import MeshData

F32 = np.float32

MATERIAL_TYPES = {"None": 0, "Regular": 1, "Parallax": 2, "Clearcoat": 3, "Anisotropic": 4, "Sheen": 5,
                  "Iridescent": 6, "Subsurface": 7, "Emissive": 8, "Transparent": 9, "TransparentIrid": 10}
TRUES = ("true", "True", "yes", "Yes", "1", "y", "Y")


@dataclass
class KeyValues:
    key: str
    values: List[str] = field(default_factory=list)
    elements: List["KeyValues"] = field(default_factory=list)


def decode(text: str) -> List[KeyValues]:
    """Codable::decode: the root objects of a scene description."""
    raw: List[KeyValues] = []
    for line in text.splitlines():
        h = line.find("#")
        if h >= 0:
            line = line[:h]
        line = line.strip(" \t\r")
        if not line:
            continue
        c = line.find(":")
        if c < 0:
            continue
        tok = KeyValues(line[:c].strip(" \t"))
        raw.append(tok)
        prev, nxt = c + 1, line.find(":", c + 1)
        while nxt >= 0:
            sub = line[prev:nxt].strip(" \t")
            if sub:
                child = KeyValues(sub)
                tok.elements.append(child)
                tok = child
            prev, nxt = nxt + 1, line.find(":", nxt + 1)
        vals = line[prev:].replace(",", " ").strip(" \t")
        tok.values = [v for v in vals.split(" ") if v]
    roots: List[KeyValues] = []
    i = 0
    while i < len(raw):
        t = raw[i]
        if t.key.startswith("*"):
            roots.append(KeyValues(t.key[1:].strip("\t "), t.values))
            i += 1
            continue
        if not roots:
            i += 1
            continue
        roots[-1].elements.append(t)
        arr = roots[-1].elements[-1]
        i += 1
        while i < len(raw) and raw[i].key.startswith("-"):
            item = raw[i]
            item.key = item.key[1:].strip("\t ")
            arr.elements.append(item)
            i += 1
    return roots


def _vec(kv: KeyValues, n: int, default=0.0):
    out = [default] * n
    for k in range(min(n, len(kv.values))):
        out[k] = float(kv.values[k])
    return np.array(out, F32)


def _bool(kv: KeyValues) -> bool:
    return bool(kv.values) and kv.values[0] in TRUES


def transformation(elements: Sequence[KeyValues]) -> np.ndarray:
    """Codable::decodeTransformation as a row-major 4 x 4 float32 matrix (T * R * S)."""
    axis, trans, angle, scale = np.zeros(3, F32), np.zeros(3, F32), F32(0.0), F32(1.0)
    for p in elements:
        if p.key == "orientation":
            axis = _vec(p, 3)
            axis = (axis / F32(np.sqrt(np.dot(axis, axis)))).astype(F32)
            if len(p.values) >= 4:
                angle = F32(float(p.values[3]))
        elif p.key == "translation":
            trans = _vec(p, 3)
        elif p.key == "scaling" and p.values:
            scale = F32(float(p.values[0]))
    m = np.eye(4, dtype=F32)
    if angle != 0.0:  # glm::rotate (ext/matrix_transform.inl), in float
        c, s = F32(math.cos(angle)), F32(math.sin(angle))
        a = axis
        t = (F32(1.0) - c) * a
        rot = np.array([[c + t[0] * a[0], t[0] * a[1] + s * a[2], t[0] * a[2] - s * a[1]],
                        [t[1] * a[0] - s * a[2], c + t[1] * a[1], t[1] * a[2] + s * a[0]],
                        [t[2] * a[0] + s * a[1], t[2] * a[1] - s * a[0], c + t[2] * a[2]]], F32)  # [col][row]
        m[:3, :3] = rot.T
    m[:3, :3] = m[:3, :3] * scale
    m[:3, 3] = trans
    return m


def _stoi(s: str) -> int:
    """std::stoi: optional leading whitespace and sign, then the longest run of digits."""
    s = s.lstrip(" \t")
    k = 1 if s[:1] in ("+", "-") else 0
    e = k
    while e < len(s) and s[e].isdigit():
        e += 1
    if e == k:
        raise ValueError(f"no integer in {s!r}")
    return int(s[:e])


def load_obj(text: str) -> MeshData:
    """Mesh(stream, Mesh::Load::Indexed) (Mesh.cpp:100-260)."""
    pos, nrm, uvs, faces = [], [], [], []
    for line in text.split("\n"):
        if not line or line[0] == "#" or len(line) < 2:
            continue
        tok = line.strip("\r").split()
        if not tok:
            continue
        if tok[0] == "v" and len(tok) >= 4:
            pos.append((float(tok[1]), float(tok[2]), float(tok[3])))
        elif tok[0] == "vn" and len(tok) >= 4:
            nrm.append((float(tok[1]), float(tok[2]), float(tok[3])))
        elif tok[0] == "vt" and len(tok) >= 3:
            uvs.append((float(tok[1]), 1.0 - np.float32(float(tok[2]))))
        elif tok[0] == "f" and len(tok) >= 4:
            faces.extend(tok[1:4])
    if not pos:
        return MeshData(np.zeros((0, 3), F32), np.zeros(0, np.uint32))
    P, N, T = np.array(pos, F32), np.array(nrm, F32), np.array(uvs, F32)
    has_uv, has_n = len(uvs) > 0, len(nrm) > 0
    used: Dict[str, int] = {}
    idx, op, on, ot = [], [], [], []
    for f in faces:
        if f in used:
            idx.append(used[f])
            continue
        a, b = f.find("/"), f.rfind("/")
        op.append(P[_stoi(f if a < 0 else f[:a]) - 1])
        if has_uv:
            # str.substr(foundF + 1, foundL): foundL characters from after the first slash.
            ot.append(T[_stoi(f if a < 0 else f[a + 1:a + 1 + b]) - 1])
        if has_n:
            on.append(N[_stoi(f if b < 0 else f[b + 1:]) - 1])
        used[f] = len(op) - 1
        idx.append(len(op) - 1)
    mesh = MeshData(np.array(op, F32), np.array(idx, np.uint32),
                    np.array(on, F32) if has_n else None, None, None, np.array(ot, F32) if has_uv else None)
    return mesh


def compute_normals(mesh: MeshData) -> None:
    """Mesh::computeNormals: sum of cross(normalize(v1 - v0), normalize(v2 - v0)), normalised."""
    P = mesh.positions.astype(F32)
    I = mesh.indices.reshape(-1, 3)
    n = np.zeros_like(P)

    def nz(v):
        return (v / np.sqrt((v * v).sum(axis=1, keepdims=True))).astype(F32)

    d01, d02 = nz(P[I[:, 1]] - P[I[:, 0]]), nz(P[I[:, 2]] - P[I[:, 0]])
    fn = np.cross(d01, d02).astype(F32)
    for k in range(3):  # face order, as the reference accumulates
        np.add.at(n, I[:, k], fn)
    with np.errstate(invalid="ignore", divide="ignore"):
        mesh.normals = nz(n)


def compute_tangents(mesh: MeshData) -> None:
    """Mesh::computeTangentsAndBitangents(force = true) (Mesh.cpp:356-508) through the library's
    rdu_mesh_tangent_frames: MikkTSpace corner frames, then Rendu's vertex split (positions, normals
    and texture coordinates of split vertices are copies). A mesh without texture coordinates gets
    frames from the constant (0.5, 0.5) and keeps no texture coordinates (its objects do not use
    them, Object.cpp:7)."""
    from . import _lib
    P = np.ascontiguousarray(mesh.positions, F32).reshape(-1, 3)
    N = np.ascontiguousarray(mesh.normals, F32).reshape(-1, 3)
    T = None if mesh.texcoords is None else np.ascontiguousarray(mesh.texcoords, F32).reshape(-1, 2)
    I = np.ascontiguousarray(mesh.indices, np.uint32).reshape(-1).copy()
    nv, ni = P.shape[0], I.size
    cap = nv + ni
    src = np.zeros(cap, np.uint32)
    tan = np.zeros((cap, 3), F32)
    bit = np.zeros((cap, 3), F32)
    count = ctypes.c_size_t(0)
    _lib.check(_lib.lib.rdu_mesh_tangent_frames(
        P.ctypes.data, N.ctypes.data, None if T is None else T.ctypes.data, nv, I.ctypes.data, ni, src.ctypes.data,
        tan.ctypes.data, bit.ctypes.data, cap, ctypes.byref(count)))
    n = count.value
    src = src[:n]
    mesh.positions = P[src]
    mesh.normals = N[src]
    if T is not None:
        mesh.texcoords = T[src]
    mesh.indices = I
    mesh.tangents = tan[:n].copy()
    mesh.bitangents = bit[:n].copy()


class Resources:
    """File index of one or more resource roots (Resources::parseDirectory: base names, first wins)."""

    def __init__(self, roots: Sequence[str]):
        self.files: Dict[str, str] = {}
        for root in roots:
            for d, _, names in sorted(os.walk(root)):
                for n in sorted(names):
                    self.files.setdefault(n, os.path.join(d, n))
        self._meshes: Dict[str, MeshData] = {}
        self._textures: Dict[str, object] = {}

    def path(self, name: str) -> Optional[str]:
        return self.files.get(name)

    def image_path(self, name: str) -> Optional[str]:
        for ext in (".png", ".jpg", ".jpeg", ".bmp", ".tga", ".exr"):
            if name + ext in self.files:
                return self.files[name + ext]
        return None

    def mesh(self, name: str) -> MeshData:
        if name not in self._meshes:
            p = self.path(name + ".obj")
            if p is None:
                raise FileNotFoundError(f"Unable to load mesh named {name}.")
            with open(p) as f:
                m = load_obj(f.read())
            if m.normals is None and m.positions.shape[0]:
                compute_normals(m)
            if m.positions.shape[0]:
                compute_tangents(m)
            self._meshes[name] = m
        return self._meshes[name]


def _read_image(path: str, flip: bool) -> np.ndarray:
    if path.lower().endswith(".exr"):
        # Image::loadHDR (Image.cpp:335-420): channels picked by name R, G, B, A (missing colour
        # channels 0, missing alpha 1), half read as float, rows flipped like the LDR images.
        with open(path, "rb") as f:
            chans = read_exr(f.read())
        h, w = next(iter(chans.values())).shape
        out = np.zeros((h, w, 4), F32)
        out[..., 3] = 1.0
        for k, name in enumerate("RGBA"):
            if name in chans:
                out[..., k] = chans[name]
        return out[::-1].copy() if flip else out
    low = path.lower()
    if low.endswith((".jpg", ".jpeg", ".png")):
        px = _read_ldr(path)
    else:
        raise NotImplementedError(f"only PNG, JPEG and uncompressed EXR textures are decoded ({os.path.basename(path)})")
    if px.ndim == 2:
        px = px[:, :, None]
    c = px.shape[2]
    out = np.ones(px.shape[:2] + (4,), F32)
    if c in (1, 2):  # grey (+ alpha): stb expands grey to RGB
        out[..., :3] = px[..., :1] / F32(255.0)
        if c == 2:
            out[..., 3] = px[..., 1] / F32(255.0)
    else:
        out[..., :c] = px[..., :c].astype(F32) / F32(255.0)
    return out[::-1].copy() if flip else out


def _read_ldr(path: str) -> np.ndarray:
    """8-bit pixels (h, w, c) of a PNG or JPEG. Pillow decodes when present (palette PNGs expand
    to RGB, or RGBA with their transparency chunk, as stb_image expands them); without it, the
    library's own reader handles 8-bit non-paletted PNGs."""
    try:
        from PIL import Image as PilImage
    except ImportError:  # pragma: no cover - Pillow is part of the image
        if path.lower().endswith(".png"):
            with open(path, "rb") as f:
                return read_png(f.read())
        raise NotImplementedError(f"JPEG textures need Pillow ({os.path.basename(path)})")
    with PilImage.open(path) as im:
        if im.mode == "P":
            im = im.convert("RGBA" if "transparency" in im.info else "RGB")
        elif im.mode not in ("L", "LA", "RGB", "RGBA"):
            im = im.convert("RGB")
        return np.asarray(im, dtype=np.uint8)


def texture(res: Resources, name: str, cube: bool = False):
    """Resources::getTexture for a 2D texture ((h, w, 4) float32) or a cubemap (6 of them)."""
    key = (name, cube)
    if key in res._textures:
        return res._textures[key]
    if name and all(ch in "0123456789,.-" for ch in name):
        col = [0.0, 0.0, 0.0, 1.0]
        toks = [t for t in name.split(",") if t]
        for i in range(min(4, len(toks))):
            col[i] = float(toks[i])
        img = np.ascontiguousarray(np.broadcast_to(np.array(col, F32), (8, 8, 4)))
        out = [img] * 6 if cube else img
    elif cube:
        faces = None
        for base in (name, name + "_0"):  # the cubemap, else the first level of a mipmapped one
            paths = [res.image_path(base + s) for s in ("_px", "_nx", "_ny", "_py", "_pz", "_nz")]
            if all(f is not None for f in paths):
                faces = paths
                break
        if faces is None:
            raise FileNotFoundError(f"Unable to find cubemap {name}.")
        out = [_read_image(f, True) for f in faces]
    else:
        p = res.image_path(name) or res.image_path(name + "_0")
        if p is None:
            raise FileNotFoundError(f"Unable to find texture {name}.")
        out = _read_image(p, False)
    res._textures[key] = out
    return out


def _texture_of(kv: KeyValues):
    """Codable::decodeTexture: (name, is_cube) of a 'srgb: name' style token."""
    if kv.key not in ("srgb", "rgb", "rgb32", "srgbcube", "rgbcube", "rgb32cube") or not kv.values:
        return "", False
    return ",".join(kv.values), kv.key.endswith("cube")


def load_scene(path: str, roots: Sequence[str]):
    """(Scene, camera factory(width, height)) of a .scene file; meshes and textures from `roots`."""
    with open(path) as f:
        tokens = decode(f.read())
    res = Resources(roots)
    scn = S.Scene()
    tex_ids: Dict[int, int] = {}

    def tex_id(img):
        k = id(img)
        if k not in tex_ids:
            tex_ids[k] = scn.add_texture(img)
        return tex_ids[k]

    objects, materials, lights = [], {}, []
    cam = {"position": np.array([0, 0, 1], F32), "center": np.zeros(3, F32), "up": np.array([0, 1, 0], F32),
           "fov": 1.3}
    for tok in tokens:
        if tok.key == "object":
            mesh_name, mat_name, shadows, skip_uv = None, None, True, False
            for p in tok.elements:
                if p.key == "mesh" and p.values:
                    mesh_name = p.values[0]
                elif p.key == "material" and p.values:
                    mat_name = p.values[0]
                elif p.key == "shadows":
                    shadows = _bool(p)
                elif p.key == "skipuvs":
                    skip_uv = _bool(p)
            if mesh_name is None:
                raise ValueError("Unable to load object (no mesh).")
            mesh = res.mesh(mesh_name)
            objects.append((mesh_name, mesh, mat_name, transformation(tok.elements),
                            not skip_uv and mesh.texcoords is not None))
        elif tok.key == "material":
            mtype, texs, masked, name = 0, [], False, ""
            for p in tok.elements:
                if p.key == "type" and p.values:
                    if p.values[0] not in MATERIAL_TYPES:
                        raise ValueError(f"Unknown material type: {p.values[0]}")
                    mtype = MATERIAL_TYPES[p.values[0]]
                elif p.key == "textures":
                    for t in p.elements:
                        n, cube = _texture_of(t)
                        texs.append(texture(res, n, cube))
                elif p.key == "masked":
                    masked = _bool(p)
                elif p.key == "name" and p.values:
                    name = p.values[0]
            if not name:
                raise ValueError("Material has no name.")
            materials[name] = (mtype, texs, masked)
        elif tok.key in ("point", "directional", "spot"):
            l = S.Light(type={"point": S.LIGHT_POINT, "directional": S.LIGHT_DIRECTIONAL, "spot": S.LIGHT_SPOT}[tok.key],
                        intensity=(1.0, 1.0, 1.0), radius=1.0, angles=(math.pi / 4, math.pi / 2),
                        direction=(0.0, 0.0, 1.0) if tok.key == "directional" else (1.0, 0.0, 0.0))
            for p in tok.elements:
                if p.key == "intensity":
                    l.intensity = tuple(_vec(p, 3))
                elif p.key == "shadows":
                    l.casts_shadow = _bool(p)
                elif p.key == "position":
                    l.position = tuple(_vec(p, 3))
                elif p.key == "direction":
                    d = _vec(p, 3)
                    l.direction = tuple(d / F32(np.sqrt(np.dot(d, d))))
                elif p.key == "radius" and p.values:
                    l.radius = float(p.values[0])
                elif p.key == "cone" and len(p.values) >= 2:
                    l.angles = (0.5 * float(p.values[0]), 0.5 * float(p.values[1]))
            lights.append(l)
        elif tok.key == "background":
            for p in tok.elements:
                if p.key == "color":
                    scn.background = S.Background(S.BG_COLOR, tuple(_vec(p, 3)))
                elif p.key == "image" and p.elements:
                    n, _ = _texture_of(p.elements[0])
                    scn.background = S.Background(S.BG_IMAGE, texture=tex_id(texture(res, n)))
                elif p.key == "cube" and p.elements:
                    n, _ = _texture_of(p.elements[0])
                    faces = texture(res, n, cube=True)
                    scn.background = S.Background(S.BG_SKYBOX, cube=tuple(scn.add_texture(f) for f in faces))
                elif p.key == "sun":
                    # Scene.cpp:191-200 + Sky::decode (Sky.cpp:7-20): the precomputed scattering
                    # table as the background texture, the sun direction normalised (zero refused).
                    sun = _vec(p, 3)
                    if not np.any(sun):
                        raise ValueError("Invalid null sun direction.")
                    sun = sun / np.float32(np.sqrt(np.float32(np.dot(sun, sun))))
                    lut = texture(res, "scattering-precomputed")
                    scn.background = S.Background(S.BG_ATMOSPHERE, texture=tex_id(lut),
                                                  sun_direction=tuple(float(v) for v in sun))
        elif tok.key == "camera":
            for p in tok.elements:
                if p.key in ("position", "center", "up"):
                    cam[p.key] = _vec(p, 3)
                elif p.key == "fov" and p.values:
                    cam["fov"] = float(p.values[0])
    # Materials resolved by name, objects sorted by material type (stable).
    mat_ids: Dict[str, int] = {}
    resolved = []
    for mesh_name, mesh, mat_name, model, use_uv in objects:
        if mat_name not in materials:
            raise ValueError(f"Missing material {mat_name}.")
        resolved.append((materials[mat_name][0], mesh_name, mesh, mat_name, model, use_uv))
    resolved.sort(key=lambda r: r[0])
    mesh_ids: Dict[str, int] = {}
    for mtype, mesh_name, mesh, mat_name, model, use_uv in resolved:
        if mat_name not in mat_ids:
            mt, texs, masked = materials[mat_name]
            ids = [tex_id(t) for t in texs[:3]] + [tex_id(texture(res, "0,0,0,1"))] * max(0, 3 - len(texs))
            mat_ids[mat_name] = scn.add_material(S.Material(mt, tuple(ids), masked))
        if mesh_name not in mesh_ids:
            mesh_ids[mesh_name] = scn.add_mesh(mesh)
        scn.objects.append(S.Object(mesh_ids[mesh_name], mat_ids[mat_name], model, use_uv))
    scn.lights.extend(lights)

    def camera(width: int, height: int) -> Camera:
        c = Camera()
        c.pose(cam["position"], cam["center"], cam["up"])
        c.projection(ratio=width / height, fov=cam["fov"])
        return c

    return scn, camera

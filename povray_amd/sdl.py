"""A reader for the subset of POV-Ray's scene description language that pvx
renders: scene files go straight to a `scenes.Scene` (and from there to
`pvx_scene_create` / `pvx_scene_set_shading`).

It follows the reference parser's reading of each statement it covers
(parser/parser.cpp: Parse_Camera 1926-1969, Parse_Sphere 5715, Parse_Box
1147, Parse_Plane 4979, Parse_Torus 5936, Parse_Cone / Parse_Cylinder
2255-2315, Parse_Disc, Parse_Quadric, Parse_Triangle / Parse_Smooth_Triangle
5539-5590, Parse_Mesh2 4048-4600, Parse_CSG 2197-2245, Parse_Light_Source,
Parse_Object_Mods; parser_expressions.cpp for expressions):

* values: floats, vectors `<a, b, c>`, `x y z`, `pi`, `+ - * /`, unary
  minus, parentheses, `sqrt sin cos tan abs pow radians degrees`,
  identifiers from `#declare` / `#local` (floats, vectors, colours, objects);
* camera (location, look_at, sky, right/up/direction lengths; `angle` sets
  the direction length from the right vector's), light_source (point,
  area_light without jitter), background, global_settings (ambient_light,
  max_trace_level, adc_bailout, radiosity count / recursion_limit /
  error_bound / nearest_count);
* sphere, box, plane, cylinder, cone, disc, torus (sturm), quadric, triangle,
  smooth_triangle, polygon, mesh2 (vertex_vectors, normal_vectors, face_indices),
  union / intersection / difference / merge, object { IDENT };
* object modifiers: pigment / color / rgb / rgbf / rgbt, finish (ambient,
  diffuse, brilliance, phong, phong_size, reflection, emission), interior
  (ior), texture { pigment finish }, translate / rotate / scale / matrix,
  inverse, open, sturm, no_shadow / no_image / no_reflection /
  no_radiosity;
* `#include "colors.inc"` (its common colour names are known); other
  includes, `#version` and unknown directives are errors.

Anything else raises `SdlError` naming the line; nothing is silently
dropped.  Transforms compose as Compose_Transforms does (row vectors); a
sphere takes translations, rotations and uniform scales (Sphere::Translate /
Rotate / Scale without an ellipsoid), meshes and triangles transform their
vertices (Mesh::Transform, Triangle::Transform).
"""
from __future__ import annotations

import math
import re

import numpy as np

from . import _abi as A
from . import scenes

__all__ = ["SdlError", "parse", "parse_file"]


class SdlError(ValueError):
    pass


_TOKEN = re.compile(r"""
    (?P<ws>\s+) | (?P<c1>//[^\n]*) | (?P<c2>/\*.*?\*/) |
    (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?) |
    (?P<str>"[^"\n]*") |
    (?P<dir>\#[A-Za-z_]+) |
    (?P<id>[A-Za-z_][A-Za-z_0-9]*) |
    (?P<sym>[{}<>,()+\-*/=;])
""", re.S | re.X)

# colors.inc (the names scenes use most)
_COLOURS = {
    "White": (1, 1, 1), "Black": (0, 0, 0), "Red": (1, 0, 0), "Green": (0, 1, 0), "Blue": (0, 0, 1),
    "Yellow": (1, 1, 0), "Cyan": (0, 1, 1), "Magenta": (1, 0, 1), "Gray10": (0.1,) * 3, "Gray20": (0.2,) * 3,
    "Gray25": (0.25,) * 3, "Gray30": (0.3,) * 3, "Gray40": (0.4,) * 3, "Gray50": (0.5,) * 3, "Gray60": (0.6,) * 3,
    "Gray70": (0.7,) * 3, "Gray75": (0.75,) * 3, "Gray80": (0.8,) * 3, "Gray90": (0.9,) * 3, "Orange": (1, 0.5, 0),
    "Grey": (0.752941, 0.752941, 0.752941), "Gray": (0.752941, 0.752941, 0.752941),
}

_SHAPES = ("sphere", "box", "plane", "cylinder", "cone", "disc", "torus", "quadric", "triangle",
           "smooth_triangle", "polygon", "mesh2", "union", "intersection", "difference", "merge", "object")


def _identity():
    return np.eye(4), np.eye(4)


def _compose(a, b):  # Compose_Transforms: a then b (row vectors)
    return a[0] @ b[0], b[1] @ a[1]


def _scale(v):
    v = np.asarray(v, np.float64)
    if np.any(v == 0.0):
        raise SdlError("scale by 0")
    m = np.eye(4)
    m[0, 0], m[1, 1], m[2, 2] = v
    inv = np.eye(4)
    inv[0, 0], inv[1, 1], inv[2, 2] = 1.0 / v
    return m, inv


class _Obj:
    """A parsed object before flattening: kind, shape data, transform,
    flags and material."""

    def __init__(self, kind, data, line):
        self.kind = kind
        self.data = data
        self.line = line
        self.tr = None           # (matrix, inverse) or None
        self.flags = 0
        self.inverted = False
        self.children = []
        self.mat = {}

    def copy(self):
        o = _Obj(self.kind, dict(self.data), self.line)
        o.tr = self.tr
        o.flags = self.flags
        o.inverted = self.inverted
        o.children = [c.copy() for c in self.children]
        o.mat = dict(self.mat)
        return o


class _Material(dict):
    """A declared texture / pigment / finish / interior: material fields."""


class _Parser:
    def __init__(self, text: str, width: int, height: int):
        self.toks = []
        line = 1
        pos = 0
        while pos < len(text):
            m = _TOKEN.match(text, pos)
            if not m:
                raise SdlError(f"line {line}: cannot read {text[pos:pos + 20]!r}")
            kind = m.lastgroup
            val = m.group()
            if kind not in ("ws", "c1", "c2"):
                self.toks.append((kind, val, line))
            line += val.count("\n")
            pos = m.end()
        self.i = 0
        self.width, self.height = width, height
        self.decl = {}
        self.objects = []
        self.camera = None
        self.lights = []
        self.settings = dict(max_trace_level=5, adc_bailout=1.0 / 255.0)
        self.background = np.zeros(3)
        self.ambient_light = np.ones(3)
        self.radiosity = None

    # -- token helpers ----------------------------------------------------------------
    def peek(self, k=0):
        j = self.i + k
        return self.toks[j] if j < len(self.toks) else ("eof", "", self.toks[-1][2] if self.toks else 1)

    def next(self):
        t = self.peek()
        self.i += 1
        return t

    def line(self):
        return self.peek()[2]

    def err(self, msg):
        raise SdlError(f"line {self.line()}: {msg}")

    def expect(self, val):
        t = self.next()
        if t[1] != val:
            raise SdlError(f"line {t[2]}: expected {val!r}, found {t[1]!r}")

    def accept(self, val):
        if self.peek()[1] == val:
            self.i += 1
            return True
        return False

    def comma(self):
        self.accept(",")

    # -- expressions --------------------------------------------------------------------
    def expr(self):
        v = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.next()[1]
            w = self.term()
            v = v + w if op == "+" else v - w
        return v

    def term(self):
        v = self.factor()
        while self.peek()[1] in ("*", "/"):
            op = self.next()[1]
            w = self.factor()
            v = v * w if op == "*" else v / w
        return v

    _FUNCS = {"sqrt": np.sqrt, "sin": np.sin, "cos": np.cos, "tan": np.tan, "abs": np.abs,
              "radians": np.radians, "degrees": np.degrees}

    def factor(self):
        kind, val, ln = self.peek()
        if val == "-":
            self.next()
            return -self.factor()
        if val == "+":
            self.next()
            return self.factor()
        if kind == "num":
            self.next()
            return float(val)
        if val == "(":
            self.next()
            v = self.expr()
            self.expect(")")
            return v
        if val == "<":
            self.next()
            comps = [self.expr()]
            while self.accept(","):
                comps.append(self.expr())
            self.expect(">")
            return np.array([float(c) for c in comps])
        if kind == "id":
            self.next()
            if val in ("x", "y", "z"):
                return np.eye(3)["xyz".index(val)].copy()
            if val == "pi":
                return math.pi
            if val in self._FUNCS:
                self.expect("(")
                a = self.expr()
                self.expect(")")
                return self._FUNCS[val](a)
            if val == "pow":
                self.expect("(")
                a = self.expr()
                self.expect(",")
                b = self.expr()
                self.expect(")")
                return a ** b
            if val in ("rgb", "color", "colour"):
                self.i -= 1
                return self.colour()
            if val in self.decl and not isinstance(self.decl[val], (_Obj, _Material)):
                return self.decl[val]
            raise SdlError(f"line {ln}: unknown identifier {val!r}")
        raise SdlError(f"line {ln}: expected a value, found {val!r}")

    def vector(self):
        v = self.expr()
        if np.isscalar(v) or np.ndim(v) == 0:
            return np.full(3, float(v))
        v = np.asarray(v, np.float64)
        if v.shape[0] < 3:
            raise SdlError(f"line {self.line()}: expected a 3-vector")
        return v[:3]

    def flt(self):
        v = self.expr()
        if np.ndim(v) != 0:
            raise SdlError(f"line {self.line()}: expected a float")
        return float(v)

    def colour(self):
        """color [rgb|rgbf|rgbt] <..> or a colour identifier: (rgb, filter, transmit)."""
        self.accept("color") or self.accept("colour")
        kind = self.peek()[1]
        filt = trans = 0.0
        if kind in ("rgb", "rgbf", "rgbt", "rgbft"):
            self.next()
            v = self.expr()
            v = np.full(3, float(v)) if np.ndim(v) == 0 else np.asarray(v, np.float64)
            n = {"rgb": 3, "rgbf": 4, "rgbt": 4, "rgbft": 5}[kind]
            if v.shape[0] < n:
                v = np.concatenate([v, np.zeros(n - v.shape[0])])
            if kind == "rgbf":
                filt = v[3]
            elif kind == "rgbt":
                trans = v[3]
            elif kind == "rgbft":
                filt, trans = v[3], v[4]
            return np.asarray(v[:3]), filt, trans
        t = self.next()
        if t[1] in _COLOURS and t[1] not in self.decl:
            return np.asarray(_COLOURS[t[1]], np.float64), 0.0, 0.0
        if t[1] in self.decl:
            v = self.decl[t[1]]
            if isinstance(v, tuple):
                return v
            return np.asarray(v, np.float64)[:3], 0.0, 0.0
        if t[1] == "<":
            self.i -= 1
            return self.vector(), 0.0, 0.0
        raise SdlError(f"line {t[2]}: expected a colour, found {t[1]!r}")

    # -- blocks ---------------------------------------------------------------------------
    def skip_block_to_end(self):
        self.expect("}")

    def parse(self):
        while self.peek()[0] != "eof":
            kind, val, ln = self.peek()
            if kind == "dir":
                self.directive()
            elif val == "camera":
                self.next()
                self.camera_block()
            elif val == "light_source":
                self.next()
                self.light_block()
            elif val == "background":
                self.next()
                self.expect("{")
                self.background = self.colour()[0]
                self.expect("}")
            elif val == "global_settings":
                self.next()
                self.global_settings()
            elif val in _SHAPES:
                self.objects.append(self.shape())
            elif val == ";":
                self.next()
            else:
                self.err(f"unsupported statement {val!r}")
        return self

    def directive(self):
        _, val, ln = self.next()
        if val in ("#declare", "#local"):
            name = self.next()
            if name[0] != "id":
                raise SdlError(f"line {ln}: expected an identifier after {val}")
            self.expect("=")
            if self.peek()[1] in _SHAPES:
                self.decl[name[1]] = self.shape()
            elif self.peek()[1] in ("texture", "pigment", "finish", "interior"):
                holder = _Obj("material", {}, ln)
                self.material(holder, self.next()[1])
                self.decl[name[1]] = _Material(holder.mat)
            elif self.peek()[1] in ("color", "colour", "rgb", "rgbf", "rgbt", "rgbft"):
                self.decl[name[1]] = self.colour()
            else:
                self.decl[name[1]] = self.expr()
            self.accept(";")
        elif val == "#include":
            f = self.next()
            if f[1].strip('"') != "colors.inc":
                raise SdlError(f"line {ln}: #include {f[1]} is not supported (colors.inc is)")
        elif val == "#version":
            self.expr()
            self.accept(";")
        else:
            raise SdlError(f"line {ln}: unsupported directive {val}")

    def camera_block(self):
        self.expect("{")
        loc, look, sky = np.zeros(3), np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0])
        angle = None
        while not self.accept("}"):
            _, val, ln = self.next()
            if val == "location":
                loc = self.vector()
            elif val == "look_at":
                look = self.vector()
            elif val == "sky":
                sky = self.vector()
            elif val == "angle":
                angle = self.flt()
            elif val in ("right", "up", "direction"):
                self.vector()  # the default aspect / unit lengths are used
            elif val == "perspective":
                pass
            else:
                raise SdlError(f"line {ln}: unsupported camera keyword {val!r}")
        cam = scenes.look_at_camera(loc, look, self.width, self.height, sky)
        if angle is not None:  # direction length from the right vector (parser.cpp, Parse_Camera)
            rl = np.linalg.norm(cam["right"])
            cam["direction"] = cam["direction"] / np.linalg.norm(cam["direction"]) * (
                0.5 * rl / math.tan(math.radians(angle) * 0.5))
        self.camera = cam

    def light_block(self):
        self.expect("{")
        pos = self.vector()
        self.comma()
        col = np.ones(3)
        light = dict(kind="point", position=pos, colour=col)
        while not self.accept("}"):
            kind, val, ln = self.peek()
            if val in ("color", "colour", "rgb") or (kind == "id" and val in _COLOURS):
                light["colour"] = self.colour()[0]
            elif val == "area_light":
                self.next()
                a1 = self.vector()
                self.comma()
                a2 = self.vector()
                self.comma()
                n1 = int(self.flt())
                self.comma()
                n2 = int(self.flt())
                light.update(kind="area", axis1=a1, axis2=a2, size=(n1, n2))
            elif val == "adaptive":
                self.next()
                if self.flt() < 100:
                    raise SdlError(f"line {ln}: adaptive below 100 is not supported (every grid point is traced)")
            elif val in ("jitter", "orient", "circular"):
                raise SdlError(f"line {ln}: area light {val} is not supported")
            else:
                raise SdlError(f"line {ln}: unsupported light_source keyword {val!r}")
        self.lights.append(light)

    def global_settings(self):
        self.expect("{")
        while not self.accept("}"):
            _, val, ln = self.next()
            if val == "ambient_light":
                self.ambient_light = self.colour()[0] if self.peek()[1] in ("color", "colour", "rgb") \
                    else self.vector()
            elif val == "max_trace_level":
                self.settings["max_trace_level"] = int(self.flt())
            elif val == "adc_bailout":
                self.settings["adc_bailout"] = self.flt()
            elif val == "assumed_gamma":
                self.flt()
            elif val == "radiosity":
                self.radiosity = {}
                self.expect("{")
                while not self.accept("}"):
                    _, k, l2 = self.next()
                    if k in ("count", "recursion_limit", "nearest_count"):
                        self.radiosity[k] = int(self.flt())
                    elif k in ("error_bound", "minimum_reuse", "maximum_reuse", "pretrace_start", "pretrace_end"):
                        self.radiosity[k] = self.flt()
                    else:
                        raise SdlError(f"line {l2}: unsupported radiosity keyword {k!r}")
            else:
                raise SdlError(f"line {ln}: unsupported global_settings keyword {val!r}")

    # -- objects ------------------------------------------------------------------------------
    def shape(self) -> _Obj:
        _, kw, ln = self.next()
        self.expect("{")
        if kw == "object":
            name = self.next()
            if name[1] not in self.decl or not isinstance(self.decl[name[1]], _Obj):
                raise SdlError(f"line {ln}: {name[1]!r} is not a declared object")
            o = self.decl[name[1]].copy()
        elif kw == "sphere":
            c = self.vector()
            self.comma()
            o = _Obj("sphere", dict(c=c, r=self.flt()), ln)
        elif kw == "box":
            a = self.vector()
            self.comma()
            b = self.vector()
            o = _Obj("box", dict(c1=np.minimum(a, b), c2=np.maximum(a, b)), ln)
        elif kw == "plane":
            n = self.vector()
            self.comma()
            d = self.flt()
            nl = np.linalg.norm(n)  # Parse_Plane: the normal normalised, Distance = -d
            if nl < 1e-10:
                raise SdlError(f"line {ln}: degenerate plane normal")
            o = _Obj("plane", dict(n=n / nl, d=-d), ln)
        elif kw == "cylinder":
            p1 = self.vector()
            self.comma()
            p2 = self.vector()
            self.comma()
            o = _Obj("cone", dict(apex=p1, base=p2, r=(self.flt(), 0.0), flags=A.PVX_CONE_CYLINDER), ln)
        elif kw == "cone":
            p1 = self.vector()
            self.comma()
            r1 = self.flt()
            self.comma()
            p2 = self.vector()
            self.comma()
            o = _Obj("cone", dict(apex=p1, base=p2, r=(r1, self.flt()), flags=0), ln)
        elif kw == "disc":
            c = self.vector()
            self.comma()
            n = self.vector()
            self.comma()
            r = self.flt()
            hole = 0.0
            if self.accept(","):
                hole = self.flt()
            elif self.peek()[0] == "num":
                hole = self.flt()
            o = _Obj("disc", dict(c=c, n=n, r=(r, hole)), ln)
        elif kw == "torus":
            R = self.flt()
            self.comma()
            o = _Obj("torus", dict(R=R, r=self.flt(), sturm=0), ln)
        elif kw == "quadric":
            a = self.vector()
            self.comma()
            b = self.vector()
            self.comma()
            c = self.vector()
            self.comma()
            o = _Obj("quadric", dict(coef=np.concatenate([a, b, c, [self.flt()]])), ln)
        elif kw == "triangle":
            p = [self.vector()]
            for _ in range(2):
                self.comma()
                p.append(self.vector())
            o = _Obj("triangle", dict(p=np.array(p), n=None), ln)
        elif kw == "smooth_triangle":
            p, n = [], []
            for k in range(3):
                if k:
                    self.comma()
                p.append(self.vector())
                self.comma()
                n.append(self.vector())
            o = _Obj("triangle", dict(p=np.array(p), n=np.array(n)), ln)
        elif kw == "polygon":
            cnt = int(self.flt())
            if cnt < 3:
                raise SdlError(f"line {ln}: polygon needs at least three points")
            pts = []
            for _ in range(cnt):
                self.comma()
                pts.append(self.vector())
            self.accept(",")
            o = _Obj("polygon", dict(p=np.array(pts)), ln)
        elif kw == "mesh2":
            o = self.mesh2(ln)
        else:  # CSG
            o = _Obj("csg", dict(op=kw), ln)
            while self.peek()[1] in _SHAPES:
                o.children.append(self.shape())
            if not o.children:
                raise SdlError(f"line {ln}: {kw} without objects")
        self.mods(o)
        return o

    def mesh2(self, ln):
        verts = norms = faces = None
        while self.peek()[1] in ("vertex_vectors", "normal_vectors", "face_indices"):
            _, k, l2 = self.next()
            self.expect("{")
            cnt = int(self.flt())
            items = []
            for _ in range(cnt):
                self.comma()
                items.append(self.vector())
            self.accept(",")
            self.expect("}")
            arr = np.array(items).reshape(-1, 3)
            if k == "vertex_vectors":
                verts = arr
            elif k == "normal_vectors":
                norms = arr
            else:
                faces = arr.astype(np.int64)
        if verts is None or faces is None:
            raise SdlError(f"line {ln}: mesh2 needs vertex_vectors and face_indices")
        if norms is not None and norms.shape[0] != verts.shape[0]:
            raise SdlError(f"line {ln}: normal_vectors must match vertex_vectors (one normal per vertex)")
        return _Obj("mesh", dict(v=verts, n=norms, f=faces), ln)

    def mods(self, o: _Obj):
        while not self.accept("}"):
            kind, val, ln = self.next()
            if val == "translate":
                v = self.vector()
                self.transform(o, scenes.translation_transform(v))
            elif val == "rotate":
                self.transform(o, scenes.rotation_transform(self.vector()))
            elif val == "scale":
                self.transform(o, _scale(self.vector()))
            elif val == "matrix":
                v = self.expr()
                v = np.asarray(v, np.float64)
                if v.shape != (12,):
                    raise SdlError(f"line {ln}: matrix takes 12 values")
                m = np.eye(4)
                m[:, :3] = v.reshape(4, 3)
                self.transform(o, (m, np.linalg.inv(m)))
            elif val == "inverse":
                o.inverted = not o.inverted
            elif val in ("open", "sturm"):
                if val == "open" and o.kind == "cone":
                    o.data["flags"] |= A.PVX_CONE_OPEN
                elif val == "sturm" and o.kind == "torus":
                    o.data["sturm"] = 1
                else:
                    raise SdlError(f"line {ln}: {val} does not apply to {o.kind}")
            elif val in ("no_shadow", "no_image", "no_reflection", "no_radiosity"):
                o.flags |= {"no_shadow": A.PVX_FLAG_NO_SHADOW, "no_image": A.PVX_FLAG_NO_IMAGE,
                            "no_reflection": A.PVX_FLAG_NO_REFLECTION,
                            "no_radiosity": A.PVX_FLAG_NO_RADIOSITY}[val]
            elif val in ("pigment", "finish", "interior", "texture"):
                self.material(o, val)
            elif val in ("color", "colour", "rgb"):
                self.i -= 1
                self.set_colour(o)
            else:
                raise SdlError(f"line {ln}: unsupported object modifier {val!r}")

    def set_colour(self, o):
        rgb, filt, trans = self.colour()
        o.mat["pigment"] = rgb
        o.mat["filter"] = filt
        o.mat["transmit"] = trans

    def material(self, o, block):
        self.expect("{")
        while not self.accept("}"):
            kind, val, ln = self.peek()
            if block == "texture" and val in ("pigment", "finish"):
                self.next()
                self.material(o, val)
            elif kind == "id" and isinstance(self.decl.get(val), _Material):
                self.next()
                o.mat.update(self.decl[val])
            elif block == "pigment" and (val in ("color", "colour", "rgb", "rgbf", "rgbt", "rgbft")
                                         or val in _COLOURS or val in self.decl):
                self.set_colour(o)
            elif block == "finish" and val in ("ambient", "diffuse", "brilliance", "phong", "phong_size",
                                               "reflection", "emission"):
                self.next()
                if val == "reflection" and self.accept("{"):
                    r = self.flt()
                    self.expect("}")
                else:
                    r = self.expr()
                o.mat[val] = r
            elif block == "interior" and val == "ior":
                self.next()
                o.mat["ior"] = self.flt()
            else:
                raise SdlError(f"line {ln}: unsupported {block} keyword {val!r}")

    def transform(self, o, tr):
        if o.kind == "csg":
            for c in o.children:
                self.transform(c, tr)
            return
        if o.kind == "sphere":  # Sphere::Translate / Rotate / Scale (no ellipsoid)
            m = tr[0]
            lin = m[:3, :3]
            s = np.linalg.norm(lin[0])
            if not np.allclose(lin @ lin.T, s * s * np.eye(3), rtol=1e-12, atol=1e-12):
                raise SdlError(f"line {o.line}: a sphere takes only translate, rotate and uniform scale")
            o.data["c"] = o.data["c"] @ lin + m[3, :3]
            o.data["r"] = o.data["r"] * s
            return
        if o.kind == "plane":  # Plane::Transform: normal by MTransNormal, a point on the plane moved
            n, d = o.data["n"], o.data["d"]
            p = -d * n
            p2 = p @ tr[0][:3, :3] + tr[0][3, :3]
            n2 = tr[1][:3, :3] @ n
            n2 = n2 / np.linalg.norm(n2)
            o.data["n"], o.data["d"] = n2, -float(n2 @ p2)
            return
        if o.kind in ("mesh", "triangle"):  # vertices (and normals) transformed
            o.data["p" if o.kind == "triangle" else "v"] = \
                o.data["p" if o.kind == "triangle" else "v"] @ tr[0][:3, :3] + tr[0][3, :3]
            key = "n"
            if o.data.get(key) is not None:
                n2 = o.data[key] @ tr[1][:3, :3].T
                o.data[key] = n2 / np.linalg.norm(n2, axis=1, keepdims=True)
            return
        o.tr = tr if o.tr is None else _compose(o.tr, tr)

    # -- flattening into a Scene ------------------------------------------------------------
    def scene(self) -> scenes.Scene:
        if self.camera is None:
            self.camera = scenes.look_at_camera((0, 0, 0), (0, 0, 1), self.width, self.height)
        arrays = dict(sph_c=[], sph_r=[], pl_n=[], pl_d=[], box_c1=[], box_c2=[], box_tr=[], tor_R=[], tor_r=[],
                      tor_tr=[], tor_s=[], xf=[], cone_a=[], cone_b=[], cone_r=[], cone_f=[], cone_t=[],
                      disc_c=[], disc_n=[], disc_r=[], disc_t=[], q=[], q_t=[], tri_p=[], tri_n=[], meshes=[],
                      poly_first=[0], poly_pts=[], poly_t=[],
                      csg=[], kids=[])

        def xf_index(o):
            if o.tr is None:
                return -1
            arrays["xf"].append(np.concatenate([o.tr[0].ravel(), o.tr[1].ravel()]))
            return len(arrays["xf"]) - 1

        def leaf(o, flags, top):
            f = flags | (A.PVX_FLAG_INVERTED if o.inverted else 0)
            k, dd = o.kind, o.data
            if k == "sphere":
                arrays["sph_c"].append(dd["c"])
                arrays["sph_r"].append(dd["r"])
                return [A.PVX_SPHERE, len(arrays["sph_r"]) - 1, f, 0]
            if k == "plane":
                arrays["pl_n"].append(dd["n"])
                arrays["pl_d"].append(dd["d"])
                return [A.PVX_PLANE, len(arrays["pl_d"]) - 1, f, 0]
            if k == "box":
                arrays["box_c1"].append(dd["c1"])
                arrays["box_c2"].append(dd["c2"])
                arrays["box_tr"].append(xf_index(o))
                return [A.PVX_BOX, len(arrays["box_c1"]) - 1, f, 0]
            if k == "torus":
                arrays["tor_R"].append(dd["R"])
                arrays["tor_r"].append(dd["r"])
                arrays["tor_s"].append(dd["sturm"])
                arrays["tor_tr"].append(xf_index(o))
                return [A.PVX_TORUS, len(arrays["tor_R"]) - 1, f, 0]
            if k == "cone":
                arrays["cone_a"].append(dd["apex"])
                arrays["cone_b"].append(dd["base"])
                arrays["cone_r"].append(dd["r"])
                arrays["cone_f"].append(dd["flags"])
                arrays["cone_t"].append(xf_index(o))
                return [A.PVX_CONE, len(arrays["cone_a"]) - 1, f, 0]
            if k == "disc":
                arrays["disc_c"].append(dd["c"])
                arrays["disc_n"].append(dd["n"])
                arrays["disc_r"].append(dd["r"])
                arrays["disc_t"].append(xf_index(o))
                return [A.PVX_DISC, len(arrays["disc_c"]) - 1, f, 0]
            if k == "quadric":
                arrays["q"].append(dd["coef"])
                arrays["q_t"].append(xf_index(o))
                return [A.PVX_QUADRIC, len(arrays["q"]) - 1, f, 0]
            if not top:
                raise SdlError(f"line {o.line}: {k} cannot be a CSG operand")
            if k == "triangle":
                arrays["tri_p"].append(dd["p"].ravel())
                arrays["tri_n"].append(dd["n"].ravel() if dd["n"] is not None else np.zeros(9))
                return [A.PVX_SMOOTH_TRIANGLE if dd["n"] is not None else A.PVX_TRIANGLE,
                        len(arrays["tri_p"]) - 1, f, 0]
            if k == "polygon":
                arrays["poly_pts"].extend(dd["p"])
                arrays["poly_first"].append(len(arrays["poly_pts"]))
                arrays["poly_t"].append(xf_index(o))
                return [A.PVX_POLYGON, len(arrays["poly_t"]) - 1, f, 0]
            if k == "mesh":
                arrays["meshes"].append((dd["v"].astype(np.float32), dd["f"].astype(np.int32),
                                         None if dd["n"] is None else dd["n"].astype(np.float32)))
                return [A.PVX_MESH, len(arrays["meshes"]) - 1, f, 0]
            raise SdlError(f"line {o.line}: unknown object kind {k}")

        def csg(o, flags):
            """CSG node index; a difference becomes an intersection whose
            operands 2..n are inverted (parser.cpp:2223-2225); an inverted CSG
            carries Invert's operator swap and child inversion (csg.cpp:629-645)."""
            op = o.data["op"]
            kids = list(o.children)
            inv_kids = [False] * len(kids)
            if op == "difference":
                op = "intersection"
                inv_kids = [False] + [True] * (len(kids) - 1)
            if o.inverted:
                op = {"union": "intersection", "intersection": "merge", "merge": "intersection"}[op]
                inv_kids = [not v for v in inv_kids]
            idx = len(arrays["csg"])
            arrays["csg"].append(None)
            first = len(arrays["kids"])
            arrays["kids"].extend([None] * len(kids))
            for j, (c, iv) in enumerate(zip(kids, inv_kids)):
                c = c.copy()
                if iv:
                    c.inverted = not c.inverted
                if c.kind == "csg":
                    arrays["kids"][first + j] = [A.PVX_CSG, csg(c, 0), 0, 0]
                else:
                    arrays["kids"][first + j] = leaf(c, c.flags, False)
            code = {"union": A.PVX_CSG_UNION, "intersection": A.PVX_CSG_INTERSECTION, "merge": A.PVX_CSG_MERGE}[op]
            arrays["csg"][idx] = [code, first, len(kids), 0]
            return idx

        objs, mats = [], []
        for o in self.objects:
            if o.kind == "csg":
                objs.append([A.PVX_CSG, csg(o, 0), o.flags, 0])
            else:
                objs.append(leaf(o, o.flags, True))
            mats.append(o.mat)
        n = len(objs)

        def matarr(key, default, shape=()):
            out = np.empty((n,) + shape)
            for i, m in enumerate(mats):
                v = m.get(key, default)
                out[i] = np.full(shape, float(v)) if (shape and np.ndim(v) == 0) else v
            return out

        materials = dict(pigment=matarr("pigment", np.zeros(3), (3,)), diffuse=matarr("diffuse", 0.6),
                         ambient=matarr("ambient", 0.1), phong=matarr("phong", 0.0),
                         phong_size=matarr("phong_size", 40.0), reflection=matarr("reflection", 0.0),
                         ior=matarr("ior", 1.0), transmit=matarr("transmit", 0.0), filter=matarr("filter", 0.0),
                         emission=matarr("emission", np.zeros(3), (3,)), brilliance=matarr("brilliance", 1.0))

        def arr(key, shape, dtype=np.float64):
            v = arrays[key]
            return np.array(v, dtype).reshape(shape) if v else None

        settings = dict(self.settings)
        settings["background"] = np.asarray(self.background, np.float64)
        settings["ambient_light"] = np.asarray(self.ambient_light, np.float64)
        if self.radiosity is not None:
            settings["radiosity"] = dict(self.radiosity)
        return scenes.Scene(
            name="sdl", objects=np.array(objs, np.uint32).reshape(-1, 4),
            sphere_center=arr("sph_c", (-1, 3)), sphere_radius=arr("sph_r", (-1,)),
            plane_normal=arr("pl_n", (-1, 3)), plane_distance=arr("pl_d", (-1,)),
            box_corner1=arr("box_c1", (-1, 3)), box_corner2=arr("box_c2", (-1, 3)),
            box_transform=arr("box_tr", (-1,), np.int32),
            torus_major=arr("tor_R", (-1,)), torus_minor=arr("tor_r", (-1,)),
            torus_transform=arr("tor_tr", (-1,), np.int32), torus_sturm=arr("tor_s", (-1,), np.uint8),
            transforms=arr("xf", (-1, 32)), meshes=arrays["meshes"],
            csg=arr("csg", (-1, 4), np.uint32), csg_children=arr("kids", (-1, 4), np.uint32),
            triangle_vertices=arr("tri_p", (-1, 9)), triangle_normals=arr("tri_n", (-1, 9)),
            cone_apex=arr("cone_a", (-1, 3)), cone_base=arr("cone_b", (-1, 3)), cone_radius=arr("cone_r", (-1, 2)),
            cone_flags=arr("cone_f", (-1,), np.uint8), cone_transform=arr("cone_t", (-1,), np.int32),
            disc_center=arr("disc_c", (-1, 3)), disc_normal=arr("disc_n", (-1, 3)),
            disc_radius=arr("disc_r", (-1, 2)), disc_transform=arr("disc_t", (-1,), np.int32),
            quadric_coef=arr("q", (-1, 10)), quadric_transform=arr("q_t", (-1,), np.int32),
            polygon_first=np.array(arrays["poly_first"], np.uint32) if arrays["poly_t"] else None,
            polygon_points=arr("poly_pts", (-1, 3)), polygon_transform=arr("poly_t", (-1,), np.int32),
            camera=self.camera, lights=self.lights, materials=materials, settings=settings)


def parse(text: str, width: int = 320, height: int = 240) -> scenes.Scene:
    """Read scene-description text into a Scene rendered at width x height."""
    return _Parser(text, width, height).parse().scene()


def parse_file(path, width: int = 320, height: int = 240) -> scenes.Scene:
    with open(path, encoding="utf-8") as f:
        return parse(f.read(), width, height)

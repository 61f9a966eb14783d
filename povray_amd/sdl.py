"""A reader for the subset of POV-Ray's scene description language that pvx
renders: scene files go straight to a `scenes.Scene` (and from there to
`pvx_scene_create` / `pvx_scene_set_shading`).

It follows the reference parser's reading of each statement it covers
(parser/parser.cpp: Parse_Camera 1926-1969, Parse_Sphere 5715, Parse_Box
1147, Parse_Plane 4979, Parse_Torus 5936, Parse_Cone / Parse_Cylinder
2255-2315, Parse_Disc, Parse_Quadric, Parse_Triangle / Parse_Smooth_Triangle
5539-5590, Parse_Mesh2 4048-4600, Parse_CSG 2197-2245, Parse_Light_Source,
Parse_Object_Mods; parser_expressions.cpp for expressions):

* the directive layer (Parse_Directive, parser_tokenizer.cpp:952):
  #declare / #local (a macro call opens a local scope), #undef, #while,
  #if / #elseif / #else, #ifdef / #ifndef, #macro with arguments, #default,
  #version, #debug / #warning / #error, #include (colors.inc's common names
  are built in; other include files are skipped with a warning);
* values: floats, vectors `<a, b, c>` (commas optional), `x y z`, `pi`,
  `+ - * /`, comparisons, `& | !`, `?:`, member access (`.x`, `.red`),
  `sqrt sin cos tan abs pow exp log ln floor ceil int mod div min max atan2
  vlength vnormalize vdot vcross select`, `rand` / `seed` as POV_rand
  (core/math/randomsequence.cpp:113-118), `image_width` / `image_height`,
  identifiers (floats, vectors, colours, objects, materials, lights);
* camera (Parse_Camera for v3.5+: location, look_at, sky, right / up /
  direction, angle, transforms), light_source (point, area_light without
  jitter), background, global_settings (ambient_light, max_trace_level,
  adc_bailout, radiosity count / recursion_limit / error_bound /
  nearest_count);
* sphere, box, plane, cylinder, cone, disc, torus (sturm), quadric, triangle,
  smooth_triangle, polygon, mesh2 (vertex_vectors, normal_vectors,
  face_indices; uv and texture lists read and dropped), lathe (linear,
  quadratic, cubic and bezier splines), sor (open),
  union / intersection / difference / merge, object { IDENT };
* object modifiers: pigment / color / rgb / rgbf / rgbt, finish (ambient,
  diffuse, brilliance, phong, phong_size, reflection, emission), interior
  (ior), texture { pigment finish }, translate / rotate / scale / matrix,
  inverse, open, sturm, no_shadow / no_image / no_reflection /
  no_radiosity.

What pvx does not render is reduced with a warning, collected in the
scene's `settings["warnings"]`: a pattern pigment becomes the mean of its
colour map, normals and finish extras (specular, roughness, metallic,
irid ...) are dropped, as are photons and special camera and light options.
A top-level union that holds triangles, meshes or polygons is split into its
operands as Link_To_Frame splits unions (parser.cpp:9425-9440).

Anything else raises `SdlError` naming the line; nothing else is silently
dropped.  Transforms compose as Compose_Transforms does (row vectors); a
sphere takes translations, rotations and uniform scales (Sphere::Translate /
Rotate / Scale without an ellipsoid), triangles transform their vertices
(Triangle::Transform) and a mesh keeps its FP32 vertices in its own space
with the composed transform as its placement (Mesh::Transform).  A mesh's
object box is then the mesh box through that one composed transform, where
the reference recomputes it once per modifier (a looser box that only culls
differently at grazing rays).
"""
from __future__ import annotations

import collections
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
    (?P<str>"(?:[^"\\\n]|\\.)*") |
    (?P<dir>\#[ \t]*[A-Za-z_]+) |
    (?P<id>[A-Za-z_][A-Za-z_0-9]*) |
    (?P<sym2><=|>=|!=) |
    (?P<sym>[{}<>,()+\-*/=;.!&|?:\[\]])
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
           "smooth_triangle", "polygon", "mesh2", "lathe", "sor", "union", "intersection", "difference", "merge",
           "object")
# the reference's other shapes: read and dropped with a warning when the
# reader is asked to skip unsupported shapes, else an error
_OTHER_SHAPES = ("blob", "superellipsoid", "julia_fractal", "height_field", "text", "prism", "bicubic_patch",
                 "isosurface", "parametric", "sphere_sweep", "lemon", "ovus", "poly", "polynomial", "cubic",
                 "quartic", "light_group", "mesh")
_LATHE_SPLINES = {"linear_spline": A.PVX_LATHE_LINEAR, "quadratic_spline": A.PVX_LATHE_QUADRATIC,
                  "cubic_spline": A.PVX_LATHE_CUBIC, "bezier_spline": A.PVX_LATHE_BEZIER}


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


class _Function:
    """A declared user function: parameter names and the body's tokens."""

    def __init__(self, params, start, end):
        self.params, self.start, self.end = params, start, end

    def copy(self):
        return self


class _ColourMap:
    """A declared color_map: its entries' colours."""

    def __init__(self, cols):
        self.cols = cols

    def copy(self):
        return self


class _Xf:
    """A declared transform: (matrix, inverse)."""

    def __init__(self, tr):
        self.tr = tr

    def copy(self):
        return self


class _Array:
    """A declared array: nested lists of values (None: not yet set)."""

    def __init__(self, dims):
        self.dims = dims

        def make(level):
            return [make(level + 1) if level + 1 < len(dims) else None for _ in range(dims[level])]
        self.data = make(0)

    def fill(self, values, ln):
        def put(dst, src, level):
            if len(src) > self.dims[level]:
                raise SdlError(f"line {ln}: too many array elements")
            for k, v in enumerate(src):
                if level + 1 < len(self.dims):
                    put(dst[k], v, level + 1)
                else:
                    dst[k] = v
        put(self.data, values, 0)

    def _cell(self, idx, ln):
        if len(idx) != len(self.dims) or any(not 0 <= i < n for i, n in zip(idx, self.dims)):
            raise SdlError(f"line {ln}: array index {idx} out of range {self.dims}")
        d = self.data
        for i in idx[:-1]:
            d = d[i]
        return d

    def get(self, idx, ln):
        v = self._cell(idx, ln)[idx[-1]]
        if v is None:
            raise SdlError(f"line {ln}: array element {idx} not initialised")
        return v

    def set(self, idx, value, ln):
        self._cell(idx, ln)[idx[-1]] = value

    def copy(self):
        return self


class _Opaque:
    """A declared block pvx does not render (maps, media, sky spheres ...)."""

    def __init__(self, kind):
        self.kind = kind

    def copy(self):
        return self


class _Light(dict):
    """A declared light_source."""

    def copy(self):
        return _Light({k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.items()})


def _promote(a, b):
    """Operands of + - * /: vectors of different lengths (a colour's five
    components and a 3-vector) are padded with zeros to the longer one."""
    if np.ndim(a) and np.ndim(b) and np.shape(a)[0] != np.shape(b)[0]:
        n = max(np.shape(a)[0], np.shape(b)[0])
        a = np.concatenate([a, np.zeros(n - np.shape(a)[0])])
        b = np.concatenate([b, np.zeros(n - np.shape(b)[0])])
    return a, b


def _truth(v):
    return bool(np.any(np.asarray(v) != 0.0))


def _tokenize(text: str, tag: str = ""):
    """Tokens (kind, value, line) of scene text; `tag` prefixes the line
    numbers of an included file ("colors.inc:12")."""
    toks = []
    line = 1
    pos = 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise SdlError(f"line {tag}{line}: cannot read {text[pos:pos + 20]!r}")
        kind = m.lastgroup
        val = m.group()
        if kind == "sym2":
            kind = "sym"
        if kind not in ("ws", "c1", "c2"):
            tok = "#" + val[1:].strip() if kind == "dir" else val
            toks.append((kind, tok, f"{tag}{line}" if tag else line))
        line += val.count("\n")
        pos = m.end()
    return toks


def _read_text(path):
    """Scene text: UTF-8, else Latin-1 (older scene files)."""
    with open(path, "rb") as f:
        raw = f.read()
    try:
        return raw.decode("utf-8")
    except UnicodeDecodeError:
        return raw.decode("latin-1")


class _Parser:
    def __init__(self, text: str, width: int, height: int, include_dirs=(), skip_unsupported=False):
        self.toks = _tokenize(text)
        self.include_dirs = list(include_dirs)  # the scene's folder, then library paths (+L)
        self.include_depth = 0
        self.version = 3.8
        self.skip_unsupported = skip_unsupported
        self.i = 0
        self.width, self.height = width, height
        # identifiers: #declare writes the global map, #local the innermost
        # (a macro call opens one)
        self.decl = collections.ChainMap({})
        self.ctrl = []          # open #while / #if / macro frames
        self.no_pump = 0        # > 0 while a directive's own expression is read
        self.macros = {}        # name -> (parameter names, body start index)
        self.rand_state = []    # seed() streams (Parser::stream_seed)
        self.defaults = {}      # #default material fields
        self.warnings = []      # what the reader reduced (textures, special cameras, light options)
        self.objects = []
        self.camera = None
        self.lights = []
        self.settings = dict(max_trace_level=5, adc_bailout=1.0 / 255.0)
        self.background = np.zeros(3)
        self.ambient_light = np.ones(3)
        self.radiosity = None

    # -- token helpers ----------------------------------------------------------------
    def raw(self):
        """The token at the cursor without directive processing."""
        j = self.i
        return self.toks[j] if j < len(self.toks) else ("eof", "", self.toks[-1][2] if self.toks else 1)

    def peek(self):
        if not self.no_pump:
            self.pump()
        return self.raw()

    def peek_at(self, k):
        """The raw token k places after the cursor."""
        j = self.i + k
        return self.toks[j] if j < len(self.toks) else ("eof", "", 0)

    def next(self):
        t = self.peek()
        self.i += 1
        return t

    def warn(self, msg):
        self.warnings.append(msg)

    # -- the directive layer (Parser::Parse_Directive, parser_tokenizer.cpp:952):
    # directives and macro calls are handled as the tokens arrive, so they may
    # appear between statements and inside blocks alike --------------------------------
    def pump(self):
        while True:
            kind, val, ln = self.raw()
            if val == "\x00end_include":
                self.i += 1
                self.include_depth -= 1
            elif kind == "dir":
                self.i += 1
                self.directive(val, ln)
            elif kind == "id" and val in self.macros:
                self.i += 1
                self.call_macro(val, ln)
            else:
                return

    _OPENERS = ("#while", "#if", "#ifdef", "#ifndef", "#macro", "#switch")

    def skip_block(self, stop_at_else, stops=("#else", "#elseif")):
        """Skip tokens to the #end (or, when stop_at_else, one of `stops`)
        closing the current block at this nesting depth; returns the directive
        that stopped the skip (consumed)."""
        depth = 0
        while self.i < len(self.toks):
            kind, val, ln = self.toks[self.i]
            self.i += 1
            if val == "\x00end_include":
                self.include_depth -= 1
                continue
            if kind != "dir":
                continue
            if val in self._OPENERS:
                depth += 1
            elif val == "#end":
                if depth == 0:
                    return val
                depth -= 1
            elif depth == 0 and stop_at_else and val in stops:
                return val
        raise SdlError("unterminated #while / #if / #macro block")

    def cond_paren_value(self):
        return self.quiet(self.cond)

    def switch_to_case(self, v, ln):
        """Skip the clauses of a #switch until one holds (#case: |v - c| <
        EPSILON, #range: lo <= v <= hi) or #else; then run it."""
        while True:
            stop = self.skip_block(True, ("#case", "#range", "#else"))
            if stop == "#end":
                return
            if stop == "#else":
                self.ctrl.append(("switch", v))
                return
            if stop == "#case":
                ok = abs(float(self.quiet(self.cond)) - v) < 1e-10
            else:
                lo, hi = self.quiet(self.range_pair)
                ok = lo <= v <= hi
            if ok:
                self.ctrl.append(("switch", v))
                return

    def range_pair(self):
        """#range's ( low, high )."""
        self.expect("(")
        lo = float(self.cond())
        self.expect(",")
        hi = float(self.cond())
        self.expect(")")
        return lo, hi

    def cond_paren(self):
        """A directive's condition: read without processing the directives
        that follow it (its lookahead must not run them early)."""
        return _truth(self.quiet(self.cond))

    def quiet(self, fn):
        self.no_pump += 1
        try:
            return fn()
        finally:
            self.no_pump -= 1

    def directive(self, val, ln):
        if val in ("#declare", "#local"):
            if self.raw()[1] == "deprecated":  # deprecated [once] "message" NAME = ...
                self.i += 1
                if self.raw()[1] == "once":
                    self.i += 1
                if self.raw()[0] == "str":
                    self.i += 1
            kind, name, l2 = self.raw()
            if kind != "id":
                raise SdlError(f"line {ln}: expected an identifier after {val}")
            self.i += 1
            idx = []
            while self.raw()[1] == "[":  # an array element: #declare A[i][j] = ...
                self.i += 1
                idx.append(int(float(self.quiet(self.cond))))
                self.quiet(lambda: self.expect("]"))
            self.expect("=")
            value = self.decl_value(ln)
            if idx:
                arr = self.decl.get(name)
                if not isinstance(arr, _Array):
                    raise SdlError(f"line {ln}: {name!r} is not an array")
                arr.set(idx, value, ln)
                self.accept_raw(";")
                return
            if val == "#local":
                self.decl[name] = value
            else:  # an existing local is changed, else the global
                for m in self.decl.maps:
                    if name in m:
                        m[name] = value
                        break
                else:
                    self.decl.maps[-1][name] = value
            self.accept_raw(";")
        elif val == "#undef":
            name = self.raw()[1]
            self.i += 1
            for m in self.decl.maps:
                m.pop(name, None)
        elif val == "#while":
            start = self.i - 1
            if self.cond_paren():
                self.ctrl.append(("while", start))
            else:
                self.skip_block(False)
        elif val in ("#if", "#ifdef", "#ifndef"):
            if val == "#if":
                ok = self.cond_paren()
            else:
                self.expect("(")
                name = self.raw()[1]
                self.i += 1
                self.expect(")")
                ok = (name in self.decl) == (val == "#ifdef")
            while not ok:
                stop = self.skip_block(True)
                if stop == "#end":
                    return
                ok = True if stop == "#else" else self.cond_paren()
            self.ctrl.append(("if", None))
        elif val in ("#else", "#elseif"):  # the branch taken ends here
            if not self.ctrl or self.ctrl[-1][0] not in ("if", "switch"):
                raise SdlError(f"line {ln}: {val} without #if")
            self.ctrl.pop()
            self.skip_block(False)
        elif val == "#switch":  # the first #case / #range that holds, falling through to #break
            v = float(self.cond_paren_value())
            self.switch_to_case(v, ln)
        elif val in ("#case", "#range"):  # met while a clause runs: fall through
            if not self.ctrl or self.ctrl[-1][0] != "switch":
                raise SdlError(f"line {ln}: {val} without #switch")
            self.quiet(self.cond if val == "#case" else self.range_pair)
        elif val == "#break":
            if not self.ctrl or self.ctrl[-1][0] != "switch":
                raise SdlError(f"line {ln}: #break outside a #switch")
            self.ctrl.pop()
            self.skip_block(False)
        elif val == "#end":
            if not self.ctrl:
                raise SdlError(f"line {ln}: #end without an open block")
            what, data = self.ctrl.pop()
            if what == "while":
                self.i = data  # back to the #while: its condition is read again
            elif what == "macro":
                self.i = data
                self.decl = self.decl.parents
        elif val == "#macro":
            kind, name, l2 = self.raw()
            self.i += 1
            self.expect("(")
            params = []
            while not self.accept(")"):
                params.append(self.raw()[1])
                self.i += 1
                self.accept(",")
            self.macros[name] = (params, self.i)
            self.skip_block(False)
        elif val == "#include":
            f = self.quiet(self.next)[1].strip('"')
            self.include(f, ln)
        elif val == "#version":
            self.version = float(self.quiet(self.cond))
            self.accept_raw(";")
        elif val == "#default":
            self.default_block(ln)
        elif val in ("#debug", "#warning"):
            self.next()
        elif val == "#error":
            raise SdlError(f"line {ln}: #error {self.next()[1]}")
        else:
            raise SdlError(f"line {ln}: unsupported directive {val}")

    def include(self, name, ln):
        """#include: the file's tokens are read in place of the directive
        (the scene's folder first, then the library paths).  A file that is
        not found is skipped with a warning (the common colours of colors.inc
        are built in)."""
        import os
        for d in self.include_dirs:
            path = os.path.join(d, name)
            if os.path.isfile(path):
                break
        else:
            self.warn(f"line {ln}: #include {name!r} not found, skipped")
            return
        if self.include_depth >= 32:
            raise SdlError(f"line {ln}: #include nested too deeply")
        new = _tokenize(_read_text(path), os.path.basename(name) + ":")
        new.append(("sym", "\x00end_include", ln))
        pos, k = self.i, len(new)
        self.toks[pos:pos] = new
        # indices kept by open blocks and macros move with the tokens after them
        self.ctrl = [(w, (d + k if isinstance(d, int) and d >= pos else d)) for w, d in self.ctrl]
        self.macros = {m: (p, b + k if b >= pos else b) for m, (p, b) in self.macros.items()}
        self.include_depth += 1

    def default_block(self, ln):
        """#default { texture / pigment / finish / normal / interior }."""
        holder = _Obj("material", {}, ln)
        self.expect("{")
        while not self.accept("}"):
            k = self.next()[1]
            if k not in ("texture", "pigment", "finish", "normal", "interior"):
                raise SdlError(f"line {ln}: unsupported #default block {k!r}")
            self.material(holder, k)
        self.defaults.update(holder.mat)

    def texture_layers(self, ln, raw=False):
        """Further texture blocks right after a texture (a layered texture):
        read and dropped with a warning; the bottom layer is kept.  `raw`:
        look ahead without running directives (after a #declare's value)."""
        dropped = False
        while (self.raw() if raw else self.peek())[1] == "texture":
            self.next()
            self.material(_Obj("material", {}, ln), "texture")
            dropped = True
        if dropped:
            self.warn(f"line {ln}: layered texture reduced to its bottom layer")

    def call_macro(self, name, ln):
        params, body = self.macros[name]
        self.expect("(")
        args = []
        while not self.accept(")"):
            args.append(self.decl_value(ln))
            self.accept(",")
        if len(args) != len(params):
            raise SdlError(f"line {ln}: macro {name} takes {len(params)} arguments")
        self.ctrl.append(("macro", self.i))
        self.decl = self.decl.new_child(dict(zip(params, args)))
        self.i = body

    def decl_value(self, ln):
        """The right-hand side of #declare / #local, or a macro argument."""
        kind, val, _ = self.peek()
        if val == "function":
            return self.function_decl(ln)
        if val == "array":
            return self.array_decl(ln)
        if kind == "id" and isinstance(self.decl.get(val), _Array) and self.peek_at(1)[1] == "[":
            self.next()
            v = self.array_element(self.decl[val], ln)
            return v.copy() if isinstance(v, _Obj) else v
        if val in _OTHER_SHAPES:
            return self.shape()
        if val in ("color_map", "colour_map"):
            self.next()
            return _ColourMap(self.colour_map_entries())
        if val in self._OPAQUE_BLOCKS:
            self.next()
            self.skip_braces()
            return _Opaque(val)
        if val in _SHAPES:
            return self.shape()
        if val == "light_source":
            self.next()
            return _Light(self.light_block())
        if val == "transform":
            self.next()
            if self.peek()[0] == "id" and isinstance(self.decl.get(self.peek()[1]), _Xf):
                return self.decl[self.next()[1]]
            return _Xf(self.transform_block(ln))
        if val == "material":
            holder = _Obj("material", {}, ln)
            self.next()
            self.expect("{")
            while not self.accept("}"):
                k = self.next()[1]
                if k in ("texture", "pigment", "finish", "normal", "interior"):
                    self.material(holder, k)
                elif isinstance(self.decl.get(k), _Material):
                    holder.mat.update(self.decl[k])
                elif self.transform_mod(k, ln) is None:
                    raise SdlError(f"line {ln}: unsupported material keyword {k!r}")
            return _Material(holder.mat)
        if val in ("texture", "pigment", "finish", "interior", "normal"):
            holder = _Obj("material", {}, ln)
            self.material(holder, self.next()[1])
            if val == "texture":
                self.texture_layers(ln, raw=True)
            return _Material(holder.mat)
        if val in ("color", "colour", "rgb", "rgbf", "rgbt", "rgbft"):
            rgb, f, t = self.quiet(self.colour)
            return np.concatenate([np.asarray(rgb, np.float64), [f, t]])
        if kind == "id" and isinstance(self.decl.get(val), (_Obj, _Material, _Light, _Function, _ColourMap, _Opaque, _Xf, _Array)) and \
                not (isinstance(self.decl.get(val), _Function) and self.peek_at(1)[1] == "("):
            self.next()
            v = self.decl[val]
            return v.copy() if isinstance(v, _Obj) else v
        return self.quiet(self.cond)

    def function_decl(self, ln):
        """function [(params)] { expression }: kept as its parameter names and
        the body's token range, evaluated when called (parameters x, y, z
        when none are listed).  Pattern and object functions are kept but
        cannot be called."""
        self.next()
        params = ["x", "y", "z"]
        if self.accept("("):
            params = []
            while not self.accept(")"):
                params.append(self.raw()[1])
                self.i += 1
                self.accept(",")
        if self.peek()[1] != "{":
            raise SdlError(f"line {ln}: expected a function body")
        start = self.i + 1
        self.skip_braces()
        return _Function(params, start, self.i - 1)

    def array_decl(self, ln):
        """array[n]...[m] (optionally `mixed`) with an optional { ... }
        initializer of values, objects or nested lists."""
        self.next()
        self.accept("mixed")
        dims = []
        while self.accept("["):
            dims.append(int(float(self.cond())))
            self.expect("]")
        if not dims:
            raise SdlError(f"line {ln}: array without a size")
        arr = _Array(dims)
        if self.peek()[1] == "{":
            def read(level):
                self.expect("{")
                out = []
                while not self.accept("}"):
                    out.append(read(level + 1) if level + 1 < len(dims) else self.decl_value(ln))
                    self.accept(",")
                return out
            arr.fill(read(0), ln)
        return arr

    def array_element(self, arr, ln):
        """A[i][j] after the array's name."""
        idx = []
        while self.accept("["):
            idx.append(int(float(self.cond())))
            self.expect("]")
        return arr.get(idx, ln)

    def call_function(self, fn, args, ln):
        if len(args) != len(fn.params):
            raise SdlError(f"line {ln}: the function takes {len(fn.params)} arguments")
        saved_i, saved_decl = self.i, self.decl
        self.decl = self.decl.new_child(dict(zip(fn.params, args)))
        self.i = fn.start
        try:
            v = self.quiet(self.cond)
            if self.i != fn.end:
                raise SdlError(f"line {ln}: function body is not a plain expression (pattern or object)")
        finally:
            self.i, self.decl = saved_i, saved_decl
        return v

    _OPAQUE_BLOCKS = ("pigment_map", "normal_map", "texture_map", "slope_map", "density_map", "material_map",
                      "media", "density", "fog", "rainbow", "sky_sphere", "photons", "camera")

    def colour_map_entries(self):
        """The colours of a color_map { [v colour] ... } (or of a declared one)."""
        cols = []
        self.expect("{")
        while not self.accept("}"):
            if self.accept("["):
                self.flt()
                self.accept(",")
                if self.peek()[1] != "]":
                    cols.append(self.colour())
                while not self.accept("]"):
                    self.next()
            else:
                t = self.next()
                if t[0] == "id" and isinstance(self.decl.get(t[1]), _ColourMap):
                    cols.extend(self.decl[t[1]].cols)
        return cols

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

    def accept_raw(self, val):
        """accept() without running the directives that may follow: a
        directive's own trailing token (the pump loop then carries on with
        the next directive instead of recursing into it)."""
        if self.raw()[1] == val:
            self.i += 1
            return True
        return False

    def comma(self):
        self.accept(",")

    # -- expressions --------------------------------------------------------------------
    def cond(self):
        """Full expression: ?:, |, &, comparisons, then arithmetic
        (parser_expressions.cpp Parse_Float precedence)."""
        v = self.lor()
        if self.accept("?"):
            a = self.cond()
            self.expect(":")
            b = self.cond()
            return a if _truth(v) else b
        return v

    def lor(self):
        v = self.land()
        while self.accept("|"):
            w = self.land()
            v = float(_truth(v) or _truth(w))
        return v

    def land(self):
        v = self.rel()
        while self.accept("&"):
            w = self.rel()
            v = float(_truth(v) and _truth(w))
        return v

    def rel(self):
        v = self.expr()
        while self.peek()[1] in ("<", "<=", "=", "!=", ">=", ">"):
            op = self.next()[1]
            w = self.expr()
            a, b = float(np.max(v)) if np.ndim(v) else float(v), float(np.max(w)) if np.ndim(w) else float(w)
            v = float({"<": a < b, "<=": a <= b, "=": a == b, "!=": a != b, ">=": a >= b, ">": a > b}[op])
        return v

    def expr(self):
        v = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.next()[1]
            w = self.term()
            v, w = _promote(v, w)
            v = v + w if op == "+" else v - w
        return v

    def term(self):
        v = self.factor()
        while self.peek()[1] in ("*", "/"):
            op = self.next()[1]
            w = self.factor()
            v, w = _promote(v, w)
            v = v * w if op == "*" else v / w
        return v

    _FUNCS = {"sqrt": np.sqrt, "sin": np.sin, "cos": np.cos, "tan": np.tan, "abs": np.abs,
              "radians": np.radians, "degrees": np.degrees, "exp": np.exp, "log": np.log10, "ln": np.log,
              "asin": np.arcsin, "acos": np.arccos, "atan": np.arctan, "floor": np.floor, "ceil": np.ceil,
              "vlength": np.linalg.norm}
    _FUNCS2 = ("int", "mod", "min", "max", "rand", "seed", "atan2", "vnormalize", "vdot", "vcross", "select", "vrotate",
               "div")

    def func2(self, name, a, ln):
        if name == "int":
            return float(math.trunc(float(a[0])))
        if name == "div":
            return float(math.trunc(float(a[0]) / float(a[1])))
        if name == "mod":
            return float(math.fmod(float(a[0]), float(a[1])))
        if name == "min":
            return float(min(float(v) for v in a))
        if name == "max":
            return float(max(float(v) for v in a))
        if name == "atan2":
            return float(math.atan2(float(a[0]), float(a[1])))
        if name == "vnormalize":
            v = np.asarray(a[0], np.float64)
            return v / np.linalg.norm(v)
        if name == "vdot":
            return float(np.dot(np.asarray(a[0])[:3], np.asarray(a[1])[:3]))
        if name == "vrotate":  # rotate a point by <x, y, z> degrees, as the rotate modifier
            m = scenes.rotation_transform(np.asarray(a[1], np.float64)[:3])[0]
            return np.asarray(a[0], np.float64)[:3] @ m[:3, :3]
        if name == "vcross":
            return np.cross(np.asarray(a[0])[:3], np.asarray(a[1])[:3])
        if name == "select":
            return a[1] if float(a[0]) < 0 else (a[2] if len(a) == 3 or float(a[0]) == 0 else a[3])
        if name == "seed":  # Parser::stream_seed (parser_expressions.cpp:3344-3351)
            self.rand_state.append(int(float(a[0])) & 0xFFFFFFFF)
            return float(len(self.rand_state) - 1)
        if name == "rand":  # POV_rand (core/math/randomsequence.cpp:113-118)
            k = int(float(a[0]))
            if k < 0 or k >= len(self.rand_state):
                raise SdlError(f"line {ln}: illegal random number generator")
            self.rand_state[k] = (self.rand_state[k] * 1812433253 + 12345) & 0xFFFFFFFF
            return self.rand_state[k] / 0xFFFFFFFF
        raise SdlError(f"line {ln}: unknown function {name}")

    def factor(self):
        v = self.primary()
        # member access: .x .y .z (.u .v, .red .green .blue .filter .transmit)
        while self.peek()[1] == ".":
            self.next()
            m = self.next()[1]
            idx = {"x": 0, "u": 0, "red": 0, "y": 1, "v": 1, "green": 1, "z": 2, "blue": 2, "filter": 3,
                   "transmit": 4, "t": 3}.get(m)
            if idx is None or np.ndim(v) == 0 or idx >= np.shape(v)[0]:
                raise SdlError(f"line {self.line()}: bad member .{m}")
            v = float(np.asarray(v)[idx])
        return v

    _CONST = {"true": 1.0, "yes": 1.0, "on": 1.0, "false": 0.0, "no": 0.0, "off": 0.0, "clock": 0.0,
              "clock_delta": 1.0, "clock_on": 0.0, "frame_number": 0.0, "initial_frame": 1.0, "final_frame": 1.0,
              "initial_clock": 0.0, "final_clock": 1.0}

    def primary(self):
        kind, val, ln = self.peek()
        if val == "!":
            self.next()
            return float(not _truth(self.factor()))
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
            v = self.cond()
            self.expect(")")
            return v
        if val == "<":  # components separated by commas, which may be left out (Parse_Comma)
            self.next()
            comps = [self.expr()]
            while not self.accept(">"):
                self.accept(",")
                if self.peek()[0] == "eof":
                    raise SdlError(f"line {ln}: unterminated vector")
                comps.append(self.expr())
            return np.array([float(c) for c in comps])
        if kind == "id":
            self.next()
            if isinstance(self.decl.get(val), _Function) and self.peek()[1] == "(":
                self.next()
                args = [] if self.peek()[1] == ")" else [self.cond()]
                while self.accept(","):
                    args.append(self.cond())
                self.expect(")")
                return self.call_function(self.decl[val], args, ln)
            if isinstance(self.decl.get(val), _Array):
                v = self.array_element(self.decl[val], ln)
                if isinstance(v, (_Obj, _Material, _Light)):
                    raise SdlError(f"line {ln}: array element of {val!r} is not a value")
                return v
            if val in ("dimensions", "dimension_size"):
                self.expect("(")
                a = self.decl.get(self.next()[1])
                if not isinstance(a, _Array):
                    raise SdlError(f"line {ln}: {val} of something that is not an array")
                d = 1
                if val == "dimension_size":
                    self.expect(",")
                    d = int(float(self.cond()))
                self.expect(")")
                return float(len(a.dims)) if val == "dimensions" else float(a.dims[d - 1])
            if val in ("x", "y", "z") and val not in self.decl:
                return np.eye(3)["xyz".index(val)].copy()
            if val == "pi":
                return math.pi
            if val in self._CONST and val not in self.decl:
                return self._CONST[val]
            if val == "version":
                return self.version
            if val == "image_width":
                return float(self.width)
            if val == "image_height":
                return float(self.height)
            if val in self._FUNCS2:
                self.expect("(")
                args = [self.cond()]
                while self.accept(","):
                    args.append(self.cond())
                self.expect(")")
                return self.func2(val, args, ln)
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
            if val in self.decl and not isinstance(self.decl[val], (_Obj, _Material, _Light, _Function, _ColourMap, _Opaque, _Xf, _Array)):
                return self.decl[val]
            if val in _COLOURS:
                return np.array(list(_COLOURS[val]) + [0.0, 0.0], np.float64)
            raise SdlError(f"line {ln}: unknown identifier {val!r}")
        raise SdlError(f"line {ln}: expected a value, found {val!r}")

    def vector(self):
        """A 3-vector; a float stands for all three components and a
        2-vector gets z = 0 (the language's vector promotion)."""
        v = self.expr()
        if np.isscalar(v) or np.ndim(v) == 0:
            return np.full(3, float(v))
        v = np.asarray(v, np.float64)
        if v.shape[0] == 2:
            return np.array([v[0], v[1], 0.0])
        if v.shape[0] < 3:
            raise SdlError(f"line {self.line()}: expected a 3-vector")
        return v[:3]

    def flt(self):
        v = self.expr()
        if np.ndim(v) != 0:
            raise SdlError(f"line {self.line()}: expected a float")
        return float(v)

    def colour(self):
        """color [rgb|rgbf|rgbt|rgbft] <..>, a colour expression (a named or
        declared colour, possibly scaled), or `red r green g ...`, each
        optionally followed by component keywords (Parse_Colour): (rgb,
        filter, transmit).  Colours are kept as 5-vectors r, g, b, filter,
        transmit."""
        self.accept("color") or self.accept("colour")
        kind = self.peek()[1]
        v5 = np.zeros(5)
        if kind in self._COMPONENTS:  # color red 1 green 0.5 ... (the older form)
            pass
        elif kind in ("rgb", "rgbf", "rgbt", "rgbft"):
            self.next()
            v = self.expr()
            v = np.full(3, float(v)) if np.ndim(v) == 0 else np.asarray(v, np.float64)
            n = {"rgb": 3, "rgbf": 4, "rgbt": 4, "rgbft": 5}[kind]
            if v.shape[0] < n:
                v = np.concatenate([v, np.zeros(n - v.shape[0])])
            v5[:3] = v[:3]
            if kind == "rgbf":
                v5[3] = v[3]
            elif kind == "rgbt":
                v5[4] = v[3]
            elif kind == "rgbft":
                v5[3:5] = v[3:5]
        else:
            v = self.expr()
            v = np.full(3, float(v)) if np.ndim(v) == 0 else np.asarray(v, np.float64)
            if v.shape[0] < 3:
                raise SdlError(f"line {self.line()}: expected a colour")
            v5[:min(5, v.shape[0])] = v[:5]
        self.colour_components(v5)
        return v5[:3].copy(), float(v5[3]), float(v5[4])

    _COMPONENTS = {"red": 0, "green": 1, "blue": 2, "filter": 3, "transmit": 4}

    def colour_components(self, v):
        """Trailing `red 1 green 0.5 ... filter f transmit t` components
        (Parse_Colour), written into the 5-vector v."""
        while self.peek()[1] in self._COMPONENTS:
            k = self._COMPONENTS[self.next()[1]]
            v[k] = self.flt()

    def skip_braces(self):
        """Skip a { ... } block (the opening brace next)."""
        self.expect("{")
        depth = 1
        while depth:
            t = self.next()
            if t[0] == "eof":
                raise SdlError("unterminated block")
            depth += (t[1] == "{") - (t[1] == "}")

    # -- blocks ---------------------------------------------------------------------------
    def skip_block_to_end(self):
        self.expect("}")

    def parse(self):
        self.run()
        if self.ctrl:
            raise SdlError(f"unterminated {self.ctrl[-1][0]} block at the end of the scene")
        return self

    def run(self):
        while self.peek()[0] != "eof":
            kind, val, ln = self.peek()
            if kind == "eof":
                break
            if val == "camera":
                self.next()
                self.camera_block()
            elif val == "light_source":
                self.next()
                self.lights.append(self.light_block())
            elif val == "background":
                self.next()
                self.expect("{")
                self.background = self.colour()[0]
                self.expect("}")
            elif val == "global_settings":
                self.next()
                self.global_settings()
            elif val in _SHAPES or val in _OTHER_SHAPES:
                o = self.shape()
                if o.kind != "skipped":
                    self.objects.append(o)
            elif val == "default":  # the #default block written without '#'
                self.next()
                self.default_block(ln)
            elif val in ("sky_sphere", "fog", "rainbow", "media"):
                self.next()
                self.skip_braces()
                self.warn(f"line {ln}: {val} ignored")
            elif val == ";":
                self.next()
            else:
                self.err(f"unsupported statement {val!r}")

    def transform_mod(self, val, ln):
        """translate / rotate / scale / matrix as a (matrix, inverse) pair, or None."""
        if val == "translate":
            return scenes.translation_transform(self.vector())
        if val == "rotate":
            return scenes.rotation_transform(self.vector())
        if val == "scale":
            return _scale(self.vector())
        if val == "matrix":
            v = np.asarray(self.expr(), np.float64)
            if v.shape != (12,):
                raise SdlError(f"line {ln}: matrix takes 12 values")
            m = np.eye(4)
            m[:, :3] = v.reshape(4, 3)
            return m, np.linalg.inv(m)
        if val == "transform":  # transform IDENT | transform { ... } (Parse_Transform)
            if self.peek()[0] == "id" and isinstance(self.decl.get(self.peek()[1]), _Xf):
                return self.decl[self.next()[1]].tr
            return self.transform_block(ln)
        return None

    def transform_block(self, ln):
        """{ translate / rotate / scale / matrix / transform / inverse ... }
        composed in order."""
        tr = _identity()
        self.expect("{")
        while not self.accept("}"):
            k = self.next()[1]
            if k == "inverse":
                tr = (tr[1], tr[0])
                continue
            if isinstance(self.decl.get(k), _Xf):
                t = self.decl[k].tr
            else:
                t = self.transform_mod(k, ln)
            if t is None:
                raise SdlError(f"line {ln}: unsupported transform keyword {k!r}")
            tr = _compose(tr, t)
        return tr

    def camera_block(self):
        """Parse_Camera for language version 3.5 and later (parser.cpp:1335-1830):
        the modifiers are collected, then applied in a fixed order — angle
        (direction length from the right vector's), look_at (with the
        handedness of the supplied up / direction / right), then the camera's
        transforms (Camera::Transform)."""
        self.expect("{")
        loc, dirv, up = np.zeros(3), np.array([0.0, 0.0, 1.0]), np.array([0.0, 1.0, 0.0])
        right, sky = np.array([1.33, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
        look, angle = None, None
        tr = _identity()
        while not self.accept("}"):
            _, val, ln = self.next()
            t = self.transform_mod(val, ln)
            if t is not None:
                tr = _compose(tr, t)
            elif val == "location":
                loc = self.vector()
            elif val == "look_at":
                look = self.vector()
            elif val == "sky":
                sky = self.vector()
            elif val == "angle":
                angle = self.flt()
            elif val == "right":
                right = self.vector()
            elif val == "up":
                up = self.vector()
            elif val == "direction":
                dirv = self.vector()
            elif val == "perspective":
                pass
            elif val in ("orthographic", "fisheye", "ultra_wide_angle", "omnimax", "panoramic", "spherical",
                         "cylinder"):
                if val == "cylinder":
                    self.flt()
                self.warn(f"line {ln}: {val} camera read as perspective")
            elif val in ("focal_point", "aperture", "blur_samples", "confidence", "variance"):
                self.cond()
                self.warn(f"line {ln}: camera {val} ignored (no focal blur)")
            else:
                raise SdlError(f"line {ln}: unsupported camera keyword {val!r}")
        if angle is not None and angle > 0.0:
            if angle >= 180.0:
                raise SdlError("viewing angle has to be smaller than 180 degrees")
            dirv = dirv / np.linalg.norm(dirv) * (np.linalg.norm(right) / math.tan(angle * math.pi / 360.0) / 2.0)
        if look is not None:
            dl, ul, rl = np.linalg.norm(dirv), np.linalg.norm(up), np.linalg.norm(right)
            hand = float(np.dot(np.cross(up, dirv), right))
            d = look - loc
            if float(np.dot(d, d)) < 1e-10:
                raise SdlError("camera location and look_at point must be different")
            d = d / np.linalg.norm(d)
            r = np.cross(sky, d)
            if np.all(np.abs(r) < 1e-10):
                self.warn("camera look_at direction parallel to sky: the supplied right vector is kept")
                r = right
            r = r / np.linalg.norm(r)
            up = np.cross(d, r) * ul
            dirv = d * dl
            right = r * (rl if hand > 0.0 else -rl)
        m = tr[0]
        loc = loc @ m[:3, :3] + m[3, :3]
        dirv, up, right = dirv @ m[:3, :3], up @ m[:3, :3], right @ m[:3, :3]
        self.camera = dict(location=loc, direction=dirv, right=right, up=up, width=int(self.width),
                           height=int(self.height))

    def light_block(self):
        """Parse_Light_Source: the light (a declared one copied, or a position
        and colour), area_light, modifiers; returns the light dict.  The
        renderer's area lights trace every grid point without jitter: jitter,
        adaptive, circular and orient are read and reduced with a warning, as
        are spotlight / cylinder shapes and fading (the light becomes a plain
        point or area light)."""
        self.expect("{")
        kind0, val0, _ = self.peek()
        if kind0 == "id" and isinstance(self.decl.get(val0), _Light):
            self.next()
            light = self.decl[val0].copy()
        else:
            light = dict(kind="point", position=self.vector(), colour=np.ones(3))
            self.comma()
        while not self.accept("}"):
            kind, val, ln = self.peek()
            if val in ("translate", "rotate", "scale", "matrix"):
                self.next()
                m = self.transform_mod(val, ln)[0]
                light["position"] = light["position"] @ m[:3, :3] + m[3, :3]
                for k in ("axis1", "axis2"):
                    if k in light:
                        light[k] = light[k] @ m[:3, :3]
            elif val in ("color", "colour", "rgb") or (kind == "id" and val in _COLOURS):
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
                    self.warn(f"line {ln}: area light adaptive read as full sampling (every grid point is traced)")
            elif val in ("jitter", "orient", "circular", "spotlight", "cylinder", "parallel", "shadowless",
                         "media_interaction", "media_attenuation"):
                self.next()
                if val in ("media_interaction", "media_attenuation"):
                    self.cond()
                self.warn(f"line {ln}: light_source {val} reduced (plain point / area light)")
            elif val in ("radius", "falloff", "tightness", "fade_power", "fade_distance"):
                self.next()
                self.cond()
                self.warn(f"line {ln}: light_source {val} ignored")
            elif val == "point_at":
                self.next()
                self.vector()
            elif val in ("looks_like", "photons", "projected_through"):
                self.next()
                self.skip_braces()
                self.warn(f"line {ln}: light_source {val} ignored")
            else:
                raise SdlError(f"line {ln}: unsupported light_source keyword {val!r}")
        return light

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
                    elif k in ("error_bound", "minimum_reuse", "maximum_reuse", "pretrace_start", "pretrace_end",
                               "brightness", "adc_bailout"):
                        self.radiosity[k] = self.flt()
                    elif k in ("low_error_factor", "gray_threshold", "max_sample", "always_sample", "normal",
                               "media", "subsurface", "nearest_count_min"):
                        self.cond()
                        self.warn(f"line {l2}: radiosity {k} ignored")
                    else:
                        raise SdlError(f"line {l2}: unsupported radiosity keyword {k!r}")
            elif val in ("noise_generator", "charset", "hf_gray_16", "number_of_waves", "irid_wavelength",
                         "max_intersections", "mm_per_unit"):
                if val == "charset":
                    self.next()
                else:
                    self.cond()
                self.warn(f"line {ln}: global_settings {val} ignored")
            elif val in ("photons", "subsurface"):
                self.skip_braces()
                self.warn(f"line {ln}: global_settings {val} ignored")
            else:
                raise SdlError(f"line {ln}: unsupported global_settings keyword {val!r}")

    # -- objects ------------------------------------------------------------------------------
    def shape(self) -> _Obj:
        _, kw, ln = self.next()
        if kw in _OTHER_SHAPES:
            if not self.skip_unsupported:
                raise SdlError(f"line {ln}: unsupported shape {kw!r}")
            self.skip_braces()
            self.warn(f"line {ln}: {kw} not supported, dropped")
            return _Obj("skipped", {}, ln)
        self.expect("{")
        if kw == "object":
            if self.peek()[1] in _SHAPES or self.peek()[1] in _OTHER_SHAPES:
                o = self.shape()  # object { shape modifiers }
            else:
                name = self.next()
                v = self.decl.get(name[1])
                if isinstance(v, _Array):
                    v = self.array_element(v, ln)
                if not isinstance(v, _Obj):
                    raise SdlError(f"line {ln}: {name[1]!r} is not a declared object")
                o = v.copy()
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
            self.accept(",")  # Parse_Disc: Parse_Comma, then a hole radius only if a float starts here
            if self.starts_float():
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
        elif kw in ("lathe", "sor"):  # Parse_Lathe (parser.cpp:2812-2995), Parse_Sor (5628-5714)
            kind = A.PVX_SOR_CUBIC if kw == "sor" else A.PVX_LATHE_LINEAR
            if kw == "lathe" and self.peek()[1] in _LATHE_SPLINES:
                kind = _LATHE_SPLINES[self.next()[1]]
            cnt = int(self.flt())
            self.comma()
            pts = []
            for _ in range(cnt):
                pts.append(self.uv())
                self.comma()
            o = _Obj(kw, dict(p=np.array(pts, np.float64).reshape(-1, 2), kind=kind, flags=0), ln)
        else:  # CSG
            o = _Obj("csg", dict(op=kw), ln)
            dropped = False
            while self.peek()[1] in _SHAPES or self.peek()[1] in _OTHER_SHAPES:
                c = self.shape()
                if c.kind == "skipped":
                    dropped = True
                    if kw == "difference" and not o.children:  # the base shape is gone
                        o.data["lost_base"] = True
                else:
                    o.children.append(c)
            if dropped and (not o.children or o.data.get("lost_base")):
                self.mods(o)
                return _Obj("skipped", {}, ln)
            if not o.children:
                raise SdlError(f"line {ln}: {kw} without objects")
        self.mods(o)
        return o

    def uv(self):
        """Parse_UV_Vect: a 2-vector (a float stands for both components)."""
        v = self.expr()
        if np.ndim(v) == 0:
            return np.array([float(v), float(v)])
        v = np.asarray(v, np.float64)
        if v.shape[0] < 2:
            raise SdlError(f"line {self.line()}: expected a 2-vector")
        return v[:2]

    def starts_float(self):
        """Whether the next token can begin a float expression (CASE_FLOAT_UNGET)."""
        kind, val, _ = self.peek()
        if kind == "num" or val in ("-", "+", "(", "!"):
            return True
        if kind == "id":
            if val in self._FUNCS or val in self._FUNCS2 or val in self._CONST or val in ("pi", "pow", "image_width",
                                                                                        "image_height"):
                return True
            v = self.decl.get(val)
            return v is not None and not isinstance(v, (_Obj, _Material, _Light)) and np.ndim(v) == 0
        return False

    def mesh2(self, ln):
        verts = norms = faces = nidx = None
        blocks = ("vertex_vectors", "normal_vectors", "uv_vectors", "texture_list", "face_indices",
                  "normal_indices", "uv_indices", "inside_vector")
        while self.peek()[1] in blocks:
            _, k, l2 = self.next()
            if k == "inside_vector":
                self.vector()
                self.warn(f"line {l2}: mesh2 inside_vector ignored (the mesh is not a solid)")
                continue
            if k == "texture_list":
                self.skip_braces()
                self.warn(f"line {l2}: mesh2 texture_list ignored (the object's texture is used)")
                continue
            self.expect("{")
            cnt = int(self.flt())
            items = []
            for _ in range(cnt):
                self.comma()
                items.append(self.expr())
                if k == "face_indices":  # per-face texture indices: one or three (Parse_Mesh2)
                    while self.peek()[1] == "," and self.i + 1 < len(self.toks) and \
                            self.toks[self.i + 1][0] == "num":
                        self.next()
                        self.flt()
            self.accept(",")
            self.expect("}")
            if k == "uv_vectors":
                self.warn(f"line {l2}: mesh2 uv_vectors ignored (no uv mapping)")
                continue
            if k == "uv_indices":
                continue
            arr = np.array([np.asarray(v, np.float64)[:3] if np.ndim(v) else np.full(3, float(v))
                            for v in items]).reshape(-1, 3)
            if k == "vertex_vectors":
                verts = arr
            elif k == "normal_vectors":
                norms = arr
            elif k == "normal_indices":
                nidx = arr.astype(np.int64)
            else:
                faces = arr.astype(np.int64)
        if verts is None or faces is None:
            raise SdlError(f"line {ln}: mesh2 needs vertex_vectors and face_indices")
        if norms is not None and nidx is not None and not np.array_equal(nidx, faces):
            self.warn(f"line {ln}: mesh2 normal_indices differ from face_indices: normals dropped (flat faces)")
            norms = None
        if norms is not None and norms.shape[0] != verts.shape[0]:
            self.warn(f"line {ln}: mesh2 normal count differs from the vertex count: normals dropped (flat faces)")
            norms = None
        return _Obj("mesh", dict(v=verts, n=norms, f=faces), ln)

    def mods(self, o: _Obj):
        if o.kind == "skipped":  # a dropped shape: its modifiers are not read
            depth = 0
            while True:
                t = self.next()
                if t[0] == "eof":
                    raise SdlError(f"line {o.line}: unterminated object")
                if t[1] == "}" and depth == 0:
                    return
                depth += (t[1] == "{") - (t[1] == "}")
        while not self.accept("}"):
            kind, val, ln = self.next()
            t = self.transform_mod(val, ln)
            if t is not None:
                self.transform(o, t)
            elif val == "inverse":
                o.inverted = not o.inverted
            elif val in ("open", "sturm"):
                if val == "open" and o.kind == "cone":
                    o.data["flags"] |= A.PVX_CONE_OPEN
                elif val == "open" and o.kind == "sor":
                    o.data["flags"] |= A.PVX_REV_OPEN
                elif val == "sturm" and o.kind == "torus":
                    o.data["sturm"] = 1
                elif val == "sturm" and o.kind in ("lathe", "sor"):
                    o.data["flags"] |= A.PVX_REV_STURM
                elif val == "sturm" and o.kind in ("csg", "mesh", "sphere", "box", "plane", "triangle"):
                    pass  # no polynomial to solve there
                else:
                    raise SdlError(f"line {ln}: {val} does not apply to {o.kind}")
            elif val in ("no_shadow", "no_image", "no_reflection", "no_radiosity"):
                o.flags |= {"no_shadow": A.PVX_FLAG_NO_SHADOW, "no_image": A.PVX_FLAG_NO_IMAGE,
                            "no_reflection": A.PVX_FLAG_NO_REFLECTION,
                            "no_radiosity": A.PVX_FLAG_NO_RADIOSITY}[val]
            elif val in ("pigment", "finish", "interior", "texture", "normal"):
                self.material(o, val)
                if val == "texture":
                    self.texture_layers(ln)
            elif val in ("color", "colour", "rgb"):
                self.i -= 1
                self.set_colour(o)
            elif val in ("hollow", "double_illuminate", "uv_mapping", "split_union", "cutaway_textures"):
                if self.peek()[1] in ("on", "off", "true", "false", "yes", "no") or self.starts_float():
                    self.cond()
                if val not in ("hollow", "split_union"):
                    self.warn(f"line {ln}: {val} ignored")
            elif val in ("photons", "interior_texture", "radiosity"):
                self.skip_braces()
                self.warn(f"line {ln}: object {val} ignored")
            elif val == "bounded_by":  # a manual bound: the object's own box bounds it here
                self.skip_braces()
                self.warn(f"line {ln}: bounded_by ignored")
            elif val == "clipped_by":
                if not self.skip_unsupported:
                    raise SdlError(f"line {ln}: clipped_by is not supported (it changes which hits exist)")
                self.skip_braces()
                self.warn(f"line {ln}: clipped_by ignored (the object is not clipped)")
            elif val == "material":  # material { texture { } interior { } }
                self.expect("{")
                while not self.accept("}"):
                    k = self.next()[1]
                    if k in ("texture", "pigment", "finish", "normal", "interior"):
                        self.material(o, k)
                    elif isinstance(self.decl.get(k), _Material):
                        o.mat.update(self.decl[k])
                    else:
                        t = self.transform_mod(k, ln)
                        if t is None:
                            raise SdlError(f"line {ln}: unsupported material keyword {k!r}")
            else:
                raise SdlError(f"line {ln}: unsupported object modifier {val!r}")

    def set_colour(self, o):
        rgb, filt, trans = self.colour()
        o.mat["pigment"] = rgb
        o.mat["filter"] = filt
        o.mat["transmit"] = trans

    # pigment patterns: read, reduced to the mean of their colour map (or of
    # their listed colours) with a warning
    _PATTERNS = ("checker", "bozo", "spotted", "agate", "granite", "marble", "wood", "gradient", "radial",
                 "leopard", "onion", "spiral1", "spiral2", "ripples", "waves", "wrinkles", "dents", "bumps",
                 "crackle", "hexagon", "brick", "mandel", "julia", "spherical", "cylindrical", "boxed",
                 "planar", "cells", "average", "image_map", "bump_map", "quilted", "facets", "tiling", "pavement",
                 "function", "pigment_pattern", "object", "slope", "aoi", "brick", "square", "triangular",
                 "cubic")
    _PATTERN_MODS = ("turbulence", "octaves", "omega", "lambda", "frequency", "phase", "scallop_wave",
                     "sine_wave", "ramp_wave", "triangle_wave", "cubic_wave", "poly_wave", "noise_generator",
                     "uv_mapping", "once", "interpolate", "map_type", "mortar", "brick_size")

    def pattern_pigment(self, o, first, ln):
        """A pattern pigment: read to its closing brace; the plain colour is the
        mean of the colour map entries or listed colours."""
        cols = []
        if first in ("checker", "hexagon", "brick", "square", "triangular"):
            while self.peek()[1] in ("color", "colour", "rgb", "rgbf", "rgbt", "rgbft") or \
                    (self.peek()[0] == "id" and self.peek()[1] in _COLOURS):
                cols.append(self.colour())
                self.accept(",")
        depth = 0
        while True:
            kind, val, l2 = self.peek()
            if val == "}" and depth == 0:
                break
            if val in ("color_map", "colour_map"):
                self.next()
                if self.peek()[0] == "id" and isinstance(self.decl.get(self.peek()[1]), _ColourMap):
                    cols.extend(self.decl[self.next()[1]].cols)
                else:
                    cols.extend(self.colour_map_entries())
                continue
            self.next()
            depth += (val == "{") - (val == "}")
        if cols:
            o.mat["pigment"] = np.mean([c[0] for c in cols], axis=0)
            o.mat["filter"] = float(np.mean([c[1] for c in cols]))
            o.mat["transmit"] = float(np.mean([c[2] for c in cols]))
        self.warn(f"line {ln}: pigment pattern {first!r} reduced to a plain colour")

    def material(self, o, block):
        if block == "normal":
            ln = self.line()
            self.skip_braces()
            self.warn(f"line {ln}: normal perturbation ignored")
            return
        self.expect("{")
        while not self.accept("}"):
            kind, val, ln = self.peek()
            if block == "texture" and val in ("pigment", "finish", "normal"):
                self.next()
                self.material(o, val)
            elif kind == "id" and isinstance(self.decl.get(val), _Material):
                self.next()
                o.mat.update(self.decl[val])
            elif block == "pigment" and val in self._PATTERNS:
                self.next()
                self.pattern_pigment(o, val, ln)
            elif block == "pigment" and (val in ("color", "colour", "rgb", "rgbf", "rgbt", "rgbft")
                                         or val in _COLOURS or val in self.decl or val == "<"):
                self.set_colour(o)
            elif val in ("translate", "rotate", "scale", "matrix", "warp", "quick_color", "quick_colour",
                         "uv_mapping"):
                self.next()
                if val == "warp":
                    self.skip_braces()
                elif val in ("quick_color", "quick_colour"):
                    self.colour()
                elif val != "uv_mapping":
                    self.cond()
            elif block == "finish" and val in ("ambient", "diffuse", "brilliance", "phong", "phong_size",
                                               "reflection", "emission"):
                self.next()
                if val == "reflection" and self.accept("{"):
                    r = self.colour()[0]
                    r2 = r
                    if self.accept(","):
                        r2 = self.colour()[0]
                    while not self.accept("}"):
                        _, k2, l3 = self.next()
                        if k2 in ("fresnel", "metallic", "falloff", "exponent"):
                            if self.starts_float() or self.peek()[1] in ("on", "off"):
                                self.cond()
                            self.warn(f"line {l3}: reflection {k2} ignored")
                    if not np.allclose(r, r2):
                        self.warn(f"line {ln}: variable reflection read as its maximum")
                    o.mat[val] = float(np.max(r2))
                elif val in ("ambient", "emission", "diffuse", "reflection") and \
                        self.peek()[1] in ("color", "colour", "rgb", "rgbf", "rgbt", "rgbft"):
                    c = self.colour()[0]
                    o.mat[val] = c if val in ("ambient", "emission") else float(np.mean(c))
                else:
                    if val == "diffuse" and self.peek()[1] == "albedo":
                        self.next()
                    r = self.expr()
                    if val in ("ambient", "emission"):
                        o.mat[val] = r
                    else:
                        o.mat[val] = float(np.mean(r)) if np.ndim(r) else float(r)
            elif block == "finish" and val in ("specular", "roughness", "metallic", "crand", "irid",
                                               "conserve_energy", "fresnel", "subsurface", "reflection_exponent",
                                               "phong_albedo", "specular_albedo"):
                self.next()
                if val in ("irid", "subsurface"):
                    self.skip_braces()
                elif self.starts_float() or self.peek()[1] in ("on", "off"):
                    self.cond()
                self.warn(f"line {ln}: finish {val} ignored")
            elif block == "interior" and val == "ior":
                self.next()
                o.mat["ior"] = self.flt()
            elif block == "interior" and val in ("caustics", "dispersion", "dispersion_samples", "fade_distance",
                                                 "fade_power", "fade_color", "fade_colour", "media"):
                self.next()
                if val == "media":
                    self.skip_braces()
                elif val in ("fade_color", "fade_colour"):
                    self.colour()
                else:
                    self.cond()
                self.warn(f"line {ln}: interior {val} ignored")
            else:  # the rest of this block is not read (a texture feature pvx does not render)
                depth = 0
                while True:
                    t = self.next()
                    if t[0] == "eof":
                        raise SdlError(f"line {ln}: unterminated {block} block")
                    if t[1] == "}" and depth == 0:
                        break
                    depth += (t[1] == "{") - (t[1] == "}")
                self.warn(f"line {ln}: {block} keyword {val!r} and the rest of its block ignored")
                return

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
                # an ellipsoid (Sphere::Transform's Do_Ellipsoid): read as the
                # same surface in quadric form |p - c|^2 - r^2 = 0, transformed
                c, r = o.data["c"], o.data["r"]
                o.kind = "quadric"
                o.data = dict(coef=np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0, -2 * c[0], -2 * c[1], -2 * c[2],
                                             float(c @ c) - r * r]))
                o.tr = tr
                self.warn(f"line {o.line}: sphere with a non-uniform scale read as a quadric")
                return
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
        if o.kind == "triangle":  # Triangle::Transform: vertices (and normals) transformed
            o.data["p"] = o.data["p"] @ tr[0][:3, :3] + tr[0][3, :3]
            if o.data.get("n") is not None:
                n2 = o.data["n"] @ tr[1][:3, :3].T
                o.data["n"] = n2 / np.linalg.norm(n2, axis=1, keepdims=True)
            return
        o.tr = tr if o.tr is None else _compose(o.tr, tr)

    # -- flattening into a Scene ------------------------------------------------------------
    def scene(self) -> scenes.Scene:
        if self.camera is None:
            self.camera = scenes.look_at_camera((0, 0, 0), (0, 0, 1), self.width, self.height)
        arrays = dict(sph_c=[], sph_r=[], pl_n=[], pl_d=[], box_c1=[], box_c2=[], box_tr=[], tor_R=[], tor_r=[],
                      tor_tr=[], tor_s=[], xf=[], cone_a=[], cone_b=[], cone_r=[], cone_f=[], cone_t=[],
                      disc_c=[], disc_n=[], disc_r=[], disc_t=[], q=[], q_t=[], tri_p=[], tri_n=[], meshes=[],
                      poly_first=[0], poly_pts=[], poly_t=[], mesh_t=[],
                      rev_first=[0], rev_pts=[], rev_kind=[], rev_flags=[], rev_t=[],
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
            if k in ("lathe", "sor"):
                arrays["rev_pts"].extend(dd["p"])
                arrays["rev_first"].append(len(arrays["rev_pts"]))
                arrays["rev_kind"].append(dd["kind"])
                arrays["rev_flags"].append(dd["flags"])
                arrays["rev_t"].append(xf_index(o))
                return [A.PVX_SOR if k == "sor" else A.PVX_LATHE, len(arrays["rev_t"]) - 1, f, 0]
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
            if k == "mesh":  # FP32 vertices stay in mesh space; the placement is kept
                arrays["meshes"].append((dd["v"].astype(np.float32), dd["f"].astype(np.int32),
                                         None if dd["n"] is None else dd["n"].astype(np.float32)))
                arrays["mesh_t"].append(xf_index(o))
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

        def needs_split(o):
            return any(c.kind in ("triangle", "mesh", "polygon") or (c.kind == "csg" and needs_split(c))
                       for c in o.children)

        def split(o, mat, flags, out):
            """A top-level union whose operands include triangles, meshes or
            polygons (which are not CSG operands here) becomes its operands as
            top-level objects: a union's closest hit is the closest of its
            operands' (CSGUnion::All_Intersections keeps every operand hit);
            the union's texture and flags pass to them."""
            for c in o.children:
                m = {**mat, **c.mat}
                if c.kind == "csg" and c.data["op"] == "union" and not c.inverted and needs_split(c):
                    split(c, m, flags | c.flags, out)
                else:
                    c = c.copy()
                    c.mat = m
                    c.flags |= flags
                    out.append(c)

        top = []
        for o in self.objects:
            if o.kind == "csg" and o.data["op"] == "union" and not o.inverted and needs_split(o):
                split(o, o.mat, o.flags, top)
            else:
                top.append(o)
        objs, mats = [], []
        for o in top:
            if o.kind == "csg":
                objs.append([A.PVX_CSG, csg(o, 0), o.flags, 0])
            else:
                objs.append(leaf(o, o.flags, True))
            mats.append({**self.defaults, **o.mat})
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
        settings["warnings"] = list(self.warnings)
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
            mesh_transform=arr("mesh_t", (-1,), np.int32),
            revolution_first=np.array(arrays["rev_first"], np.uint32) if arrays["rev_t"] else None,
            revolution_points=arr("rev_pts", (-1, 2)), revolution_kind=arr("rev_kind", (-1,), np.uint8),
            revolution_flags=arr("rev_flags", (-1,), np.uint8), revolution_transform=arr("rev_t", (-1,), np.int32),
            camera=self.camera, lights=self.lights, materials=materials, settings=settings)


def parse(text: str, width: int = 320, height: int = 240, include_dirs=(), skip_unsupported=False) -> scenes.Scene:
    """Read scene-description text into a Scene rendered at width x height;
    #include files are looked for in `include_dirs`.  With
    `skip_unsupported`, shapes pvx does not have are dropped with a warning
    (else they are an error)."""
    return _Parser(text, width, height, include_dirs, skip_unsupported).parse().scene()


def parse_file(path, width: int = 320, height: int = 240, library_paths=(), skip_unsupported=False) -> scenes.Scene:
    """Read a scene file; #include files are looked for in its folder, then
    in `library_paths` (the reference's Library_Path / +L option)."""
    import os
    return parse(_read_text(path), width, height, [os.path.dirname(os.path.abspath(path))] + list(library_paths),
                 skip_unsupported)

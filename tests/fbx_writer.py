"""Binary FBX writer for the loader tests, written from the published record
layout (not from any reader): a 27-byte header ("Kaydara FBX Binary  ", 0x00,
0x1A, 0x00, little-endian u32 version), then node records

    end offset, property count, property-list bytes   (u32 each; u64 from 7500)
    name length (u8), name, properties, nested records, an all-zero record
    closing a non-empty nested list

and an all-zero record after the top-level list.  Properties are a type byte
and a value: Y i16, C bool, I i32, F f32, D f64, L i64, S / R u32 length +
bytes, and arrays f d l i b as u32 count, u32 encoding (0 raw, 1 zlib), u32
byte length, data.

Plain Python values: int -> I, float -> D, str / bytes -> S; wrap with L(),
F(), Y(), C() or Arr() for the other types.

fbx_ascii writes the same tree in the ASCII form: "Name: v, v { children }",
';' comments, arrays as "*count { a: v,v,... }" over several lines, object
names "Class::Name" for the binary "Name\x00\x01Class"."""
import struct
import zlib


class L(int):
    pass


class F(float):
    pass


class Y(int):
    pass


class C:
    def __init__(self, v):
        self.v = bool(v)


class Arr:
    def __init__(self, kind, values, compress=False):
        assert kind in "fdlib"
        self.kind, self.values, self.compress = kind, list(values), compress


class Node:
    def __init__(self, name, *props, kids=()):
        self.name, self.props, self.kids = name, list(props), list(kids)


def P(name, typ, label, flags, *values):
    """A Properties70 entry."""
    return Node("P", name, typ, label, flags, *values)


def props70(*entries):
    return Node("Properties70", kids=list(entries))


def vec(name, x, y, z, typ=None):
    return P(name, typ or name, "", "A", float(x), float(y), float(z))


def encode_prop(p):
    if isinstance(p, Arr):
        fmt = {"f": "<f", "d": "<d", "l": "<q", "i": "<i", "b": "<B"}[p.kind]
        raw = b"".join(struct.pack(fmt, v) for v in p.values)
        data = zlib.compress(raw) if p.compress else raw
        return p.kind.encode() + struct.pack("<III", len(p.values), 1 if p.compress else 0, len(data)) + data
    if isinstance(p, C):
        return b"C" + bytes([1 if p.v else 0])
    if isinstance(p, L):
        return b"L" + struct.pack("<q", int(p))
    if isinstance(p, Y):
        return b"Y" + struct.pack("<h", int(p))
    if isinstance(p, F):
        return b"F" + struct.pack("<f", float(p))
    if isinstance(p, bool):
        return b"C" + bytes([1 if p else 0])
    if isinstance(p, int):
        return b"I" + struct.pack("<i", p)
    if isinstance(p, float):
        return b"D" + struct.pack("<d", p)
    if isinstance(p, str):
        p = p.encode("latin-1")
    if isinstance(p, (bytes, bytearray)):
        return b"S" + struct.pack("<I", len(p)) + bytes(p)
    raise TypeError(type(p))


def encode_node(node, offset, wide):
    hs = 25 if wide else 13
    name = node.name.encode()
    props = b"".join(encode_prop(p) for p in node.props)
    body = b""
    cur = offset + hs + len(name) + len(props)
    for k in node.kids:
        kb = encode_node(k, cur, wide)
        body += kb
        cur += len(kb)
    if node.kids:
        body += b"\0" * hs
        cur += hs
    head = struct.pack("<QQQB" if wide else "<IIIB", cur, len(node.props), len(props), len(name))
    return head + name + props + body


def fbx_bytes(top, version=7400):
    wide = version >= 7500
    out = b"Kaydara FBX Binary  \x00\x1a\x00" + struct.pack("<I", version)
    for n in top:
        out += encode_node(n, len(out), wide)
    out += b"\0" * (25 if wide else 13)
    out += b"\xfa\xbc\xab\x09\xd0\xc8\xd4\x66\xb1\x76\xfb\x83\x1c\xf7\x26\x7e" + b"\0" * 4  # footer
    return out


def ascii_value(p):
    if isinstance(p, Arr):
        vals = [repr(float(v)) if p.kind in "fd" else str(int(v)) for v in p.values]
        # values wrap over lines after a comma, eight to a line, as exporters write them
        body = ",\n\t\t".join(",".join(vals[i:i + 8]) for i in range(0, len(vals), 8))
        return f"*{len(p.values)} {{\n\t\ta: {body}\n\t}}"
    if isinstance(p, C):
        return "T" if p.v else "F"
    if isinstance(p, (L, Y)) or (isinstance(p, int) and not isinstance(p, bool)):
        return str(int(p))
    if isinstance(p, bool):
        return "T" if p else "F"
    if isinstance(p, (F, float)):
        return repr(float(p))
    if isinstance(p, (bytes, bytearray)):
        p = bytes(p).decode("latin-1")
    if "\x00\x01" in p:  # binary object name "Name\x00\x01Class" -> "Class::Name"
        name, cls = p.split("\x00\x01")
        p = f"{cls}::{name}"
    return '"' + p + '"'


def ascii_node(node, depth):
    pad = "\t" * depth
    vals = ", ".join(ascii_value(p) for p in node.props)
    out = f"{pad}{node.name}: {vals}"
    if node.kids:
        out += " {\n" + "".join(ascii_node(k, depth + 1) for k in node.kids) + f"{pad}}}\n"
    else:
        out += "\n"
    return out


def fbx_ascii(top, version=7400):
    out = f"; FBX {version // 1000}.{version // 100 % 10}.{version % 100} project file\n; written by the tests\n\n"
    return (out + "".join(ascii_node(n, 0) for n in top)).encode("latin-1")


# ------------------------------------------------------------ scene helpers
def obj_name(name, cls):
    return name.encode() + b"\x00\x01" + cls.encode()


def layer_element(kind, data_name, values, mapping, ref, index_name=None, index=None, typed=0, compress=False):
    kids = [Node("Version", 101), Node("Name", ""), Node("MappingInformationType", mapping),
            Node("ReferenceInformationType", ref), Node(data_name, Arr("d", values, compress))]
    if index is not None:
        kids.append(Node(index_name, Arr("i", index, compress)))
    return Node(kind, typed, kids=kids)


def material_layer(indices, mapping="ByPolygon", ref="IndexToDirect"):
    return Node("LayerElementMaterial", 0, kids=[Node("Version", 101), Node("Name", ""),
                                                 Node("MappingInformationType", mapping),
                                                 Node("ReferenceInformationType", ref),
                                                 Node("Materials", Arr("i", indices))])


def layer(*types):
    kids = [Node("Version", 100)]
    for t in types:
        kids.append(Node("LayerElement", kids=[Node("Type", t), Node("TypedIndex", 0)]))
    return Node("Layer", 0, kids=kids)


def geometry(gid, name, points, polys, normals=None, uvs=None, materials=None, compress=False):
    """points: [(x, y, z)]; polys: lists of control-point indices; normals /
    uvs: (values, mapping, ref, index or None); materials: per-polygon
    indices (ByPolygon) or ('AllSame', [m])."""
    pvi = []
    for p in polys:
        pvi += list(p[:-1]) + [-p[-1] - 1]
    kids = [Node("Vertices", Arr("d", [c for pt in points for c in pt], compress)),
            Node("PolygonVertexIndex", Arr("i", pvi, compress)), Node("GeometryVersion", 124)]
    types = []
    if normals is not None:
        vals, mapping, ref, idx = normals
        kids.append(layer_element("LayerElementNormal", "Normals", vals, mapping, ref, "NormalsIndex", idx,
                                  compress=compress))
        types.append("LayerElementNormal")
    if uvs is not None:
        vals, mapping, ref, idx = uvs
        kids.append(layer_element("LayerElementUV", "UV", vals, mapping, ref, "UVIndex", idx, compress=compress))
        types.append("LayerElementUV")
    if materials is not None:
        if isinstance(materials, tuple):
            kids.append(material_layer(materials[1], mapping=materials[0], ref="IndexToDirect"))
        else:
            kids.append(material_layer(materials))
        types.append("LayerElementMaterial")
    if types:
        kids.append(layer(*types))
    return Node("Geometry", L(gid), obj_name(name, "Geometry"), "Mesh", kids=kids)


def model(mid, name, props=()):
    return Node("Model", L(mid), obj_name(name, "Model"), "Mesh",
                kids=[Node("Version", 232), props70(*props)])


def material(mid, name, props=(), shading="phong"):
    return Node("Material", L(mid), obj_name(name, "Material"), "",
                kids=[Node("Version", 102), Node("ShadingModel", shading), Node("MultiLayer", 0), props70(*props)])


def texture(tid, name, relative, absolute="C:/elsewhere/x.png"):
    return Node("Texture", L(tid), obj_name(name, "Texture"), "",
                kids=[Node("Type", "TextureVideoClip"), Node("FileName", absolute), Node("RelativeFilename", relative)])


def oo(src, dst):
    return Node("C", "OO", L(src), L(dst))


def op(src, dst, prop):
    return Node("C", "OP", L(src), L(dst), prop)


def document(objects, connections, templates=(), version=7400, ascii=False):
    """templates: [(object type, template name, [P entries])]."""
    defs = [Node("Version", 100)]
    for otype, tname, entries in templates:
        defs.append(Node("ObjectType", otype, kids=[Node("Count", 1),
                                                   Node("PropertyTemplate", tname, kids=[props70(*entries)])]))
    top = [Node("FBXHeaderExtension", kids=[Node("FBXHeaderVersion", 1003), Node("FBXVersion", version)]),
           Node("GlobalSettings", kids=[Node("Version", 1000), props70(P("UpAxis", "int", "Integer", "", 1))]),
           Node("Definitions", kids=defs), Node("Objects", kids=list(objects)),
           Node("Connections", kids=list(connections))]
    return fbx_ascii(top, version) if ascii else fbx_bytes(top, version)

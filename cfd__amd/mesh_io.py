"""Mesh ingestion for the voxelizer (src/geom/parse.js): STL (binary, found
by arithmetic, and ASCII), OBJ, PLY, glTF 2.0 (.gltf with data URIs) and GLB.
Returns float32 positions and uint32 triangle indices, plus the list of
tolerated defects (`avisos`); `ler` picks the format from the content.
"""
from __future__ import annotations

import base64
import json
import re
import struct

import numpy as np

TRI_BYTES = 50      # 12 floats + uint16 attribute
HEADER_BYTES = 84   # 80-byte header + uint32 count


def _stl_binary_count(buf: bytes):
    """Binary STL detection by size arithmetic (parse.js:59-98): the count at
    offset 80 closing the file exactly, else a count at 76..84 that closes it
    exactly (a 79-byte header exists in the wild), else a count at 80 with the
    file at most one triangle short or 4 bytes long."""
    n_bytes = len(buf)
    if n_bytes < HEADER_BYTES:
        return None

    def tentar(base):
        if base + 4 > n_bytes:
            return None
        n = struct.unpack_from("<I", buf, base)[0]
        if n == 0 or n > 200e6:
            return None
        return n, base + 4 + n * TRI_BYTES

    padrao = tentar(80)
    if padrao and padrao[1] == n_bytes:
        return {"n": padrao[0], "base": 84, "truncado": 0, "deslocado": 0}
    for base in range(76, 85):
        if base == 80:
            continue
        t = tentar(base)
        if t and t[1] == n_bytes:
            return {"n": t[0], "base": base + 4, "truncado": 0, "deslocado": base - 80}
    if padrao:
        folga = n_bytes - padrao[1]
        disponivel = (n_bytes - HEADER_BYTES) // TRI_BYTES
        if -TRI_BYTES <= folga <= 4 and disponivel > 0:
            n = min(padrao[0], disponivel)
            return {"n": n, "base": 84, "truncado": padrao[0] - n, "deslocado": 0}
    return None


def parseSTL(buf: bytes, avisos: list | None = None) -> dict:
    """parse.js:141-175"""
    avisos = [] if avisos is None else avisos
    b = _stl_binary_count(buf)
    if b:
        n, base = b["n"], b["base"]
        rec = np.frombuffer(buf, dtype=np.uint8, count=n * TRI_BYTES, offset=base)
        rec = rec.reshape(n, TRI_BYTES)
        pos = rec[:, 12:48].copy().view("<f4").reshape(-1).astype(np.float32)
        if b["deslocado"]:
            avisos.append(f"{base - 4}-byte header instead of 80: triangle count at offset "
                          f"{base - 4}, where the size arithmetic closes exactly")
        if b["truncado"]:
            avisos.append(f"file ends early: {b['truncado']} declared triangle(s) missing")
        head = buf[:min(80, base - 4)].split(b"\0")[0].decode("latin1").strip()
        if head.lower().startswith("solid"):
            avisos.append('header starts with "solid" but the file is binary')
        return {"positions": pos, "indices": np.arange(n * 3, dtype=np.uint32),
                "formato": "stl-binario", "nome": head, "deslocado": b["deslocado"],
                "avisos": avisos}
    text = buf.decode("utf-8", errors="replace")
    if not re.match(r"^\s*solid", text, re.I):
        raise ValueError("not STL: the binary count does not close with the file size and "
                         'the content does not start with "solid"')
    nums = re.findall(r"vertex\s+(-?[\d.eE+-]+)\s+(-?[\d.eE+-]+)\s+(-?[\d.eE+-]+)", text)
    if not nums:
        raise ValueError('ASCII STL without any "vertex"')
    v = [float(x) for t in nums for x in t]
    if len(v) % 9:
        avisos.append(f"{(len(v) % 9) // 3} loose vertex(es) at the end, dropped")
        v = v[: len(v) - len(v) % 9]
    pos = np.array(v, dtype=np.float32)
    m = re.match(r"^\s*solid[ \t]*(.*)$", text, re.M)
    return {"positions": pos, "indices": np.arange(pos.size // 3, dtype=np.uint32),
            "formato": "stl-ascii", "nome": (m.group(1) if m else "").strip(),
            "deslocado": 0, "avisos": avisos}


def parseOBJ(text: str, avisos: list | None = None) -> dict:
    """parse.js:177-222: 1-based and negative indices, polygon fans."""
    avisos = [] if avisos is None else avisos
    vx: list[float] = []
    tris: list[int] = []
    nome = ""

    def resolver(s):
        try:
            i = int(s)
        except ValueError:
            return -1
        return i - 1 if i > 0 else len(vx) // 3 + i

    for linha in text.split("\n"):
        if linha[:1] == "#":
            continue
        t = linha.strip()
        if not t:
            continue
        if t.startswith("v "):
            p = t.split()
            vx += [float(p[1]), float(p[2]), float(p[3])]
        elif t.startswith("f "):
            idx = [resolver(s.split("/")[0]) for s in t.split()[1:]]
            idx = [i for i in idx if i >= 0]
            for k in range(1, len(idx) - 1):
                tris += [idx[0], idx[k], idx[k + 1]]
        elif t.startswith(("o ", "g ")) and not nome:
            nome = t[2:].strip()
    if not tris:
        raise ValueError("OBJ without triangulable faces")
    nv = len(vx) // 3
    fora = sum(1 for i in tris if i < 0 or i >= nv)
    if fora:
        avisos.append(f"{fora} face index(es) out of range, dropped")
    return {"positions": np.array(vx, dtype=np.float32),
            "indices": np.array([i for i in tris if 0 <= i < nv], dtype=np.uint32),
            "formato": "obj", "nome": nome, "avisos": avisos}


_PLY_T = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2",
          "int16": "i2", "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4",
          "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4", "double": "f8",
          "float64": "f8"}


def parsePLY(buf: bytes, avisos: list | None = None) -> dict:
    """parse.js:224-330: ascii and binary little/big endian."""
    avisos = [] if avisos is None else avisos
    head = buf[:65536].decode("latin1")
    fim = head.find("end_header")
    if fim < 0:
        raise ValueError("PLY without end_header in the first 64 KB")
    data_start = head.index("\n", fim) + 1
    formato, els = None, []
    for line in re.split(r"\r?\n", head[:fim]):
        p = line.strip().split()
        if not p:
            continue
        if p[0] == "format":
            formato = p[1]
        elif p[0] == "element":
            els.append({"nome": p[1], "n": int(p[2]), "props": []})
        elif p[0] == "property" and els:
            if p[1] == "list":
                els[-1]["props"].append({"lista": True, "tt": p[2], "ti": p[3], "nome": p[4]})
            else:
                els[-1]["props"].append({"lista": False, "t": p[1], "nome": p[2]})
    ev = next((e for e in els if e["nome"] == "vertex"), None)
    ef = next((e for e in els if e["nome"] == "face"), None)
    if ev is None:
        raise ValueError('PLY without a "vertex" element')
    ix = next(k for k, q in enumerate(ev["props"]) if q["nome"] == "x")
    pos = np.zeros(ev["n"] * 3, dtype=np.float32)
    tris: list[int] = []
    if formato == "ascii":
        tok = buf.decode("latin1")[data_start:].split()
        p = 0
        for i in range(ev["n"]):
            for k in range(len(ev["props"])):
                v = float(tok[p])
                p += 1
                if ix <= k < ix + 3:
                    pos[i * 3 + k - ix] = v
        if ef:
            for _ in range(ef["n"]):
                n = int(float(tok[p]))
                p += 1
                face = [int(float(t)) for t in tok[p:p + n]]
                p += n
                for k in range(1, n - 1):
                    tris += [face[0], face[k], face[k + 1]]
    else:
        end = "<" if formato == "binary_little_endian" else ">"
        p = data_start

        def ler(t):
            nonlocal p
            dt = np.dtype(end + _PLY_T[t])
            v = np.frombuffer(buf, dtype=dt, count=1, offset=p)[0]
            p += dt.itemsize
            return v

        for i in range(ev["n"]):
            for k, q in enumerate(ev["props"]):
                v = ler(q["t"])
                if ix <= k < ix + 3:
                    pos[i * 3 + k - ix] = v
        if ef:
            pl = next((q for q in ef["props"] if q["lista"]), None)
            for _ in range(ef["n"]):
                for q in ef["props"]:
                    if q is not pl:
                        ler(q["t"])
                        continue
                    n = int(ler(q["tt"]))
                    face = [int(ler(q["ti"])) for _ in range(n)]
                    for k in range(1, n - 1):
                        tris += [face[0], face[k], face[k + 1]]
    if not tris:
        raise ValueError("PLY without faces")
    return {"positions": pos, "indices": np.array(tris, dtype=np.uint32), "formato": "ply",
            "nome": "", "avisos": avisos}


def extentos(positions) -> dict:
    """Bounding box of the finite vertices and the count of non-finite ones
    (parse.js:588-605)."""
    p = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    ok = np.isfinite(p.sum(axis=1))
    q = p[ok]
    if len(q):
        lo, hi = q.min(axis=0), q.max(axis=0)
    else:
        lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    return {"min": lo.tolist(), "max": hi.tolist(), "tamanho": (hi - lo).tolist(),
            "centro": ((lo + hi) / 2).tolist(), "naoFinitos": int((~ok).sum())}


# ---------------------------------------------------------------- glTF / GLB

_COMPONENTE = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16,
               5125: np.uint32, 5126: np.float32}
_N_COMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT4": 16}


def _mat4_mul(a, b):
    """Column-major 4x4 product a*b, float64 in the reference's summation
    order (parse.js:348-357)."""
    o = [0.0] * 16
    for c in range(4):
        for r in range(4):
            o[c * 4 + r] = (a[r] * b[c * 4] + a[4 + r] * b[c * 4 + 1] +
                            a[8 + r] * b[c * 4 + 2] + a[12 + r] * b[c * 4 + 3])
    return o


def _trs(t, r, s):
    """Translation, unit quaternion (x, y, z, w), scale -> column-major
    matrix (parse.js:359-371)."""
    x, y, z, w = (float(v) for v in r)
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz = x * x2, x * y2, x * z2
    yy, yz, zz = y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    s = [float(v) for v in s]
    return [(1 - (yy + zz)) * s[0], (xy + wz) * s[0], (xz - wy) * s[0], 0.0,
            (xy - wz) * s[1], (1 - (xx + zz)) * s[1], (yz + wx) * s[1], 0.0,
            (xz + wy) * s[2], (yz - wx) * s[2], (1 - (xx + yy)) * s[2], 0.0,
            float(t[0]), float(t[1]), float(t[2]), 1.0]


def _montar_gltf(doc: dict, buffers: list, avisos: list) -> dict:
    """Walk the scene, apply node transforms and concatenate every triangle
    primitive into one mesh (parse.js:388-498)."""
    pos_acc, idx_acc = [], []
    total_v = 0

    def dados(i_acc):
        acc = doc["accessors"][i_acc]
        dt = _COMPONENTE.get(acc.get("componentType"))
        n = _N_COMP.get(acc.get("type"))
        if dt is None or n is None:
            raise ValueError(f"accessor {i_acc}: unsupported type")
        count = int(acc["count"])
        if acc.get("bufferView") is None:
            return np.zeros(count * n, dtype=dt)     # empty sparse accessor
        bv = doc["bufferViews"][acc["bufferView"]]
        buf = buffers[bv.get("buffer", 0)] if bv.get("buffer", 0) < len(buffers) else None
        if buf is None:
            raise ValueError("bufferView points to a missing buffer")
        base = bv.get("byteOffset", 0) + acc.get("byteOffset", 0)
        tam = np.dtype(dt).itemsize
        passo = bv.get("byteStride", 0) or 0
        if not passo or passo == tam * n:
            return np.frombuffer(buf, dtype=np.dtype(dt).newbyteorder("<"), count=count * n,
                                 offset=base).astype(dt)
        # interleaved: de-interleave
        out = np.empty(count * n, dtype=dt)
        mv = memoryview(buf)
        for i in range(count):
            for k in range(n):
                o = base + i * passo + k * tam
                out[i * n + k] = np.frombuffer(mv[o:o + tam], dtype=np.dtype(dt).newbyteorder("<"))[0]
        return out

    def visitar(i_node, pai):
        nodes = doc.get("nodes") or []
        if i_node >= len(nodes) or nodes[i_node] is None:
            return
        node = nodes[i_node]
        local = node["matrix"] if "matrix" in node else _trs(
            node.get("translation", [0, 0, 0]), node.get("rotation", [0, 0, 0, 1]),
            node.get("scale", [1, 1, 1]))
        M = _mat4_mul(pai, [float(v) for v in local])
        nonlocal total_v
        if node.get("mesh") is not None:
            for prim in doc["meshes"][node["mesh"]].get("primitives", []):
                if prim.get("mode") is not None and prim.get("mode") != 4:
                    avisos.append(f"primitive with mode={prim['mode']} (not TRIANGLES) skipped")
                    continue
                if (prim.get("extensions") or {}).get("KHR_draco_mesh_compression") is not None:
                    avisos.append("Draco-compressed primitive skipped — export without Draco")
                    continue
                i_pos = (prim.get("attributes") or {}).get("POSITION")
                if i_pos is None:
                    continue
                p = dados(i_pos).astype(np.float64).reshape(-1, 3)
                x, y, z = p[:, 0], p[:, 1], p[:, 2]
                t = np.empty_like(p)
                t[:, 0] = M[0] * x + M[4] * y + M[8] * z + M[12]
                t[:, 1] = M[1] * x + M[5] * y + M[9] * z + M[13]
                t[:, 2] = M[2] * x + M[6] * y + M[10] * z + M[14]
                pos_acc.append(t.astype(np.float32).reshape(-1))
                nv = p.shape[0]
                if prim.get("indices") is not None:
                    src = dados(prim["indices"])
                    ind = (src.astype(np.uint64) + total_v).astype(np.uint32)
                else:
                    ind = np.arange(total_v, total_v + nv, dtype=np.uint32)
                idx_acc.append(ind)
                total_v += nv
        for filho in node.get("children", []) or []:
            visitar(filho, M)

    cenas = doc.get("scenes")
    cena = cenas[doc.get("scene", 0)] if cenas else None
    raizes = (cena or {}).get("nodes")
    if raizes is None:
        raizes = list(range(len(doc.get("nodes") or [])))
    ident = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]
    for r in raizes:
        visitar(r, ident)
    if not pos_acc:
        raise ValueError("glTF without triangle geometry")
    return {"positions": np.concatenate(pos_acc).astype(np.float32),
            "indices": np.concatenate(idx_acc).astype(np.uint32)}


def _data_uri(uri: str) -> bytes:
    return base64.b64decode(uri.split(",", 1)[1])


def parseGLB(buf: bytes, avisos: list | None = None) -> dict:
    """Binary glTF (parse.js:500-532): a JSON chunk and BIN chunks."""
    avisos = [] if avisos is None else avisos
    if len(buf) < 12 or struct.unpack_from("<I", buf, 0)[0] != 0x46546C67:
        raise ValueError('no "glTF" magic')
    versao = struct.unpack_from("<I", buf, 4)[0]
    if versao != 2:
        avisos.append(f"GLB version {versao}; the reader is for glTF 2.0")
    doc, bins = None, []
    p = 12
    while p + 8 <= len(buf):
        tam, tipo = struct.unpack_from("<II", buf, p)
        dados = buf[p + 8:p + 8 + tam]
        if tipo == 0x4E4F534A:
            doc = json.loads(dados.decode("utf-8"))
        elif tipo == 0x004E4942:
            bins.append(dados)
        p += 8 + tam + ((4 - (tam % 4)) % 4)
    if doc is None:
        raise ValueError("GLB without a JSON chunk")
    buffers = []
    for i, b in enumerate(doc.get("buffers") or []):
        if b.get("uri") is None:
            buffers.append(bins[0] if bins else (bins[i] if i < len(bins) else None))
        elif b["uri"].startswith("data:"):
            buffers.append(_data_uri(b["uri"]))
        else:
            avisos.append(f'external buffer "{b["uri"]}" could not be loaded')
            buffers.append(None)
    if not buffers:
        buffers.append(bins[0] if bins else None)
    cenas = doc.get("scenes")
    nome = ((cenas[doc.get("scene", 0)] if cenas else {}) or {}).get("name") or \
        (doc.get("asset") or {}).get("generator", "")
    return {**_montar_gltf(doc, buffers, avisos), "formato": "glb", "nome": nome,
            "avisos": avisos}


def parseGLTF(texto: str, avisos: list | None = None) -> dict:
    """JSON glTF with embedded data URIs (parse.js:534-545)."""
    avisos = [] if avisos is None else avisos
    doc = json.loads(texto)
    buffers = []
    for b in doc.get("buffers") or []:
        if (b.get("uri") or "").startswith("data:"):
            buffers.append(_data_uri(b["uri"]))
        else:
            avisos.append(f'external buffer "{b.get("uri")}" could not be loaded — use .glb, '
                          "which embeds the data")
            buffers.append(None)
    cenas = doc.get("scenes")
    nome = ((cenas[doc.get("scene", 0)] if cenas else {}) or {}).get("name", "")
    return {**_montar_gltf(doc, buffers, avisos), "formato": "gltf", "nome": nome,
            "avisos": avisos}


def ler(buf: bytes, nomeArquivo: str = "") -> dict:
    """Any supported format, chosen by content with the file name as the
    tie-breaker (parse.js:555-586)."""
    avisos: list = []
    m = re.search(r"\.([a-z0-9]+)$", nomeArquivo, re.I)
    ext = m.group(1).lower() if m else ""
    magia = struct.unpack_from("<I", buf, 0)[0] if len(buf) >= 4 else 0
    inicio = buf[:1024].decode("latin-1")
    if magia == 0x46546C67:
        r = parseGLB(buf, avisos)
    elif re.match(r"^ply\s", inicio):
        r = parsePLY(buf, avisos)
    elif ext == "gltf" or re.match(r'^\s*\{[\s\S]*"asset"', inicio):
        r = parseGLTF(buf.decode("utf-8"), avisos)
    elif ext == "obj" or re.search(r"^\s*(v\s|vn\s|vt\s|#|mtllib|o\s|g\s)", inicio, re.M):
        r = parseSTL(buf, avisos) if _stl_binary_count(buf) else \
            parseOBJ(buf.decode("utf-8", errors="replace"), avisos)
    else:
        r = parseSTL(buf, avisos)
    n_tri = len(r["indices"]) // 3
    if not n_tri:
        raise ValueError(f"{r['formato']}: no triangles")
    return {**r, "avisos": avisos, "nTriangulos": n_tri, "nVertices": len(r["positions"]) // 3}

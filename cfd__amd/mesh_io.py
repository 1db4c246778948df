"""Mesh ingestion for the voxelizer (src/geom/parse.js): STL (binary, found
by arithmetic, and ASCII), OBJ and PLY.  Returns float32 positions and
uint32 triangle indices, plus the list of tolerated defects (`avisos`).
"""
from __future__ import annotations

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

"""OBJ and camera.cfg restatement -- TEST INFRASTRUCTURE ONLY (the checker
for sg_obj_load / sg_camera_load; nothing in the product imports it).

Follows FWObject::read (src/Loaders/FWObjectReader.cpp:181-502) for what
decides triangles and CameraManager::read (src/Loaders/CameraManager.cpp:93-165).
Numbers are converted with the C library's own strtod / strtol / strtof
(through ctypes), which is what the reference's reader calls, so coordinate
bits are those of `(float)strtod(text)`.

Pinned by: tests/test_obj.py hand-written OBJ cases with known triangles
(fan triangulation, v/vt/vn reference forms, rejected lines, the 'fo'
prefix, comments), the ruins scene's counts (38,665 triangles, 24,625
vertices) and its camera.cfg values.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_libc = C.CDLL(None)
_libc.strtod.restype = C.c_double
_libc.strtod.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]
_libc.strtol.restype = C.c_long
_libc.strtol.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), C.c_int]
_libc.strtof.restype = C.c_float
_libc.strtof.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]

_WS = b" \t\n\v\f\r"


class _Cursor:
    """a position in one line (bytes), with the reader's primitives"""

    def __init__(self, line: bytes):
        self.b = line
        self.i = 0

    def ch(self) -> int:
        return self.b[self.i] if self.i < len(self.b) else 0

    def skip_ws(self):
        while self.i < len(self.b) and self.b[self.i] in _WS:
            self.i += 1

    def _call(self, fn, *extra):
        buf = C.create_string_buffer(self.b[self.i:])
        end = C.c_char_p()
        v = fn(buf, C.byref(end), *extra)
        used = C.cast(end, C.c_void_p).value - C.addressof(buf)
        return v, used

    def strtod(self):
        v, used = self._call(_libc.strtod)
        self.i += used
        return v, used > 0

    def strtol(self, advance=True):
        v, used = self._call(_libc.strtol, 10)
        if advance:
            self.i += used
        return v, used


def _vertex(cur: _Cursor, verts: list) -> bool:
    """FWObjectReader.cpp:224-271"""
    c = chr(cur.ch()).lower() if cur.ch() else ""
    if c == "n":
        kind = 1
        cur.i += 1
    elif c == "t":
        kind = 2
        cur.i += 1
    elif cur.ch() and cur.ch() in _WS:
        kind = 0
    else:
        return False
    vals = []
    while True:
        cur.skip_ws()
        if cur.ch() == 0:
            break
        v, ok = cur.strtod()
        if not ok:
            return False
        if len(vals) >= 3:
            return False
        vals.append(np.float32(v))     # (float)strtod: round the double once more
    if (kind != 2 and len(vals) != 3) or (kind == 2 and len(vals) < 2):
        return False
    if kind == 0:
        verts.append(vals)
    return True


def _face(cur: _Cursor) -> list | None:
    """FWObjectReader.cpp:273-364: vertex indices of one face, or None"""
    idx = []
    cur.skip_ws()
    while True:
        if len(idx) >= 4096:
            return None
        v, used = cur.strtol()
        if not used:
            return None
        idx.append(v - 1)
        cur.skip_ws()
        if cur.ch() == ord("/"):
            cur.i += 1
            cur.skip_ws()
            if cur.ch() != ord("/"):
                _, used = cur.strtol()
                if not used:
                    return None
            cur.skip_ws()
            if cur.ch() == ord("/"):
                cur.i += 1
                cur.skip_ws()
                _, used = cur.strtol(advance=False)
                j = cur.i + used
                while j < len(cur.b) and cur.b[j] in _WS:
                    j += 1
                nxt = cur.b[j] if j < len(cur.b) else 0
                if (nxt and chr(nxt).isdigit()) or ((nxt == 0 or nxt == ord("#")) and used):
                    if not used:
                        return None
                    cur.i += used
        cur.skip_ws()
        if cur.ch() == 0:
            break
    return idx if len(idx) >= 3 else None


def read_obj(path: str):
    """-> (soa9 float32 (9, T), n_vertices, n_faces, n_bad_lines)"""
    verts, tris = [], []
    faces = bad = 0
    with open(path, "rb") as f:
        data = f.read()
    lines = data.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    for line in lines:
        cur = _Cursor(line)
        cur.skip_ws()
        c = chr(cur.ch()).lower() if cur.ch() else ""
        ok = True
        if c in ("", "o", "g", "s", "#"):
            pass
        elif c == "v":
            cur.i += 1
            ok = _vertex(cur, verts)
        elif c == "f":
            cur.i += 1
            if cur.ch() and chr(cur.ch()).lower() == "o":
                cur.i += 1
            idx = _face(cur)
            ok = idx is not None
            if ok:
                faces += 1
                for k in range(1, len(idx) - 1):
                    tris.append((idx[0], idx[k], idx[k + 1]))
        else:
            rest = line[cur.i:]
            if rest[:6].lower() in (b"usemtl", b"mtllib"):
                ok = rest[6:].strip(_WS) != b""
        if not ok:
            bad += 1
    V = np.array(verts, np.float32).reshape(-1, 3)
    T = np.array(tris, np.int64).reshape(-1, 3)
    if T.size and (T.min() < 0 or T.max() >= len(V)):
        raise ValueError("face index outside the vertex list")
    soa9 = np.empty((9, len(T)), np.float32)
    for c in range(3):
        soa9[3 * c:3 * c + 3] = V[T[:, c]].T
    return soa9, len(V), faces, bad


def read_camera(path: str) -> dict:
    """CameraManager::read: lines holding '#' are skipped; floats by strtof"""
    cam = {"position": [0.0] * 3, "orientation": [0.0] * 3, "up": [0.0] * 3, "rotation": [0.0] * 3,
           "resX": 0, "resY": 0, "FOV": 0.0}
    with open(path, "rb") as f:
        for line in f.read().split(b"\n"):
            if b"#" in line:
                continue
            tok = line.split()
            if not tok:
                continue
            key = tok[0].decode()
            if key in ("position", "orientation", "up", "rotation"):
                cam[key] = [float(_libc.strtof(t, None)) for t in tok[1:4]]
            elif key in ("resX", "resY"):
                cam[key] = int(tok[1])
            elif key == "FOV":
                cam[key] = float(_libc.strtof(tok[1], None))
    return cam

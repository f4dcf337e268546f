"""Scene ingestion and frame files.

read_obj follows hussar::WavefrontFile (libhussar/include/hussar/io/wavefront.h:12-64): only
``v x y z`` and ``f a b c`` lines are used; face indices are 1-based and relative to the
vertices of the file being read (they are offset by the mesh's vertex count before the read);
a face uses its first three indices; ``#`` comment lines, ``s`` lines and empty lines are
skipped; any other command is reported and skipped. Where the reference's stream parsing
would stop making progress (a ``v/vt/vn`` face token), this reader raises instead.

The .SIM files are the raw layout of examples/src/utils.h:29-40: each frame's complex64
samples in index order ([chirp][sample][channel], radar.h:487-494), frames appended.
"""
from __future__ import annotations

import sys
from typing import Optional

import numpy as np

from .types import FrameConfig, TriangleMesh


def read_obj(path: str, mesh: Optional[TriangleMesh] = None, log=sys.stdout) -> TriangleMesh:
    mesh = mesh if mesh is not None else TriangleMesh()
    offset = int(mesh.vertexBuffer.shape[0])
    verts, faces = [], []
    with open(path, "r") as f:
        for lineno, line in enumerate(f, 1):
            if line.startswith("#"):
                continue
            tok = line.split()
            if not tok:
                continue
            cmd = tok[0]
            if cmd == "v":
                verts.append([np.float32(x) for x in tok[1:4]])
            elif cmd == "f":
                try:
                    idx = [int(x) for x in tok[1:4]]
                except ValueError:
                    raise ValueError(f"{path}:{lineno}: face indices must be plain integers "
                                     f"(no v/vt/vn syntax): {line.strip()!r}") from None
                if len(idx) < 3:
                    raise ValueError(f"{path}:{lineno}: a face needs three indices")
                faces.append([i + offset - 1 for i in idx])
            elif cmd == "s":
                pass
            else:
                print(f"unsupported wavefront command: {cmd}", file=log)
    if verts or faces:
        v = np.asarray(verts, np.float32).reshape(-1, 3)
        i = np.asarray(faces, np.int64).reshape(-1, 3)
        # TriangleMesh.append offsets indices by the current vertex count itself
        mesh.append(v, (i - offset).astype(np.int32))
    return mesh


def write_frame(file, frame: np.ndarray) -> None:
    """writeFrameToFile (examples/src/utils.h:29-35): append one frame as raw complex64."""
    np.ascontiguousarray(frame, np.complex64).tofile(file)
    file.flush()


def save_frame(frame: np.ndarray, path: str) -> None:
    """saveFrame (examples/src/utils.h:37-40): one frame to ``path + '.SIM'``."""
    with open(path + ".SIM", "wb") as f:
        write_frame(f, frame)


def read_sim(path: str, config: FrameConfig) -> np.ndarray:
    """All frames of a .SIM file as complex64 [frame, chirp, sample, channel]."""
    a = np.fromfile(path, np.complex64)
    per = config.chirpCount * config.samplesPerChirp * config.channelCount
    if a.size % per:
        raise ValueError(f"{path}: {a.size} samples is not a whole number of {per}-sample frames")
    return a.reshape(-1, config.chirpCount, config.samplesPerChirp, config.channelCount)

"""Image files: raw interleaved complex64, row-major [ny][nx] -- the bytes the
reference's SAR_Processor::saveResult writes (code/SAR_BP/CUDA/BPA_CUDA.cpp:92-100)
-- plus a JSON sidecar (<path>.json) with shape, dtype, layout, SHA-256 and
optional grid / radar metadata."""
from __future__ import annotations

import hashlib
import json
import os
import struct
import zlib
from typing import Optional

import numpy as np


def save_image(path: str, image: np.ndarray, meta: Optional[dict] = None) -> dict:
    img = np.ascontiguousarray(image, dtype=np.complex64)
    if img.ndim != 2:
        raise ValueError("image must be 2-D [ny][nx]")
    raw = img.tobytes()
    with open(path, "wb") as f:
        f.write(raw)
    side = {"format": "raw complex64 interleaved (re, im), row-major [ny][nx]",
            "ny": int(img.shape[0]), "nx": int(img.shape[1]), "dtype": "complex64",
            "bytes": len(raw), "sha256": hashlib.sha256(raw).hexdigest()}
    if meta:
        side["meta"] = meta
    with open(path + ".json", "w") as f:
        json.dump(side, f, indent=1)
    return side


def load_image(path: str, shape: Optional[tuple] = None, verify: bool = True) -> np.ndarray:
    """Reads a raw complex64 image.  The shape comes from the sidecar if present
    (verified against its SHA-256), else from `shape`."""
    side = None
    if os.path.exists(path + ".json"):
        with open(path + ".json") as f:
            side = json.load(f)
    with open(path, "rb") as f:
        raw = f.read()
    if side is not None:
        if verify and hashlib.sha256(raw).hexdigest() != side["sha256"]:
            raise ValueError(f"{path}: SHA-256 does not match its sidecar")
        shape = (side["ny"], side["nx"])
    if shape is None:
        raise ValueError(f"{path}: no sidecar, pass shape=(ny, nx)")
    arr = np.frombuffer(raw, dtype=np.complex64)
    if arr.size != shape[0] * shape[1]:
        raise ValueError(f"{path}: {arr.size} samples, expected {shape[0]}x{shape[1]}")
    return arr.reshape(shape).copy()


def magnitude_db(image: np.ndarray, floor_db: float = -60.0) -> np.ndarray:
    """20 log10 |I| normalised to the peak, clipped at floor_db (display)."""
    a = np.abs(image)
    peak = a.max() if a.size else 0.0
    if peak <= 0:
        return np.full(a.shape, floor_db)
    return np.maximum(20.0 * np.log10(np.maximum(a / peak, 1e-30)), floor_db)


def save_png(path: str, image: np.ndarray, floor_db: float = -40.0) -> None:
    """Writes |image| in dB (0 dB = the peak, clipped at floor_db) as an 8-bit
    greyscale PNG, rows top to bottom.  Uses only zlib (no imaging library)."""
    db = magnitude_db(np.asarray(image), floor_db)
    g = np.clip(np.rint((db - floor_db) * (255.0 / -floor_db)), 0, 255).astype(np.uint8)
    h, w = g.shape
    raw = b"".join(b"\x00" + g[r].tobytes() for r in range(h))

    def chunk(kind: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 0, 0, 0, 0)) + \
        chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)

"""Film output: the reference's 8-bit PNG path and an FP32 PFM dump.

FrameBuffer::writeToFile (base/framebuffer.cpp:27-58): NaN -> 0, pow(c, 1/gamma),
clamp to [0,1], *255 truncated to uint8, image flipped vertically on write
(row j = 0 of the film is the bottom row of the picture).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np


def to_srgb8(film: np.ndarray, gamma: float = 2.0) -> np.ndarray:
    c = np.array(film, np.float32, copy=True)
    nan = np.isnan(c)
    c[nan] = 0.0
    with np.errstate(invalid="ignore"):
        c = np.power(np.maximum(c, 0.0), 1.0 / gamma, where=~nan, out=c)
    c = np.clip(c, 0.0, 1.0)
    return (c * 255).astype(np.uint8)[::-1]


def write_png(path: str, film: np.ndarray, gamma: float = 2.0) -> None:
    img = to_srgb8(film, gamma)
    h, w, _ = img.shape
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(h))

    def chunk(tag, data):
        c = struct.pack(">I", len(data)) + tag + data
        return c + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
    png += chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def write_pfm(path: str, film: np.ndarray) -> None:
    """Little-endian colour PFM; PFM rows run bottom to top, as the film's."""
    h, w, _ = film.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(np.ascontiguousarray(film, "<f4").tobytes())


def read_pfm(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = (int(x) for x in f.readline().split())
        scale = float(f.readline())
        dt = "<f4" if scale < 0 else ">f4"
        return np.frombuffer(f.read(), dt).reshape(h, w, 3).astype(np.float32)


def rel_mse(img: np.ndarray, ref: np.ndarray, eps: float = 1e-2) -> float:
    """Relative MSE: mean over pixels and channels of (img-ref)^2 / (ref^2 + eps)."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + eps)))

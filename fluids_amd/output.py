"""Frame files and FPS logs (include/sph.h, sph_frames_* / sph_perf_*): the
output boundary in place of the reference's GL vertex buffer and -o option."""
import ctypes as C
import struct

import numpy as np

from ._lib import check, lib

MAGIC = 0x46485053  # "SPHF"


class FrameWriter:
    def __init__(self, sim, path, positions=True, velocities=False):
        fields = (1 if positions else 0) | (2 if velocities else 0)
        self._h = lib().sph_frames_open(sim._h, str(path).encode(), fields)
        if not self._h:
            raise RuntimeError(lib().sph_last_error().decode())

    def write(self):
        check(lib().sph_frames_write(self._h))

    def close(self) -> int:
        if not getattr(self, "_h", None):
            return 0
        h, self._h = self._h, None
        return check(lib().sph_frames_close(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PerfLog:
    def __init__(self, sim, path):
        self._h = lib().sph_perf_open(sim._h, str(path).encode())
        if not self._h:
            raise RuntimeError(lib().sph_last_error().decode())

    def frame(self):
        check(lib().sph_perf_frame(self._h))

    def close(self) -> int:
        if not getattr(self, "_h", None):
            return 0
        h, self._h = self._h, None
        return check(lib().sph_perf_close(h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_frames(path):
    """Yields dicts (frame, time, pos, vel) from a frame file."""
    with open(path, "rb") as f:
        head = f.read(16)
        if len(head) != 16:
            raise ValueError("not a frame file")
        magic, version, fields, _ = struct.unpack("<4I", head)
        if magic != MAGIC or version != 1:
            raise ValueError("not a version-1 frame file")
        while True:
            rec = f.read(24)
            if not rec:
                return
            if len(rec) != 24:
                raise ValueError("truncated frame header")
            frame, time, n, flds = struct.unpack("<QdII", rec)
            out = {"frame": frame, "time": time, "pos": None, "vel": None}
            for bit, key in ((1, "pos"), (2, "vel")):
                if flds & bit:
                    data = f.read(16 * n)
                    if len(data) != 16 * n:
                        raise ValueError("truncated frame")
                    out[key] = np.frombuffer(data, np.float32).reshape(n, 4).copy()
            yield out

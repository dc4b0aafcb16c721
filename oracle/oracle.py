"""ctypes binding for the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  It loads oracle/build/librdp_oracle.so, building it with make
when it is missing.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "librdp_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


class ScanoutOptions(ctypes.Structure):
    """rdpo_scanout_options (the reference's ScanoutOptions without upscaling)."""
    _fields_ = [(n, ctypes.c_uint) for n in ("crop_overscan_pixels", "crop_left", "crop_right", "crop_top",
                                             "crop_bottom")] + \
               [(n, ctypes.c_int) for n in ("crop_enable", "persist_frame_on_invalid_input", "blend_previous_frame",
                                            "upscale_deinterlacing", "vi_aa", "vi_scale", "vi_serrate",
                                            "vi_dither_filter", "vi_divot_filter", "vi_gamma_dither")]

    @classmethod
    def default(cls, **kw):
        o = cls()
        _load().rdpo_scanout_options_default(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


_other_libs = {}


def _load(path: "str | None" = None):
    """The built oracle library, or with path another build of the same
    sources (the planted-mistake builds of the pin tests)."""
    global _lib
    if path is not None:
        if path not in _other_libs:
            _other_libs[path] = _bind(ctypes.CDLL(path))
        return _other_libs[path]
    if _lib is not None:
        return _lib
    build()
    _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(lib):
    vp, u32, u64, sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t
    lib.rdpo_create.restype = vp
    lib.rdpo_create.argtypes = [u32]
    lib.rdpo_destroy.argtypes = [vp]
    lib.rdpo_rdram.restype = vp
    lib.rdpo_rdram.argtypes = [vp]
    lib.rdpo_hidden_rdram.restype = vp
    lib.rdpo_hidden_rdram.argtypes = [vp]
    lib.rdpo_tmem.restype = vp
    lib.rdpo_tmem.argtypes = [vp]
    lib.rdpo_process.restype = sz
    lib.rdpo_process.argtypes = [vp, vp, sz]
    lib.rdpo_flush.argtypes = [vp]
    lib.rdpo_scanout_ex.restype = ctypes.c_int
    lib.rdpo_scanout_ex.argtypes = [vp, ctypes.POINTER(ScanoutOptions), vp, ctypes.POINTER(ctypes.c_uint),
                                    ctypes.POINTER(ctypes.c_uint)]
    lib.rdpo_scanout_options_default.argtypes = [ctypes.POINTER(ScanoutOptions)]
    lib.rdpo_begin_vi_register_per_scanline.argtypes = [vp, u32]
    lib.rdpo_set_vi_register_for_scanline.argtypes = [vp, u32, u32]
    lib.rdpo_latch_vi_register_for_scanline.argtypes = [vp, ctypes.c_uint]
    lib.rdpo_end_vi_register_per_scanline.argtypes = [vp]
    lib.rdpo_pass_ends.restype = sz
    lib.rdpo_pass_ends.argtypes = [vp, vp, sz]
    lib.rdpo_set_vi_register.argtypes = [vp, ctypes.c_uint, u32]
    lib.rdpo_scanout.restype = ctypes.c_int
    lib.rdpo_scanout.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_uint), ctypes.POINTER(ctypes.c_uint)]
    lib.rdpo_fragment_count.restype = u64
    lib.rdpo_fragment_count.argtypes = [vp]
    lib.rdpo_fragment_counts.argtypes = [vp, vp]
    lib.rdpo_primitive_count.restype = u64
    lib.rdpo_primitive_count.argtypes = [vp]
    lib.rdpo_state_digest.restype = u64
    lib.rdpo_state_digest.argtypes = [vp]
    lib.rdpo_command_length.restype = ctypes.c_uint
    lib.rdpo_command_length.argtypes = [u32]
    lib.rdpo_enable_upscaling.restype = ctypes.c_int
    lib.rdpo_enable_upscaling.argtypes = [vp, ctypes.c_uint]
    lib.rdpo_enable_super_sampled_readback.restype = ctypes.c_int
    lib.rdpo_enable_super_sampled_readback.argtypes = [vp, ctypes.c_int]
    lib.rdpo_upscaled_rdram.restype = vp
    lib.rdpo_upscaled_rdram.argtypes = [vp]
    lib.rdpo_upscaled_hidden_rdram.restype = vp
    lib.rdpo_upscaled_hidden_rdram.argtypes = [vp]
    lib.rdpo_blender_divider_table.argtypes = [vp]
    lib.rdpo_gamma_table.argtypes = [vp]
    lib.rdpo_perspective_table.argtypes = [vp]
    return lib


class Oracle:
    """One emulated RDP: RDRAM, hidden RDRAM, TMEM and register state."""

    def __init__(self, rdram_size: int = 8 << 20, lib_path: "str | None" = None):
        self._lib = _load(lib_path)
        self._h = self._lib.rdpo_create(rdram_size)
        if not self._h:
            raise RuntimeError("rdpo_create failed")
        self.rdram_size = rdram_size
        buf = (ctypes.c_uint8 * rdram_size).from_address(self._lib.rdpo_rdram(self._h))
        self.rdram = np.frombuffer(buf, dtype=np.uint32)
        hbuf = (ctypes.c_uint8 * (rdram_size // 2)).from_address(self._lib.rdpo_hidden_rdram(self._h))
        self.hidden = np.frombuffer(hbuf, dtype=np.uint8)
        tbuf = (ctypes.c_uint16 * 2048).from_address(self._lib.rdpo_tmem(self._h))
        self.tmem = np.frombuffer(tbuf, dtype=np.uint16)

    def enable_upscaling(self, factor: int):
        """Render qualifying passes a second time at factor (2, 4 or 8) times
        the resolution (before the first command).  upscaled_rdram /
        upscaled_hidden then view the factor^2 sample planes: (planes,
        rdram_size // 4) words and (planes, rdram_size // 2) bytes, plane
        sy * factor + sx holding sample (sx, sy) of each pixel."""
        if self._lib.rdpo_enable_upscaling(self._h, factor) != 0:
            raise ValueError(f"upscaling factor {factor} refused")
        planes = factor * factor
        n = self.rdram_size
        buf = (ctypes.c_uint8 * (planes * n)).from_address(self._lib.rdpo_upscaled_rdram(self._h))
        self.upscaled_rdram = np.frombuffer(buf, dtype=np.uint32).reshape(planes, n // 4)
        hbuf = (ctypes.c_uint8 * (planes * (n // 2))).from_address(self._lib.rdpo_upscaled_hidden_rdram(self._h))
        self.upscaled_hidden = np.frombuffer(hbuf, dtype=np.uint8).reshape(planes, n // 2)
        self.upscaling = factor

    def enable_super_sampled_readback(self, dither: bool = False):
        """After enable_upscaling: passes render at scale only and resolve
        into native RDRAM (the reference's SUPER_SAMPLED_READ_BACK flag)."""
        if self._lib.rdpo_enable_super_sampled_readback(self._h, int(dither)) != 0:
            raise ValueError("super-sampled readback needs upscaling, before the first command")

    def close(self):
        if self._h:
            self._lib.rdpo_destroy(self._h)
            self._h = None
            self.rdram = self.hidden = self.tmem = None
            self.upscaled_rdram = self.upscaled_hidden = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def process(self, words: np.ndarray) -> int:
        """Decode and render commands.  Pixels of a render pass that is still
        open stay in the pass cache until flush() or a command that ends the
        pass (SyncFull, a new colour/depth image, a flush rule)."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        return int(self._lib.rdpo_process(self._h, w.ctypes.data, w.size))

    def flush(self):
        """End the open render pass (the reference's idle())."""
        self._lib.rdpo_flush(self._h)

    def set_vi_register(self, reg: int, value: int):
        self._lib.rdpo_set_vi_register(self._h, reg, value & 0xFFFFFFFF)

    def scanout(self, options: "ScanoutOptions | None" = None):
        """VI scanout: (height, width) uint32 RGBA8 words (R in the low byte)."""
        opts = options if options is not None else ScanoutOptions.default()
        out = np.zeros(640 * 576 * getattr(self, "upscaling", 1) ** 2, dtype=np.uint32)
        w = ctypes.c_uint(0)
        h = ctypes.c_uint(0)
        self._lib.rdpo_scanout_ex(self._h, ctypes.byref(opts), out.ctypes.data, ctypes.byref(w), ctypes.byref(h))
        return out[: w.value * h.value].reshape(h.value, w.value) if w.value and h.value else out[:0].reshape(0, 0)

    def begin_vi_register_per_scanline(self, flags: int = 3):
        self._lib.rdpo_begin_vi_register_per_scanline(self._h, flags)

    def set_vi_register_for_scanline(self, reg_bit: int, value: int):
        self._lib.rdpo_set_vi_register_for_scanline(self._h, reg_bit, value & 0xFFFFFFFF)

    def latch_vi_register_for_scanline(self, vi_line: int):
        self._lib.rdpo_latch_vi_register_for_scanline(self._h, vi_line)

    def end_vi_register_per_scanline(self):
        self._lib.rdpo_end_vi_register_per_scanline(self._h)

    def pass_ends(self) -> np.ndarray:
        """Primitive counter at the end of every render pass that held a draw."""
        n = int(self._lib.rdpo_pass_ends(self._h, None, 0))
        out = np.zeros(n, dtype=np.uint64)
        self._lib.rdpo_pass_ends(self._h, out.ctypes.data, n)
        return out

    def digest(self) -> int:
        """Digest of RDRAM, hidden RDRAM and TMEM (the library's state digest)."""
        return int(self._lib.rdpo_state_digest(self._h))

    @property
    def fragments(self) -> int:
        return int(self._lib.rdpo_fragment_count(self._h))

    FRAGMENT_MODES = ("fill", "copy", "1cycle", "1cycle_tex", "2cycle")

    def fragments_by_mode(self) -> dict:
        """Fragments split by the draw's mode (FRAGMENT_MODES)."""
        out = np.zeros(len(self.FRAGMENT_MODES), dtype=np.uint64)
        self._lib.rdpo_fragment_counts(self._h, out.ctypes.data)
        return {m: int(v) for m, v in zip(self.FRAGMENT_MODES, out)}

    @property
    def primitives(self) -> int:
        return int(self._lib.rdpo_primitive_count(self._h))


def command_length(word: int) -> int:
    return int(_load().rdpo_command_length(word & 0xFFFFFFFF))


def blender_divider_table() -> np.ndarray:
    out = np.zeros(0x8000, dtype=np.uint8)
    _load().rdpo_blender_divider_table(out.ctypes.data)
    return out


def gamma_table() -> np.ndarray:
    out = np.zeros(256 + 256 * 64, dtype=np.uint8)
    _load().rdpo_gamma_table(out.ctypes.data)
    return out


def perspective_table() -> np.ndarray:
    out = np.zeros((64, 2), dtype=np.int16)
    _load().rdpo_perspective_table(out.ctypes.data)
    return out

"""Python binding of librdp_mi.so, shaped like RDP::CommandProcessor.

The methods mirror the reference's surface (parallel-rdp/rdp_device.hpp:111-198):
enqueue_command, flush, signal_timeline, wait_for_timeline, idle,
begin_frame_context, set_vi_register, scanout_sync, begin_read_rdram, ...

The library must be built (python parallel-rdp_amd/build.py).  Loading
fails loudly when it is missing; there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librdp_mi.so")

_lib = None

VALIDATION_FN = ctypes.CFUNCTYPE(None, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_void_p)


class RDPError(RuntimeError):
    pass


ERROR_TIMEOUT = -5
ERROR_LIMIT = -6

# rdp_mi_create flags (rdp_device.hpp:43-51).
FLAG_UPSCALING = {2: 1 << 2, 4: 1 << 3, 8: 1 << 4}
FLAG_SUPER_SAMPLED_READ_BACK = 1 << 5
FLAG_SUPER_SAMPLED_DITHER = 1 << 6

# Words per command by opcode (the RDP's command lengths).
COMMAND_LENGTHS = np.array([
    2, 0, 0, 0, 2, 0, 0, 0, 8, 12, 24, 28, 24, 28, 40, 44,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 4, 4, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
    2, 0, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
], dtype=np.int64)


class ScanoutOptions(ctypes.Structure):
    """rdp_mi_scanout_options (the reference's ScanoutOptions without upscaling)."""
    _fields_ = [(n, ctypes.c_uint) for n in ("crop_overscan_pixels", "crop_left", "crop_right", "crop_top",
                                             "crop_bottom")] + \
               [(n, ctypes.c_int) for n in ("crop_enable", "persist_frame_on_invalid_input", "blend_previous_frame",
                                            "upscale_deinterlacing", "vi_aa", "vi_scale", "vi_serrate",
                                            "vi_dither_filter", "vi_divot_filter", "vi_gamma_dither")]

    @classmethod
    def default(cls, **kw):
        o = cls()
        load_library().rdp_mi_scanout_options_default(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


class DumpEvent(ctypes.Structure):
    """rdp_mi_dump_event: one record of an RDPDUMP2 capture."""
    _fields_ = [("kind", ctypes.c_uint32), ("op", ctypes.c_uint32), ("num_words", ctypes.c_uint32),
                ("words", ctypes.POINTER(ctypes.c_uint32)), ("vi_register", ctypes.c_uint32),
                ("vi_value", ctypes.c_uint32), ("offset", ctypes.c_uint32), ("size", ctypes.c_uint32)]


# Record kinds of RDPDUMP2 (rdp_dump.cpp:28-40); EOF is the player's end marker.
DUMP_EOF, DUMP_UPDATE_DRAM, DUMP_RDP_COMMAND, DUMP_SET_VI_REGISTER, DUMP_END_FRAME = 0, 1, 2, 3, 4
DUMP_SIGNAL_COMPLETE, DUMP_END_OF_FILE, DUMP_UPDATE_DRAM_FLUSH = 5, 6, 7
DUMP_UPDATE_HIDDEN_DRAM, DUMP_UPDATE_HIDDEN_DRAM_FLUSH = 8, 9

# Draw commands (replayer_driver.hpp:92-111, command_is_draw_call).
DRAW_OPS = frozenset((0x08, 0x09, 0x0a, 0x0b, 0x0c, 0x0d, 0x0e, 0x0f, 0x24, 0x25, 0x36))


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RDPError(f"{LIB_PATH} is not built; run python parallel-rdp_amd/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, sz, cu = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_uint
    lib.rdp_mi_create.restype = ctypes.c_int
    lib.rdp_mi_create.argtypes = [ctypes.POINTER(vp), vp, sz, sz, sz, u32]
    lib.rdp_mi_destroy.argtypes = [vp]
    lib.rdp_mi_is_supported.restype = ctypes.c_int
    lib.rdp_mi_is_supported.argtypes = [vp]
    lib.rdp_mi_enqueue_command.argtypes = [vp, cu, vp]
    lib.rdp_mi_enqueue_command_direct.argtypes = [vp, cu, vp]
    lib.rdp_mi_enqueue_command_list.restype = sz
    lib.rdp_mi_enqueue_command_list.argtypes = [vp, sz, vp]
    lib.rdp_mi_set_quirks.argtypes = [vp, u32]
    lib.rdp_mi_set_hold_submissions.argtypes = [vp, ctypes.c_int]
    lib.rdp_mi_flush.argtypes = [vp]
    lib.rdp_mi_signal_timeline.restype = u64
    lib.rdp_mi_signal_timeline.argtypes = [vp]
    lib.rdp_mi_wait_for_timeline.restype = ctypes.c_int
    lib.rdp_mi_wait_for_timeline.argtypes = [vp, u64]
    lib.rdp_mi_wait_for_timeline_timeout.restype = ctypes.c_int
    lib.rdp_mi_wait_for_timeline_timeout.argtypes = [vp, u64, u64]
    lib.rdp_mi_completed_timeline.restype = u64
    lib.rdp_mi_completed_timeline.argtypes = [vp]
    lib.rdp_mi_idle.restype = ctypes.c_int
    lib.rdp_mi_idle.argtypes = [vp]
    lib.rdp_mi_begin_frame_context.argtypes = [vp]
    for name in ("rdp_mi_begin_read_rdram", "rdp_mi_begin_read_hidden_rdram", "rdp_mi_get_tmem"):
        getattr(lib, name).restype = vp
        getattr(lib, name).argtypes = [vp]
    lib.rdp_mi_get_upscaling_factor.restype = cu
    lib.rdp_mi_get_upscaling_factor.argtypes = [vp]
    for name in ("rdp_mi_read_upscaled_rdram", "rdp_mi_read_upscaled_hidden_rdram"):
        getattr(lib, name).restype = ctypes.c_int
        getattr(lib, name).argtypes = [vp, sz, vp, sz]
    lib.rdp_mi_end_write_rdram.argtypes = [vp]
    lib.rdp_mi_end_write_hidden_rdram.argtypes = [vp]
    lib.rdp_mi_get_rdram_size.restype = sz
    lib.rdp_mi_get_rdram_size.argtypes = [vp]
    lib.rdp_mi_get_hidden_rdram_size.restype = sz
    lib.rdp_mi_get_hidden_rdram_size.argtypes = [vp]
    lib.rdp_mi_set_vi_register.argtypes = [vp, cu, u32]
    lib.rdp_mi_scanout_sync.restype = ctypes.c_int
    lib.rdp_mi_scanout_sync.argtypes = [vp, vp, ctypes.POINTER(cu), ctypes.POINTER(cu)]
    lib.rdp_mi_scanout_options_default.argtypes = [ctypes.POINTER(ScanoutOptions)]
    lib.rdp_mi_scanout_sync_ex.restype = ctypes.c_int
    lib.rdp_mi_scanout_sync_ex.argtypes = [vp, ctypes.POINTER(ScanoutOptions), vp, ctypes.POINTER(cu),
                                           ctypes.POINTER(cu)]
    lib.rdp_mi_scanout_buffer_create.restype = ctypes.c_int
    lib.rdp_mi_scanout_buffer_create.argtypes = [ctypes.POINTER(vp)]
    lib.rdp_mi_scanout_buffer_destroy.argtypes = [vp]
    lib.rdp_mi_scanout_async_buffer.restype = ctypes.c_int
    lib.rdp_mi_scanout_async_buffer.argtypes = [vp, ctypes.POINTER(ScanoutOptions), vp]
    lib.rdp_mi_scanout_buffer_wait.restype = ctypes.c_int
    lib.rdp_mi_scanout_buffer_wait.argtypes = [vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint32)),
                                               ctypes.POINTER(cu), ctypes.POINTER(cu)]
    lib.rdp_mi_begin_vi_register_per_scanline.argtypes = [vp, u32]
    lib.rdp_mi_set_vi_register_for_scanline.argtypes = [vp, u32, u32]
    lib.rdp_mi_latch_vi_register_for_scanline.argtypes = [vp, cu]
    lib.rdp_mi_end_vi_register_per_scanline.argtypes = [vp]
    lib.rdp_mi_set_validation_callback.argtypes = [vp, VALIDATION_FN, vp]
    lib.rdp_mi_batch_flush.restype = ctypes.c_int
    lib.rdp_mi_batch_flush.argtypes = [ctypes.POINTER(vp), cu]
    lib.rdp_mi_batch_flush_noalloc.restype = ctypes.c_int
    lib.rdp_mi_batch_flush_noalloc.argtypes = [ctypes.POINTER(vp), cu]
    lib.rdp_mi_last_error.restype = ctypes.c_char_p
    lib.rdp_mi_last_error.argtypes = []
    lib.rdp_mi_primitive_count.restype = u64
    lib.rdp_mi_primitive_count.argtypes = [vp]
    lib.rdp_mi_state_digest.restype = ctypes.c_int
    lib.rdp_mi_state_digest.argtypes = [vp, ctypes.POINTER(u64)]
    # RDPDUMP2 captures
    lib.rdp_mi_begin_dump.restype = ctypes.c_int
    lib.rdp_mi_begin_dump.argtypes = [vp, ctypes.c_char_p]
    lib.rdp_mi_end_dump.argtypes = [vp]
    lib.rdp_mi_dump_writer_open.restype = ctypes.c_int
    lib.rdp_mi_dump_writer_open.argtypes = [ctypes.c_char_p, u32, u32, ctypes.POINTER(vp)]
    lib.rdp_mi_dump_writer_flush_rdram.argtypes = [vp, vp, u32]
    lib.rdp_mi_dump_writer_flush_hidden_rdram.argtypes = [vp, vp, u32]
    lib.rdp_mi_dump_writer_command.argtypes = [vp, u32, vp, u32]
    lib.rdp_mi_dump_writer_set_vi_register.argtypes = [vp, u32, u32]
    lib.rdp_mi_dump_writer_end_frame.argtypes = [vp]
    lib.rdp_mi_dump_writer_signal_complete.argtypes = [vp]
    lib.rdp_mi_dump_writer_close.argtypes = [vp]
    lib.rdp_mi_dump_player_open.restype = ctypes.c_int
    lib.rdp_mi_dump_player_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    lib.rdp_mi_dump_player_rdram_size.restype = u32
    lib.rdp_mi_dump_player_rdram_size.argtypes = [vp]
    lib.rdp_mi_dump_player_hidden_rdram_size.restype = u32
    lib.rdp_mi_dump_player_hidden_rdram_size.argtypes = [vp]
    lib.rdp_mi_dump_player_rewind.restype = ctypes.c_int
    lib.rdp_mi_dump_player_rewind.argtypes = [vp]
    lib.rdp_mi_dump_player_next.restype = ctypes.c_int
    lib.rdp_mi_dump_player_next.argtypes = [vp, ctypes.POINTER(DumpEvent)]
    lib.rdp_mi_dump_player_rdram.restype = vp
    lib.rdp_mi_dump_player_rdram.argtypes = [vp]
    lib.rdp_mi_dump_player_hidden_rdram.restype = vp
    lib.rdp_mi_dump_player_hidden_rdram.argtypes = [vp]
    lib.rdp_mi_dump_player_apply.restype = ctypes.c_int
    lib.rdp_mi_dump_player_apply.argtypes = [vp, ctypes.POINTER(DumpEvent), vp]
    lib.rdp_mi_dump_player_destroy.argtypes = [vp]
    lib.rdp_mi_framebuffer_view.restype = ctypes.c_int
    lib.rdp_mi_framebuffer_view.argtypes = [vp, u32, u32, u32, u32, u32, u32, vp]
    _lib = lib
    return lib


def last_error() -> str:
    return load_library().rdp_mi_last_error().decode(errors="replace")


def _check(ret: int, what: str):
    if ret != 0:
        raise RDPError(f"{what} failed ({ret}): {last_error()}")


class CommandProcessor:
    """One emulated RDP with device-resident RDRAM, hidden RDRAM and TMEM."""

    def __init__(self, rdram_size: int = 8 << 20, flags: int = 0, host_rdram: Optional[np.ndarray] = None):
        """host_rdram: a writable uint32 array of rdram_size bytes that becomes
        the processor's RDRAM (host-coherent mode); it must outlive the
        processor."""
        self._lib = load_library()
        h = ctypes.c_void_p()
        ptr = None
        if host_rdram is not None:
            if host_rdram.nbytes != rdram_size or not host_rdram.flags.c_contiguous or not host_rdram.flags.writeable:
                raise ValueError("host_rdram must be a writable contiguous array of rdram_size bytes")
            ptr = host_rdram.ctypes.data
        _check(self._lib.rdp_mi_create(ctypes.byref(h), ptr, 0, rdram_size, rdram_size // 2, flags), "rdp_mi_create")
        self._h = h
        self.host_rdram = host_rdram
        self.rdram_size = rdram_size
        self._validation_cb = None
        self.validation_errors = []

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            self._lib.rdp_mi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_is_supported(self) -> bool:
        return bool(self._lib.rdp_mi_is_supported(self._h))

    # -- commands ---------------------------------------------------------
    def enqueue_command(self, words: Sequence[int]):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        self._lib.rdp_mi_enqueue_command(self._h, w.size, w.ctypes.data)

    def enqueue_commands(self, words: np.ndarray) -> int:
        w = np.ascontiguousarray(words, dtype=np.uint32)
        return int(self._lib.rdp_mi_enqueue_command_list(self._h, w.size, w.ctypes.data))

    def set_quirks(self, quirks: int):
        self._lib.rdp_mi_set_quirks(self._h, quirks)

    def hold_submissions(self, hold: bool = True):
        """rdp_mi_set_hold_submissions: keep completed passes on the host
        until a recording, a batched submission or an explicit flush."""
        self._lib.rdp_mi_set_hold_submissions(self._h, 1 if hold else 0)

    # -- synchronisation --------------------------------------------------
    def flush(self):
        self._lib.rdp_mi_flush(self._h)

    def enqueue_commands_direct(self, words: np.ndarray) -> int:
        """enqueue_command_direct for each complete command (decoded on this
        thread, no ring).  Returns the number of words consumed."""
        w = np.ascontiguousarray(words, dtype=np.uint32)
        i = 0
        while i < w.size:
            n = int(COMMAND_LENGTHS[(int(w[i]) >> 24) & 63])
            if n == 0:
                i += 1
                continue
            if i + n > w.size:
                break
            self._lib.rdp_mi_enqueue_command_direct(self._h, n, w[i:].ctypes.data)
            i += n
        return i

    def signal_timeline(self) -> int:
        return int(self._lib.rdp_mi_signal_timeline(self._h))

    def wait_for_timeline(self, value: int, timeout_us: Optional[int] = None):
        """Blocks until `value` is published (also for a value not yet
        signalled).  With timeout_us, raises TimeoutError when it is not
        published in time."""
        if timeout_us is None:
            _check(self._lib.rdp_mi_wait_for_timeline(self._h, value), "wait_for_timeline")
            return
        ret = self._lib.rdp_mi_wait_for_timeline_timeout(self._h, value, int(timeout_us))
        if ret == ERROR_TIMEOUT:
            raise TimeoutError(last_error())
        _check(ret, "wait_for_timeline")

    @property
    def completed_timeline(self) -> int:
        return int(self._lib.rdp_mi_completed_timeline(self._h))

    def idle(self):
        _check(self._lib.rdp_mi_idle(self._h), "idle")

    def begin_frame_context(self):
        self._lib.rdp_mi_begin_frame_context(self._h)

    # -- memory -----------------------------------------------------------
    def read_rdram(self) -> np.ndarray:
        """Copy of RDRAM as host-endian uint32 words (valid after idle())."""
        p = self._lib.rdp_mi_begin_read_rdram(self._h)
        if not p:
            raise RDPError("begin_read_rdram failed: " + last_error())
        buf = (ctypes.c_uint8 * self.rdram_size).from_address(p)
        return np.frombuffer(buf, dtype=np.uint32).copy()

    def read_hidden_rdram(self) -> np.ndarray:
        p = self._lib.rdp_mi_begin_read_hidden_rdram(self._h)
        if not p:
            raise RDPError("begin_read_hidden_rdram failed: " + last_error())
        buf = (ctypes.c_uint8 * (self.rdram_size // 2)).from_address(p)
        return np.frombuffer(buf, dtype=np.uint8).copy()

    def write_rdram(self, words: np.ndarray):
        p = self._lib.rdp_mi_begin_read_rdram(self._h)
        if not p:
            raise RDPError("begin_read_rdram failed: " + last_error())
        buf = (ctypes.c_uint8 * self.rdram_size).from_address(p)
        np.frombuffer(buf, dtype=np.uint32)[:] = np.asarray(words, dtype=np.uint32)
        self._lib.rdp_mi_end_write_rdram(self._h)

    def write_hidden_rdram(self, data: np.ndarray):
        p = self._lib.rdp_mi_begin_read_hidden_rdram(self._h)
        if not p:
            raise RDPError("begin_read_hidden_rdram failed: " + last_error())
        buf = (ctypes.c_uint8 * (self.rdram_size // 2)).from_address(p)
        np.frombuffer(buf, dtype=np.uint8)[:] = np.asarray(data, dtype=np.uint8)
        self._lib.rdp_mi_end_write_hidden_rdram(self._h)

    # -- upscaled domain -----------------------------------------------------
    @property
    def upscaling_factor(self) -> int:
        return int(self._lib.rdp_mi_get_upscaling_factor(self._h))

    def read_upscaled_rdram(self) -> np.ndarray:
        """The factor^2 RDRAM sample planes as (planes, rdram_size // 4)
        words; plane sy * factor + sx holds sample (sx, sy) of each pixel.
        Call after idle() or a timeline wait."""
        planes = self.upscaling_factor ** 2
        out = np.empty((planes, self.rdram_size // 4), dtype=np.uint32)
        _check(self._lib.rdp_mi_read_upscaled_rdram(self._h, 0, out.ctypes.data, out.nbytes), "read_upscaled_rdram")
        return out

    def read_upscaled_hidden_rdram(self) -> np.ndarray:
        planes = self.upscaling_factor ** 2
        out = np.empty((planes, self.rdram_size // 2), dtype=np.uint8)
        _check(self._lib.rdp_mi_read_upscaled_hidden_rdram(self._h, 0, out.ctypes.data, out.nbytes),
               "read_upscaled_hidden_rdram")
        return out

    def read_tmem(self) -> np.ndarray:
        p = self._lib.rdp_mi_get_tmem(self._h)
        if not p:
            raise RDPError("get_tmem failed: " + last_error())
        return np.frombuffer((ctypes.c_uint16 * 2048).from_address(p), dtype=np.uint16).copy()

    # -- VI ---------------------------------------------------------------
    def set_vi_register(self, reg: int, value: int):
        self._lib.rdp_mi_set_vi_register(self._h, reg, value & 0xFFFFFFFF)

    def scanout_sync(self, options: Optional[ScanoutOptions] = None) -> np.ndarray:
        """VI scanout as a (height, width) array of RGBA8 words (R in the low byte);
        (0, 0) for a blank or invalid frame."""
        opts = options if options is not None else ScanoutOptions.default()
        out = np.zeros(640 * 576 * self.upscaling_factor ** 2, dtype=np.uint32)
        w = ctypes.c_uint(0)
        h = ctypes.c_uint(0)
        _check(self._lib.rdp_mi_scanout_sync_ex(self._h, ctypes.byref(opts), out.ctypes.data, ctypes.byref(w),
                                                ctypes.byref(h)), "scanout")
        if not w.value or not h.value:
            return out[:0].reshape(0, 0)
        return out[: w.value * h.value].reshape(h.value, w.value)

    def scanout_async(self, buffer: "ScanoutBuffer", options: Optional[ScanoutOptions] = None):
        """Queue a scanout into `buffer` and return at once
        (CommandProcessor::scanout_async_buffer)."""
        opts = options if options is not None else ScanoutOptions.default()
        _check(self._lib.rdp_mi_scanout_async_buffer(self._h, ctypes.byref(opts), buffer.handle), "scanout_async")

    def begin_vi_register_per_scanline(self, flags: int = 3):
        self._lib.rdp_mi_begin_vi_register_per_scanline(self._h, flags)

    def set_vi_register_for_scanline(self, reg_bit: int, value: int):
        self._lib.rdp_mi_set_vi_register_for_scanline(self._h, reg_bit, value & 0xFFFFFFFF)

    def latch_vi_register_for_scanline(self, vi_line: int):
        self._lib.rdp_mi_latch_vi_register_for_scanline(self._h, vi_line)

    def end_vi_register_per_scanline(self):
        self._lib.rdp_mi_end_vi_register_per_scanline(self._h)

    # -- validation -------------------------------------------------------
    def set_validation_callback(self, record: bool = True):
        def cb(err, msg, user):
            self.validation_errors.append((int(err), msg.decode(errors="replace") if msg else ""))
        self._validation_cb = VALIDATION_FN(cb) if record else VALIDATION_FN()
        self._lib.rdp_mi_set_validation_callback(self._h, self._validation_cb, None)

    @property
    def primitive_count(self) -> int:
        return int(self._lib.rdp_mi_primitive_count(self._h))

    VIEW_COLOR, VIEW_DEPTH, VIEW_COVERAGE = 0, 1, 2

    def framebuffer_view(self, mode: int, fb_addr: int, fb_size: int, depth_addr: int, width: int,
                         height: int = 0) -> np.ndarray:
        """(height, width) RGBA8 words (R in the low byte) of the colour,
        depth or coverage view of a framebuffer (rdp_mi_framebuffer_view)."""
        h = height or max(1, width * 3 // 4)
        out = np.zeros(width * h, np.uint32)
        _check(self._lib.rdp_mi_framebuffer_view(self._h, mode, fb_addr, fb_size, depth_addr, width, h,
                                                 out.ctypes.data), "framebuffer_view")
        return out.reshape(h, width)

    def begin_dump(self, path: str):
        """Capture this processor's commands, VI registers and scanouts to an
        RDPDUMP2 file (the reference's PARALLEL_RDP_DUMP_PATH capture)."""
        _check(self._lib.rdp_mi_begin_dump(self._h, os.fsencode(path)), "begin_dump")

    def end_dump(self):
        self._lib.rdp_mi_end_dump(self._h)

    def state_digest(self) -> int:
        """Device-side digest of RDRAM, hidden RDRAM and TMEM after the work
        submitted so far (call after idle())."""
        d = ctypes.c_uint64()
        _check(self._lib.rdp_mi_state_digest(self._h, ctypes.byref(d)), "state_digest")
        return int(d.value)


class ScanoutBuffer:
    """VIScanoutBuffer: a host image of one scanout plus its fence."""

    def __init__(self):
        self._lib = load_library()
        h = ctypes.c_void_p()
        _check(self._lib.rdp_mi_scanout_buffer_create(ctypes.byref(h)), "scanout_buffer_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def wait(self) -> np.ndarray:
        """Waits for the fence; the image as a (height, width) array."""
        px = ctypes.POINTER(ctypes.c_uint32)()
        w = ctypes.c_uint(0)
        h = ctypes.c_uint(0)
        _check(self._lib.rdp_mi_scanout_buffer_wait(self._h, ctypes.byref(px), ctypes.byref(w), ctypes.byref(h)),
               "scanout_buffer_wait")
        if not w.value or not h.value:
            return np.zeros((0, 0), np.uint32)
        return np.ctypeslib.as_array(px, shape=(h.value * w.value,)).copy().reshape(h.value, w.value)

    def close(self):
        if self._h:
            self._lib.rdp_mi_scanout_buffer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_launch_config(xcd_chunk: int = 64, min_class: int = 0, tile_size: int = 8, persistent: bool = False):
    """Raster launch shape for immediate submissions (flush, batch flush, serving)."""
    lib = load_library()
    lib.rdp_mi_set_launch_config.restype = ctypes.c_int
    lib.rdp_mi_set_launch_config.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
    _check(lib.rdp_mi_set_launch_config(xcd_chunk, min_class, tile_size, int(persistent)), "set_launch_config")


def batch_flush(processors: Sequence[CommandProcessor]):
    lib = load_library()
    arr = (ctypes.c_void_p * len(processors))(*[p.handle.value for p in processors])
    _check(lib.rdp_mi_batch_flush(arr, len(processors)), "batch_flush")


def batch_flush_noalloc(processors: Sequence[CommandProcessor]) -> int:
    """rdp_mi_batch_flush_noalloc: the batched submission through the
    library's static device scratch array (no allocation); returns when the
    work is done.  Returns 0, or ERROR_LIMIT when the submission does not fit
    the array (its passes then stay pending); raises on other errors."""
    lib = load_library()
    arr = (ctypes.c_void_p * len(processors))(*[p.handle.value for p in processors])
    ret = lib.rdp_mi_batch_flush_noalloc(arr, len(processors))
    if ret != ERROR_LIMIT:
        _check(ret, "batch_flush_noalloc")
    return ret


class Recording:
    """Device-resident recorded submission of pending passes (see rdp_mi_record)."""

    def __init__(self, processors: Sequence[CommandProcessor], batched: bool = True):
        lib = load_library()
        lib.rdp_mi_record.restype = ctypes.c_int
        lib.rdp_mi_record.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_void_p)]
        lib.rdp_mi_recording_replay.restype = ctypes.c_int
        lib.rdp_mi_recording_replay.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_void_p]
        lib.rdp_mi_recording_profile.restype = ctypes.c_int
        lib.rdp_mi_recording_profile.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p]
        lib.rdp_mi_recording_raster_bytes.restype = ctypes.c_uint64
        lib.rdp_mi_recording_raster_bytes.argtypes = [ctypes.c_void_p]
        lib.rdp_mi_recording_rounds.restype = ctypes.c_uint
        lib.rdp_mi_recording_rounds.argtypes = [ctypes.c_void_p]
        lib.rdp_mi_recording_passes.restype = ctypes.c_uint
        lib.rdp_mi_recording_passes.argtypes = [ctypes.c_void_p]
        lib.rdp_mi_recording_destroy.argtypes = [ctypes.c_void_p]
        self._lib = lib
        arr = (ctypes.c_void_p * len(processors))(*[p.handle.value for p in processors])
        h = ctypes.c_void_p()
        _check(lib.rdp_mi_record(arr, len(processors), int(batched), ctypes.byref(h)), "rdp_mi_record")
        self._h = h
        self.processors = list(processors)

    def replay(self, iterations: int = 1, use_graph: bool = True) -> np.ndarray:
        ms = np.zeros(iterations, dtype=np.float32)
        _check(self._lib.rdp_mi_recording_replay(self._h, iterations, int(use_graph), ms.ctypes.data), "replay")
        return ms

    def autotune(self, iterations: int = 3) -> dict:
        """Times every raster launch shape on this submission and keeps the
        fastest; returns it with its time per replay."""
        self._lib.rdp_mi_recording_autotune.restype = ctypes.c_int
        self._lib.rdp_mi_recording_autotune.argtypes = [ctypes.c_void_p, ctypes.c_uint,
                                                        ctypes.POINTER(ctypes.c_int32),
                                                        ctypes.POINTER(ctypes.c_uint32),
                                                        ctypes.POINTER(ctypes.c_uint32),
                                                        ctypes.POINTER(ctypes.c_int32),
                                                        ctypes.POINTER(ctypes.c_float)]
        chunk, cls, tile, pers, ms = (ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32(),
                                      ctypes.c_float())
        _check(self._lib.rdp_mi_recording_autotune(self._h, iterations, ctypes.byref(chunk), ctypes.byref(cls),
                                                   ctypes.byref(tile), ctypes.byref(pers), ctypes.byref(ms)),
               "autotune")
        return {"xcd_chunk": chunk.value, "min_class": cls.value, "tile_size": tile.value,
                "persistent": bool(pers.value), "ms": ms.value}

    def autotune_results(self) -> list:
        """Every candidate of the last autotune with its ms per replay."""
        f = self._lib.rdp_mi_recording_autotune_results
        f.restype = ctypes.c_uint
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint] + [ctypes.c_void_p] * 5
        n = int(f(self._h, 0, None, None, None, None, None))
        chunk, cls, tile = (np.zeros(n, np.int32), np.zeros(n, np.uint32), np.zeros(n, np.uint32))
        pers, ms = np.zeros(n, np.int32), np.zeros(n, np.float32)
        f(self._h, n, chunk.ctypes.data, cls.ctypes.data, tile.ctypes.data, pers.ctypes.data, ms.ctypes.data)
        return [{"xcd_chunk": int(a), "min_class": int(b), "tile_size": int(c), "persistent": bool(p),
                 "ms": float(d)} for a, b, c, p, d in zip(chunk, cls, tile, pers, ms)]

    def profile(self, iterations: int = 1) -> np.ndarray:
        ms = np.zeros(4, dtype=np.float32)
        _check(self._lib.rdp_mi_recording_profile(self._h, iterations, ms.ctypes.data), "profile")
        return ms

    def profile_raster(self, iterations: int = 3) -> list:
        """Each raster launch timed alone: dicts of kind ('k_raster<c>' or
        'k_raster_fill'), ms, algorithmic bytes and tiles, in launch order."""
        lib = self._lib
        lib.rdp_mi_recording_profile_raster.restype = ctypes.c_int
        lib.rdp_mi_recording_profile_raster.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint] + \
            [ctypes.c_void_p] * 4 + [ctypes.POINTER(ctypes.c_uint)]
        cap = 64
        ms = np.zeros(cap, np.float32)
        by = np.zeros(cap, np.uint64)
        kind = np.zeros(cap, np.uint32)
        tiles = np.zeros(cap, np.uint32)
        n = ctypes.c_uint()
        _check(lib.rdp_mi_recording_profile_raster(self._h, iterations, cap, ms.ctypes.data, by.ctypes.data,
                                                   kind.ctypes.data, tiles.ctypes.data, ctypes.byref(n)),
               "profile_raster")
        return [{"kernel": "k_raster_fill" if int(kind[i]) == 3 else f"k_raster<{int(kind[i])}>",
                 "ms": float(ms[i]), "bytes": int(by[i]), "tiles": int(tiles[i])} for i in range(min(n.value, cap))]

    def profile_rounds(self, iterations: int = 1) -> np.ndarray:
        self._lib.rdp_mi_recording_profile_rounds.restype = ctypes.c_int
        self._lib.rdp_mi_recording_profile_rounds.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p]
        ms = np.zeros((self.rounds, 4), dtype=np.float32)
        _check(self._lib.rdp_mi_recording_profile_rounds(self._h, iterations, ms.ctypes.data), "profile_rounds")
        return ms

    @property
    def raster_bytes(self) -> int:
        return int(self._lib.rdp_mi_recording_raster_bytes(self._h))

    @property
    def rounds(self) -> int:
        return int(self._lib.rdp_mi_recording_rounds(self._h))

    @property
    def raster_launches(self) -> int:
        self._lib.rdp_mi_recording_raster_launches.restype = ctypes.c_uint
        self._lib.rdp_mi_recording_raster_launches.argtypes = [ctypes.c_void_p]
        return int(self._lib.rdp_mi_recording_raster_launches(self._h))

    @property
    def passes(self) -> int:
        return int(self._lib.rdp_mi_recording_passes(self._h))

    def close(self):
        if self._h:
            self._lib.rdp_mi_recording_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching frame server over a fixed set of processors.

    Each request is one frame of complete commands for one processor with an
    arrival time (microseconds after start()).  Arrived requests are admitted
    whenever their processor is idle and the admitted, unsubmitted frames fit
    the device-byte budget; each is decoded at once on a host thread, and
    decoded frames are submitted together whenever one of the submission
    streams is free, while earlier batches still run.  Per processor, frames
    render in arrival order and the results equal rendering them alone.
    """

    def __init__(self, processors: Sequence[CommandProcessor], hbm_budget_bytes: int = 0, max_batch: int = 0,
                 decode_threads: int = 0, streams: int = 0):
        lib = load_library()
        vp, sz, cu = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint
        dp = ctypes.POINTER(ctypes.c_double)
        lib.rdp_mi_server_create.restype = ctypes.c_int
        lib.rdp_mi_server_create.argtypes = [ctypes.POINTER(vp), cu, sz, cu, cu, cu, ctypes.POINTER(vp)]
        lib.rdp_mi_server_submit.restype = ctypes.c_int
        lib.rdp_mi_server_submit.argtypes = [vp, cu, vp, sz, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64)]
        lib.rdp_mi_server_start.restype = ctypes.c_int
        lib.rdp_mi_server_start.argtypes = [vp]
        lib.rdp_mi_server_drain.restype = ctypes.c_int
        lib.rdp_mi_server_drain.argtypes = [vp]
        lib.rdp_mi_server_results.restype = sz
        lib.rdp_mi_server_results.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32),
                                              dp, dp, dp, dp, sz]
        lib.rdp_mi_server_stats.argtypes = [vp, vp]
        lib.rdp_mi_server_reset.restype = ctypes.c_int
        lib.rdp_mi_server_reset.argtypes = [vp, ctypes.c_double]
        lib.rdp_mi_server_destroy.argtypes = [vp]
        lib.rdp_mi_server_set_digests.restype = ctypes.c_int
        lib.rdp_mi_server_set_batch_window.restype = ctypes.c_int
        lib.rdp_mi_server_set_batch_window.argtypes = [vp, ctypes.c_double]
        lib.rdp_mi_server_set_digests.argtypes = [vp, ctypes.c_int]
        lib.rdp_mi_server_digests.restype = sz
        lib.rdp_mi_server_digests.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.POINTER(ctypes.c_int32), sz]
        self._lib = lib
        self.processors = list(processors)
        arr = (vp * len(processors))(*[p.handle.value for p in processors])
        h = vp()
        _check(lib.rdp_mi_server_create(arr, len(processors), hbm_budget_bytes, max_batch, decode_threads, streams,
                                        ctypes.byref(h)), "rdp_mi_server_create")
        self._h = h

    def submit(self, instance: int, words: np.ndarray, arrival_us: float = 0.0) -> int:
        w = np.ascontiguousarray(words, dtype=np.uint32)
        t = ctypes.c_uint64()
        _check(self._lib.rdp_mi_server_submit(self._h, instance, w.ctypes.data, w.size, float(arrival_us),
                                              ctypes.byref(t)), "rdp_mi_server_submit")
        return int(t.value)

    def start(self):
        _check(self._lib.rdp_mi_server_start(self._h), "rdp_mi_server_start")

    def drain(self):
        _check(self._lib.rdp_mi_server_drain(self._h), "rdp_mi_server_drain")

    def reset(self, start_in_us: float = 0.0):
        """After drain(): clear results and counters; arrival times of later
        requests count from start_in_us microseconds from now."""
        _check(self._lib.rdp_mi_server_reset(self._h, float(start_in_us)), "rdp_mi_server_reset")

    def results(self) -> dict:
        """Completed requests in completion order: arrays ticket, instance,
        arrival_us, admit_us, submit_us, complete_us."""
        n = int(self._lib.rdp_mi_server_results(self._h, None, None, None, None, None, None, 0))
        out = {
            "ticket": np.zeros(n, np.uint64),
            "instance": np.zeros(n, np.uint32),
            "arrival_us": np.zeros(n, np.float64),
            "admit_us": np.zeros(n, np.float64),
            "submit_us": np.zeros(n, np.float64),
            "complete_us": np.zeros(n, np.float64),
        }
        dp = ctypes.POINTER(ctypes.c_double)
        got = int(self._lib.rdp_mi_server_results(
            self._h, out["ticket"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            out["instance"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            out["arrival_us"].ctypes.data_as(dp), out["admit_us"].ctypes.data_as(dp),
            out["submit_us"].ctypes.data_as(dp), out["complete_us"].ctypes.data_as(dp), n))
        return {k: v[:min(n, got)] for k, v in out.items()}

    def set_batch_window(self, window_us: float):
        """A free submission stream waits until max_batch frames are ready, no
        admitted frame is still decoding, or the oldest has waited window_us."""
        _check(self._lib.rdp_mi_server_set_batch_window(self._h, float(window_us)), "rdp_mi_server_set_batch_window")

    def set_digests(self, enable: bool = True):
        """Digest every served processor's state after each frame (before start())."""
        _check(self._lib.rdp_mi_server_set_digests(self._h, int(enable)), "rdp_mi_server_set_digests")

    def digests(self) -> dict:
        """Per completed request, in the order of results(): ticket, the state
        digest after the frame and the batch status (0 or an error code)."""
        n = int(self._lib.rdp_mi_server_digests(self._h, None, None, None, 0))
        t = np.zeros(n, np.uint64)
        d = np.zeros(n, np.uint64)
        st = np.zeros(n, np.int32)
        got = int(self._lib.rdp_mi_server_digests(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                  d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                  st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n))
        m = min(n, got)
        return {"ticket": t[:m], "digest": d[:m], "status": st[:m]}

    def stats(self) -> dict:
        v = np.zeros(7, np.uint64)
        self._lib.rdp_mi_server_stats(self._h, v.ctypes.data)
        return dict(zip(("batches", "frames", "max_batch_frames", "max_estimated_bytes", "max_arena_bytes",
                         "submit_us_total", "submit_us_max"), (int(x) for x in v)))

    def close(self):
        if self._h:
            self._lib.rdp_mi_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DumpWriter:
    """RDPDumpWriter: writes an RDPDUMP2 capture (rdp_dump_write.hpp:31-57)."""

    def __init__(self, path: str, rdram_size: int = 8 << 20, hidden_rdram_size: int = 4 << 20):
        self._lib = load_library()
        h = ctypes.c_void_p()
        _check(self._lib.rdp_mi_dump_writer_open(os.fsencode(path), rdram_size, hidden_rdram_size, ctypes.byref(h)),
               "dump_writer_open")
        self._h = h

    def flush_rdram(self, rdram: np.ndarray):
        a = np.ascontiguousarray(rdram)
        self._lib.rdp_mi_dump_writer_flush_rdram(self._h, a.ctypes.data, a.nbytes)

    def flush_hidden_rdram(self, hidden: np.ndarray):
        a = np.ascontiguousarray(hidden, dtype=np.uint8)
        self._lib.rdp_mi_dump_writer_flush_hidden_rdram(self._h, a.ctypes.data, a.nbytes)

    def command(self, words: np.ndarray):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        self._lib.rdp_mi_dump_writer_command(self._h, (int(w[0]) >> 24) & 63, w.ctypes.data, w.size)

    def set_vi_register(self, reg: int, value: int):
        self._lib.rdp_mi_dump_writer_set_vi_register(self._h, reg, value & 0xFFFFFFFF)

    def end_frame(self):
        self._lib.rdp_mi_dump_writer_end_frame(self._h)

    def signal_complete(self):
        self._lib.rdp_mi_dump_writer_signal_complete(self._h)

    def close(self):
        if self._h:
            self._lib.rdp_mi_dump_writer_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DumpPlayer:
    """DumpPlayer (rdp_dump.hpp:49-68) over the library's reader."""

    def __init__(self, path: str):
        self._lib = load_library()
        h = ctypes.c_void_p()
        _check(self._lib.rdp_mi_dump_player_open(os.fsencode(path), ctypes.byref(h)), "dump_player_open")
        self._h = h
        self.rdram_size = int(self._lib.rdp_mi_dump_player_rdram_size(h))
        self.hidden_rdram_size = int(self._lib.rdp_mi_dump_player_hidden_rdram_size(h))
        self.event = DumpEvent()

    def next(self) -> Optional[DumpEvent]:
        """The next record, or None at the end of the capture."""
        r = self._lib.rdp_mi_dump_player_next(self._h, ctypes.byref(self.event))
        if r < 0:
            raise RDPError("malformed capture: " + last_error())
        return None if r == 0 else self.event

    def words(self) -> np.ndarray:
        e = self.event
        return np.ctypeslib.as_array(e.words, shape=(e.num_words,)).copy() if e.num_words else np.zeros(0, np.uint32)

    def apply(self, proc: CommandProcessor):
        """Drive `proc` with the current record as the reference's replay driver does."""
        _check(self._lib.rdp_mi_dump_player_apply(self._h, ctypes.byref(self.event), proc.handle), "dump_player_apply")

    def rdram(self) -> np.ndarray:
        p = self._lib.rdp_mi_dump_player_rdram(self._h)
        return np.frombuffer((ctypes.c_uint8 * self.rdram_size).from_address(p), dtype=np.uint32).copy()

    def hidden_rdram(self) -> np.ndarray:
        p = self._lib.rdp_mi_dump_player_hidden_rdram(self._h)
        return np.frombuffer((ctypes.c_uint8 * self.hidden_rdram_size).from_address(p), dtype=np.uint8).copy()

    def rewind(self):
        _check(self._lib.rdp_mi_dump_player_rewind(self._h), "dump_player_rewind")

    def close(self):
        if self._h:
            self._lib.rdp_mi_dump_player_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

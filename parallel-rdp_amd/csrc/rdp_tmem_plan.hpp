// TMEM upload planning, tmem_update.comp:108-577: for one TMEM halfword and
// one upload, where in RDRAM the value it writes comes from (no memory
// access).  Shared by the TMEM gather kernel and by host-side checks.
#pragma once

#include <stdint.h>

#include "rdp_types.hpp"

#if defined(__HIPCC__)
#define RDP_TP __host__ __device__ __forceinline__
#else
#define RDP_TP inline
#endif

namespace rdp_mi {

RDP_TP int tp_min(int a, int b) { return a < b ? a : b; }
RDP_TP int tp_max(int a, int b) { return a > b ? a : b; }

RDP_TP int div_floor(int num, int den)
{
	if (den <= 0) return 0;
	if (num >= 0) return num / den;
	return -((-num + den - 1) / den);
}

// Where the halfword a TMEM upload writes comes from: the planning half of
// the gather (no memory access), so a thread can plan several uploads and
// then issue their RDRAM reads together.
enum : uint32_t { TMEM_LOAD_16 = 0, TMEM_LOAD_32 = 1 };
struct TmemLoad {
	uint32_t addr;
	uint32_t kind;
};

RDP_TP bool tmem_plan_32(const UploadInfo &u, int idx, TmemLoad &ld)
{
	const bool yuv = u.tmem_fmt == TF_YUV;
	const int off16 = (u.tmem_offset & 0x7ff) >> 1;
	const int stride = u.tmem_stride_words;
	int pixel_offset = (idx - off16) & 0x3ff;
	int ux, uy, ux_xor = 0;
	if (u.mode == UP_BLOCK) {
		int word_offset = pixel_offset >> 1;
		if (stride == 0) {
			int c0 = word_offset & ~1, c1 = c0 + 1;
			int t0 = (c0 * u.dxt) >> 16, t1 = (c1 * u.dxt) >> 16;
			if (t0 != t1) {
				int w0 = c0 ^ (t0 & 1), w1 = c1 ^ (t1 & 1);
				if (w1 == word_offset) ux_xor = (t1 & 1) << 1;
				else if (w0 == word_offset) ux_xor = (t0 & 1) << 1;
				else return false;
			} else
				ux_xor ^= (t0 & 1) << 1;
		} else {
			int min_t = div_floor(word_offset & ~1, u.max_line_stride);
			int max_t = div_floor(word_offset | 1, u.min_line_stride);
			int max_cand = (word_offset | 1) - stride * min_t;
			int min_cand = (word_offset & ~1) - stride * max_t;
			min_t = tp_max(min_t, (min_cand * u.dxt) >> 16);
			max_t = tp_min(max_t, (max_cand * u.dxt) >> 16);
			bool found = false;
			for (int t = max_t; t >= min_t; t--) {
				int cs0 = (word_offset & ~1) - stride * t, cs1 = (word_offset | 1) - stride * t;
				int ct0 = (cs0 * u.dxt) >> 16, ct1 = (cs1 * u.dxt) >> 16;
				if (((cs1 + ct1 * stride) ^ (ct1 & 1)) == word_offset) {
					found = true;
					pixel_offset = (cs1 << 1) + (pixel_offset & 1);
					break;
				} else if (((cs0 + ct0 * stride) ^ (ct0 & 1)) == word_offset) {
					found = true;
					pixel_offset = (cs0 << 1) + (pixel_offset & 1);
					break;
				}
			}
			if (!found) return false;
		}
		ux = pixel_offset;
		uy = 0;
	} else if (stride == 0) {
		ux = pixel_offset;
		uy = u.height - 1;
	} else {
		uy = div_floor(pixel_offset, stride);
		ux = pixel_offset - uy * stride;
		if (uy >= u.height) {
			ux += stride * (uy - u.height + 1);
			uy = u.height - 1;
		}
	}
	int last_line_x = ux ^ ((uy & 1) << 1);
	if (last_line_x >= u.width && uy > 0) {
		uy--;
		ux += stride;
	}
	int iter_off = 0;
	ux ^= ((uy & 1) << 1) | ux_xor;
	if (u.vram_size == 3 || yuv)
		iter_off = 4 * (ux & ~1);
	else if (u.vram_size == 2) {
		if (ux & 2) {
			if (uy > 0) {
				uy--;
				ux += stride;
				ux ^= 2;
			} else
				return false;
		}
		iter_off = 2 * (ux & ~1);
	} else if (u.vram_size == 1) {
		for (int i = 0; i < 4 && uy > 0 && (ux & 6) != 0; i++) {
			uy--;
			ux += stride;
			ux ^= 2;
		}
		if (ux & 6) return false;
		iter_off = ux & ~1;
	}
	if (ux >= u.width) return false;
	int line_addr = u.vram_addr + ((uy * u.vram_width) << (u.vram_size - 1));
	ld.addr = (uint32_t)(line_addr + iter_off + 4 * (ux & 1));
	ld.kind = TMEM_LOAD_32;
	return true;
}

RDP_TP bool tmem_plan_16(const UploadInfo &u, int idx, TmemLoad &ld)
{
	const int off16 = (u.tmem_offset & 0xfff) >> 1;
	const int stride = u.tmem_stride_words;
	int pixel_offset = (idx - off16) & 0x7ff;
	int ux, uy, ux_xor = 0;
	if (u.mode == UP_BLOCK) {
		int word_offset = pixel_offset >> 2;
		if (stride == 0)
			ux_xor = (((word_offset * u.dxt) >> 16) & 1) << 1;
		else {
			int min_t = div_floor(word_offset, u.max_line_stride);
			int max_t = div_floor(word_offset, u.min_line_stride);
			int max_cand = word_offset - stride * min_t;
			int min_cand = word_offset - stride * max_t;
			min_t = tp_max(min_t, (min_cand * u.dxt) >> 16);
			max_t = tp_min(max_t, (max_cand * u.dxt) >> 16);
			bool found = false;
			for (int t = max_t; t >= min_t; t--) {
				int cand = word_offset - stride * t;
				int ct = (cand * u.dxt) >> 16;
				if (cand + ct * stride == word_offset) {
					found = true;
					ux_xor = (ct & 1) << 1;
					pixel_offset = (cand << 2) + (pixel_offset & 3);
				}
			}
			if (!found) return false;
		}
		ux = pixel_offset;
		uy = 0;
	} else if (stride == 0) {
		ux = pixel_offset;
		uy = u.height - 1;
	} else {
		uy = div_floor(pixel_offset, stride);
		ux = pixel_offset - uy * stride;
		if (uy >= u.height) {
			ux += stride * (uy - u.height + 1);
			uy = u.height - 1;
		}
	}
	int iter_off = 0;
	if (u.tmem_size != u.vram_size) {
		if (u.vram_size - u.tmem_size == 1) {
			iter_off = (ux & ~3) * 4;
			if ((ux & ~3) + 2 < (u.vram_effective_width >> (3 - u.vram_size)))
				iter_off += 8;
		} else if (u.tmem_size == 2 && u.vram_size == 1) {
			if (ux & 4) {
				if ((stride & 4) != 0 && uy > 0) {
					uy--;
					ux += stride;
				} else
					return false;
			}
			iter_off = ux & ~3;
		}
	} else
		iter_off = (ux & ~3) * 2;
	if (ux >= u.width) return false;
	int line_addr = u.vram_addr + ((uy * u.vram_width) << (u.vram_size - 1));
	ux ^= ((uy & 1) << 1) | ux_xor;
	ld.addr = (uint32_t)(line_addr + iter_off + 2 * (ux & 3));
	ld.kind = TMEM_LOAD_16;
	return true;
}

RDP_TP bool tmem_plan_lut(const UploadInfo &u, int idx, TmemLoad &ld)
{
	const int off16 = (u.tmem_offset & 0xfff) >> 1;
	const int p = (idx - off16) & 0x7ff;
	const int diff = u.vram_size - u.tmem_size;
	int splat;
	if (diff == 2) {
		splat = (p >> 2) << (u.vram_size - 2);
		if (splat >= u.vram_effective_width) return false;
	} else if (diff == 1) {
		if (p & 4) return false;
		splat = (p & ~7) >> (u.tmem_size + (u.vram_size == 2 ? 2 : 0));
		if (splat >= u.vram_effective_width) return false;
	} else if (diff == 0) {
		if (p & 0xc) return false;
		splat = (p & ~3) >> (u.tmem_size + (u.vram_size == 2 ? 2 : 0));
		if (splat >= u.vram_effective_width) return false;
	} else if (diff == -1) {
		if (p & 0x1c) return false;
		splat = (p >> u.tmem_size) & ~7;
		if (splat >= u.vram_effective_width) return false;
	} else {
		int span_pixel = (p >> 2) * 4;
		if (span_pixel + 2 < u.vram_effective_width) span_pixel += 2;
		if (span_pixel >= u.vram_effective_width) return false;
		splat = span_pixel;
	}
	uint32_t addr = (uint32_t)(u.vram_addr + (splat << (u.vram_size - 1)));
	addr += 2 * (addr & 1) * (uint32_t)(p & 3);
	ld.addr = addr;
	ld.kind = TMEM_LOAD_16;
	return true;
}

// The read half: the halfword value from RDRAM.  32-bit sources keep the
// high halfword for the lower TMEM half and the low one for the upper half;
// YUV splits luma and chroma bytes the same way.

RDP_TP bool tmem_plan(const UploadInfo &u, int idx, bool upper, TmemLoad &ld)
{
	if (u.mode == UP_TLUT) return tmem_plan_lut(u, idx, ld);
	if (u.tmem_size == 3 || u.tmem_fmt == TF_YUV) return tmem_plan_32(u, idx & 0x3ff, ld);
	return tmem_plan_16(u, idx, ld);
}

}  // namespace rdp_mi

// Host-side bounds for ordering a round's TMEM replay against the previous
// round's raster: what a pass's raster can write and what a TMEM upload can
// read, as RDRAM byte ranges.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <vector>

#include "rdp_frontend.hpp"
#include "rdp_types.hpp"

namespace rdp_mi {

// RDRAM byte range [lo, hi) before the wrap at the RDRAM size.
struct ByteRange {
	uint64_t lo, hi;
};

// Every RDRAM byte the raster stage of a pass can write: colour image and
// depth image rows 0 .. fb_height-1 of fb_width pixels, word aligned.
inline void pass_write_ranges(const HostPass &hp, std::vector<ByteRange> &out)
{
	if (hp.prims.empty()) return;
	const uint64_t px = (uint64_t)hp.fb_width * hp.fb_height;
	uint64_t base, bytes;
	switch (hp.fb_fmt) {
	case FB_I4:
	case FB_I8: base = hp.fb_addr; bytes = px; break;
	case FB_RGBA5551:
	case FB_IA88: base = hp.fb_addr & ~1u; bytes = 2 * px; break;
	default: base = hp.fb_addr & ~3u; bytes = 4 * px; break;
	}
	out.push_back({ base & ~3ull, (base + bytes + 3) & ~3ull });
	if (hp.depth_addr != hp.fb_addr) {
		const uint64_t d = hp.depth_addr & ~1u;
		out.push_back({ d & ~3ull, (d + 2 * px + 3) & ~3ull });
	}
}

// A bound on the RDRAM bytes a TMEM upload reads: rows 0 .. height-1 of the
// source image (vram_width texels apart, at most 4 bytes each) and, for
// TLUT loads, vram_effective_width entries from vram_addr.  False when the
// record does not allow a bound.
inline bool upload_read_range(const UploadInfo &u, ByteRange &r)
{
	if (u.vram_addr < 0 || u.width < 0 || u.height < 0 || u.vram_width < 0 || u.vram_effective_width < 0) return false;
	const uint64_t w = (uint64_t)std::max(std::max(u.vram_width, u.width), u.vram_effective_width);
	r.lo = (uint64_t)u.vram_addr & ~3ull;
	r.hi = (uint64_t)u.vram_addr + (((uint64_t)u.height + 1) * w + w + 16) * 4 + 64;
	return true;
}

inline bool ranges_overlap(const ByteRange &a, const ByteRange &b, uint64_t size)
{
	if (a.hi - a.lo >= size || b.hi - b.lo >= size) return true;
	ByteRange pa[2], pb[2];
	auto split = [size](const ByteRange &x, ByteRange *out) {
		const uint64_t lo = x.lo % size, hi = lo + (x.hi - x.lo);
		if (hi <= size) {
			out[0] = { lo, hi };
			return 1;
		}
		out[0] = { lo, size };
		out[1] = { 0, hi - size };
		return 2;
	};
	const int na = split(a, pa), nb = split(b, pb);
	for (int i = 0; i < na; i++)
		for (int j = 0; j < nb; j++)
			if (pa[i].lo < pb[j].hi && pb[j].lo < pa[i].hi) return true;
	return false;
}

}  // namespace rdp_mi

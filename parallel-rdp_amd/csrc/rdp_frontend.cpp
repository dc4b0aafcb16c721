// Host front end: see rdp_frontend.hpp.
#include "rdp_frontend.hpp"

#include <string.h>

#include <algorithm>

namespace rdp_mi {

static inline int32_t sext32(int32_t v, int bits) { return (int32_t)((uint32_t)v << (32 - bits)) >> (32 - bits); }

Frontend::Frontend(uint32_t size) : rdram_size(size) {}

unsigned Frontend::command_length(uint32_t first_word)
{
	static const uint8_t lengths[64] = {
		2, 0, 0, 0, 2, 0, 0, 0, 8, 12, 24, 28, 24, 28, 40, 44,
		0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
		0, 0, 0, 0, 4, 4, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
		2, 0, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2,
	};
	return lengths[(first_word >> 24) & 63];
}

void Frontend::flush()
{
	if (!cur.empty()) {
		cur.fb_addr = fb_addr;
		cur.fb_width = fb_width;
		cur.fb_fmt = fb_fmt;
		cur.depth_addr = depth_addr;
		done.push_back(std::move(cur));
	}
	cur = HostPass();
	state_dirty = true;
	color_write_pending = false;
	depth_write_pending = false;
}

// ---------------------------------------------------------------------------
// Draw state resolution: build_derived_attributes / build_combiner_constants
// (rdp_renderer.cpp:754-948), deduce_static_texture_state (1333-1400),
// deduce_noise_state (1136-1169), normalize_static_state (1294-1331).
// ---------------------------------------------------------------------------
static bool accesses_texel0(const uint8_t c[8])
{
	return c[0] == 1 || c[1] == 1 || c[2] == 1 || c[3] == 1 || c[2] == 8 || c[4] == 1 || c[5] == 1 || c[6] == 1 ||
	       c[7] == 1;
}

static bool accesses_texel1(const uint8_t c[8])
{
	return c[0] == 2 || c[1] == 2 || c[2] == 2 || c[3] == 2 || c[2] == 9 || c[4] == 2 || c[5] == 2 || c[6] == 2 ||
	       c[7] == 2;
}

static bool accesses_lod_frac(const uint8_t c[8]) { return c[2] == 13 || c[6] == 0; }

static void rgb3(uint8_t *dst, uint32_t col)
{
	dst[0] = (uint8_t)(col >> 24);
	dst[1] = (uint8_t)(col >> 16);
	dst[2] = (uint8_t)(col >> 8);
}

void Frontend::resolve_state(DrawState &d) const
{
	memset(&d, 0, sizeof(d));
	memcpy(d.scissor, scissor, sizeof(scissor));
	memcpy(d.comb, comb, sizeof(comb));
	memcpy(d.blend, blend, sizeof(blend));
	d.db_flags = db_flags;
	d.coverage_mode = coverage_mode;
	d.z_mode = z_mode;
	d.dither = dither;
	d.raw_ras_flags = ras_flags;
	d.quirks = quirks;
	d.use_prim_depth = use_prim_depth ? 1 : 0;
	d.prim_depth = prim_depth;
	d.prim_dz = prim_dz;
	d.fill_color = fill_color;
	d.min_lod = min_level;
	for (int i = 0; i < 4; i++) {
		d.factors[i] = (int32_t)(int16_t)convert[i];
		d.fog[i] = (uint8_t)(fog_color >> (24 - 8 * i));
		d.blendc[i] = (uint8_t)(blend_color >> (24 - 8 * i));
	}
	memcpy(d.tiles, tiles, sizeof(tiles));

	for (int cyc = 0; cyc < 2; cyc++) {
		const uint8_t *c = comb[cyc];
		uint8_t(*k)[4] = d.constants[cyc];
		switch (c[0]) {
		case 5: rgb3(k[0], env_color); break;
		case 3: rgb3(k[0], prim_color); break;
		default: break;
		}
		switch (c[1]) {
		case 5: rgb3(k[1], env_color); break;
		case 3: rgb3(k[1], prim_color); break;
		case 7: rgb3(k[1], (uint32_t)convert[4] << 8); break;
		case 6:
			for (int i = 0; i < 3; i++) k[1][i] = (uint8_t)key_center[i];
			break;
		default: break;
		}
		switch (c[2]) {
		case 3: rgb3(k[2], prim_color); break;
		case 5: rgb3(k[2], env_color); break;
		case 10: rgb3(k[2], 0x01010101u * (prim_color & 0xff)); break;
		case 12: rgb3(k[2], 0x01010101u * (env_color & 0xff)); break;
		case 14: rgb3(k[2], 0x01010101u * prim_lod_frac); break;
		case 15: rgb3(k[2], (uint32_t)convert[5] << 8); break;
		case 6:
			for (int i = 0; i < 3; i++) k[2][i] = (uint8_t)key_scale[i];
			break;
		default: break;
		}
		switch (c[3]) {
		case 3: rgb3(k[3], prim_color); break;
		case 5: rgb3(k[3], env_color); break;
		default: break;
		}
		if (c[4] == 3) k[0][3] = (uint8_t)prim_color;
		else if (c[4] == 5) k[0][3] = (uint8_t)env_color;
		if (c[5] == 3) k[1][3] = (uint8_t)prim_color;
		else if (c[5] == 5) k[1][3] = (uint8_t)env_color;
		if (c[6] == 3) k[2][3] = (uint8_t)prim_color;
		else if (c[6] == 5) k[2][3] = (uint8_t)env_color;
		else if (c[6] == 6) k[2][3] = (uint8_t)prim_lod_frac;
		if (c[7] == 3) k[3][3] = (uint8_t)prim_color;
		else if (c[7] == 5) k[3][3] = (uint8_t)env_color;
	}

	uint32_t f = ras_flags;
	const bool multi = (f & RAS_MULTI_CYCLE) != 0;
	bool t0 = false, t1 = false, pt1 = false, lodf = false;
	if (!(f & (RAS_FILL | RAS_COPY))) {
		if (multi) {
			t0 = accesses_texel0(comb[0]) || accesses_texel1(comb[1]);
			t1 = accesses_texel1(comb[0]) || accesses_texel0(comb[1]);
			lodf = accesses_lod_frac(comb[0]) || accesses_lod_frac(comb[1]);
		} else {
			t0 = accesses_texel0(comb[1]);
			pt1 = accesses_texel1(comb[1]);
		}
		if (t1 && (f & RAS_CONVERT_ONE)) t0 = true;
	}
	bool need_noise = false, need_dual = false;
	if ((dither & 3) == 2 || ((dither >> 2) & 3) == 2)
		need_noise = true;
	else if (!(f & (RAS_COPY | RAS_FILL))) {
		if (multi && comb[0][0] == 7) need_noise = true;
		if (comb[1][0] == 7) need_noise = true;
		if (multi && comb[0][0] == 7 && comb[1][0] == 7) need_dual = true;
		if ((f & (RAS_ALPHA_TEST | RAS_ALPHA_TEST_DITHER)) == (RAS_ALPHA_TEST | RAS_ALPHA_TEST_DITHER))
			need_noise = true;
	}
	if (f & RAS_FILL) {
		f = RAS_FILL;
		d.dither = 0;
	} else if (f & RAS_COPY) {
		f &= RAS_COPY | RAS_TLUT | RAS_TLUT_TYPE | RAS_TEX_LOD | RAS_DETAIL_LOD | RAS_PERSPECTIVE | RAS_ALPHA_TEST;
		d.dither = 0;
	} else {
		if (!multi && !pt1) f &= ~(RAS_BILERP1 | RAS_CONVERT_ONE);
		if (t0) f |= RAS_USES_TEXEL0;
		if (t1) f |= RAS_USES_TEXEL1;
		if (pt1) f |= RAS_USES_PIPELINED_TEXEL1;
		if (lodf || (ras_flags & RAS_TEX_LOD)) f |= RAS_USES_LOD;
		if (need_noise) f |= RAS_NEED_NOISE;
		if (need_dual) f |= RAS_NEED_NOISE_DUAL;
	}
	d.ras_flags = f;
}

// Renderer::validate_draw_state, rdp_renderer.cpp:1425-1462.
void Frontend::validate_draw() const
{
	if (!validation) return;
	if (ras_flags & RAS_FILL) {
		if (fb_fmt == FB_I4) report(VALIDATION_FILL_4BPP, "Attempted to use Fill mode on 4bpp surface.");
		if (db_flags & DB_DEPTH_TEST) report(VALIDATION_FILL_DEPTH_TEST, "Attempted to use Fill mode with depth test.");
		if (db_flags & DB_IMAGE_READ)
			report(VALIDATION_FILL_IMAGE_READ, "Attempted to use Fill mode with image read enable.");
		if ((db_flags & DB_DEPTH_UPDATE) && !use_prim_depth)
			report(VALIDATION_FILL_DEPTH_WRITE, "Attempted to use Fill mode with depth write enabled.");
	} else if (ras_flags & RAS_COPY) {
		if (fb_fmt == FB_RGBA8888) report(VALIDATION_COPY_32BPP, "Attempted to use Copy mode on 32bpp surface.");
	}
}

void Frontend::draw(unsigned op, const uint32_t *w, unsigned len)
{
	validate_draw();
	const uint32_t prim_index = (uint32_t)prim_counter++;
	if (fb_width == 0) return;

	int32_t yl, yh;
	if (op >= 0x08 && op <= 0x0f) {
		yl = sext32((int32_t)w[0], 14);
		yh = sext32((int32_t)w[1], 14);
	} else {
		uint32_t l = w[0] & 0xfff;
		if (ras_flags & (RAS_COPY | RAS_FILL)) l |= 3;
		yl = (int32_t)l;
		yh = (int32_t)(w[1] & 0xfff);
	}
	// allocate_span_jobs, rdp_renderer.cpp:1523-1554
	const int32_t ylo = std::max(yh, (int32_t)scissor[1]) >> 2;
	const int32_t yhi = std::min(yl - 1, (int32_t)scissor[3] - 1) >> 2;
	if (yhi < ylo) return;

	const uint32_t lines = (uint32_t)(yhi - ylo + 2);
	const uint32_t rows = (uint32_t)((yhi >> 3) - (ylo >> 3) + 1);
	if (!cur.prims.empty() && cur.num_spans + lines > max_spans_per_pass) flush();

	if (state_dirty || cur.states.empty()) {
		DrawState d;
		resolve_state(d);
		cur.states.push_back(d);
		state_dirty = false;
	}

	PrimRecord r;
	memset(&r, 0, sizeof(r));
	r.word_offset = (uint32_t)cur.words.size();
	r.state_index = (uint16_t)(cur.states.size() - 1);
	r.tmem_instance = (uint8_t)cur.uploads.size();
	r.opcode = (uint8_t)op;
	r.ylo = ylo;
	r.yhi = yhi;
	r.span_offset = cur.num_spans;
	r.tilerow_offset = cur.num_tilerows;
	r.prim_index = prim_index;
	cur.words.insert(cur.words.end(), w, w + len);
	cur.prims.push_back(r);
	cur.num_spans += lines;
	cur.num_tilerows += rows;
	cur.fb_height = std::max(cur.fb_height, (uint32_t)(yhi + 1));
	color_write_pending = true;
	if (db_flags & DB_DEPTH_UPDATE) depth_write_pending = true;

	if (cur.prims.size() >= max_prims_per_pass || cur.states.size() >= 65535) flush();
}

// ---------------------------------------------------------------------------
// TMEM uploads: Renderer::load_tile / load_tile_iteration
// (rdp_renderer.cpp:3099-3486); float reciprocals become integer divisors.
// ---------------------------------------------------------------------------

// True when the load reads RDRAM bytes that a draw already recorded in the
// open pass writes (colour or depth image), so the pass must complete first.
// This is a full byte-range overlap test; the reference only tests the start
// address (rdp_renderer.cpp:3051-3097).
bool Frontend::load_needs_flush(uint32_t mode, uint32_t slo, uint32_t tlo, uint32_t shi, uint32_t thi) const
{
	if (!color_write_pending && !depth_write_pending) return false;
	if (teximg.size == 0) return true;
	uint64_t pixel_count, offset_pixels;
	if (mode == UP_BLOCK) {
		pixel_count = (shi - slo + 1) & 0xfff;
		offset_pixels = slo + (uint64_t)teximg.width * tlo;
	} else {
		uint64_t max_x = ((shi >> 2) - (slo >> 2)) & 0xfff;
		int64_t max_y = (int64_t)(thi >> 2) - (int64_t)(tlo >> 2);
		if (max_y < 0) max_y = 0;
		pixel_count = (uint64_t)max_y * teximg.width + max_x + 1;
		offset_pixels = (slo >> 2) + (uint64_t)teximg.width * (tlo >> 2);
	}
	const uint32_t shift = teximg.size - 1;
	uint64_t start = teximg.addr + (offset_pixels << shift);
	uint64_t end = start + (((pixel_count << shift) + 7) & ~7ull) + 16;
	const uint64_t mask = rdram_size - 1;

	auto overlaps = [&](uint64_t a0, uint64_t a1) {
		if (a1 - a0 >= rdram_size || end - start >= rdram_size) return true;
		// Compare modulo the RDRAM size by testing both wrapped copies.
		uint64_t s = start & mask, e = s + (end - start);
		uint64_t b0 = a0 & mask, b1 = b0 + (a1 - a0);
		for (int k = -1; k <= 1; k++) {
			uint64_t o = (uint64_t)((int64_t)rdram_size * k);
			if (s + o < b1 && b0 < e + o) return true;
		}
		return false;
	};

	uint32_t bpp_shift = fb_fmt == FB_RGBA8888 ? 2 : (fb_fmt == FB_RGBA5551 || fb_fmt == FB_IA88) ? 1 : 0;
	uint64_t pixels = (uint64_t)fb_width * cur.fb_height;
	if (color_write_pending && overlaps(fb_addr, fb_addr + ((pixels << bpp_shift) | 1))) return true;
	if (depth_write_pending && overlaps(depth_addr, depth_addr + (pixels << 1) + 1)) return true;
	return false;
}

void Frontend::load(uint32_t mode, const uint32_t *w)
{
	const uint32_t tile = (w[1] >> 24) & 7;
	const uint32_t slo = (w[0] >> 12) & 0xfff, shi = (w[1] >> 12) & 0xfff;
	const uint32_t tlo = w[0] & 0xfff, thi = w[1] & 0xfff;

	if (mode == UP_TLUT && (thi >> 2) > (tlo >> 2))
		report(VALIDATION_MULTILINE_TLUT, "Attempting to load multiple lines in TLUT.");

	if (load_needs_flush(mode, slo, tlo, shi, thi)) flush();

	if (mode != UP_BLOCK) {
		if ((thi >> 2) < (tlo >> 2)) return;
		if (((((shi >> 2) - (slo >> 2)) + 1) & 0xfff) == 0) return;
	} else {
		unsigned pc = ((shi - slo) + 1) & 0xfff;
		if (!pc || pc > 2048) return;
	}

	state_dirty = true;  // tile sizes change below
	TileDesc &t = tiles[tile];
	if (mode == UP_TILE) {
		unsigned px_per_line = (((shi >> 2) - (slo >> 2)) + 1) & 0xfff;
		unsigned qw = ((px_per_line << t.size) + 15) >> 4;
		if (t.size > teximg.size) qw <<= t.size - teximg.size;
		else if (t.size < teximg.size) qw >>= teximg.size - t.size;
		unsigned bytes_per_line = std::max<unsigned>(qw * 8, t.stride);
		unsigned max_bytes = t.fmt == TF_YUV ? 0x800 : 0x1000;
		unsigned num_lines = ((thi >> 2) - (tlo >> 2)) + 1;
		if (bytes_per_line * num_lines > max_bytes) {
			unsigned max_lines = (max_bytes / bytes_per_line) & ~1u;
			if (max_lines == 0) return;
			for (unsigned line = 0; line < num_lines; line += max_lines) {
				unsigned n = std::min(num_lines - line, max_lines);
				uint32_t tl = tlo + (line << 2);
				load_tile_iteration(tile, mode, slo, tl, shi, tl + ((n - 1) << 2), line * t.stride);
			}
			t.slo = slo;
			t.shi = shi;
			t.tlo = tlo;
			t.thi = thi;
			return;
		}
	}
	load_tile_iteration(tile, mode, slo, tlo, shi, thi, 0);
}

void Frontend::load_tile_iteration(uint32_t tile, uint32_t mode, uint32_t slo, uint32_t tlo, uint32_t shi, uint32_t thi,
                                   uint32_t tmem_offset)
{
	TileDesc &t = tiles[tile];
	t.slo = slo;
	t.shi = shi;
	t.tlo = tlo;
	t.thi = thi;
	const uint32_t isize = teximg.size;
	if (t.fmt == TF_YUV && (t.size != 2 || isize != 2)) return;
	if (isize == 0) {
		report(VALIDATION_LOAD_TILE_4BPP, "4-bit VRAM pointer crashes the RDP.");
		return;
	}
	if (t.size == 3 && t.fmt != TF_RGBA) return;
	if (mode == UP_TLUT && t.size == 3) return;
	if (mode != UP_TLUT) {
		if ((isize == 3 && t.size == 1) || (isize == 2 && t.size == 0) || (isize == 3 && t.size == 0)) return;
	}

	UploadInfo u;
	memset(&u, 0, sizeof(u));
	u.tmem_stride_words = (int32_t)(t.stride >> 1);
	uint32_t ux = 0, uy = 0;
	uint32_t m = mode;
	if (m == UP_BLOCK) {
		ux = slo;
		uy = tlo;
		unsigned pixel_count = (shi - slo + 1) & 0xfff;
		unsigned dt = thi;
		unsigned max_iter = (pixel_count - 1) >> (4u - isize);
		unsigned max_t = (max_iter * dt) >> 11;
		if (max_t != 0) {
			unsigned max_elems = ((1u << 11) + dt - 1u) / dt;
			unsigned min_elems = (1u << 11) / dt;
			bool uneven = max_elems != min_elems;
			if (uneven) {
				unsigned overflow_amt = (dt * max_elems - (1u << 11)) * max_t;
				if (overflow_amt < dt) {
					min_elems = max_elems;
					uneven = false;
				}
			}
			u.dxt = (int32_t)(dt << 5);
			if (t.size == 3 || t.fmt == TF_YUV) {
				u.tmem_stride_words <<= 1;
				if (uneven && t.size != isize) return;
			}
			if (t.size > isize) {
				unsigned sh = t.size - isize;
				max_elems <<= sh;
				min_elems <<= sh;
				u.dxt >>= sh;
			} else if (isize > t.size)
				return;
			u.max_line_stride = (int32_t)(max_elems + ((unsigned)u.tmem_stride_words >> 2));
			u.min_line_stride = (int32_t)(min_elems + ((unsigned)u.tmem_stride_words >> 2));
			u.width = (int32_t)pixel_count;
			u.height = 1;
			u.tmem_stride_words >>= 2;
		} else {
			u.width = (int32_t)pixel_count;
			u.height = 1;
			u.tmem_stride_words = 0;
			m = UP_TILE;
		}
	} else {
		ux = slo >> 2;
		uy = tlo >> 2;
		u.width = (int32_t)((((shi >> 2) - (slo >> 2)) + 1) & 0xfff);
		u.height = (int32_t)(((thi >> 2) - (tlo >> 2)) + 1);
	}
	if (!u.width) return;

	switch (isize) {
	case 1: u.vram_effective_width = (u.width + 7) & ~7; break;
	case 2: u.vram_effective_width = m == UP_TLUT ? u.width : ((u.width + 3) & ~3); break;
	case 3: u.vram_effective_width = (u.width + 1) & ~1; break;
	default: break;
	}
	switch (t.size) {
	case 0: u.width = ((u.width + 15) & ~15) >> 2; break;
	case 1: u.width = ((u.width + 7) & ~7) >> 1; break;
	case 2:
		u.width = (u.width + 3) & ~3;
		if (t.fmt == TF_YUV) u.width >>= 1;
		break;
	default: u.width = (u.width + 1) & ~1; break;
	}
	if (u.height > 1 && m == UP_TLUT) return;

	u.vram_addr = (int32_t)(teximg.addr + ((teximg.width * uy + ux) << (isize - 1)));
	u.vram_width = m == UP_BLOCK ? u.vram_effective_width : (int32_t)teximg.width;
	u.vram_size = (int32_t)isize;
	u.tmem_offset = (int32_t)((t.offset + tmem_offset) & 0xfff);
	u.tmem_size = (int32_t)t.size;
	u.tmem_fmt = (int32_t)t.fmt;
	u.mode = (int32_t)m;
	cur.uploads.push_back(u);
	state_dirty = true;  // draws after this upload see a new TMEM snapshot
	if (cur.uploads.size() + 1 >= (size_t)kMaxUploadsPerPass + 1) flush();
}

// ---------------------------------------------------------------------------
// Command dispatch, rdp_device.cpp:421-959.
// ---------------------------------------------------------------------------
#define SET_FLAG(var, cond, bit)              \
	do {                                      \
		if (cond) (var) |= (bit);             \
		else (var) &= ~(uint32_t)(bit);       \
	} while (0)

void Frontend::command(const uint32_t *w)
{
	const unsigned op = (w[0] >> 24) & 63;
	switch (op) {
	case 0x08: case 0x09: case 0x0a: case 0x0b: case 0x0c: case 0x0d: case 0x0e: case 0x0f:
	case 0x24: case 0x25: case 0x36:
		draw(op, w, command_length(w[0]));
		return;
	case 0x29:  // SyncFull ends the frame's work
		flush();
		return;
	case 0x26: case 0x27: case 0x28:  // SyncLoad / SyncPipe / SyncTile
		return;
	case 0x04:  // MetaSetQuirks
		quirks = w[1];
		break;
	case 0x2a:
		key_width[1] = (w[0] >> 12) & 0xfff;
		key_width[2] = w[0] & 0xfff;
		key_center[1] = (w[1] >> 24) & 0xff;
		key_scale[1] = (w[1] >> 16) & 0xff;
		key_center[2] = (w[1] >> 8) & 0xff;
		key_scale[2] = w[1] & 0xff;
		break;
	case 0x2b:
		key_width[0] = (w[1] >> 16) & 0xfff;
		key_center[0] = (w[1] >> 8) & 0xff;
		key_scale[0] = w[1] & 0xff;
		break;
	case 0x2c: {
		uint64_t mm = ((uint64_t)w[0] << 32) | w[1];
		for (int i = 0; i < 6; i++) {
			int32_t k = (int32_t)((mm >> (45 - 9 * i)) & 0x1ff);
			convert[i] = i < 4 ? 2 * sext32(k, 9) + 1 : k;
		}
		break;
	}
	case 0x2d:
		scissor[0] = (w[0] >> 12) & 0xfff;
		scissor[1] = w[0] & 0xfff;
		scissor[2] = (w[1] >> 12) & 0xfff;
		scissor[3] = w[1] & 0xfff;
		SET_FLAG(ras_flags, w[1] & (1u << 25), RAS_INTERLACE_FIELD);
		SET_FLAG(ras_flags, w[1] & (1u << 24), RAS_INTERLACE_KEEP_ODD);
		break;
	case 0x2e:
		prim_depth = (int32_t)((w[1] >> 16) & 0x7fff) << 16;
		prim_dz = w[1] & 0xffff;
		break;
	case 0x2f: {
		const uint32_t w0 = w[0], w1 = w[1];
		SET_FLAG(ras_flags, w0 & (1u << 19), RAS_PERSPECTIVE);
		SET_FLAG(ras_flags, w0 & (1u << 18), RAS_DETAIL_LOD);
		SET_FLAG(ras_flags, w0 & (1u << 17), RAS_SHARPEN_LOD);
		SET_FLAG(ras_flags, w0 & (1u << 16), RAS_TEX_LOD);
		SET_FLAG(ras_flags, w0 & (1u << 15), RAS_TLUT);
		SET_FLAG(ras_flags, w0 & (1u << 14), RAS_TLUT_TYPE);
		SET_FLAG(ras_flags, w0 & (1u << 13), RAS_SAMPLE_QUAD);
		SET_FLAG(ras_flags, w0 & (1u << 12), RAS_MID_TEXEL);
		SET_FLAG(ras_flags, w0 & (1u << 11), RAS_BILERP0);
		SET_FLAG(ras_flags, w0 & (1u << 10), RAS_BILERP1);
		SET_FLAG(ras_flags, w0 & (1u << 9), RAS_CONVERT_ONE);
		SET_FLAG(db_flags, w1 & (1u << 14), DB_FORCE_BLEND);
		SET_FLAG(ras_flags, w1 & (1u << 13), RAS_ALPHA_CVG_SELECT);
		SET_FLAG(ras_flags, w1 & (1u << 12), RAS_CVG_TIMES_ALPHA);
		SET_FLAG(db_flags, w1 & (1u << 7), DB_COLOR_ON_CVG);
		SET_FLAG(db_flags, w1 & (1u << 6), DB_IMAGE_READ);
		SET_FLAG(db_flags, w1 & (1u << 5), DB_DEPTH_UPDATE);
		SET_FLAG(db_flags, w1 & (1u << 4), DB_DEPTH_TEST);
		SET_FLAG(ras_flags, w1 & (1u << 3), RAS_AA);
		SET_FLAG(db_flags, w1 & (1u << 3), DB_AA);
		SET_FLAG(ras_flags, w1 & (1u << 1), RAS_ALPHA_TEST_DITHER);
		SET_FLAG(ras_flags, w1 & (1u << 0), RAS_ALPHA_TEST);
		dither = (w0 >> 4) & 0xf;
		SET_FLAG(db_flags, (dither >> 2) != 3, DB_DITHER);
		coverage_mode = (w1 >> 8) & 3;
		z_mode = (w1 >> 10) & 3;
		ras_flags &= ~(uint32_t)(RAS_MULTI_CYCLE | RAS_FILL | RAS_COPY);
		db_flags &= ~(uint32_t)DB_MULTI_CYCLE;
		switch ((w0 >> 20) & 3) {
		case 1: ras_flags |= RAS_MULTI_CYCLE; db_flags |= DB_MULTI_CYCLE; break;
		case 2: ras_flags |= RAS_COPY; break;
		case 3: ras_flags |= RAS_FILL; break;
		default: break;
		}
		blend[0][0] = (w1 >> 30) & 3;
		blend[1][0] = (w1 >> 28) & 3;
		blend[0][1] = (w1 >> 26) & 3;
		blend[1][1] = (w1 >> 24) & 3;
		blend[0][2] = (w1 >> 22) & 3;
		blend[1][2] = (w1 >> 20) & 3;
		blend[0][3] = (w1 >> 18) & 3;
		blend[1][3] = (w1 >> 16) & 3;
		use_prim_depth = ((w1 >> 2) & 1) != 0;
		break;
	}
	case 0x30: load(UP_TLUT, w); return;
	case 0x32: {
		TileDesc &t = tiles[(w[1] >> 24) & 7];
		t.slo = (w[0] >> 12) & 0xfff;
		t.shi = (w[1] >> 12) & 0xfff;
		t.tlo = w[0] & 0xfff;
		t.thi = w[1] & 0xfff;
		break;
	}
	case 0x33: load(UP_BLOCK, w); return;
	case 0x34: load(UP_TILE, w); return;
	case 0x35: {
		TileDesc &t = tiles[(w[1] >> 24) & 7];
		t.offset = ((w[0] >> 0) & 511) << 3;
		t.stride = ((w[0] >> 9) & 511) << 3;
		t.size = (w[0] >> 19) & 3;
		t.fmt = (w[0] >> 21) & 7;
		t.palette = (w[1] >> 20) & 15;
		t.shift_s = (w[1] >> 0) & 15;
		t.mask_s = (w[1] >> 4) & 15;
		t.shift_t = (w[1] >> 10) & 15;
		t.mask_t = (w[1] >> 14) & 15;
		t.flags = 0;
		if (w[1] & (1u << 8)) t.flags |= TILE_MIRROR_S;
		if (w[1] & (1u << 9)) t.flags |= TILE_CLAMP_S;
		if (w[1] & (1u << 18)) t.flags |= TILE_MIRROR_T;
		if (w[1] & (1u << 19)) t.flags |= TILE_CLAMP_T;
		if (t.mask_s > 10) t.mask_s = 10;
		else if (t.mask_s == 0) t.flags |= TILE_CLAMP_S;
		if (t.mask_t > 10) t.mask_t = 10;
		else if (t.mask_t == 0) t.flags |= TILE_CLAMP_T;
		break;
	}
	case 0x37: fill_color = w[1]; break;
	case 0x38: fog_color = w[1]; break;
	case 0x39: blend_color = w[1]; break;
	case 0x3a:
		min_level = (w[0] >> 8) & 31;
		prim_lod_frac = w[0] & 0xff;
		prim_color = w[1];
		break;
	case 0x3b: env_color = w[1]; break;
	case 0x3c: {
		const uint32_t w0 = w[0], w1 = w[1];
		comb[0][0] = (w0 >> 20) & 0xf;
		comb[0][2] = (w0 >> 15) & 0x1f;
		comb[0][1] = (w1 >> 28) & 0xf;
		comb[0][3] = (w1 >> 15) & 0x7;
		comb[0][4] = (w0 >> 12) & 0x7;
		comb[0][5] = (w1 >> 12) & 0x7;
		comb[0][6] = (w0 >> 9) & 0x7;
		comb[0][7] = (w1 >> 9) & 0x7;
		comb[1][0] = (w0 >> 5) & 0xf;
		comb[1][2] = (w0 >> 0) & 0x1f;
		comb[1][1] = (w1 >> 24) & 0xf;
		comb[1][3] = (w1 >> 6) & 0x7;
		comb[1][4] = (w1 >> 21) & 0x7;
		comb[1][5] = (w1 >> 3) & 0x7;
		comb[1][6] = (w1 >> 18) & 0x7;
		comb[1][7] = (w1 >> 0) & 0x7;
		break;
	}
	case 0x3d:
		teximg.fmt = (w[0] >> 21) & 7;
		teximg.size = (w[0] >> 19) & 3;
		teximg.width = (w[0] & 0x3ff) + 1;
		teximg.addr = w[1] & 0x00ffffffu;
		return;  // does not affect draw state
	case 0x3e: {
		uint32_t a = w[1] & 0xffffff;
		if (a != depth_addr) flush();
		depth_addr = a;
		return;
	}
	case 0x3f: {
		const unsigned fmt = (w[0] >> 21) & 7, size = (w[0] >> 19) & 3;
		uint32_t width = (w[0] & 1023) + 1, addr = w[1] & 0xffffff;
		uint32_t ff;
		switch (size) {
		case 0: ff = FB_I4; break;
		case 1: ff = FB_I8; break;
		case 2: ff = fmt ? FB_IA88 : FB_RGBA5551; break;
		default: ff = FB_RGBA8888; break;
		}
		if (addr != fb_addr || width != fb_width || ff != fb_fmt) flush();
		fb_addr = addr;
		fb_width = width;
		fb_fmt = ff;
		return;
	}
	default:
		return;
	}
	state_dirty = true;
}

}  // namespace rdp_mi

// Host front end: see rdp_frontend.hpp.
#include "rdp_frontend.hpp"

#include <string.h>

#include <algorithm>

namespace rdp_mi {

static inline int32_t sext32(int32_t v, int bits) { return (int32_t)((uint32_t)v << (32 - bits)) >> (32 - bits); }

static bool accesses_texel0(const uint8_t c[8]);
static bool accesses_texel1(const uint8_t c[8]);
static bool accesses_lod_frac(const uint8_t c[8]);

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

// The framebuffer rows the upscaled domain updates for the reference pass
// now ending (0: none).  Without super-sampled readback, a pass that renders
// (render_pass_is_upscaled / submit_render_pass, rdp_renderer.cpp:2223-2246,
// 2988-3010); with it, every pass with span jobs or TMEM uploads (flush_queues,
// rdp_renderer.cpp:2908-2916, then submit_render_pass_upscaled).
uint32_t Frontend::ref_domain_height() const
{
	if (!scale_log2 || !fb_width) return 0;
	const bool work = ssaa ? (ref.span_jobs || ref.uploads) : (ref.span_jobs != 0);
	return work ? ref.deduced_height : 0;
}

// Hands a completed device pass over.  domain_height: the largest region of
// the reference passes it holds (ref_domain_height).  Merged
// reference passes of one framebuffer update the union of their regions:
// each one's update leaves the rows an earlier one already synchronised
// unchanged, and a pass writes no row past its own region.
void Frontend::complete_pass(HostPass &&p, uint32_t domain_height)
{
	p.fb_addr = fb_addr;
	p.fb_width = fb_width;
	p.fb_fmt = fb_fmt;
	p.depth_addr = depth_addr;
	p.marks.clear();
	if (scale_log2 && ssaa) {
		// Super-sampled readback: the pre update over the region, the draws
		// at scale only (with write masks), then the resolve.
		p.domain_height = domain_height;
		p.domain_pre = p.ssaa_clear = domain_height != 0;
		HostPass up;
		if (domain_height && !p.prims.empty()) {
			up = upscaled_copy(p);
			up.domain_height = domain_height;
			up.ssaa_resolve = true;
		}
		// The native pass keeps its uploads (the TMEM snapshots the upscaled
		// pass reads) and draws nothing.
		p.prims.clear();
		p.words.clear();
		p.states.clear();
		p.num_spans = p.num_tilerows = p.fb_height = p.raster_class = 0;
		done.push_back(std::move(p));
		if (!up.prims.empty()) done.push_back(std::move(up));
	} else if (scale_log2) {
		// A pass that renders runs the domain's pre update over its region; a
		// 16/32-bit colour pass at least 256 pixels wide is then rendered
		// again at scale (should_render_upscaled, rdp_renderer.cpp:2999-3010).
		const bool upscaled =
		    domain_height && (fb_fmt == FB_RGBA5551 || fb_fmt == FB_RGBA8888) && fb_width >= 256;
		p.domain_height = domain_height;
		p.domain_pre = domain_height != 0;
		p.domain_post = upscaled;
		done.push_back(std::move(p));
		if (upscaled) done.push_back(upscaled_copy(done.back()));
	} else
		done.push_back(std::move(p));
}

void Frontend::flush()
{
	const uint32_t dh = std::max(domain_height, ref_domain_height());
	if (!cur.empty()) complete_pass(std::move(cur), dh);
	domain_height = 0;
	cur = HostPass();
	state_dirty = true;
	color_write_pending = false;
	depth_write_pending = false;
	if (pass_end_log && ref.prims) pass_end_log->push_back(prim_counter);
	ref.reset();
}

// The same draws at S = 2^scale_log2 times the resolution
// (submit_render_pass_upscaled, rdp_renderer.cpp:2287-2342): lines
// S * ylo .. S * yhi + S - 1 per primitive (span_setup.comp:89-95), an
// S-times framebuffer, no TMEM uploads of its own (it reads the native
// pass's TMEM snapshots).
HostPass Frontend::upscaled_copy(const HostPass &n) const
{
	const int32_t S = 1 << scale_log2;
	HostPass u;
	u.words = n.words;
	u.states = n.states;
	u.prims = n.prims;
	u.fb_addr = n.fb_addr;
	u.fb_width = n.fb_width * (uint32_t)S;
	u.fb_fmt = n.fb_fmt;
	u.depth_addr = n.depth_addr;
	u.raster_class = n.raster_class;
	u.scale_log2 = scale_log2;
	u.native_uploads = (uint32_t)n.uploads.size();
	uint32_t spans = 0, rows = 0, height = 0;
	for (PrimRecord &r : u.prims) {
		r.ylo *= S;
		r.yhi = r.yhi * S + (S - 1);
		r.span_offset = spans;
		r.tilerow_offset = rows;
		spans += (uint32_t)(r.yhi - r.ylo + 2);
		rows += (uint32_t)((r.yhi >> 3) - (r.ylo >> 3) + 1);
		height = std::max(height, (uint32_t)(r.yhi + 1));
	}
	u.num_spans = spans;
	u.num_tilerows = rows;
	u.fb_height = height;
	return u;
}

// ---------------------------------------------------------------------------
// Render-pass boundaries.
//
// The reference renders each pass tile by tile and keeps every pixel's colour
// and depth in registers from the pass's first primitive to its last
// (depth_blend.comp:86-149).  Ending a pass writes those registers back and
// the next pass reads RDRAM again.  For RGBA5551, IA88, RGBA8888 and I4 that
// round trip leaves every value a later primitive reads unchanged, so passes
// can be merged; for I8 colour images (odd pixels store the green channel)
// and for colour/depth aliasing it does not, and a TMEM load issued after a
// boundary sees what the earlier passes wrote.  The front end therefore
// tracks where the reference ends its passes (RefPassState) and either ends
// its own pass there (I8, aliasing) or records a PassMark that can later be
// turned into a real boundary (a load that reads pixels written before the
// mark, or a device-side size limit).
// ---------------------------------------------------------------------------
void RefPassState::reset()
{
	statics.clear();
	depth_blends.clear();
	tiles.clear();
	prims = span_jobs = uploads = max_shaded_tiles = deduced_height = 0;
	color_write_pending = depth_write_pending = false;
	key_gen = ~0ull;
}

template <typename T>
static void cache_add(std::vector<T> &cache, const T &key)
{
	for (size_t i = cache.size(); i-- > 0;)
		if (memcmp(&cache[i], &key, sizeof(T)) == 0) return;
	cache.push_back(key);
}

static uint8_t keep_or(uint8_t v, uint32_t keep_mask, uint8_t zero)
{
	return (v < 32 && ((keep_mask >> v) & 1)) ? v : zero;
}

struct EdgeSetup {
	int32_t xh, xm, xl, yh, ym, yl, dxhdy, dxmdy, dxldy;
	bool flip;
};

// interpolate_x, rdp_renderer.cpp:962-992: the edges at sub-scanline y of
// the S-times domain (S = 1 without upscaling).
static void interpolate_x_64(const EdgeSetup &s, int y, int64_t S, int64_t &left, int64_t &right)
{
	const int64_t yh_base = (int64_t)(s.yh & ~3) * S, ym_base = (int64_t)s.ym * S;
	const int64_t xh = S * s.xh + ((int64_t)y - yh_base) * s.dxhdy;
	const int64_t xm = S * s.xm + ((int64_t)y - yh_base) * s.dxmdy;
	int64_t xl = S * s.xl + ((int64_t)y - ym_base) * s.dxldy;
	if (y < S * s.ym) xl = xm;
	const int64_t hs = xh >> 15, ls = xl >> 15;
	left = std::min<int64_t>(std::max<int64_t>(s.flip ? hs : ls, INT32_MIN), INT32_MAX);
	right = std::min<int64_t>(std::max<int64_t>(s.flip ? ls : hs, INT32_MIN), INT32_MAX);
}

// Renderer::compute_conservative_max_num_tiles, rdp_renderer.cpp:994-1060:
// 8x8 tiles of the S-times domain, S the upscaling factor.
static uint32_t conservative_tile_count(const EdgeSetup &s, const uint32_t scissor[4], int S)
{
	if (s.yl <= s.yh) return 0;
	const int start_y = std::max(s.yh & ~3, (int)scissor[1]) * S;
	const int end_y = std::min((s.yl - 1) | 3, (int)scissor[3] - 1) * S;
	if (end_y < start_y) return 0;
	int64_t l[4], r[4];
	interpolate_x_64(s, start_y, S, l[0], r[0]);
	interpolate_x_64(s, end_y, S, l[1], r[1]);
	l[2] = l[3] = l[0];
	r[2] = r[3] = r[0];
	const int ym = S * s.ym;
	if (ym > start_y && ym < end_y) {
		interpolate_x_64(s, ym, S, l[2], r[2]);
		interpolate_x_64(s, ym - 1, S, l[3], r[3]);
	}
	int64_t sx = std::min(std::min(l[0], l[1]), std::min(l[2], l[3]));
	int64_t ex = std::max(std::max(r[0], r[1]), std::max(r[2], r[3]));
	if (std::max(sx < 0 ? -sx : sx, ex < 0 ? -ex : ex) > 2047 * (int64_t)S) {
		sx = 0;
		ex = 0x7fffffff;
	}
	sx = std::max<int64_t>(sx, (int64_t)S * (scissor[0] >> 2));
	ex = std::min<int64_t>(ex, (int64_t)S * ((scissor[2] + 3) >> 2) - 1);
	if (ex < sx) return 0;
	return (uint32_t)((ex / 8 - sx / 8 + 1) * (end_y / 32 - start_y / 32 + 1));
}

// The per-draw half of Renderer::draw_shaded_primitive that feeds
// need_flush (rdp_renderer.cpp:1464-1521, 1556-1590).  Returns true when the
// reference ends its render pass after this draw.
bool Frontend::ref_account_draw(unsigned op, const uint32_t *w)
{
	// Edge setup as decoded by rdp_device.cpp:268-294 and 713-821, after
	// fixup_triangle_setup (rdp_renderer.cpp:1402-1409).
	EdgeSetup e;
	if (op >= 0x08 && op <= 0x0f) {
		e.flip = (w[0] & 0x800000u) != 0;
		e.yl = sext32((int32_t)w[0], 14);
		e.ym = sext32((int32_t)(w[1] >> 16), 14);
		e.yh = sext32((int32_t)w[1], 14);
		e.xl = sext32((int32_t)w[2], 28) >> 1;
		e.dxldy = sext32((int32_t)(w[3] >> 2), 28) >> 1;
		e.xh = sext32((int32_t)w[4], 28) >> 1;
		e.dxhdy = sext32((int32_t)(w[5] >> 2), 28) >> 1;
		e.xm = sext32((int32_t)w[6], 28) >> 1;
		e.dxmdy = sext32((int32_t)(w[7] >> 2), 28) >> 1;
	} else {
		uint32_t yl = w[0] & 0xfff;
		if (ras_flags & (RAS_COPY | RAS_FILL)) yl |= 3;
		e.flip = true;
		e.xh = (int32_t)(((w[1] >> 12) & 0xfff) << 13);
		e.xl = e.xm = (int32_t)(((w[0] >> 12) & 0xfff) << 13);
		e.yl = e.ym = (int32_t)yl;
		e.yh = (int32_t)(w[1] & 0xfff);
		e.dxhdy = e.dxmdy = e.dxldy = 0;
	}
	if (e.ym < (e.yh & ~3)) e.ym = 0x7fff;

	const int S = 1 << scale_log2;
	ref.max_shaded_tiles += conservative_tile_count(e, scissor, S);
	const int max_sub = std::min(e.yl - 1, (int)scissor[3] - 1);
	ref.deduced_height = std::max<uint32_t>(ref.deduced_height, (uint32_t)std::max((max_sub >> 2) + 1, 0));
	const int min_line = std::max(e.yh, (int)scissor[1]) >> 2, max_line = max_sub >> 2;
	if (max_line >= min_line) ref.span_jobs += (uint32_t)((std::min(max_line - min_line + 2, 1024) + 63) / 64);

	const unsigned setup_tile = op >= 0x08 && op <= 0x0f ? (w[0] >> 16) & 63 : (op == 0x36 ? 0 : (w[1] >> 24) & 7);
	if (ref.key_gen != state_gen || ref.key_setup_tile != setup_tile) {
		add_ref_keys(setup_tile);
		ref.key_gen = state_gen;
		ref.key_setup_tile = setup_tile;
	}

	ref.color_write_pending = true;
	if (db_flags & DB_DEPTH_UPDATE) ref.depth_write_pending = true;
	ref.prims++;

	// The tile-instance capacity grows with the factor squared (rdp_renderer.cpp:63-67).
	const uint32_t tiles_side = 128u * (uint32_t)S;
	return ref.statics.size() == 64 || ref.depth_blends.size() == 64 || ref.tiles.size() + 8 > 256 ||
	       ref.prims == 256 || ref.span_jobs * 64 + 1024 > 32 * 1024 ||
	       ref.max_shaded_tiles + tiles_side * tiles_side > (uint32_t)(S * S) * 0x8000u;
}

// Adds the draw's static, depth/blend and tile keys to the pass caches:
// normalize_static_state after deduce_static_texture_state and
// deduce_noise_state (rdp_renderer.cpp:1136-1400), and the depth/blend and
// tile caches of draw_shaded_primitive.  The keys depend only on the draw
// state and the setup tile.
void Frontend::add_ref_keys(unsigned setup_tile)
{
	RefPassState::StaticKey k;
	memset(&k, 0, sizeof(k));
	const unsigned tile = setup_tile & 7;
	uint32_t f = ras_flags;
	if (f & RAS_FILL) {
		k.flags = RAS_FILL;
	} else if (f & RAS_COPY) {
		k.flags = (f & (RAS_COPY | RAS_TLUT | RAS_TLUT_TYPE | RAS_TEX_LOD | RAS_DETAIL_LOD | RAS_PERSPECTIVE |
		                RAS_ALPHA_TEST)) |
		          RAS_USE_STATIC_TEXTURE_SIZE_FORMAT;
		k.texture_fmt = tiles[tile].fmt;
		k.texture_size = tiles[tile].size;
	} else {
		const bool multi = (f & RAS_MULTI_CYCLE) != 0;
		bool t0, t1 = false, pt1 = false, lodf = false;
		if (multi) {
			t0 = accesses_texel0(comb[0]) || accesses_texel1(comb[1]);
			t1 = accesses_texel1(comb[0]) || accesses_texel0(comb[1]);
			lodf = accesses_lod_frac(comb[0]) || accesses_lod_frac(comb[1]);
		} else {
			t0 = accesses_texel0(comb[1]);
			pt1 = accesses_texel1(comb[1]);
		}
		if (t1 && (f & RAS_CONVERT_ONE)) t0 = true;
		if (t0) f |= RAS_USES_TEXEL0;
		if (t1) f |= RAS_USES_TEXEL1;
		if (pt1) f |= RAS_USES_PIPELINED_TEXEL1;
		if (lodf || (f & RAS_TEX_LOD)) f |= RAS_USES_LOD;
		if (t0 || t1 || pt1) {
			const bool physical_t1 = t1 && (!(f & RAS_CONVERT_ONE) || (f & RAS_BILERP1));
			unsigned max_lod = setup_tile >> 3;
			if (!(f & RAS_TEX_LOD)) max_lod = physical_t1 ? 1 : 0;
			if (f & RAS_DETAIL_LOD) max_lod++;
			max_lod = std::min(max_lod, 7u);
			bool same = true;
			for (unsigned i = 1; i <= max_lod && same; i++) {
				const TileDesc &t = tiles[(tile + i) & 7];
				same = t.fmt == tiles[tile].fmt && t.size == tiles[tile].size;
			}
			if (same) {
				f |= RAS_USE_STATIC_TEXTURE_SIZE_FORMAT;
				k.texture_fmt = tiles[tile].fmt;
				k.texture_size = tiles[tile].size;
			}
		}
		if ((dither & 3) == 2 || ((dither >> 2) & 3) == 2)
			f |= RAS_NEED_NOISE;
		else {
			if (multi && comb[0][0] == 7) f |= RAS_NEED_NOISE;
			if (comb[1][0] == 7) f |= RAS_NEED_NOISE;
			if (multi && comb[0][0] == 7 && comb[1][0] == 7) f |= RAS_NEED_NOISE_DUAL;
			if ((f & (RAS_ALPHA_TEST | RAS_ALPHA_TEST_DITHER)) == (RAS_ALPHA_TEST | RAS_ALPHA_TEST_DITHER))
				f |= RAS_NEED_NOISE;
		}
		if (!(f & (RAS_MULTI_CYCLE | RAS_USES_PIPELINED_TEXEL1))) f &= ~(RAS_BILERP1 | RAS_CONVERT_ONE);
		// normalize_combiner, rdp_renderer.cpp:1171-1292.
		for (int c = 0; c < 2; c++) {
			k.comb[c][0] = keep_or(comb[c][0], 0xd7u, 8);
			k.comb[c][1] = keep_or(comb[c][1], 0x97u, 8);
			k.comb[c][2] = keep_or(comb[c][2], 0xab97u, 16);
			k.comb[c][3] = keep_or(comb[c][3], 0x57u, 7);
			k.comb[c][4] = keep_or(comb[c][4], 0x57u, 7);
			k.comb[c][5] = keep_or(comb[c][5], 0x57u, 7);
			k.comb[c][6] = keep_or(comb[c][6], 0x17u, 7);
			k.comb[c][7] = keep_or(comb[c][7], 0x57u, 7);
		}
		k.flags = f;
		k.dither = dither;
	}
	cache_add(ref.statics, k);

	RefPassState::DepthBlendKey dk;
	memset(&dk, 0, sizeof(dk));
	memcpy(dk.blend, blend, sizeof(dk.blend));
	dk.flags = db_flags;
	dk.coverage_mode = coverage_mode;
	dk.z_mode = z_mode;
	cache_add(ref.depth_blends, dk);
	for (int i = 0; i < 8; i++) cache_add(ref.tiles, tiles[i]);
}

// The reference ends a render pass here (flush_queues).
void Frontend::ref_pass_end()
{
	if (fb_fmt == FB_I8 || fb_addr == depth_addr) {
		flush();
		return;
	}
	domain_height = std::max(domain_height, ref_domain_height());
	if (!cur.empty()) {
		PassMark m;
		m.domain_height = domain_height;
		m.prims = (uint32_t)cur.prims.size();
		m.uploads = (uint32_t)cur.uploads.size();
		m.words = (uint32_t)cur.words.size();
		m.states = (uint32_t)cur.states.size();
		m.spans = cur.num_spans;
		m.tilerows = cur.num_tilerows;
		m.fb_height = cur.fb_height;
		m.color_written = color_write_pending;
		m.depth_written = depth_write_pending;
		cur.marks.push_back(m);
	}
	if (pass_end_log && ref.prims) pass_end_log->push_back(prim_counter);
	ref.reset();
}

// Turn the last PassMark into a real boundary: everything before it becomes
// a completed pass, the rest stays open.
void Frontend::split_at_last_mark()
{
	const PassMark m = cur.marks.back();
	HostPass head;
	head.words.assign(cur.words.begin(), cur.words.begin() + m.words);
	head.states.assign(cur.states.begin(), cur.states.begin() + m.states);
	head.prims.assign(cur.prims.begin(), cur.prims.begin() + m.prims);
	head.uploads.assign(cur.uploads.begin(), cur.uploads.begin() + m.uploads);
	head.fb_addr = fb_addr;
	head.fb_width = fb_width;
	head.fb_fmt = fb_fmt;
	head.depth_addr = depth_addr;
	head.fb_height = m.fb_height;
	head.num_spans = m.spans;
	head.num_tilerows = m.tilerows;
	for (const PrimRecord &r : head.prims) head.raster_class = std::max<uint32_t>(head.raster_class, r.raster_class);
	if (!head.empty()) complete_pass(std::move(head), m.domain_height);
	// What stays open is the reference pass after the mark.
	domain_height = 0;

	HostPass tail;
	tail.words.assign(cur.words.begin() + m.words, cur.words.end());
	tail.uploads.assign(cur.uploads.begin() + m.uploads, cur.uploads.end());
	std::vector<int32_t> remap(cur.states.size(), -1);
	color_write_pending = depth_write_pending = false;
	for (size_t i = m.prims; i < cur.prims.size(); i++) {
		PrimRecord r = cur.prims[i];
		if (remap[r.state_index] < 0) {
			remap[r.state_index] = (int32_t)tail.states.size();
			tail.states.push_back(cur.states[r.state_index]);
		}
		if (cur.states[r.state_index].db_flags & DB_DEPTH_UPDATE) depth_write_pending = true;
		color_write_pending = true;
		r.state_index = (uint16_t)remap[r.state_index];
		r.word_offset -= m.words;
		r.tmem_instance = (uint16_t)(r.tmem_instance - m.uploads);
		r.span_offset -= m.spans;
		r.tilerow_offset -= m.tilerows;
		tail.fb_height = std::max(tail.fb_height, (uint32_t)(r.yhi + 1));
		tail.raster_class = std::max<uint32_t>(tail.raster_class, r.raster_class);
		tail.prims.push_back(r);
	}
	tail.num_spans = cur.num_spans - m.spans;
	tail.num_tilerows = cur.num_tilerows - m.tilerows;
	cur = std::move(tail);
	state_dirty = true;
	state_serial++;
}

void Frontend::split_or_flush()
{
	if (!cur.marks.empty())
		split_at_last_mark();
	else
		flush();
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
	if (fb_width != 0 && yhi >= ylo) {
		const uint32_t lines = (uint32_t)(yhi - ylo + 2);
		const uint32_t rows = (uint32_t)((yhi >> 3) - (ylo >> 3) + 1);
		if (!cur.prims.empty() && cur.num_spans + lines > max_spans_per_pass) split_or_flush();

		if (state_dirty || cur.states.empty()) {
			DrawState d;
			resolve_state(d);
			cur.states.push_back(d);
			state_dirty = false;
			state_serial++;
		}
		// The raster variant for this state and setup tile (k_raster
		// dispatches on it) and its occupancy class.
		const uint32_t setup_tile = (op >= 0x08 && op <= 0x0f) ? (w[0] >> 16) & 63 : (op == 0x36 ? 0 : (w[1] >> 24) & 7);
		if (class_state_serial != state_serial || class_setup_tile != setup_tile) {
			const DrawState &st = cur.states.back();
			variant_value = select_raster_variant(st.ras_flags, st.comb, st.tiles, setup_tile, st.db_flags, st.blend,
			                                      st.z_mode, st.coverage_mode);
			class_value = raster_variant_class(variant_value);
			class_state_serial = state_serial;
			class_setup_tile = setup_tile;
		}

		PrimRecord r;
		memset(&r, 0, sizeof(r));
		r.word_offset = (uint32_t)cur.words.size();
		r.state_index = (uint16_t)(cur.states.size() - 1);
		r.tmem_instance = (uint16_t)cur.uploads.size();
		r.opcode = (uint8_t)op;
		r.ylo = ylo;
		r.yhi = yhi;
		r.span_offset = cur.num_spans;
		r.tilerow_offset = cur.num_tilerows;
		r.prim_index = prim_index;
		r.raster_class = (uint8_t)class_value;
		r.variant = (uint16_t)variant_value;
		cur.raster_class = std::max(cur.raster_class, class_value);
		cur.words.insert(cur.words.end(), w, w + len);
		cur.prims.push_back(r);
		cur.num_spans += lines;
		cur.num_tilerows += rows;
		cur.fb_height = std::max(cur.fb_height, (uint32_t)(yhi + 1));
		color_write_pending = true;
		if (db_flags & DB_DEPTH_UPDATE) depth_write_pending = true;
	}

	if (ref_account_draw(op, w)) ref_pass_end();
	if (cur.prims.size() >= max_prims_per_pass || cur.states.size() >= 65535) split_or_flush();
}

// ---------------------------------------------------------------------------
// TMEM uploads: Renderer::load_tile / load_tile_iteration
// (rdp_renderer.cpp:3099-3486); float reciprocals become integer divisors.
// ---------------------------------------------------------------------------

// Renderer::tmem_upload_needs_flush, rdp_renderer.cpp:3051-3097: only the
// texture image's start address is compared with the pending images.
bool Frontend::ref_upload_needs_flush(uint32_t addr) const
{
	const uint32_t mask = rdram_size - 1;
	const uint32_t pending = ref.deduced_height * fb_width;
	if (ref.color_write_pending) {
		uint32_t offset = (addr - fb_addr) & mask;
		if (fb_fmt == FB_RGBA5551 || fb_fmt == FB_I8)
			offset >>= 1;
		else if (fb_fmt == FB_RGBA8888)
			offset >>= 2;
		if (offset < pending) return true;
	}
	if (ref.depth_write_pending && (((addr - depth_addr) & mask) >> 1) < pending) return true;
	return false;
}

// True when the load may read RDRAM bytes written by primitives recorded
// before the open pass's last PassMark (a conservative byte-range test,
// modulo the RDRAM size).  The reference has written those pixels back by the
// time it runs the upload, so the pass has to be cut at that mark.
bool Frontend::load_reads_marked_writes(uint32_t mode, uint32_t slo, uint32_t tlo, uint32_t shi, uint32_t thi) const
{
	const PassMark &m = cur.marks.back();
	if (!m.color_written && !m.depth_written) return false;
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
	const uint64_t start = teximg.addr + (offset_pixels << shift);
	const uint64_t end = start + (((pixel_count << shift) + 7) & ~7ull) + 16;
	const uint64_t mask = rdram_size - 1;

	auto overlaps = [&](uint64_t a0, uint64_t a1) {
		if (a1 - a0 >= rdram_size || end - start >= rdram_size) return true;
		// Compare modulo the RDRAM size by testing the wrapped copies.
		uint64_t s = start & mask, e = s + (end - start);
		uint64_t b0 = a0 & mask, b1 = b0 + (a1 - a0);
		for (int k = -1; k <= 1; k++) {
			uint64_t o = (uint64_t)((int64_t)rdram_size * k);
			if (s + o < b1 && b0 < e + o) return true;
		}
		return false;
	};

	uint32_t bpp_shift = fb_fmt == FB_RGBA8888 ? 2 : (fb_fmt == FB_RGBA5551 || fb_fmt == FB_IA88) ? 1 : 0;
	uint64_t pixels = (uint64_t)fb_width * m.fb_height;
	if (m.color_written && overlaps(fb_addr, fb_addr + ((pixels << bpp_shift) | 1))) return true;
	if (m.depth_written && overlaps(depth_addr, depth_addr + (pixels << 1) + 1)) return true;
	return false;
}

void Frontend::load(uint32_t mode, const uint32_t *w)
{
	const uint32_t tile = (w[1] >> 24) & 7;
	const uint32_t slo = (w[0] >> 12) & 0xfff, shi = (w[1] >> 12) & 0xfff;
	const uint32_t tlo = w[0] & 0xfff, thi = w[1] & 0xfff;

	if (mode == UP_TLUT && (thi >> 2) > (tlo >> 2))
		report(VALIDATION_MULTILINE_TLUT, "Attempting to load multiple lines in TLUT.");

	if (ref_upload_needs_flush(teximg.addr))
		flush();
	else if (!cur.marks.empty() && load_reads_marked_writes(mode, slo, tlo, shi, thi))
		split_at_last_mark();

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
	if (++ref.uploads + 1 >= 256) ref_pass_end();  // rdp_renderer.cpp:3484-3485
	if (cur.uploads.size() >= (size_t)kMaxUploadsPerPass) split_or_flush();
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
	if (!((op >= 0x08 && op <= 0x0f) || op == 0x24 || op == 0x25 || op == 0x36)) state_gen++;
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

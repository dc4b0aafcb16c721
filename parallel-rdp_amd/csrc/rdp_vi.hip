// Video interface scanout: host register decode + two gfx950 kernels.
//
//   k_vi_filter  one lane per texel of the fetch-filter image: extracts the
//                VRAM pixels it needs straight from RDRAM (extract_vram.comp)
//                and applies the AA or dither reconstruction filter
//                (vi_fetch.frag), writing the normal and the fetch-bug layer.
//   k_vi_scale   one lane per output pixel: HStart/VStart scissor and
//                serrate selection, divot median (vi_divot.frag) on the fly
//                from the filter image, bilinear resampling, gamma / gamma
//                dither (vi_scale.frag) and the blend with the previous frame
//                (vi_blend_fields.frag).
//
// Images are small (at most ~2600 x 1200 filter texels, 640 x 576 output),
// so both kernels are bound by launch and L2 latency, not HBM.
#include "rdp_vi.hpp"

#include <vector>

#include <math.h>
#include <string.h>

#include <algorithm>

namespace rdp_mi {

enum : uint32_t {
	VI_TYPE_MASK = 3,
	VI_TYPE_RGBA5551 = 2,
	VI_TYPE_RGBA8888 = 3,
	VI_GAMMA_DITHER = 1u << 2,
	VI_GAMMA = 1u << 3,
	VI_DIVOT = 1u << 4,
	VI_SERRATE = 1u << 6,
	VI_AA_MASK = 3u << 8,
	VI_AA_RESAMP_ONLY = 2u << 8,
	VI_AA_RESAMP_REPLICATE = 3u << 8,
	VI_DITHER_FILTER = 1u << 16,
	VI_META_AA = 1u << 17,
	VI_META_SCALE = 1u << 18,
};

// rdp_common.hpp:375-391
constexpr int kVSyncNtsc = 525, kHOffsetNtsc = 108, kHOffsetPal = 128, kVOffsetNtsc = 34, kVOffsetPal = 44;
constexpr int kVResNtsc = 480, kVResPal = 576, kScanoutWidth = 640;
constexpr int kVEndPal = (kVOffsetPal + kVResPal) | 1, kVEndNtsc = (kVOffsetNtsc + kVResNtsc) | 1;

enum { REG_CONTROL, REG_ORIGIN, REG_WIDTH, REG_INTR, REG_V_CURRENT_LINE, REG_TIMING, REG_V_SYNC, REG_H_SYNC,
       REG_LEAP, REG_H_START, REG_V_START, REG_V_BURST, REG_X_SCALE, REG_Y_SCALE };

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
void vi_gamma_table(uint8_t *out)
{
	auto isqrt = [](unsigned v) {
		unsigned r = 0;
		while ((r + 1) * (r + 1) <= v) r++;
		return r;
	};
	for (unsigned i = 0; i < 256; i++) out[i] = (uint8_t)(isqrt(i * 256) & ~1u);
	for (unsigned j = 0; j < 256 * 64; j++) out[256 + j] = (uint8_t)(isqrt(j * 4) & ~1u);
}

// Per-scanline registers, video_interface.cpp:1126-1213.
void ViState::begin_per_scanline(const uint32_t *regs, uint32_t flags)
{
	per_line_flags = flags;
	h_start_latched = regs[REG_H_START];
	x_scale_latched = regs[REG_X_SCALE];
	h_start_lines[0] = regs[REG_H_START];
	x_scale_lines[0] = regs[REG_X_SCALE];
	per_line_line = 0;
	per_line_ended = false;
}

void ViState::set_for_scanline(uint32_t reg_bit, uint32_t value)
{
	if (!(per_line_flags & reg_bit)) return;
	if (reg_bit == 1)
		h_start_latched = value;
	else if (reg_bit == 2)
		x_scale_latched = value;
}

void ViState::latch_for_scanline(unsigned vi_line)
{
	vi_line = std::min(vi_line, (unsigned)kViVEndMax - 1);
	if (vi_line <= per_line_line) return;
	unsigned n = per_line_line;
	while (++n < vi_line) {
		h_start_lines[n] = h_start_lines[per_line_line];
		x_scale_lines[n] = x_scale_lines[per_line_line];
	}
	h_start_lines[n] = h_start_latched;
	x_scale_lines[n] = x_scale_latched;
	per_line_line = n;
}

void ViState::end_per_scanline()
{
	if (per_line_flags == 0 || per_line_ended) return;
	unsigned n = per_line_line;
	while (++n < (unsigned)kViVEndMax) {
		h_start_lines[n] = h_start_lines[per_line_line];
		x_scale_lines[n] = x_scale_lines[per_line_line];
	}
	per_line_ended = true;
}

void ViState::release()
{
	if (params_done) {
		if (params_busy) (void)hipEventSynchronize(params_done);
		(void)hipEventDestroy(params_done);
		params_done = nullptr;
		params_busy = false;
	}
	for (auto &p : d_images)
		if (p) (void)hipFree(p), p = nullptr;
	if (d_filter) (void)hipFree(d_filter), d_filter = nullptr;
	if (d_params) (void)hipFree(d_params), d_params = nullptr;
	if (h_params) (void)hipHostFree(h_params), h_params = nullptr;
	filter_words = 0;
}

namespace {

struct Decoded {
	uint32_t status;
	int vi_width, vi_offset, v_current_line, v_start, init_y_add, v_res;
	bool is_pal;
	int max_x, max_y, h_start, h_res, h_start_clamp, h_res_clamp;
};

// VideoInterface::decode_vi_registers, video_interface.cpp:176-367.
void decode(const ViState &s, const uint32_t *regs, Decoded &r, ViLine *lines)
{
	memset(&r, 0, sizeof(r));
	r.status = regs[REG_CONTROL];
	r.vi_width = (int)(regs[REG_WIDTH] & 0xfff);
	r.vi_offset = (int)(regs[REG_ORIGIN] & 0xffffff);
	r.v_current_line = (int)(regs[REG_V_CURRENT_LINE] & 1);
	r.v_start = (int)((regs[REG_V_START] >> 16) & 0x3ff);
	int y_start = (int)((regs[REG_Y_SCALE] >> 16) & 0xfff);
	const int y_add = (int)(regs[REG_Y_SCALE] & 0xfff);
	r.init_y_add = y_add;
	int v_end = (int)(regs[REG_V_START] & 0x3ff);
	const int v_sync = (int)(regs[REG_V_SYNC] & 0x3ff);
	r.is_pal = (unsigned)v_sync > (unsigned)(kVSyncNtsc + 25);
	const int v_end_max = r.is_pal ? kVEndPal : kVEndNtsc;
	if (v_end > v_end_max) v_end = v_end_max;
	if (r.v_start > v_end_max) r.v_start = v_end_max;
	const int v_start_offset = r.is_pal ? kVOffsetPal : kVOffsetNtsc;
	r.v_res = (v_end - r.v_start) >> 1;
	r.v_start = (r.v_start - v_start_offset) / 2;
	if (r.v_start < 0) {
		y_start -= y_add * r.v_start;
		r.v_start = 0;
	}
	r.v_res = std::min(r.v_res, kViMaxLines - r.v_start);

	int hs_clamp_lo = INT32_MAX, he_clamp_hi = 0, hs_lo = INT32_MAX, he_hi = 0;
	r.max_x = 0;
	const bool degenerate_y = r.v_res <= 0;
	if (degenerate_y)
		memset(lines, 0, sizeof(ViLine) * kViMaxLines);
	else {
		for (int l = 0; l < r.v_start; l++) memset(&lines[l], 0, sizeof(ViLine));
		for (int l = r.v_start + r.v_res; l < kViMaxLines; l++) memset(&lines[l], 0, sizeof(ViLine));
	}

	auto analyze = [&](int x_start, int x_add, int h_start, int h_end, ViLine *line) {
		bool left_clamp = false, right_clamp = false;
		h_start -= r.is_pal ? kHOffsetPal : kHOffsetNtsc;
		h_end -= r.is_pal ? kHOffsetPal : kHOffsetNtsc;
		if (h_start < 0) {
			x_start -= x_add * h_start;
			h_start = 0;
			left_clamp = true;
		}
		if (h_end > kScanoutWidth) {
			h_end = kScanoutWidth;
			right_clamp = true;
		}
		const int h_start_clamp = h_start + (left_clamp ? 0 : 8);
		const int h_end_clamp = h_end - (right_clamp ? 0 : 7);
		const int h_res = h_end - h_start;
		r.max_x = std::max(r.max_x, (x_start + h_res * x_add) >> 10);
		hs_lo = std::min(hs_lo, h_start);
		he_hi = std::max(he_hi, h_end);
		hs_clamp_lo = std::min(hs_clamp_lo, h_start_clamp);
		he_clamp_hi = std::max(he_clamp_hi, h_end_clamp);
		if (line) *line = ViLine{ h_start, h_start_clamp, h_end_clamp, x_start, x_add, y_start, y_add, 0 };
	};

	if (degenerate_y || !s.per_line_ended) {
		ViLine info;
		analyze((int)((regs[REG_X_SCALE] >> 16) & 0xfff), (int)(regs[REG_X_SCALE] & 0xfff),
		        (int)((regs[REG_H_START] >> 16) & 0x3ff), (int)(regs[REG_H_START] & 0x3ff), &info);
		for (int l = r.v_start; l < r.v_start + r.v_res; l++) lines[l] = info;
	} else {
		for (int l = r.v_start; l < r.v_start + r.v_res; l++) {
			const int eff = 2 * l + v_start_offset + (r.v_current_line == 0);
			const uint32_t xs = s.x_scale_lines[eff], hs = s.h_start_lines[eff];
			analyze((int)((xs >> 16) & 0xfff), (int)(xs & 0xfff), (int)((hs >> 16) & 0x3ff), (int)(hs & 0x3ff),
			        &lines[l]);
		}
	}
	r.max_y = (y_start + r.v_res * y_add) >> 10;
	r.h_start = hs_lo;
	r.h_res = he_hi - hs_lo;
	r.h_start_clamp = hs_clamp_lo;
	r.h_res_clamp = he_clamp_hi - hs_clamp_lo;
}

void shift_rect(int &ox, int &oy, int &w, int &h, int dx, int dy)
{
	ox += dx;
	oy += dy;
	if (ox < 0) {
		w += ox;
		ox = 0;
	}
	if (oy < 0) {
		h += oy;
		oy = 0;
	}
	if (w < 0) w = 0;
	if (h < 0) h = 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Device side.
// ---------------------------------------------------------------------------
#define DEV static __device__ __forceinline__

struct Texel {
	uint32_t r, g, b, a;
};

DEV uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
DEV uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

DEV uint32_t pack(Texel t) { return t.r | (t.g << 8) | (t.b << 16) | (t.a << 24); }
DEV Texel unpack(uint32_t v) { return Texel{ v & 0xff, (v >> 8) & 0xff, (v >> 16) & 0xff, v >> 24 }; }

// extract_vram.comp:61-106, for VRAM pixel (x, y) relative to the origin;
// upscaled, (x, y) is sample (x mod S, y mod S) of native pixel (x / S, y / S)
// in plane (y mod S) * S + x mod S.
DEV Texel fetch_vram(const ViParams &P, const uint8_t *__restrict__ rdram, const uint8_t *__restrict__ hidden, int x,
                     int y)
{
	Texel t;
	const int L = P.scale_log2, S = 1 << L;
	const size_t slice = (size_t)(((y & (S - 1)) << L) + (x & (S - 1)));
	x >>= L;
	y >>= L;
	rdram += slice * ((size_t)P.rdram_mask + 1);
	hidden += slice * (((size_t)P.rdram_mask + 1) >> 1);
	int linear = y * P.fb_width + x + P.fb_offset;
	if ((P.status & VI_TYPE_MASK) == VI_TYPE_RGBA8888) {
		linear &= (int)(P.rdram_mask >> 2);
		const uint32_t w = reinterpret_cast<const uint32_t *>(rdram)[linear];
		t = Texel{ w >> 24, (w >> 16) & 0xff, (w >> 8) & 0xff, (w >> 5) & 7 };
	} else {
		linear &= (int)(P.rdram_mask >> 1);
		const uint32_t w = reinterpret_cast<const uint16_t *>(rdram)[linear ^ 1];
		const uint32_t hid = hidden[linear];
		t = Texel{ (w >> 8) & 0xf8, (w >> 3) & 0xf8, (w << 2) & 0xf8, ((w & 1) << 2) | hid };
	}
	if (!(P.status & VI_META_AA)) t.a = 7;
	return t;
}

// Extracted-image texel (ex, ey): VRAM (ex + x_off, ey + y_off).
DEV Texel ext(const ViParams &P, const uint8_t *rdram, const uint8_t *hidden, int ex, int ey)
{
	return fetch_vram(P, rdram, hidden, ex + P.x_off, ey + P.y_off);
}

DEV void check_neighbor(const Texel &c, uint32_t lo[3], uint32_t hi[3], uint32_t slo[3], uint32_t shi[3])
{
	if (c.a != 7) return;
	const uint32_t v[3] = { c.r, c.g, c.b };
#pragma unroll
	for (int i = 0; i < 3; i++) {
		slo[i] = umin(slo[i], umax(v[i], lo[i]));
		shi[i] = umax(shi[i], umin(v[i], hi[i]));
		lo[i] = umin(v[i], lo[i]);
		hi[i] = umax(v[i], hi[i]);
	}
}

DEV int clamp1(int v) { return v < -1 ? -1 : (v > 1 ? 1 : v); }

// vi_fetch.frag:58-163: one texel of both layers.
__global__ __launch_bounds__(256) void k_vi_filter(const ViParams *__restrict__ params, const uint8_t *__restrict__ rdram,
                                                  const uint8_t *__restrict__ hidden, uint32_t *__restrict__ out)
{
	const ViParams &P = *params;
	const int x = (int)(blockIdx.x * 16 + (threadIdx.x & 15));
	const int y = (int)(blockIdx.y * 16 + (threadIdx.x >> 4));
	if (x >= P.aw || y >= P.ah) return;
	const int px = x + 2, py = y + 2;
	const Texel mid = ext(P, rdram, hidden, px, py);
	const uint32_t m[3] = { mid.r, mid.g, mid.b };
	uint32_t color[3], color_bug[3];
	if (mid.a != 7) {
		uint32_t lo[3], hi[3], slo[3], shi[3];
#pragma unroll
		for (int i = 0; i < 3; i++) lo[i] = hi[i] = slo[i] = shi[i] = m[i];
		const Texel lu = ext(P, rdram, hidden, px - 1, py - 1), ru = ext(P, rdram, hidden, px + 1, py - 1);
		const Texel tl = ext(P, rdram, hidden, px - 2, py), tr = ext(P, rdram, hidden, px + 2, py);
		const Texel ld = ext(P, rdram, hidden, px - 1, py + 1), rd = ext(P, rdram, hidden, px + 1, py + 1);
		check_neighbor(lu, lo, hi, slo, shi);
		check_neighbor(ru, lo, hi, slo, shi);
		check_neighbor(tl, lo, hi, slo, shi);
		check_neighbor(tr, lo, hi, slo, shi);
		uint32_t lob[3], hib[3], slob[3], shib[3];
#pragma unroll
		for (int i = 0; i < 3; i++) {
			lob[i] = lo[i];
			hib[i] = hi[i];
			slob[i] = slo[i];
			shib[i] = shi[i];
		}
		check_neighbor(ld, lo, hi, slo, shi);
		check_neighbor(rd, lo, hi, slo, shi);
		if (P.fetch_bug) {
			check_neighbor(tl, lob, hib, slob, shib);
			check_neighbor(tr, lob, hib, slob, shib);
#pragma unroll
			for (int i = 0; i < 3; i++) {
				if (m[i] == lo[i]) slo[i] = lo[i];
				if (m[i] == hi[i]) shi[i] = hi[i];
				if (m[i] == lob[i]) slob[i] = lob[i];
				if (m[i] == hib[i]) shib[i] = hib[i];
			}
		}
		const uint32_t coeff = 7u - mid.a;
#pragma unroll
		for (int i = 0; i < 3; i++) {
			const uint32_t off = slo[i] + shi[i] - (m[i] << 1);
			color[i] = (m[i] + ((off * coeff + 4u) >> 3)) & 0xffu;
			const uint32_t offb = slob[i] + shib[i] - (m[i] << 1);
			color_bug[i] = (m[i] + ((offb * coeff + 4u) >> 3)) & 0xffu;
		}
	} else if (P.status & VI_DITHER_FILTER) {
		int tmp[3], acc[3] = { 0, 0, 0 }, accb[3];
#pragma unroll
		for (int i = 0; i < 3; i++) tmp[i] = (int)(m[i] >> 3);
		for (int dy = -1; dy <= 0; dy++)
			for (int dx = -1; dx <= 1; dx++) {
				const Texel t = ext(P, rdram, hidden, px + dx, py + dy);
				acc[0] += clamp1((int)(t.r >> 3) - tmp[0]);
				acc[1] += clamp1((int)(t.g >> 3) - tmp[1]);
				acc[2] += clamp1((int)(t.b >> 3) - tmp[2]);
			}
#pragma unroll
		for (int i = 0; i < 3; i++) accb[i] = acc[i];
		const Texel a = ext(P, rdram, hidden, px - 1, py + 1), b = ext(P, rdram, hidden, px + 1, py + 1);
		const Texel c = ext(P, rdram, hidden, px, py + 1);
		const Texel l = ext(P, rdram, hidden, px - 1, py), rr = ext(P, rdram, hidden, px + 1, py);
		const uint32_t av[3] = { a.r, a.g, a.b }, bv[3] = { b.r, b.g, b.b }, cv[3] = { c.r, c.g, c.b };
		const uint32_t lv[3] = { l.r, l.g, l.b }, rv[3] = { rr.r, rr.g, rr.b };
#pragma unroll
		for (int i = 0; i < 3; i++) {
			acc[i] += clamp1((int)(av[i] >> 3) - tmp[i]) + clamp1((int)(bv[i] >> 3) - tmp[i]) +
			          clamp1((int)(cv[i] >> 3) - tmp[i]);
			accb[i] += clamp1((int)(lv[i] >> 3) - tmp[i]) + clamp1((int)(rv[i] >> 3) - tmp[i]);
			// An out-of-range value written to an 8-bit UINT attachment keeps its low 8 bits.
			color[i] = ((m[i] & 0xf8u) + (uint32_t)acc[i]) & 0xffu;
			color_bug[i] = ((m[i] & 0xf8u) + (uint32_t)accb[i]) & 0xffu;
		}
	} else {
#pragma unroll
		for (int i = 0; i < 3; i++) color[i] = color_bug[i] = m[i];
	}
	const size_t idx = (size_t)y * (size_t)P.aw + (size_t)x;
	out[idx] = pack(Texel{ color[0], color[1], color[2], mid.a });
	if (P.fetch_bug) out[(size_t)P.aw * (size_t)P.ah + idx] = pack(Texel{ color_bug[0], color_bug[1], color_bug[2], mid.a });
}

// Filter-image texel with the divot median applied when enabled
// (vi_divot.frag:54-92); zero outside the image.
DEV Texel filtered(const ViParams &P, const uint32_t *__restrict__ A, int layer, int x, int y)
{
	const uint32_t *L = A + (size_t)layer * (size_t)P.aw * (size_t)P.ah;
	auto at = [&](int xx) -> Texel {
		if (xx < 0 || y < 0 || xx >= P.aw || y >= P.ah) return Texel{ 0, 0, 0, 0 };
		return unpack(L[(size_t)y * (size_t)P.aw + (size_t)xx]);
	};
	if (!P.divot) return at(x);
	// The divot image is (aw - 3) wide and ah high.
	if (x < 0 || x >= P.aw - 3 || y < 0 || y >= P.ah) return Texel{ 0, 0, 0, 0 };
	const Texel l = at(x), m = at(x + 1), r = at(x + 2);
	if ((l.a & m.a & r.a) == 7) return m;
	auto med = [](uint32_t a, uint32_t b, uint32_t c) { return umax(umin(a, b), umin(umax(a, b), c)); };
	return Texel{ med(l.r, m.r, r.r), med(l.g, m.g, r.g), med(l.b, m.b, r.b), m.a };
}

DEV uint32_t reseed_noise(uint32_t x, uint32_t y, uint32_t z)
{
	const uint32_t P = 1103515245u;
	uint32_t s0 = x, s1 = y, s2 = z;
#pragma unroll
	for (int i = 0; i < 3; i++) {
		const uint32_t n0 = ((s0 >> 8) ^ s1) * P, n1 = ((s1 >> 8) ^ s2) * P, n2 = ((s2 >> 8) ^ s0) * P;
		s0 = n0;
		s1 = n1;
		s2 = n2;
	}
	return (s0 >> 16) & 0xffff;
}

DEV uint32_t vi_lerp(uint32_t a, uint32_t b, uint32_t l) { return (a + (((b - a) * l + 16u) >> 5)) & 0xffu; }

// vi_scale.frag:74-153 plus vi_blend_fields.frag.
__global__ __launch_bounds__(256) void k_vi_scale(const ViParams *__restrict__ params, const uint32_t *__restrict__ A,
                                                 const uint8_t *__restrict__ gamma, const uint32_t *__restrict__ prev,
                                                 uint32_t *__restrict__ out)
{
	const ViParams &P = *params;
	const int px = (int)(blockIdx.x * 64 + (threadIdx.x & 63));
	const int py = (int)(blockIdx.y * 4 + (threadIdx.x >> 6));
	if (px >= P.ow || py >= P.oh) return;
	uint32_t result = 0;
	if (px >= P.draw_x && px < P.draw_x + P.draw_w && py >= P.draw_y && py < P.draw_y + P.draw_h) {
		int cx = px + P.crop_l, cy = py + P.crop_t;
		const ViLine li = P.lines[cy >> P.info_y_shift];
		cy -= P.v_start;
		if (!(cx < li.h_start_clamp || cx >= li.h_end_clamp || ((cy & P.serrate_mask) != P.serrate_select))) {
			cx -= li.h_start;
			cy >>= P.serrate_shift;
			const uint32_t noise = (P.status & VI_GAMMA_DITHER) ? reseed_noise((uint32_t)cx, (uint32_t)cy, P.frame_count) : 0;
			const int x = cx * li.x_add + li.x_start;
			const int y = (cy - li.y_base) * li.y_add + li.y_start;
			const int bx = x >> 10, by = y >> 10;
			const Texel t00 = filtered(P, A, 0, bx, by);
			uint32_t c[3] = { t00.r, t00.g, t00.b };
			int bug = 0;
			if (P.fetch_bug) {
				const int prev_y = (y - P.init_y_add) >> 10, next_y = (y + P.init_y_add) >> 10;
				if (cy != 0 && by == prev_y && by != next_y) bug = 1;
			}
			if (P.status & VI_META_SCALE) {
				const uint32_t xf = (uint32_t)((x >> 5) & 31), yf = (uint32_t)((y >> 5) & 31);
				const Texel t10 = filtered(P, A, 0, bx + 1, by);
				const Texel t01 = filtered(P, A, bug, bx, by + 1);
				const Texel t11 = filtered(P, A, bug, bx + 1, by + 1);
				const uint32_t v10[3] = { t10.r, t10.g, t10.b }, v01[3] = { t01.r, t01.g, t01.b };
				const uint32_t v11[3] = { t11.r, t11.g, t11.b };
#pragma unroll
				for (int i = 0; i < 3; i++) c[i] = vi_lerp(vi_lerp(c[i], v01[i], yf), vi_lerp(v10[i], v11[i], yf), xf);
			}
			if (P.status & VI_GAMMA) {
				if (P.status & VI_GAMMA_DITHER) {
					const uint32_t n[3] = { noise & 0x3f, (noise >> 6) & 0x3f, ((noise >> 9) & 0x38) | (noise & 7) };
#pragma unroll
					for (int i = 0; i < 3; i++) c[i] = gamma[(c[i] << 6) + n[i] + 256u];
				} else {
#pragma unroll
					for (int i = 0; i < 3; i++) c[i] = gamma[c[i]];
				}
			} else if (P.status & VI_GAMMA_DITHER) {
#pragma unroll
				for (int i = 0; i < 3; i++) c[i] = umin(c[i] + ((noise >> i) & 1u), 0xffu);
			}
			result = c[0] | (c[1] << 8) | (c[2] << 16) | 0xff000000u;
		}
	}
	if (P.blend && (result >> 24) == 0) {
		bool in_rect = false;
		for (int k = 0; k < P.nrects; k++)
			in_rect |= px >= P.rects[k][0] && px < P.rects[k][0] + P.rects[k][2] && py >= P.rects[k][1] &&
			           py < P.rects[k][1] + P.rects[k][3];
		if (in_rect) {
			uint32_t p = 0;
			if (px < P.prev_w && py < P.prev_h) p = prev[(size_t)py * (size_t)P.prev_w + (size_t)px];
			result = (p >> 24) ? (p & 0x00ffffffu) : 0;
		}
	}
	out[(size_t)py * (size_t)P.ow + (size_t)px] = result;
}

#define VI_CHECK(expr)                                                                                         \
	do {                                                                                                       \
		if ((expr) != hipSuccess) return RDP_MI_ERROR_DEVICE;                                                  \
	} while (0)

// VideoInterface::scanout, video_interface.cpp:1215-1426 (scaling factor 1,
// no downscale or deinterlace pass).
// VideoInterface::scanout_memory_range (video_interface.cpp:369-396): the
// RDRAM bytes a scanout fetches, a 2-pixel border on each side for the AA
// and divot filters included; length 0 for a frame that fetches nothing.
void vi_memory_range(const ViState &s, const uint32_t *regs, uint32_t &offset, uint32_t &length)
{
	Decoded r;
	std::vector<ViLine> lines(kViMaxLines);
	decode(s, regs, r, lines.data());
	offset = length = 0;
	if (r.vi_offset == 0 || r.h_res <= 0 || r.h_start >= kScanoutWidth) return;
	const int divot = (r.status & VI_DIVOT) ? 1 : 0;
	const int aa_width = r.max_x + 2 + 4 + divot * 2, aa_height = r.max_y + 1 + 4;
	const int x_off = divot ? -3 : -2, y_off = -2;
	const int pixel = ((r.status & VI_TYPE_MASK) | VI_TYPE_RGBA5551) == VI_TYPE_RGBA8888 ? 4 : 2;
	const int64_t off = (int64_t)(r.vi_offset & ~(pixel - 1)) + (int64_t)(x_off + y_off * r.vi_width) * pixel;
	offset = (uint32_t)off;  // wraps like the fetch, modulo the RDRAM size
	length = (uint32_t)((int64_t)(aa_height * r.vi_width + aa_width) * pixel);
}

hipError_t launch_domain_range(const uint8_t *rdram, const uint8_t *hidden, uint8_t *ref, uint8_t *planes,
                               uint8_t *hplanes, uint32_t rdram_size, uint32_t num_planes, uint32_t addr,
                               uint32_t num_pixels, uint32_t size_log2, hipStream_t s);

int vi_scanout(ViState &s, const uint32_t *regs, const ViSource &src, const uint8_t *d_gamma, hipStream_t stream,
               const rdp_mi_scanout_options &opts, uint32_t *out, unsigned *width, unsigned *height, bool sync)
{
	const uint8_t *d_rdram = src.rdram, *d_hidden = src.hidden;
	const uint32_t rdram_size = src.rdram_size;
	const int L = (int)src.scale_log2, SF = 1 << L;
	if (!s.h_params) {
		VI_CHECK(hipHostMalloc((void **)&s.h_params, sizeof(ViParams), hipHostMallocDefault));
		VI_CHECK(hipMalloc((void **)&s.d_params, sizeof(ViParams)));
		VI_CHECK(hipEventCreateWithFlags(&s.params_done, hipEventDisableTiming));
	}
	const size_t image_words = (size_t)640 * 576 << (2 * L);
	if (s.image_words < image_words) {
		VI_CHECK(hipStreamSynchronize(stream));
		for (auto &img : s.d_images) {
			if (img) VI_CHECK(hipFree(img));
			img = nullptr;
		}
		for (auto &img : s.d_images) VI_CHECK(hipMalloc((void **)&img, image_words * sizeof(uint32_t)));
		s.image_words = image_words;
	}
	// The previous scanout's parameter upload reads the pinned block.
	if (s.params_busy) {
		VI_CHECK(hipEventSynchronize(s.params_done));
		s.params_busy = false;
	}
	ViParams &P = *s.h_params;
	memset(&P, 0, sizeof(P));
	Decoded r;
	decode(s, regs, r, P.lines);
	s.per_line_flags = 0;  // clear_per_scanline_state
	s.per_line_ended = false;
	*width = *height = 0;

	if (r.vi_offset == 0) {
		s.prev_valid = false;
		return 0;
	}
	if (!opts.vi_serrate) r.status &= ~VI_SERRATE;
	const bool is_aa = (r.status & VI_AA_MASK) < VI_AA_RESAMP_ONLY && opts.vi_aa;
	const bool is_bilinear = (r.status & VI_AA_MASK) < VI_AA_RESAMP_REPLICATE && opts.vi_scale;
	r.status &= ~(VI_AA_MASK | VI_META_AA | VI_META_SCALE);
	if (is_aa) r.status |= VI_META_AA;
	if (is_bilinear) r.status |= VI_META_SCALE;
	if (!opts.vi_gamma_dither) r.status &= ~VI_GAMMA_DITHER;
	if (!opts.vi_divot_filter) r.status &= ~VI_DIVOT;
	if (!opts.vi_dither_filter) r.status &= ~VI_DITHER_FILTER;

	const bool is_blank = (r.status & VI_TYPE_RGBA5551) == 0;
	if (is_blank && s.previous_frame_blank) {
		s.frame_count++;
		s.prev_valid = false;
		return 0;
	}
	if (is_blank) s.prev_valid = false;
	r.status |= VI_TYPE_RGBA5551;
	s.previous_frame_blank = is_blank;
	const int divot = (r.status & VI_DIVOT) ? 1 : 0;

	if (r.h_res <= 0 || r.h_start >= kScanoutWidth) {
		s.frame_count++;
		if (opts.persist_frame_on_invalid_input && (s.frame_count - s.last_valid_frame_count < 4)) {
			if (s.prev_valid) {
				VI_CHECK(hipMemcpyAsync(out, s.d_images[s.cur ^ 1], (size_t)s.prev_w * s.prev_h * 4, hipMemcpyDeviceToHost,
				                        stream));
				if (sync) VI_CHECK(hipStreamSynchronize(stream));
				*width = s.prev_w;
				*height = s.prev_h;
			}
		} else
			s.prev_valid = false;
		return 0;
	}
	s.last_valid_frame_count = s.frame_count;
	const bool degenerate = r.h_res <= 0 || r.v_res <= 0;

	// video_interface.cpp:1299-1318: every coordinate scales with the factor.
	r.h_start *= SF;
	r.h_start_clamp *= SF;
	r.v_start *= SF;
	r.h_res *= SF;
	r.h_res_clamp *= SF;
	r.v_res *= SF;
	r.max_x = r.max_x * SF + (SF - 1);
	r.max_y = r.max_y * SF + (SF - 1);
	for (ViLine &li : P.lines) {
		li.h_start *= SF;
		li.h_start_clamp *= SF;
		li.h_end_clamp *= SF;
		li.x_start *= SF;
		li.y_start *= SF;
		li.y_base *= SF;
	}
	if (L && !degenerate) {
		// vram_fetch_stage (video_interface.cpp:412-420): the domain's pre
		// update over the memory the fetch reads, colour only.
		uint32_t offset = 0, length = 0;
		vi_memory_range(s, regs, offset, length);
		VI_CHECK(launch_domain_range(src.native_rdram, src.native_hidden, src.ref, src.planes, src.hplanes, rdram_size,
		                             1u << (2 * L), offset, length, (r.status & VI_TYPE_MASK) == VI_TYPE_RGBA8888 ? 2 : 1,
		                             stream));
	}

	P.status = r.status;
	P.fb_offset = (r.status & VI_TYPE_MASK) == VI_TYPE_RGBA8888 ? r.vi_offset >> 2 : r.vi_offset >> 1;
	P.fb_width = r.vi_width;
	P.x_off = divot ? -3 : -2;
	P.y_off = -2;
	P.aw = r.max_x + 3 + divot * 2;
	P.ah = r.max_y + 2;
	P.divot = divot;
	P.fetch_bug = r.init_y_add < 1024 && SF == 1;
	P.init_y_add = r.init_y_add;
	P.frame_count = s.frame_count;
	P.rdram_mask = rdram_size - 1;
	P.scale_log2 = L;

	// scale_stage, video_interface.cpp:719-998.
	const bool serrate = (r.status & VI_SERRATE) && !opts.upscale_deinterlacing;
	const int W = kScanoutWidth * SF, H = ((r.is_pal ? kVResPal : kVResNtsc) >> (serrate ? 0 : 1)) * SF;
	int cl, cr, ct, cb;
	if (opts.crop_enable) {
		cl = (int)opts.crop_left;
		cr = (int)opts.crop_right;
		ct = (int)opts.crop_top;
		cb = (int)opts.crop_bottom;
		if (serrate) {
			ct *= 2;
			cb *= 2;
		}
	} else {
		const int cy = (int)opts.crop_overscan_pixels * (serrate ? 2 : 1);
		const int cx = (int)lroundf((float)cy * ((float)W / (float)H));
		cl = cr = cx;
		ct = cb = cy;
	}
	cl *= SF;
	cr *= SF;
	ct *= SF;
	cb *= SF;
	P.ow = W;
	P.oh = H;
	if (cl + cr < W && ct + cb < H) {
		P.ow = W - (cl + cr);
		P.oh = H - (ct + cb);
	}
	P.crop_l = cl;
	P.crop_t = ct;
	int v_start = r.v_start, v_res = r.v_res;
	P.info_y_shift = L;
	if (serrate) {
		v_start *= 2;
		v_res *= 2;
		P.serrate_shift = 1;
		P.serrate_mask = 1;
		P.serrate_select = r.v_current_line == 0;
		P.info_y_shift++;
	}
	P.v_start = v_start;
	int dx = r.h_start, dy = v_start, dw = r.h_res, dh = v_res;
	shift_rect(dx, dy, dw, dh, -cl, -ct);
	if (degenerate || r.h_res <= 0 || v_res <= 0) dw = dh = 0;
	P.draw_x = dx;
	P.draw_y = dy;
	P.draw_w = dw;
	P.draw_h = dh;

	if (s.prev_valid && opts.blend_previous_frame) {
		P.blend = 1;
		P.prev_w = (int)s.prev_w;
		P.prev_h = (int)s.prev_h;
		const int ph = (int)s.prev_h;
		auto add = [&](int x, int y, int w, int h) {
			shift_rect(x, y, w, h, -cl, -ct);
			P.rects[P.nrects][0] = x;
			P.rects[P.nrects][1] = y;
			P.rects[P.nrects][2] = w;
			P.rects[P.nrects][3] = h;
			P.nrects++;
		};
		if (degenerate) {
			if (r.h_res > 0) add(r.h_start, 0, r.h_res, ph);
		} else {
			if (r.h_res > 0 && v_start > 0) add(r.h_start, 0, r.h_res, v_start);
			if (r.h_res_clamp > 0 && v_res > 0) add(r.h_start_clamp, v_start, r.h_res_clamp, v_res);
			if (r.h_res > 0 && ph > v_start + v_res) add(r.h_start, v_start + v_res, r.h_res, ph - (v_start + v_res));
		}
	}

	// Fetch-filter image (both layers).
	if (!degenerate) {
		const size_t words = (size_t)P.aw * (size_t)P.ah * 2;
		if (words > s.filter_words) {
			if (s.d_filter) {
				VI_CHECK(hipStreamSynchronize(stream));
				VI_CHECK(hipFree(s.d_filter));
				s.d_filter = nullptr;
			}
			VI_CHECK(hipMalloc((void **)&s.d_filter, words * sizeof(uint32_t)));
			s.filter_words = words;
		}
	}
	VI_CHECK(hipMemcpyAsync(s.d_params, s.h_params, sizeof(ViParams), hipMemcpyHostToDevice, stream));
	VI_CHECK(hipEventRecord(s.params_done, stream));
	s.params_busy = true;
	if (!degenerate) {
		dim3 grid((unsigned)((P.aw + 15) / 16), (unsigned)((P.ah + 15) / 16));
		hipLaunchKernelGGL(k_vi_filter, grid, dim3(256), 0, stream, s.d_params, d_rdram, d_hidden, s.d_filter);
		VI_CHECK(hipGetLastError());
	}
	uint32_t *dst = s.d_images[s.cur];
	const uint32_t *prevp = s.d_images[s.cur ^ 1];
	{
		dim3 grid((unsigned)((P.ow + 63) / 64), (unsigned)((P.oh + 3) / 4));
		hipLaunchKernelGGL(k_vi_scale, grid, dim3(256), 0, stream, s.d_params, s.d_filter, d_gamma, prevp, dst);
		VI_CHECK(hipGetLastError());
	}
	VI_CHECK(hipMemcpyAsync(out, dst, (size_t)P.ow * (size_t)P.oh * 4, hipMemcpyDeviceToHost, stream));
	if (sync) VI_CHECK(hipStreamSynchronize(stream));
	*width = (unsigned)P.ow;
	*height = (unsigned)P.oh;
	s.prev_w = (unsigned)P.ow;
	s.prev_h = (unsigned)P.oh;
	s.prev_valid = true;
	s.cur ^= 1;
	s.frame_count++;
	return 0;
}

}  // namespace rdp_mi

// gfx950 kernels for the RDP rendering hot path.
//
// Four stages per render pass (any number of passes, from any number of
// streams, per launch; every block finds its pass through a work item):
//   k_setup   one wave64 per primitive: decodes the raw command words with the
//             pass draw state, writes PrimSetup, walks the edges for every
//             covered scanline (SpanSetup) and reduces the per-8-line x extent
//             into tile-row ranges.
//   k_tmem_gather / k_tmem_scan
//             one lane per (block of 32 uploads, TMEM halfword) applies the
//             block's uploads in order; then one lane per halfword carries
//             values across blocks, completing the TMEM snapshot after every
//             upload.
//   k_bin     one wave64 per (tile row, group of 64 primitives): a ballot per
//             tile column produces the ordered 64-bit primitive masks.
//   k_raster  one wave64 per 8x8 screen tile: colour/depth/coverage of the 64
//             pixels live in registers while every binned primitive is shaded
//             (coverage, interpolation, texturing, combiner) and merged (Z test,
//             blender, dither, coverage update) in submission order; RDRAM is
//             read once and written back once, masked to dirty pixels.
//
// All arithmetic is the RDP's integer fixed point, evaluated in 32-bit two's
// complement (compiled with -fwrapv).  Semantics follow the reference shaders
// cited at each function.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>

#include "rdp_types.hpp"
#include "rdp_tmem_plan.hpp"

namespace rdp_mi {

#define DEV static __device__ __forceinline__
#define MDEV __device__ __forceinline__

// Explicit address spaces: read-only per-pass tables go through the constant
// address space so uniform reads become scalar loads; everything else is
// global memory (never flat).
#define AS_GLOBAL __attribute__((address_space(1)))
#define AS_CONST __attribute__((address_space(4)))
#define AS_LDS __attribute__((address_space(3)))
template <typename T> DEV const AS_CONST T *cptr(const T *p) { return (const AS_CONST T *)p; }
template <typename T> DEV AS_GLOBAL T *gptr(T *p) { return (AS_GLOBAL T *)p; }
template <typename T> DEV const AS_GLOBAL T *gptr(const T *p) { return (const AS_GLOBAL T *)p; }

// Whole-record loads (the implicit copy constructors cannot read from a
// non-generic address space).
template <typename T> DEV T ldc(const AS_CONST T *p)
{
	static_assert(sizeof(T) % 4 == 0, "record size");
	T v;
	const AS_CONST uint32_t *src = (const AS_CONST uint32_t *)p;
	uint32_t *dst = reinterpret_cast<uint32_t *>(&v);
#pragma unroll
	for (unsigned i = 0; i < sizeof(T) / 4; i++) dst[i] = src[i];
	return v;
}

template <typename T> DEV T ldg16(const AS_GLOBAL T *p)
{
	static_assert(sizeof(T) % 16 == 0, "record size");
	T v;
	const AS_GLOBAL uint32_t *src = (const AS_GLOBAL uint32_t *)p;
	uint32_t *dst = reinterpret_cast<uint32_t *>(&v);
#pragma unroll
	for (unsigned i = 0; i < sizeof(T) / 4; i++) dst[i] = src[i];
	return v;
}

// --------------------------------------------------------------------------
// Integer helpers with GLSL semantics.
// --------------------------------------------------------------------------
DEV int32_t sext(int32_t v, int bits) { return (int32_t)((uint32_t)v << (32 - bits)) >> (32 - bits); }
DEV int find_msb(int32_t v)
{
	uint32_t u = (uint32_t)(v < 0 ? ~v : v);
	return u ? 31 - __clz((int)u) : -1;
}
DEV int find_lsb(uint32_t v) { return __ffs((int)v) - 1; }
DEV int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
DEV int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
DEV int32_t iclamp(int32_t v, int32_t lo, int32_t hi) { return imin(imax(v, lo), hi); }

DEV uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// --------------------------------------------------------------------------
// RDRAM access: host-endian 32-bit words holding big-endian data.
// --------------------------------------------------------------------------
struct Mem {
	AS_GLOBAL uint8_t *rdram;
	AS_GLOBAL uint8_t *hidden;
	uint32_t mask8;
	MDEV uint32_t rd8(uint32_t addr) const { return rdram[(addr & mask8) ^ 3]; }
	MDEV uint32_t rd16i(uint32_t index) const
	{
		return reinterpret_cast<const AS_GLOBAL uint16_t *>(rdram)[(index & (mask8 >> 1)) ^ 1];
	}
	MDEV void wr16i(uint32_t index, uint32_t v) const
	{
		reinterpret_cast<AS_GLOBAL uint16_t *>(rdram)[(index & (mask8 >> 1)) ^ 1] = (uint16_t)v;
	}
	MDEV uint32_t rd32i(uint32_t index) const { return reinterpret_cast<const AS_GLOBAL uint32_t *>(rdram)[index & (mask8 >> 2)]; }
	MDEV void wr32i(uint32_t index, uint32_t v) const { reinterpret_cast<AS_GLOBAL uint32_t *>(rdram)[index & (mask8 >> 2)] = v; }
	MDEV uint32_t rd16a(uint32_t addr) const
	{
		if ((addr & 1) == 0)
			return rd16i(addr >> 1);
		return (rd8(addr) << 8) | rd8(addr + 1);
	}
	MDEV uint32_t hid(uint32_t index) const { return hidden[index & (mask8 >> 1)]; }
	MDEV void set_hid(uint32_t index, uint32_t v) const { hidden[index & (mask8 >> 1)] = (uint8_t)v; }
};

// Write mask of a host-coherent stream: the bytes store_color / the depth
// store write, in the same layout as RDRAM, set to 0xff.
DEV void mark_color(AS_GLOBAL uint8_t *wm, uint32_t mask8, uint32_t fmt, uint32_t index)
{
	switch (fmt) {
	case FB_I4:
	case FB_I8: wm[((index & mask8) ^ 3)] = 0xff; break;
	case FB_RGBA5551:
	case FB_IA88: reinterpret_cast<AS_GLOBAL uint16_t *>(wm)[(index & (mask8 >> 1)) ^ 1] = 0xffff; break;
	default: reinterpret_cast<AS_GLOBAL uint32_t *>(wm)[index & (mask8 >> 2)] = 0xffffffffu; break;
	}
}

DEV void mark_depth(AS_GLOBAL uint8_t *wm, uint32_t mask8, uint32_t index)
{
	reinterpret_cast<AS_GLOBAL uint16_t *>(wm)[(index & (mask8 >> 1)) ^ 1] = 0xffff;
}

// ==========================================================================
// Stage 1: primitive decode + span setup.
// ==========================================================================

// Renderer::build_derived_attributes dz part, rdp_renderer.cpp:717-740, 910-928.
DEV int dz_compress_host(int dz)
{
	int val = 0;
	if (dz & 0xff00) val |= 8;
	if (dz & 0xf0f0) val |= 4;
	if (dz & 0xcccc) val |= 2;
	if (dz & 0xaaaa) val |= 1;
	return val;
}

DEV int normalize_dzpix(int dz)
{
	if (dz >= 0x8000) return 0x8000;
	if (dz == 0) return 1;
	return 1 << (find_msb(dz) + 1);
}

// CommandProcessor decoders, rdp_device.cpp:268-344 and 713-821, plus
// Renderer::fixup_triangle_setup (rdp_renderer.cpp:1402-1423) and the
// primitive-depth override (rdp_renderer.cpp:1489-1497).
template <typename WordPtr, typename State>
DEV void decode_primitive(WordPtr w, uint32_t op, const State &st, PrimSetup &s)
{
	const uint32_t raw = st.raw_ras_flags;
	for (int i = 0; i < 4; i++) {
		s.rgba[i] = s.drgba_dx[i] = s.drgba_de[i] = s.drgba_dy[i] = 0;
		s.stzw[i] = s.dstzw_dx[i] = s.dstzw_de[i] = s.dstzw_dy[i] = 0;
	}
	s.flags = 0;
	s.tile = 0;
	if (op >= 0x08 && op <= 0x0f) {
		uint32_t w0 = w[0], w1 = w[1];
		bool flip = (w0 & 0x800000u) != 0;
		bool sign_dxhdy = (w[5] & 0x80000000u) != 0;
		if (flip) s.flags |= TRI_FLIP;
		if (flip == sign_dxhdy) s.flags |= TRI_DO_OFFSET;
		if (raw & RAS_COPY) s.flags |= TRI_SKIP_XFRAC;
		if (st.quirks & 0x100) s.flags |= TRI_NATIVE_LOD;
		s.tile = (w0 >> 16) & 63;
		s.yl = sext((int32_t)w0, 14);
		s.ym = sext((int32_t)(w1 >> 16), 14);
		s.yh = sext((int32_t)w1, 14);
		s.xl = sext((int32_t)w[2], 28) >> 1;
		s.xh = sext((int32_t)w[4], 28) >> 1;
		s.xm = sext((int32_t)w[6], 28) >> 1;
		s.dxldy = sext((int32_t)(w[3] >> 2), 28) >> 1;
		s.dxhdy = sext((int32_t)(w[5] >> 2), 28) >> 1;
		s.dxmdy = sext((int32_t)(w[7] >> 2), 28) >> 1;
		WordPtr p = w + 8;
		if (op & 4) {
			for (int i = 0; i < 2; i++) {
				s.rgba[2 * i + 0] = (int32_t)((p[i] & 0xffff0000u) | ((p[4 + i] >> 16) & 0xffff));
				s.rgba[2 * i + 1] = (int32_t)((p[i] << 16) | (p[4 + i] & 0xffff));
				s.drgba_dx[2 * i + 0] = (int32_t)((p[2 + i] & 0xffff0000u) | ((p[6 + i] >> 16) & 0xffff));
				s.drgba_dx[2 * i + 1] = (int32_t)((p[2 + i] << 16) | (p[6 + i] & 0xffff));
				s.drgba_de[2 * i + 0] = (int32_t)((p[8 + i] & 0xffff0000u) | ((p[12 + i] >> 16) & 0xffff));
				s.drgba_de[2 * i + 1] = (int32_t)((p[8 + i] << 16) | (p[12 + i] & 0xffff));
				s.drgba_dy[2 * i + 0] = (int32_t)((p[10 + i] & 0xffff0000u) | ((p[14 + i] >> 16) & 0xffff));
				s.drgba_dy[2 * i + 1] = (int32_t)((p[10 + i] << 16) | (p[14 + i] & 0xffff));
			}
			p += 16;
		}
		if (op & 2) {
			s.stzw[0] = (int32_t)((p[0] & 0xffff0000u) | ((p[4] >> 16) & 0xffff));
			s.stzw[1] = (int32_t)(((p[0] << 16) & 0xffff0000u) | (p[4] & 0xffff));
			s.stzw[3] = (int32_t)((p[1] & 0xffff0000u) | ((p[5] >> 16) & 0xffff));
			s.dstzw_dx[0] = (int32_t)((p[2] & 0xffff0000u) | ((p[6] >> 16) & 0xffff));
			s.dstzw_dx[1] = (int32_t)(((p[2] << 16) & 0xffff0000u) | (p[6] & 0xffff));
			s.dstzw_dx[3] = (int32_t)((p[3] & 0xffff0000u) | ((p[7] >> 16) & 0xffff));
			s.dstzw_de[0] = (int32_t)((p[8] & 0xffff0000u) | ((p[12] >> 16) & 0xffff));
			s.dstzw_de[1] = (int32_t)(((p[8] << 16) & 0xffff0000u) | (p[12] & 0xffff));
			s.dstzw_de[3] = (int32_t)((p[9] & 0xffff0000u) | ((p[13] >> 16) & 0xffff));
			s.dstzw_dy[0] = (int32_t)((p[10] & 0xffff0000u) | ((p[14] >> 16) & 0xffff));
			s.dstzw_dy[1] = (int32_t)(((p[10] << 16) & 0xffff0000u) | (p[14] & 0xffff));
			s.dstzw_dy[3] = (int32_t)((p[11] & 0xffff0000u) | ((p[15] >> 16) & 0xffff));
			p += 16;
		}
		if (op & 1) {
			s.stzw[2] = (int32_t)p[0];
			s.dstzw_dx[2] = (int32_t)p[1];
			s.dstzw_de[2] = (int32_t)p[2];
			s.dstzw_dy[2] = (int32_t)p[3];
		}
	} else {
		uint32_t xl = (w[0] >> 12) & 0xfff, yl = w[0] & 0xfff;
		uint32_t xh = (w[1] >> 12) & 0xfff, yh = w[1] & 0xfff;
		if (raw & (RAS_COPY | RAS_FILL))
			yl |= 3;
		s.xh = (int32_t)(xh << 13);
		s.xl = (int32_t)(xl << 13);
		s.xm = (int32_t)(xl << 13);
		s.ym = (int32_t)yl;
		s.yl = (int32_t)yl;
		s.yh = (int32_t)yh;
		s.dxhdy = s.dxmdy = s.dxldy = 0;
		if (op == 0x36) {
			s.flags = TRI_FLIP | TRI_DISABLE_UPSCALING;
		} else {
			int32_t ss = (int32_t)((w[2] >> 16) & 0xffff), tt = (int32_t)(w[2] & 0xffff);
			int32_t dsdx = sext((int32_t)((w[3] >> 16) & 0xffff), 16);
			int32_t dtdy = sext((int32_t)(w[3] & 0xffff), 16);
			s.flags = TRI_FLIP;
			if (op == 0x25 || (st.quirks & 1)) s.flags |= TRI_DISABLE_UPSCALING;
			if (st.quirks & 0x100) s.flags |= TRI_NATIVE_LOD;
			s.tile = (w[1] >> 24) & 7;
			s.stzw[0] = ss << 16;
			s.stzw[1] = tt << 16;
			if (op == 0x24) {
				s.dstzw_dx[0] = dsdx << 11;
				s.dstzw_de[1] = dtdy << 11;
				s.dstzw_dy[1] = dtdy << 11;
			} else {
				s.dstzw_dx[1] = dtdy << 11;
				s.dstzw_de[0] = dsdx << 11;
				s.dstzw_dy[0] = dsdx << 11;
			}
			if (raw & RAS_COPY) s.flags |= TRI_SKIP_XFRAC;
		}
	}

	// fixup_triangle_setup
	int start_y = s.yh & ~3;
	if (s.ym < start_y) s.ym = 0x7fff;
	if (raw & RAS_INTERLACE_FIELD) {
		s.flags |= TRI_INTERLACE_FIELD;
		if (raw & RAS_INTERLACE_KEEP_ODD) s.flags |= TRI_INTERLACE_KEEP_ODD;
	}
	if (raw & (RAS_COPY | RAS_FILL)) s.flags |= TRI_FILL_COPY_RASTER;

	// dz from the decoded gradients (before any primitive-depth override)
	if (st.use_prim_depth) {
		s.dz = (int32_t)st.prim_dz;
		s.dz_compressed = dz_compress_host((int)st.prim_dz);
		s.stzw[2] = st.prim_depth;
		s.dstzw_dx[2] = s.dstzw_de[2] = s.dstzw_dy[2] = 0;
	} else {
		int dzdx = s.dstzw_dx[2] >> 16;
		int dzdy = s.dstzw_dy[2] >> 16;
		int dzpix = (dzdx < 0 ? (~dzdx & 0x7fff) : dzdx) + (dzdy < 0 ? (~dzdy & 0x7fff) : dzdy);
		dzpix = normalize_dzpix(dzpix);
		s.dz = dzpix;
		s.dz_compressed = dz_compress_host(dzpix);
	}
}

DEV int32_t quantize_x(int32_t x) { return (x >> 12) | ((x & 0xfff) != 0); }

// interpolate_snapped (span_setup.comp:80-85): whole native lines step by
// the full gradient, the lines between them by the gradient >> L.
DEV int32_t interpolate_snapped(int32_t dvalue, int dy, int L)
{
	return (dy >> L) * dvalue + (dy & ((1 << L) - 1)) * (dvalue >> L);
}

// span_setup.comp:87-227 for one scanline; L is SCALING_LOG2 (0 natively;
// otherwise y counts lines of the 2^L-times domain).
template <typename Scissor>
DEV void compute_span(const PrimSetup &s, Scissor sc, int y, SpanSetup &out, int L)
{
	const int S = 1 << L;
	const bool flip = (s.flags & TRI_FLIP) != 0;
	{
		const bool do_offset = (s.flags & TRI_DO_OFFSET) != 0;
		const bool skip_xfrac = (s.flags & TRI_SKIP_XFRAC) != 0;
		int dy = y - (s.yh >> 2) * S;
		int32_t xh = s.xh * S + dy * (s.dxhdy << 2);
		int32_t diff_c[4] = { 0, 0, 0, 0 }, diff_t[4] = { 0, 0, 0, 0 };
		if (do_offset) {
			xh += (3 * S) * s.dxhdy;
#pragma unroll
			for (int i = 0; i < 4; i++) {
				int32_t deh = s.drgba_de[i] & ~0x1ff, dyh = s.drgba_dy[i] & ~0x1ff;
				diff_c[i] = deh - (deh >> 2) - dyh + (dyh >> 2);
				deh = s.dstzw_de[i] & ~0x1ff;
				dyh = s.dstzw_dy[i] & ~0x1ff;
				diff_t[i] = deh - (deh >> 2) - dyh + (dyh >> 2);
			}
		}
		out.base_x = xh >> 15;
		int32_t xfrac = skip_xfrac ? 0 : ((xh >> 7) & 0xff);
#pragma unroll
		for (int i = 0; i < 4; i++) {
			int32_t v = s.rgba[i] + interpolate_snapped(s.drgba_de[i], dy, L);
			out.rgba[i] =
			    ((v & ~0x1ff) + diff_c[i] - interpolate_snapped((s.drgba_dx[i] >> 8) & ~1, xfrac, L)) & ~0x3ff;
			v = s.stzw[i] + interpolate_snapped(s.dstzw_de[i], dy, L);
			out.stzw[i] =
			    ((v & ~0x1ff) + diff_t[i] - interpolate_snapped((s.dstzw_dx[i] >> 8) & ~1, xfrac, L)) & ~0x3ff;
		}
	}

	const int yh_base = (s.yh & ~3) * S;
	const int ym_base = s.ym * S;
	const int ylo = imax(s.yh, (int)sc[1]) * S;
	const int yhi = imin(s.yl, (int)sc[3]) * S;
	const int32_t lo_sc = S * (int32_t)(sc[0] << 1), hi_sc = S * (int32_t)(sc[2] << 1);
	bool all_over = true, all_under = true, all_invalid = true;
	int32_t start_x = 0x7fffffff, end_x = (int32_t)0x80000000;
#pragma unroll
	for (int i = 0; i < 4; i++) {
		int ys = y * 4 + i;
		bool clip_y = ys < ylo || ys >= yhi;
		int32_t xh = s.xh * S + (ys - yh_base) * s.dxhdy;
		int32_t xm = s.xm * S + (ys - yh_base) * s.dxmdy;
		int32_t xl = s.xl * S + (ys - ym_base) * s.dxldy;
		if (ys < ym_base) xl = xm;
		xl = sext(xl, 27 + L);
		xh = sext(xh, 27 + L);
		int32_t xh_s = quantize_x(xh), xl_s = quantize_x(xl);
		int32_t l = flip ? xh_s : xl_s;
		int32_t r = flip ? xl_s : xh_s;
		bool invalid = (l >> 1) > (r >> 1);
		if (!(imin(l, r) >= hi_sc)) all_over = false;
		if (!(imax(l, r) < lo_sc)) all_under = false;
		l = imin(imax(l, lo_sc), hi_sc);
		r = imin(imax(r, lo_sc), hi_sc);
		invalid = invalid || clip_y;
		if (invalid) {
			l = 0xffff;
			r = 0;
		} else
			all_invalid = false;
		out.xleft[i] = (uint16_t)l;
		out.xright[i] = (uint16_t)r;
		start_x = imin(start_x, l);
		end_x = imax(end_x, r);
	}
	out.start_x = start_x >> 3;
	out.end_x = end_x >> 3;
	bool valid = !all_invalid && !all_over && !all_under;
	if ((s.flags & TRI_INTERLACE_FIELD) && (((y >> L) & 1) != (int)((s.flags & TRI_INTERLACE_KEEP_ODD) != 0)))
		valid = false;
	out.valid = valid ? 1 : 0;
	out.lodlength = (int16_t)(flip ? (out.end_x - out.base_x) : (out.base_x - out.start_x));
}

// Setup runs in two launches.  k_setup: one lane per primitive decodes its
// command words and draw state, writes the RasterPrim record and the decoded
// PrimSetup, and marks which primitive owns each of its tile-row entries.
// k_span: one wave per 8 tile-row entries, one lane per span row: the rows
// of small triangles pack densely into waves and a tall triangle spreads over
// many waves.
__global__ __launch_bounds__(64) void k_setup(const PassDesc *__restrict__ passes, const uint32_t *__restrict__ items)
{
	const uint32_t item = items[blockIdx.x];
	const PassDesc &P = passes[item >> 7];
	const uint32_t prim = (item & 0x7f) * 64 + threadIdx.x;
	if (prim >= P.num_prims) return;

	const PrimRecord &rec = P.prims[prim];
	const DrawState &st = P.states[rec.state_index];
	PrimSetup s;
	decode_primitive(P.words + rec.word_offset, rec.opcode, st, s);
	s.state_index = rec.state_index;
	s.tmem_instance = rec.tmem_instance;
	s.prim_index = rec.prim_index;
	s.ylo = rec.ylo;
	s.yhi = rec.yhi;
	s.span_offset = rec.span_offset;
	s.tilerow_offset = rec.tilerow_offset;
	for (int i = 0; i < 4; i++) s.scissor[i] = st.scissor[i];

	RasterPrim r;
	for (int i = 0; i < 4; i++) {
		r.drgba_dx[i] = s.drgba_dx[i];
		r.drgba_dy[i] = s.drgba_dy[i];
		r.dstzw_dx[i] = s.dstzw_dx[i];
		r.dstzw_dy[i] = s.dstzw_dy[i];
		r.factors[i] = st.factors[i];
		r.fog[i] = st.fog[i];
		r.blendc[i] = st.blendc[i];
	}
	r.flags = s.flags;
	r.tile = s.tile;
	r.ylo = s.ylo;
	r.yhi = s.yhi;
	r.span_offset = s.span_offset;
	r.tmem_instance = s.tmem_instance;
	r.prim_index = s.prim_index;
	r.state_index = s.state_index;
	r.dz = s.dz;
	r.dz_compressed = s.dz_compressed;
	r.ras_flags = st.ras_flags;
	r.dither = st.dither;
	r.db_flags = st.db_flags;
	r.coverage_mode = st.coverage_mode;
	r.z_mode = st.z_mode;
	r.fill_color = st.fill_color;
	r.min_lod = st.min_lod;
	r.variant = rec.variant;
	r.pad0[0] = r.pad0[1] = 0;
	for (int c = 0; c < 2; c++) {
		for (int i = 0; i < 8; i++) r.comb[c][i] = st.comb[c][i];
		for (int i = 0; i < 4; i++) {
			r.blend[c][i] = st.blend[c][i];
			for (int k = 0; k < 4; k++) r.constants[c][i][k] = st.constants[c][i][k];
		}
	}
	for (int i = 0; i < 8; i++) r.pad1[i] = 0;
	P.rprims[prim] = r;
	P.psetups[prim] = s;

	// Spans cover rows ylo .. yhi + 1 and k_span covers the tile rows
	// ylo / 8 .. yhi / 8; row yhi + 1 lies beyond them when it starts a new
	// tile row, and is done here.
	const int rows = (s.yhi >> 3) - (s.ylo >> 3) + 1;
	for (int k = 0; k < rows; k++) P.tilerow_prim[s.tilerow_offset + (uint32_t)k] = (uint16_t)prim;
	if (((s.yhi + 1) & 7) == 0) {
		SpanSetup sp;
		compute_span(s, st.scissor, s.yhi + 1, sp, (int)P.scale_log2);
		P.spans[s.span_offset + (uint32_t)(s.yhi + 1 - s.ylo)] = sp;
	}
}

#ifndef RDP_SPAN_WAVES
#define RDP_SPAN_WAVES 6
#endif
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RDP_SPAN_WAVES)))
void k_span(const PassDesc *__restrict__ passes, const uint32_t *__restrict__ items)
{
	const uint32_t item = items[blockIdx.x];
	const PassDesc &P = passes[item >> 17];
	const int lane = threadIdx.x;
	const uint32_t e = (item & 0x1ffff) * 8 + (uint32_t)(lane >> 3);
	const bool live = e < P.num_tilerows;
	int tlo = 0x7fffffff, thi = -1;
	if (live) {
		const PrimSetup &s = P.psetups[P.tilerow_prim[e]];
		const int y = ((s.ylo >> 3) + (int)(e - s.tilerow_offset)) * 8 + (lane & 7);
		if (y >= s.ylo && y <= s.yhi + 1) {
			SpanSetup sp;
			compute_span(s, s.scissor, y, sp, (int)P.scale_log2);
			P.spans[s.span_offset + (uint32_t)(y - s.ylo)] = sp;
			if (y <= s.yhi && sp.valid) {
				const int lo = imax(sp.start_x, 0), hi = imin(sp.end_x, (int)P.fb_width - 1);
				if (lo <= hi) {
					tlo = lo >> 3;
					thi = hi >> 3;
				}
			}
		}
	}
	// Reduce over the 8 lanes of one tile row.
#pragma unroll
	for (int m = 1; m < 8; m <<= 1) {
		tlo = imin(tlo, __shfl_xor(tlo, m, 64));
		thi = imax(thi, __shfl_xor(thi, m, 64));
	}
	if (live && (lane & 7) == 0) P.tilerows[e] = thi >= tlo ? ((uint32_t)tlo | ((uint32_t)thi << 16)) : 0xffffu;
}

// ==========================================================================
// Stage 2: TMEM snapshots, tmem_update.comp:108-577 (one lane per halfword).
// ==========================================================================
DEV uint32_t tmem_fetch(const Mem &m, const TmemLoad &ld, bool upper, bool yuv)
{
	if (ld.kind == TMEM_LOAD_16) return m.rd16a(ld.addr);
	const uint32_t addr = ld.addr;
	uint32_t word;
	if ((addr & 3) == 0)
		word = m.rd32i(addr >> 2);
	else
		word = (m.rd8(addr) << 24) | (m.rd8(addr + 1) << 16) | (m.rd8(addr + 2) << 8) | m.rd8(addr + 3);
	if (yuv) {
		if (upper)
			word = (((word >> 16) & 0xff) << 8) | (word & 0xff);
		else
			word = (((word >> 24) & 0xff) << 8) | ((word >> 8) & 0xff);
	} else {
		word >>= 16 - 16 * (uint32_t)upper;
		word &= 0xffff;
	}
	return word;
}

// Stage 2 runs in two kernels so a pass's uploads are processed in parallel.
// k_tmem_gather: one thread per (block of 32 uploads, TMEM halfword) applies
// the block's uploads in order: from the block's first write to its halfword
// on, every snapshot slot of the block is complete.  It records where that
// first write is and the halfword's value after the block.  The item is
// pass << 12 | block << 3 | 256-halfword chunk.
constexpr int kTmemBlock = 32;

__global__ __launch_bounds__(256) void k_tmem_gather(const PassDesc *__restrict__ passes,
                                                     const uint32_t *__restrict__ items)
{
	const uint32_t item = items[blockIdx.x];
	const PassDesc &P = passes[item >> 12];
	const int g = (int)((item & 7) * 256 + threadIdx.x);  // 0..2047
	const int block = (int)((item >> 3) & 0x1ff);
	const int i0 = block * kTmemBlock;
	const Mem m{ gptr(P.rdram), gptr(P.hidden), P.rdram_mask };
	const int idx = g ^ 1;
	const bool upper = idx >= 0x400;
	const int n = (int)P.num_uploads;
	const int end = imin(i0 + kTmemBlock, n);
	uint32_t cur = 0;
	int first = kTmemBlock;
	for (int i = i0; i < end; i++) {
		const UploadInfo u = P.uploads[i];
		TmemLoad ld;
		if (tmem_plan(u, idx, upper, ld)) {
			cur = tmem_fetch(m, ld, upper, u.mode != UP_TLUT && u.tmem_fmt == TF_YUV);
			if (first == kTmemBlock) first = i - i0;
		}
		if (first != kTmemBlock) P.tmem_instances[(i + 1) * 2048 + g] = (uint16_t)cur;
	}
	P.tmem_blocks[block * 2048 + g] = ((uint32_t)first << 16) | cur;
}

// k_tmem_scan: one thread per TMEM halfword walks the blocks in order,
// filling each block's slots before its first write with the value carried
// from the blocks before (the last writer wins), then writes the final TMEM
// back to the stream.  The item is pass << 12 | 256-halfword chunk.
__global__ __launch_bounds__(256) void k_tmem_scan(const PassDesc *__restrict__ passes,
                                                   const uint32_t *__restrict__ items)
{
	const uint32_t item = items[blockIdx.x];
	const PassDesc &P = passes[item >> 12];
	const int g = (int)((item & 7) * 256 + threadIdx.x);
	uint32_t carry = P.tmem[g];
	P.tmem_instances[g] = (uint16_t)carry;
	const int n = (int)P.num_uploads;
	const int blocks = (n + kTmemBlock - 1) / kTmemBlock;
	for (int b = 0; b < blocks; b++) {
		const uint32_t info = P.tmem_blocks[b * 2048 + g];
		const int i0 = b * kTmemBlock;
		const int count = imin(kTmemBlock, n - i0);
		const int first = (int)(info >> 16);
		const int gap = imin(first, count);
		for (int k = 0; k < gap; k++) P.tmem_instances[(i0 + k + 1) * 2048 + g] = (uint16_t)carry;
		if (first < count) carry = info & 0xffff;
	}
	P.tmem[g] = (uint16_t)carry;
}

// ==========================================================================
// Stage 3: binning.
// ==========================================================================
// One wave per (pass, tile row, block of 8 bin groups).  Primitive p of a
// group covers the tile columns [lo_p, hi_p] of the row, so its bit is on in
// the mask of column x iff lo_p <= x < hi_p + 1: XOR the bit into a
// difference array at lo_p and hi_p + 1, and the inclusive XOR prefix over the
// columns gives every column's mask.  Each lane then holds one column's masks
// of the 8 groups and writes them as one 64-byte record (the group count is
// padded to a multiple of 8, so the records are line aligned).
DEV uint64_t xor_prefix64(uint64_t v, int lane)
{
	uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		const uint32_t tl = __shfl_up(lo, o, 64), th = __shfl_up(hi, o, 64);
		if (lane >= o) {
			lo ^= tl;
			hi ^= th;
		}
	}
	return ((uint64_t)hi << 32) | lo;
}

__global__ __launch_bounds__(64) void k_bin(const PassDesc *__restrict__ passes, const uint32_t *__restrict__ items)
{
	__shared__ unsigned long long s_diff[65];
	const uint32_t item = items[blockIdx.x];
	const PassDesc &P = passes[item >> 17];
	const uint32_t ty = (item >> 4) & 0x1fff, gb = (item & 0xf) * 8;
	const int lane = threadIdx.x;
	const uint32_t tiles_x = P.tiles_x, G = P.num_groups;
	const uint32_t real_groups = (P.num_prims + 63) / 64;
	const uint32_t ng = real_groups > gb ? min(real_groups - gb, 8u) : 0u;

	// Column range of each of the lane's 8 primitives in this tile row
	// (lo > hi when the primitive misses the row).
	uint32_t lo[8], hi[8];
#pragma unroll
	for (uint32_t k = 0; k < 8; k++) {
		lo[k] = 0xffff;
		hi[k] = 0;
		const uint32_t prim = (gb + k) * 64 + (uint32_t)lane;
		if (k < ng && prim < P.num_prims) {
			const PrimRecord &r = P.prims[prim];
			if (r.yhi >= r.ylo) {
				const int row0 = r.ylo >> 3, row1 = r.yhi >> 3;
				if ((int)ty >= row0 && (int)ty <= row1) {
					const uint32_t v = P.tilerows[r.tilerow_offset + (ty - row0)];
					lo[k] = v & 0xffff;
					hi[k] = v >> 16;
				}
			}
		}
	}

	uint64_t *out = P.bin_fine + (size_t)ty * tiles_x * G + gb;
	for (uint32_t x0 = 0; x0 < tiles_x; x0 += 64) {
		uint64_t m[8];
#pragma unroll
		for (uint32_t k = 0; k < 8; k++) {
			m[k] = 0;
			if (k >= ng) continue;
			s_diff[lane] = 0;
			if (lane == 0) s_diff[64] = 0;
			__syncthreads();
			// The part of [lo, hi] inside columns x0 .. x0 + 63.
			const uint32_t a = max(lo[k], x0), b = min(hi[k], x0 + 63);
			if (lo[k] <= hi[k] && a <= b) {
				const unsigned long long bit = 1ull << lane;
				atomicXor(&s_diff[a - x0], bit);
				atomicXor(&s_diff[b + 1 - x0], bit);
			}
			__syncthreads();
			m[k] = xor_prefix64(s_diff[lane], lane);
			__syncthreads();
		}
		const uint32_t x = x0 + (uint32_t)lane;
		if (x < tiles_x) {
			uint64_t *o = out + (size_t)x * G;
#pragma unroll
			for (uint32_t k = 0; k < 8; k++) o[k] = m[k];
		}
	}
}

// ==========================================================================
// Stage 4: per-tile shading and depth/blend.
// ==========================================================================

// coverage.h:31-44.  Sample layout per sub-scanline: 0,4 / 2,6 / 0,4 / 2,6.
DEV int compute_coverage(const SpanSetup &sp, int x)
{
	int cov = 0;
	const int xs = x << 3;
#pragma unroll
	for (int k = 0; k < 8; k++) {
		const int line = k >> 1;
		const int off = ((k & 1) ? 4 : 0) + ((line & 1) ? 2 : 0);
		const int xv = xs + off;
		if (xv >= (int)sp.xleft[line] && xv < (int)sp.xright[line])
			cov |= 1 << k;
	}
	return cov;
}

DEV int32_t clamp_9bit(int32_t c) { return iclamp(sext(c - 0x80, 9) + 0x80, 0, 0xff); }

DEV int32_t clamp_z(int32_t z)
{
	z = sext(z - (1 << 17), 19) + (1 << 17);
	return iclamp(z, 0, 0x3ffff);
}

// perspective.h:46-112.  Reciprocal table r_i = round(2^20 / (64+i)), except
// entry 6 which the hardware table holds one lower; slopes are -4 * (r_i - r_i+1).
DEV int32_t persp_base(int i) { return ((1 << 20) + (64 + i) / 2) / (64 + i) - (i == 6); }

// The reciprocal table of perspective_divide (perspective.h:27-44), 65
// entries from the closed form above, in LDS: the raster kernels fill it at
// their start (persp_table_init), so the per-lane lookup is one LDS read
// instead of an integer division.
__shared__ int32_t s_persp_base[65];

DEV void persp_table_init()
{
	for (int k = (int)threadIdx.x; k < 65; k += (int)blockDim.x) s_persp_base[k] = persp_base(k);
	__syncthreads();
}

DEV void perspective_divide(int32_t s, int32_t t, int32_t w, int32_t &os, int32_t &ot, bool &overflow)
{
	bool w_carry = w <= 0;
	w &= 0x7fff;
	int shift = imin(14 - find_msb(w), 14);
	int32_t normout = (w << shift) & 0x3fff;
	int32_t wnorm = normout & 0xff;
	int i = normout >> 8;
	int32_t base = s_persp_base[i];
	int32_t slope = -4 * (base - s_persp_base[i + 1]);
	// s and t are 16-bit, the reciprocal below 2^15 and the slope term
	// small: every operand fits 24 bits, so the full-rate 24-bit multiply
	// gives the exact 32-bit product.
	int32_t rcp = (__mul24(slope, wnorm) >> 10) + base;
	int32_t prod[2] = { __mul24(s, rcp), __mul24(t, rcp) };
	int32_t temp_mask = ((1 << 30) - 1) & -((1 << 29) >> shift);
	int32_t out[2];
#pragma unroll
	for (int k = 0; k < 2; k++) {
		int32_t oob = prod[k] & temp_mask;
		if (shift != 14) {
			prod[k] >>= (13 - shift);
			out[k] = prod[k];
		} else
			out[k] = prod[k] << 1;
		if (oob != temp_mask && oob != 0) {
			out[k] = (prod[k] & (1 << 29)) == 0 ? 0x7fff : -0x8000;
			overflow = true;
		}
	}
	if (w_carry) {
		out[0] = out[1] = 0x7fff;
		overflow = true;
	}
	os = iclamp(out[0], -0x10000, 0xffff);
	ot = iclamp(out[1], -0x10000, 0xffff);
}

DEV void stw_to_st(int32_t s, int32_t t, int32_t w, bool persp, int32_t &os, int32_t &ot, bool &ovf)
{
	s >>= 16;
	t >>= 16;
	w >>= 16;
	if (persp)
		perspective_divide(s, t, w, os, ot, ovf);
	else {
		os = s;
		ot = t;
	}
}

// noise.h:29-37
DEV uint32_t reseed_noise(uint32_t x, uint32_t y, uint32_t prim)
{
	const uint32_t P = 1103515245u;
	uint32_t s0 = x, s1 = y, s2 = prim;
#pragma unroll
	for (int r = 0; r < 3; r++) {
		uint32_t n0 = ((s0 >> 8) ^ s1) * P;
		uint32_t n1 = ((s1 >> 8) ^ s2) * P;
		uint32_t n2 = ((s2 >> 8) ^ s0) * P;
		s0 = n0;
		s1 = n1;
		s2 = n2;
	}
	return (s0 >> 16) & 0xffff;
}

// ---- texture sampling: texture.h -----------------------------------------
struct Tex {
	const AS_LDS uint16_t *tmem;
	MDEV uint32_t t16(uint32_t i) const { return tmem[i & 0x7ff]; }
	MDEV uint32_t t8(uint32_t i) const { return reinterpret_cast<const AS_LDS uint8_t *>(tmem)[i & 0xfff]; }
};

// Branch-free on the tile's (uniform) mask and mirror bits: with no mask
// both operations reduce to the identity (mirror bit 0, mask all ones).
DEV int mask_coord(int v, uint32_t mask_bits, bool mirror)
{
	const int bit = mask_bits ? (1 << mask_bits) : 0;
	const int keep = mask_bits ? bit - 1 : -1;
	v ^= imax((v & (mirror ? bit : 0)) - 1, 0);
	return v & keep;
}

DEV void rgba16_to(uint32_t w, int32_t o[4])
{
#pragma unroll
	for (int i = 0; i < 3; i++) {
		uint32_t c = (w >> (11 - 5 * i)) & 31;
		o[i] = (int32_t)((c << 3) | (c >> 2));
	}
	o[3] = (w & 1) ? 0xff : 0;
}

DEV void ia16_to(uint32_t w, int32_t o[4])
{
	o[0] = o[1] = o[2] = (int32_t)(w >> 8);
	o[3] = (int32_t)(w & 0xff);
}

// A texel row's TMEM base: offset + stride * row.  Every TMEM address keeps
// at most its low 13 bits, which depend only on the low 24 bits of the
// operands, so the 24-bit multiply (full rate, where a 32-bit one issues at a
// quarter of it) gives the same address.
DEV uint32_t row_base(const TileDesc &t, uint32_t row) { return t.offset + __umul24(t.stride, row); }

DEV uint32_t tidx8(const TileDesc &t, uint32_t v, uint32_t wrap, uint32_t add)
{
	uint32_t b = (row_base(t, v) + add) & wrap;
	return (b ^ ((v & 1) << 2)) ^ 3;
}

DEV uint32_t tidx16(const TileDesc &t, uint32_t s, uint32_t v, uint32_t wrap)
{
	uint32_t b = (row_base(t, v) + s * 2) & wrap;
	return ((b >> 1) ^ ((v & 1) << 1)) ^ 1;
}

// sample_texel_* of texture.h:90-315 for one texel.
// TF >= 0 is the tile's format * 4 + size known at compile time.
template <int TF>
DEV void fetch_texel(const Tex &tx, const TileDesc &t, uint32_t s, uint32_t v, bool tlut, bool tlut_type,
                     uint32_t lut_offset, uint32_t addr_xor, int32_t o[4])
{
	const uint32_t size = TF >= 0 ? (uint32_t)(TF & 3) : (uint32_t)t.size;
	if (tlut) {
		uint32_t lut_entry;
		if (size == 0) {
			uint32_t w = (tx.t8(tidx8(t, v, 0x7ff, s >> 1)) >> ((~s & 1) * 4)) & 0xf;
			w |= (uint32_t)t.palette << 4;
			lut_entry = (w << 2) + lut_offset;
		} else if (size == 1) {
			lut_entry = (tx.t8(tidx8(t, v, 0x7ff, s)) << 2) + lut_offset;
		} else {
			lut_entry = ((tx.t16(tidx16(t, s, v, 0x7ff)) >> 6) & ~3u) + lut_offset;
		}
		uint32_t w = tx.t16(0x400 | (lut_entry ^ addr_xor));
		if (tlut_type) ia16_to(w, o);
		else rgba16_to(w, o);
		return;
	}
	const uint32_t fmt = TF >= 0 ? (uint32_t)(TF >> 2) : (uint32_t)t.fmt;
	int kind;  // 0: 4-bit replicate, 1: 8-bit, 2: 16-bit 2x8 (ci32), 3: rgba16, 4: rgba32, 5: ia4, 6: ia8, 7: ia16, 8: ci4
	switch (fmt) {
	case TF_RGBA: kind = size == 0 ? 0 : size == 1 ? 1 : size == 2 ? 3 : 4; break;
	case TF_CI: kind = size == 0 ? 8 : size == 1 ? 1 : 2; break;
	case TF_IA: kind = size == 0 ? 5 : size == 1 ? 6 : size == 2 ? 7 : 2; break;
	case TF_I: kind = size == 0 ? 0 : size == 1 ? 1 : 2; break;
	default: kind = -1; break;
	}
	switch (kind) {
	case 0: {
		uint32_t w = (tx.t8(tidx8(t, v, 0xfff, s >> 1)) >> ((~s & 1) * 4)) & 0xf;
		w |= w << 4;
		o[0] = o[1] = o[2] = o[3] = (int32_t)w;
		break;
	}
	case 1: {
		uint32_t w = tx.t8(tidx8(t, v, 0xfff, s));
		o[0] = o[1] = o[2] = o[3] = (int32_t)w;
		break;
	}
	case 2: {
		uint32_t w = tx.t16(tidx16(t, s, v, 0xfff));
		o[0] = o[2] = (int32_t)(w >> 8);
		o[1] = o[3] = (int32_t)(w & 0xff);
		break;
	}
	case 3: rgba16_to(tx.t16(tidx16(t, s, v, 0xfff)), o); break;
	case 4: {
		uint32_t i = tidx16(t, s, v, 0x7ff);
		uint32_t lo = tx.t16(i), hi = tx.t16(i | 0x400);
		o[0] = (int32_t)(lo >> 8);
		o[1] = (int32_t)(lo & 0xff);
		o[2] = (int32_t)(hi >> 8);
		o[3] = (int32_t)(hi & 0xff);
		break;
	}
	case 5: {
		uint32_t w = (tx.t8(tidx8(t, v, 0xfff, s >> 1)) >> ((~s & 1) * 4)) & 0xf;
		uint32_t in = w & 0xe;
		in = (in << 4) | (in << 1) | (in >> 2);
		o[0] = o[1] = o[2] = (int32_t)in;
		o[3] = (w & 1) ? 0xff : 0;
		break;
	}
	case 6: {
		uint32_t w = tx.t8(tidx8(t, v, 0xfff, s));
		uint32_t in = w >> 4, al = w & 0xf;
		o[0] = o[1] = o[2] = (int32_t)(in | (in << 4));
		o[3] = (int32_t)(al | (al << 4));
		break;
	}
	case 7: ia16_to(tx.t16(tidx16(t, s, v, 0xfff)), o); break;
	case 8: {
		uint32_t w = (tx.t8(tidx8(t, v, 0xfff, s >> 1)) >> ((~s & 1) * 4)) & 0xf;
		w |= (uint32_t)t.palette << 4;
		o[0] = o[1] = o[2] = o[3] = (int32_t)w;
		break;
	}
	default: o[0] = o[1] = o[2] = o[3] = 0; break;
	}
}

DEV void fetch_yuv16(const Tex &tx, const TileDesc &t, uint32_t s, uint32_t v, uint32_t cx, int32_t o[4])
{
	uint32_t bo = row_base(t, v);
	uint32_t bl = (bo + s) & 0x7ff, bc = (bo + cx * 2) & 0x7ff;
	uint32_t il = (bl ^ ((v & 1) << 2)) ^ 3;
	uint32_t ic = ((bc >> 1) ^ ((v & 1) << 1)) ^ 1;
	uint32_t luma = tx.t8(il | 0x800);
	uint32_t chroma = tx.t16(ic);
	o[0] = (int32_t)((chroma >> 8) & 0xff) - 0x80;
	o[1] = (int32_t)(chroma & 0xff) - 0x80;
	o[2] = o[3] = (int32_t)luma;
}

// The tile's shift (uniform): right by 0-10, or left by 1-5 with the
// result kept to 16 bits (shifts 11-15).  Both forms are computed and one
// selected, which costs fewer instructions than a branch per coordinate.
DEV int tile_shift(int c, int shift)
{
	c = iclamp(c, -0x8000, 0x7fff);
	const int right = c >> (shift < 11 ? shift : 0);
	const int left = (int)((uint32_t)c << ((32 - shift) & 31)) >> 16;
	return shift < 11 ? right : left;
}

DEV int clamp_and_shift_coord(bool clamp_bit, int c, int lo, int hi, int shift)
{
	c = tile_shift(c, shift);
	const int base = c - (lo << 3);
	const int clamped = (c >> 3) >= hi ? (((hi >> 2) - (lo >> 2)) & 0x3ff) << 5 : imax(base, 0);
	return clamp_bit ? clamped : base;
}

DEV int shift_coord(int c, int lo, int shift) { return tile_shift(c, shift) - (lo << 3); }

DEV void convert_factors(const int32_t in[4], const int32_t f[4], int32_t o[4])
{
	int32_t t0 = sext(in[0], 9), t1 = sext(in[1], 9), t2 = sext(in[2], 9);
	// 16-bit factors times 9-bit terms: exact with 24-bit multiplies.
	o[0] = t2 + ((__mul24(f[0], t1) + 0x80) >> 8);
	o[1] = t2 + ((__mul24(f[1], t0) + __mul24(f[2], t1) + 0x80) >> 8);
	o[2] = t2 + ((__mul24(f[3], t0) + 0x80) >> 8);
	o[3] = t2;
}

DEV void bilinear_3tap2(const int32_t *t00, const int32_t *t10, const int32_t *t01, const int32_t *t11, int fx, int fy,
                        int32_t *o)
{
	int sum = fx + fy;
	const int32_t *b = sum >= 32 ? t11 : t00;
	int ffx = sum >= 32 ? 32 - fy : fx;
	int ffy = sum >= 32 ? 32 - fx : fy;
#pragma unroll
	for (int i = 0; i < 2; i++)
		o[i] = (((__mul24(t10[i] - b[i], ffx) + __mul24(t01[i] - b[i], ffy)) + 0x10) >> 5) + b[i];
}

// texture.h:498-923
template <int TF>
DEV void sample_texture(const Tex &tx, const TileDesc &tile, int32_t st_s, int32_t st_t, bool tlut, bool tlut_type,
                        bool sample_quad, bool mid_state, bool convert_one, bool bilerp, const int32_t factors[4],
                        const int32_t prev[4], int32_t out[4])
{
	int s = clamp_and_shift_coord((tile.flags & TILE_CLAMP_S) != 0, st_s, (int)tile.slo, (int)tile.shi, tile.shift_s);
	int t = clamp_and_shift_coord((tile.flags & TILE_CLAMP_T) != 0, st_t, (int)tile.tlo, (int)tile.thi, tile.shift_t);
	int fx = 0, fy = 0;
	if (sample_quad || tlut) {
		fx = s & 31;
		fy = t & 31;
	}
	int sum_frac = fx + fy;
	s >>= 5;
	t >>= 5;
	const bool ms = (tile.flags & TILE_MIRROR_S) != 0, mt = (tile.flags & TILE_MIRROR_T) != 0;
	int s0 = mask_coord(s, tile.mask_s, ms);
	int t0 = mask_coord(t, tile.mask_t, mt);
	int s1 = mask_coord(s + 1, tile.mask_s, ms);
	int t1 = mask_coord(t + 1, tile.mask_t, mt);
	int tdiff = imax(t1 - t0, -255);
	t1 = (t0 & 0xff) + tdiff;
	t0 &= 0xff;

	int32_t tb[4] = { 0, 0, 0, 0 }, t10[4] = { 0, 0, 0, 0 }, t01[4] = { 0, 0, 0, 0 }, t11[4] = { 0, 0, 0, 0 };
	const bool mid_texel = mid_state && bilerp && fx == 0x10 && fy == 0x10;
	const bool upper_lut = sum_frac >= 0x20;
	if (mid_texel) sum_frac = 0;
	const bool yuv = (TF >= 0 ? (uint32_t)(TF >> 2) : (uint32_t)tile.fmt) == TF_YUV;
	int bs = sum_frac >= 0x20 ? s1 : s0;
	int bt = sum_frac >= 0x20 ? t1 : t0;
	const int chroma_frac = ((s0 & 1) << 4) | (fx >> 1);

	if (tlut) {
		if (!sample_quad) {
			bs = s0;
			bt = t0;
			s1 = s0;
			t1 = t0;
		}
		if (!yuv) {
			const bool upper = sum_frac >= 0x20;
			const uint32_t axor = upper_lut ? 2 : 1;
			fetch_texel<TF>(tx, tile, (uint32_t)bs, (uint32_t)bt, true, tlut_type, upper ? 3 : 0, axor, tb);
			if (bilerp) {
				fetch_texel<TF>(tx, tile, (uint32_t)s1, (uint32_t)t0, true, tlut_type, 1, axor, t10);
				fetch_texel<TF>(tx, tile, (uint32_t)s0, (uint32_t)t1, true, tlut_type, 2, axor, t01);
			}
			if (mid_texel)
				fetch_texel<TF>(tx, tile, (uint32_t)s1, (uint32_t)t1, true, tlut_type, 3, axor, t11);
		}
	} else if (yuv) {
		uint32_t cx0 = (uint32_t)(s0 >> 1), cx1 = (uint32_t)((s1 + (s1 - s0)) >> 1);
		fetch_yuv16(tx, tile, (uint32_t)s0, (uint32_t)t0, cx0, tb);
		if (sample_quad) {
			fetch_yuv16(tx, tile, (uint32_t)s1, (uint32_t)t0, cx1, t10);
			fetch_yuv16(tx, tile, (uint32_t)s0, (uint32_t)t1, cx0, t01);
			fetch_yuv16(tx, tile, (uint32_t)s1, (uint32_t)t1, cx1, t11);
		}
	} else {
		fetch_texel<TF>(tx, tile, (uint32_t)bs, (uint32_t)bt, false, false, 0, 0, tb);
		if (sample_quad) {
			fetch_texel<TF>(tx, tile, (uint32_t)s1, (uint32_t)t0, false, false, 0, 0, t10);
			fetch_texel<TF>(tx, tile, (uint32_t)s0, (uint32_t)t1, false, false, 0, 0, t01);
		}
		if (mid_texel)
			fetch_texel<TF>(tx, tile, (uint32_t)s1, (uint32_t)t1, false, false, 0, 0, t11);
	}

	int32_t acc[4];
	if (convert_one) {
		int32_t ps[4];
#pragma unroll
		for (int i = 0; i < 4; i++) ps[i] = sext(prev[i], 9);
		if (sample_quad) {
			const bool mid_rg = yuv ? (mid_state && chroma_frac == 0x10 && fy == 0x10) : mid_texel;
			const bool mid_ba = mid_texel;
			const bool upper_ba = sum_frac >= 32;
			const bool upper_rg = yuv ? ((chroma_frac + fy) >= 32 && !mid_rg) : upper_ba;
			int32_t frg0 = upper_rg ? ps[1] : ps[0], frg1 = upper_rg ? ps[0] : ps[1];
			int32_t fba0 = upper_ba ? ps[1] : ps[0], fba1 = upper_ba ? ps[0] : ps[1];
#pragma unroll
			for (int i = 0; i < 4; i++) {
				const bool rg = i < 2;
				const bool mid = rg ? mid_rg : mid_ba;
				const int32_t f0 = rg ? frg0 : fba0, f1 = rg ? frg1 : fba1;
				int32_t conv;
				if (mid)
					conv = f0 * (t01[i] - t11[i]) + f1 * (t10[i] - t11[i]) + ((tb[i] - t11[i]) << 6) + 0x80;
				else {
					const bool up = rg ? upper_rg : upper_ba;
					int32_t base = (up && yuv) ? t11[i] : tb[i];
					conv = f0 * (t10[i] - base) + f1 * (t01[i] - base) + 0x80;
				}
				acc[i] = (conv >> 8) + ps[2];
			}
		} else
			acc[0] = acc[1] = acc[2] = acc[3] = ps[2];
	} else if (yuv) {
		if (sample_quad) {
			int32_t ac[2], al[2];
			if (bilerp) {
				if (mid_state && chroma_frac == 0x10 && fy == 0x10) {
					ac[0] = (tb[0] + t10[0] + t11[0] + t01[0] + 2) >> 2;
					ac[1] = (tb[1] + t10[1] + t11[1] + t01[1] + 2) >> 2;
				} else
					bilinear_3tap2(tb, t10, t01, t11, chroma_frac, fy, ac);
				if (mid_texel) {
					al[0] = (tb[2] + t10[2] + t11[2] + t01[2] + 2) >> 2;
					al[1] = (tb[3] + t10[3] + t11[3] + t01[3] + 2) >> 2;
				} else
					bilinear_3tap2(tb + 2, t10 + 2, t01 + 2, t11 + 2, fx, fy, al);
			} else {
				const int32_t *l = (fx + fy >= 32) ? t11 : tb;
				const int32_t *c = (chroma_frac + fy >= 32) ? t11 : tb;
				al[0] = l[2];
				al[1] = l[3];
				ac[0] = c[0];
				ac[1] = c[1];
			}
			acc[0] = ac[0];
			acc[1] = ac[1];
			acc[2] = al[0];
			acc[3] = al[1];
		} else {
#pragma unroll
			for (int i = 0; i < 4; i++) acc[i] = tb[i];
		}
	} else if (mid_texel) {
#pragma unroll
		for (int i = 0; i < 4; i++) acc[i] = (tb[i] + t01[i] + t10[i] + t11[i] + 2) >> 2;
	} else if (bilerp && (sample_quad || tlut)) {
		const int ffx = sum_frac >= 32 ? 32 - fy : fx;
		const int ffy = sum_frac >= 32 ? 32 - fx : fy;
#pragma unroll
		for (int i = 0; i < 4; i++) acc[i] = ((((t10[i] - tb[i]) * ffx + (t01[i] - tb[i]) * ffy) + 0x10) >> 5) + tb[i];
	} else {
#pragma unroll
		for (int i = 0; i < 4; i++) acc[i] = tb[i];
	}

	if (!bilerp && !convert_one)
		convert_factors(acc, factors, out);
	else {
#pragma unroll
		for (int i = 0; i < 4; i++) out[i] = acc[i];
	}
}

// texture.h:925-1003
DEV void compute_lod(uint32_t &tile0, uint32_t &tile1, int32_t &lod_frac, uint32_t max_level, int min_lod,
                     const int32_t st[2], const int32_t stdx[2], const int32_t stdy[2], bool ovf, bool tex_lod_en,
                     bool sharpen, bool detail)
{
	bool magnify = false, distant = false;
	uint32_t tile_offset = 0;
	if (ovf) {
		distant = true;
		lod_frac = 0xff;
	} else {
		int32_t dx0 = stdx[0] - st[0], dx1 = stdx[1] - st[1];
		int32_t dy0 = stdy[0] - st[0], dy1 = stdy[1] - st[1];
		dx0 ^= dx0 >> 31;
		dx1 ^= dx1 >> 31;
		dy0 ^= dy0 >> 31;
		dy1 ^= dy1 >> 31;
		int32_t max_d = imax(imax(dx0, dy0), imax(dx1, dy1));
		if (max_d >= 0x4000) {
			distant = true;
			lod_frac = 0xff;
			tile_offset = max_level;
		} else if (max_d < 32) {
			distant = max_level == 0;
			magnify = true;
			if (!sharpen && !detail)
				lod_frac = distant ? 0xff : 0;
			else
				lod_frac = (imax(min_lod, max_d) << 3) + (sharpen ? -0x100 : 0);
		} else {
			int mip_base = imax(find_msb(max_d >> 5), 0);
			distant = (uint32_t)mip_base >= max_level;
			if (distant && !sharpen && !detail)
				lod_frac = 0xff;
			else {
				lod_frac = ((max_d << 3) >> mip_base) & 0xff;
				tile_offset = (uint32_t)mip_base;
			}
		}
	}
	if (tex_lod_en) {
		if (distant) tile_offset = max_level;
		if (!detail) {
			tile0 = (tile0 + tile_offset) & 7;
			tile1 = (distant || (!sharpen && magnify)) ? tile0 : ((tile0 + 1) & 7);
		} else {
			tile1 = (tile0 + tile_offset + ((distant || magnify) ? 1 : 2)) & 7;
			tile0 = (tile0 + tile_offset + (magnify ? 0 : 1)) & 7;
		}
	}
}

// texture.h:360-472
DEV int sample_copy_word(const Tex &tx, const TileDesc &tile, int st_s, int st_t, int s_offset, bool tlut)
{
	const bool high_word = s_offset < 2;
	const bool replicate = high_word && tile.size != 2 && !tlut;
	const int shamt = imin((int)tile.size, 2);
	const uint32_t idx_mask = (tile.size == 3 || tlut) ? 0x3ff : 0x7ff;
	const bool ms = (tile.flags & TILE_MIRROR_S) != 0, mt = (tile.flags & TILE_MIRROR_T) != 0;
	if (replicate) {
		int sx = st_s + 2 * s_offset;
		int t = mask_coord(st_t, tile.mask_t, mt);
		uint32_t tbase = row_base(tile, (uint32_t)t);
		uint32_t sam[2];
#pragma unroll
		for (int i = 0; i < 2; i++) {
			int sv = mask_coord(sx + i, tile.mask_s, ms);
			uint32_t nib = (tbase * 2 + ((uint32_t)sv << shamt)) & 0x1fff;
			nib ^= ((uint32_t)t & 1) * 8;
			uint32_t v = tx.t16(((nib >> 2) & idx_mask) ^ 1);
			if (tile.size == 1)
				v = (v >> (8 - 4 * (nib & 2))) & 0xff;
			else if (tile.size == 0)
				v = ((v >> (12 - 4 * (nib & 3))) & 0xf) * 0x11;
			else
				v >>= 8;
			sam[i] = v;
		}
		return (int)((sam[0] << 8) | sam[1]);
	}
	int sv = mask_coord(st_s + s_offset, tile.mask_s, ms);
	int t = mask_coord(st_t, tile.mask_t, mt);
	uint32_t tbase = row_base(tile, (uint32_t)t);
	uint32_t nib = (tbase * 2 + ((uint32_t)sv << shamt)) & 0x1fff;
	nib ^= ((uint32_t)t & 1) * 8;
	int samp = (int)tx.t16(((nib >> 2) & idx_mask) ^ 1);
	if (tlut) {
		if (tile.size == 0) {
			samp >>= 12 - 4 * (nib & 3);
			samp &= 0xf;
			samp |= (int)tile.palette << 4;
		} else {
			samp >>= 8 - 4 * (nib & 2);
			samp &= 0xff;
		}
		samp = (samp << 2) + s_offset;
		samp = (int)tx.t16(((uint32_t)samp | 0x400) ^ 1);
	}
	return samp;
}

// ---- combiner: combiner.h -------------------------------------------------
DEV int32_t special_expand(int32_t v) { return sext(v - 0x80, 9) + 0x80; }

struct CombIn {
	int32_t shade[4], combined[4], texel0[4], texel1[4];
	int32_t lod_frac, noise;
};

// Selector decoding follows the reference's select_* helpers: values that
// do not name a per-pixel input read the constant slot.
// The constant inputs are the primitive's resolved constants (k, per
// selector slot: prim / env colour, key, convert and LOD terms), which the
// host fills only for the selectors that name one and leaves 0 for every
// other; those selectors return 0 here without reading k, so a combiner
// known at compile time drops the loads and arithmetic of its zero terms.
DEV int32_t sel_rgb(int which, int sel, int ch, const CombIn &in, const AS_CONST uint8_t *k)
{
	switch (which) {
	case 0:  // muladd
		switch (sel) {
		case 0: return in.combined[ch];
		case 1: return in.texel0[ch];
		case 2: return in.texel1[ch];
		case 4: return in.shade[ch];
		case 6: return 0x100;
		case 7: return in.noise;
		case 3:
		case 5: return k[ch];
		default: return 0;
		}
	case 1:  // mulsub
		switch (sel) {
		case 0: return in.combined[ch];
		case 1: return in.texel0[ch];
		case 2: return in.texel1[ch];
		case 4: return in.shade[ch];
		case 7: return ((int32_t)k[1] << 8) | k[2];
		case 3:
		case 5:
		case 6: return k[ch];
		default: return 0;
		}
	case 2:  // mul
		switch (sel) {
		case 0: return in.combined[ch];
		case 1: return in.texel0[ch];
		case 2: return in.texel1[ch];
		case 4: return in.shade[ch];
		case 7: return in.combined[3];
		case 8: return in.texel0[3];
		case 9: return in.texel1[3];
		case 11: return in.shade[3];
		case 13: return in.lod_frac;
		case 15: return ((int32_t)k[1] << 8) | k[2];
		case 3:
		case 5:
		case 6:
		case 10:
		case 12:
		case 14: return k[ch];
		default: return 0;
		}
	default:  // add
		switch (sel) {
		case 0: return in.combined[ch];
		case 1: return in.texel0[ch];
		case 2: return in.texel1[ch];
		case 4: return in.shade[ch];
		case 6: return 0x100;
		case 3:
		case 5: return k[ch];
		default: return 0;
		}
	}
}

DEV int32_t sel_alpha(int which, int sel, const CombIn &in, const AS_CONST uint8_t *k)
{
	if (which == 2) {
		switch (sel) {
		case 0: return in.lod_frac;
		case 1: return in.texel0[3];
		case 2: return in.texel1[3];
		case 4: return in.shade[3];
		case 3:
		case 5:
		case 6: return k[3];
		default: return 0;
		}
	}
	switch (sel) {
	case 0: return in.combined[3];
	case 1: return in.texel0[3];
	case 2: return in.texel1[3];
	case 4: return in.shade[3];
	case 6: return 0x100;
	case 3:
	case 5: return k[3];
	default: return 0;
	}
}

// Selector sources for combine(): read from the primitive at run time, or a
// combiner known at compile time (kKnownCombiners), which lets every selector
// switch fold away.
struct RuntimeSel {
	const AS_CONST uint8_t *p;
	MDEV int operator[](int i) const { return p[i]; }
};

template <int C, int CYCLE>
struct KnownSel {
	MDEV int operator[](int i) const { return kKnownCombiners[C][CYCLE][i]; }
};

template <typename Sel>
DEV void combine(const CombIn &in, const Sel &sel, const AS_CONST uint8_t (*k)[4], int32_t out[4])
{
#pragma unroll
	for (int ch = 0; ch < 4; ch++) {
		int32_t a, b, c, d;
		if (ch < 3) {
			a = sel_rgb(0, sel[0], ch, in, k[0]);
			b = sel_rgb(1, sel[1], ch, in, k[1]);
			c = sel_rgb(2, sel[2], ch, in, k[2]);
			d = sel_rgb(3, sel[3], ch, in, k[3]);
		} else {
			a = sel_alpha(0, sel[4], in, k[0]);
			b = sel_alpha(1, sel[5], in, k[1]);
			c = sel_alpha(2, sel[6], in, k[2]);
			d = sel_alpha(3, sel[7], in, k[3]);
		}
		c = sext(c, 9);
		a = special_expand(a);
		b = special_expand(b);
		d = special_expand(d);
		out[ch] = ((__mul24(a - b, c) + 0x80) >> 8) + d;
	}
}

DEV void dither_coefficients(int x, int y, int mode_rgb, int mode_alpha, uint32_t noise, int &rgb_dith, int &alpha_dith)
{
	// dither.h:26-68: magic-square and Bayer matrices packed 3 bits per entry.
	const uint64_t magic = 0x1ull * 0 | 06ull << 3 | 01ull << 6 | 07ull << 9 | 04ull << 12 | 02ull << 15 | 05ull << 18 |
	                       03ull << 21 | 03ull << 24 | 05ull << 27 | 02ull << 30 | 04ull << 33 | 07ull << 36 |
	                       01ull << 39 | 06ull << 42 | 00ull << 45;
	const uint64_t bayer = 00ull | 04ull << 3 | 01ull << 6 | 05ull << 9 | 04ull << 12 | 00ull << 15 | 05ull << 18 |
	                       01ull << 21 | 03ull << 24 | 07ull << 27 | 02ull << 30 | 06ull << 33 | 07ull << 36 |
	                       03ull << 39 | 06ull << 42 | 02ull << 45;
	const int idx = (y & 3) * 4 + (x & 3);
	// The matrix entry every mode that uses one reads: Bayer for the odd RGB
	// modes, magic square for the even ones (the alpha pattern follows the
	// RGB matrix, and in the noise and off RGB modes the same parity rule).
	// Selects on the primitive's modes rather than branches: the modes are
	// uniform, and branches on them cost scalar instructions per pixel.
	const int m = (int)(((mode_rgb & 1) ? bayer : magic) >> (3 * idx)) & 7;
	rgb_dith = mode_rgb < 2 ? m * 0x49 : (mode_rgb == 2 ? (int)(noise & 0x1ff) : 0);
	alpha_dith = mode_alpha == 3 ? 0 : (mode_alpha == 2 ? (int)(noise & 7) : (mode_alpha == 1 ? (~m & 7) : m));
}

// ---- depth / blend: z_encode.h, depth_test.h, blender.h, coverage.h --------
DEV int32_t z_decompress(uint32_t z)
{
	int32_t e = (int32_t)(z >> 11), man = (int32_t)(z & 0x7ff);
	return (man << imax(6 - e, 0)) + (0x40000 - (0x40000 >> e));
}

DEV uint32_t z_compress(int32_t z)
{
	int32_t e = iclamp(17 - find_msb(imax(0x3ffff - z, 1)), 0, 7);
	int32_t man = (z >> imax(6 - e, 0)) & 0x7ff;
	return (uint32_t)((e << 11) + man) & 0xffff;
}

// The RasterPrim fields ylo .. variant, which the tile loop tests for every
// primitive before it shades: read together with one scalar load instead of
// one dependent load per field.
struct RasterPrimHead {
	int32_t ylo, yhi;
	uint32_t span_offset, tmem_instance, prim_index, state_index;
	int32_t dz, dz_compressed;
	uint32_t ras_flags, dither, db_flags, coverage_mode, z_mode, fill_color, min_lod, variant;
};
static_assert(offsetof(RasterPrim, variant) - offsetof(RasterPrim, ylo) == offsetof(RasterPrimHead, variant),
              "RasterPrimHead mirrors RasterPrim");

struct Pixel {
	int32_t c[4];
	uint32_t depth, dz;
	// PIXEL_COLOR_DIRTY | PIXEL_DEPTH_DIRTY, kept in a vector register
	// (raster_prim): as booleans the two flags live in scalar lane masks,
	// merged with scalar instructions at the end of every divergent branch
	// that sets them.
	uint32_t dirty;
};
#define PIXEL_COLOR_DIRTY 1u
#define PIXEL_DEPTH_DIRTY 2u

// The Z comparison terms of depth_test.h:41-144 for a fragment at depth z
// with primitive dz against the pixel's stored depth and dz code; shared by
// the depth test and the early rejection in shade_pixel.
struct DepthCompare {
	int32_t memory_z, comb_ip, comb;
	bool farther, nearer, front, max_z;
};

DEV DepthCompare depth_compare(uint32_t depth, uint32_t dz_code, int32_t z, int32_t prim_dz)
{
	DepthCompare d;
	d.memory_z = z_decompress(depth);
	int32_t memory_dz = 1 << dz_code;
	const int32_t precision = ((int32_t)depth >> 11) & 0xf;
	bool coplanar = false;
	if (precision < 3) {
		if (memory_dz != 0x8000)
			memory_dz = imax(memory_dz << 1, 16 >> precision);
		else {
			coplanar = true;
			memory_dz = 0xffff;
		}
	}
	int32_t comb = prim_dz | memory_dz;
	if (comb != 0) comb = 1 << find_msb(comb);
	d.comb_ip = comb;
	d.comb = comb << 3;
	d.farther = coplanar || ((z + d.comb) >= d.memory_z);
	d.max_z = d.memory_z == 0x3ffff;
	d.front = z < d.memory_z;
	d.nearer = coplanar || ((z - d.comb) <= d.memory_z);
	return d;
}

// True when the depth test fails whatever the fragment's coverage (which
// alpha may still change through coverage-times-alpha or alpha-coverage
// select): with every Z mode the failing terms do not involve coverage.  A
// rejected fragment changes nothing, so its shading can be skipped.
DEV bool depth_rejects(const DepthCompare &d, uint32_t z_mode)
{
	switch (z_mode) {
	case 0:
	case 1: return !d.max_z && !d.front && !d.nearer;
	case 2: return !d.front && !d.max_z;
	default: return !(d.farther && d.nearer && !d.max_z);
	}
}

DEV void write_color(Pixel &p, uint32_t fmt, int32_t r, int32_t g, int32_t b, int32_t a)
{
	p.c[0] = r;
	p.c[1] = g;
	p.c[2] = b;
	if (fmt != FB_I4) p.c[3] = a;
	p.dirty |= PIXEL_COLOR_DIRTY;
}

DEV void alias_color_to_depth(Pixel &p, uint32_t fmt)
{
	if (fmt != FB_RGBA5551 && fmt != FB_IA88) return;
	// Both forms are computed and selected, then stored once each: stores of
	// depth and dz from two branches get merged into one store at a variable
	// field offset, which keeps the whole Pixel in scratch memory.
	const uint32_t dz5551 = (((uint32_t)p.c[3] >> 3) | ((uint32_t)p.c[2] & 8)) & 0xff;
	const uint32_t depth5551 =
	    ((((uint32_t)p.c[0] & 0xf8) << 6) | (((uint32_t)p.c[1] & 0xf8) << 1) | (((uint32_t)p.c[2] & 0xf8) >> 4)) & 0xffff;
	const uint32_t w = (((uint32_t)p.c[0] & 0xff) << 8) | ((uint32_t)p.c[3] & 0xff);
	const uint32_t depth88 = (w >> 2) & 0xffff, dz88 = (((w & 3) << 2) | ((w & 1) * 3)) & 0xff;
	const bool is5551 = fmt == FB_RGBA5551;
	p.depth = is5551 ? depth5551 : depth88;
	p.dz = is5551 ? dz5551 : dz88;
}

DEV void alias_depth_to_color(Pixel &p, uint32_t fmt)
{
	uint32_t w = (p.depth << 4) | p.dz;
	if (fmt == FB_RGBA5551) {
		p.c[0] = (int32_t)((w >> 10) & 0xf8);
		p.c[1] = (int32_t)((w >> 5) & 0xf8);
		p.c[2] = (int32_t)(w & 0xf8);
		p.c[3] = (int32_t)((w & 7) << 5);
	} else if (fmt == FB_IA88) {
		p.c[0] = (int32_t)((w >> 10) & 0xff);
		p.c[3] = (int32_t)((w >> 2) & 0xff);
	}
	p.dirty |= PIXEL_COLOR_DIRTY;
}

template <typename Modes>
DEV void blender(const AS_GLOBAL uint8_t *lut, const int32_t pixel[4], const int32_t memory[4], const AS_CONST uint8_t *fog,
                 const AS_CONST uint8_t *blendc, int32_t shade_alpha, const Modes &modes, bool force_blend, bool blend_en,
                 bool color_on_cvg, bool cvg_wrap, int shift0, int shift1, bool final_cycle, int32_t out[3])
{
	int32_t src1[3], src0[3];
#pragma unroll
	for (int i = 0; i < 3; i++) {
		switch (modes[2]) {
		case 0: src1[i] = pixel[i]; break;
		case 1: src1[i] = memory[i]; break;
		case 2: src1[i] = blendc[i]; break;
		default: src1[i] = fog[i]; break;
		}
		switch (modes[0]) {
		case 0: src0[i] = pixel[i]; break;
		case 1: src0[i] = memory[i]; break;
		case 2: src0[i] = blendc[i]; break;
		default: src0[i] = fog[i]; break;
		}
	}
	if (final_cycle && color_on_cvg && !cvg_wrap) {
		out[0] = src1[0];
		out[1] = src1[1];
		out[2] = src1[2];
		return;
	}
	if (final_cycle && (!blend_en || (modes[1] == 0 && modes[3] == 0 && pixel[3] == 0xff))) {
		out[0] = src0[0];
		out[1] = src0[1];
		out[2] = src0[2];
		return;
	}
	int32_t a0, a1;
	switch (modes[1]) {
	case 0: a0 = pixel[3]; break;
	case 1: a0 = fog[3]; break;
	case 2: a0 = shade_alpha; break;
	default: a0 = 0; break;
	}
	switch (modes[3]) {
	case 0: a1 = ~a0 & 0xff; break;
	case 1: a1 = memory[3]; break;
	case 2: a1 = 0xff; break;
	default: a1 = 0; break;
	}
	a0 >>= 3;
	a1 >>= 3;
	if (modes[3] == 1) {
		a0 = (a0 >> shift0) & 0x3c;
		a1 = (a1 >> shift1) | 3;
	}
	if (!final_cycle || force_blend) {
#pragma unroll
		for (int i = 0; i < 3; i++) out[i] = ((__mul24(src0[i], a0) + __mul24(src1[i], a1 + 1)) >> 5) & 0xff;
	} else {
		const uint32_t sum = (uint32_t)((a0 >> 2) + (a1 >> 2) + 1);
#pragma unroll
		for (int i = 0; i < 3; i++) {
			uint32_t b = (uint32_t)((__mul24(src0[i], a0) + __mul24(src1[i], a1 + 1)) >> 2) & 0x7ff;
			uint32_t idx = (sum << 11) | b;
			out[i] = idx < 0x8000 ? lut[idx] : 0;
		}
	}
}

DEV void load_color(const Mem &m, uint32_t fmt, uint32_t index, Pixel &p)
{
	switch (fmt) {
	case FB_I4:
	case FB_I8: {
		uint32_t i = index & m.mask8;
		uint32_t w = m.rdram[i ^ 3];
		p.c[0] = p.c[1] = p.c[2] = (int32_t)w;
		p.c[3] = (int32_t)m.hid(i >> 1);
		break;
	}
	case FB_RGBA5551: {
		uint32_t w = m.rd16i(index);
		p.c[0] = (int32_t)((w >> 8) & 0xf8);
		p.c[1] = (int32_t)((w >> 3) & 0xf8);
		p.c[2] = (int32_t)((w << 2) & 0xf8);
		p.c[3] = (int32_t)(((m.hid(index) << 5) | ((w & 1) << 7)) & 0xff);
		break;
	}
	case FB_IA88: {
		uint32_t w = m.rd16i(index);
		p.c[0] = p.c[1] = p.c[2] = (int32_t)((w >> 8) & 0xff);
		p.c[3] = (int32_t)(w & 0xff);
		break;
	}
	default: {
		uint32_t w = m.rd32i(index);
		p.c[0] = (int32_t)(w >> 24);
		p.c[1] = (int32_t)((w >> 16) & 0xff);
		p.c[2] = (int32_t)((w >> 8) & 0xff);
		p.c[3] = (int32_t)(w & 0xff);
		break;
	}
	}
}

DEV void store_color(const Mem &m, uint32_t fmt, uint32_t index, const Pixel &p)
{
	switch (fmt) {
	case FB_I4: {
		uint32_t i = index & m.mask8;
		m.rdram[i ^ 3] = 0;
		if (i & 1) m.set_hid(i >> 1, (uint32_t)p.c[3]);
		break;
	}
	case FB_I8: {
		uint32_t i = index & m.mask8;
		uint32_t col = (uint32_t)((i & 1) ? p.c[1] : p.c[0]) & 0xff;
		m.rdram[i ^ 3] = (uint8_t)col;
		if (i & 1) m.set_hid(i >> 1, (col & 1) * 3);
		break;
	}
	case FB_RGBA5551: {
		uint32_t cov = ((uint32_t)p.c[3] & 0xff) >> 5;
		uint32_t w = (((uint32_t)p.c[0] & 0xf8) << 8) | (((uint32_t)p.c[1] & 0xf8) << 3) | (((uint32_t)p.c[2] & 0xf8) >> 2) |
		             (cov >> 2);
		m.wr16i(index, w);
		m.set_hid(index, cov & 3);
		break;
	}
	case FB_IA88: {
		uint32_t w = (((uint32_t)p.c[0] & 0xff) << 8) | ((uint32_t)p.c[3] & 0xff);
		m.wr16i(index, w);
		m.set_hid(index, ((uint32_t)p.c[3] & 1) * 3);
		break;
	}
	default: {
		uint32_t w = (((uint32_t)p.c[0] & 0xff) << 24) | (((uint32_t)p.c[1] & 0xff) << 16) | (((uint32_t)p.c[2] & 0xff) << 8) |
		             ((uint32_t)p.c[3] & 0xff);
		m.wr32i(index, w);
		uint32_t i = index & (m.mask8 >> 2);
		m.set_hid(2 * i, ((uint32_t)p.c[1] & 1) * 3);
		m.set_hid(2 * i + 1, ((uint32_t)p.c[3] & 1) * 3);
		break;
	}
	}
}

struct Shaded {
	int32_t combined[4];
	int32_t z_dith;
	int32_t coverage;
	int32_t shade_alpha;
};

#define COVERAGE_FILL 0x40
#define COVERAGE_COPY 0x20

// shading.h:38-359 for one pixel of one primitive (no upscaling).
// A raster code variant: flag bits in KNOWN are the compile-time constants in
// VALUE, everything else comes from the primitive (see RDP_RASTER_VARIANTS).
template <uint32_t KNOWN, uint32_t VALUE, int COMB, int TF, int DBP>
struct RasterVariant {
	static constexpr int comb = COMB;
	static constexpr int texfmt = TF;
	static constexpr int dbp = DBP;
	static MDEV bool ras(uint32_t f, uint32_t bit) { return (KNOWN & bit) ? ((VALUE & bit) != 0) : ((f & bit) != 0); }
	// Known to be a fill or copy primitive, or to sample texel 0.
	static constexpr bool fill_or_copy = (KNOWN & VALUE & (RAS_FILL | RAS_COPY)) != 0;
	static constexpr bool textured = (KNOWN & VALUE & RAS_USES_TEXEL0) != 0;
	// Depth/blend state: the profile's constants, or the primitive's values.
	static MDEV bool db(uint32_t f, uint32_t bit)
	{
		return DBP != DBP_ANY ? (kDepthBlendProfiles[DBP].db_flags & bit) != 0 : (f & bit) != 0;
	}
	static MDEV uint32_t z_mode(uint32_t v) { return DBP != DBP_ANY ? kDepthBlendProfiles[DBP].z_mode : v; }
	static MDEV uint32_t coverage_mode(uint32_t v) { return DBP != DBP_ANY ? kDepthBlendProfiles[DBP].coverage_mode : v; }
};
using GenericVariant = RasterVariant<0u, 0u, (int)COMB_ANY, TEXFMT_ANY, (int)DBP_ANY>;

// Blender selector sources, like the combiner's: the primitive's, or a
// profile's constants.
struct RuntimeBlend {
	const AS_CONST uint8_t *p;
	MDEV int operator[](int i) const { return p[i]; }
};
template <int DBP, int CYCLE>
struct KnownBlend {
	MDEV int operator[](int i) const { return kDepthBlendProfiles[DBP].blend[CYCLE][i]; }
};

template <class V>
// L: SCALING_LOG2 of the pass (0 natively, where every term below that
// carries it reduces to the native form at compile time).
DEV bool shade_pixel(const AS_CONST RasterPrim *ps, const RasterPrimHead &h, const AS_CONST TileDesc *tiles,
                     const Tex &tx, uint32_t fb_fmt, const SpanSetup &span, const AS_GLOBAL SpanSetup *next_span, int x,
                     int y, const Pixel &px,
                     Shaded &out, int L)
{
	const uint32_t f = h.ras_flags;
	const bool tlut = V::ras(f, RAS_TLUT), tlut_type = V::ras(f, RAS_TLUT_TYPE), sample_quad = V::ras(f, RAS_SAMPLE_QUAD);
	const bool cvg_times_alpha = V::ras(f, RAS_CVG_TIMES_ALPHA), alpha_cvg_select = V::ras(f, RAS_ALPHA_CVG_SELECT);
	const bool perspective = V::ras(f, RAS_PERSPECTIVE), aa = V::ras(f, RAS_AA), multi = V::ras(f, RAS_MULTI_CYCLE);
	const bool interlace = V::ras(f, RAS_INTERLACE_FIELD), alpha_test = V::ras(f, RAS_ALPHA_TEST);
	const bool mid_texel = V::ras(f, RAS_MID_TEXEL), convert_one = V::ras(f, RAS_CONVERT_ONE);
	const bool bilerp0 = V::ras(f, RAS_BILERP0), bilerp1 = V::ras(f, RAS_BILERP1);
	bool uses_texel1 = V::ras(f, RAS_USES_TEXEL1);

	uint32_t noise = 0;
	if (V::ras(f, RAS_NEED_NOISE))
		noise = reseed_noise((uint32_t)x, (uint32_t)y, h.prim_index);

	const bool flip = (ps->flags & TRI_FLIP) != 0;
	const int32_t factors[4] = { ps->factors[0], ps->factors[1], ps->factors[2], ps->factors[3] };

	if (V::ras(f, RAS_COPY)) {
		if (!span.valid || x < span.start_x || x > span.end_x) return false;
		// interpolate_st_copy: the samples of a pixel repeat its texel.
		int dxp = (flip ? (x - span.start_x) : (span.end_x - x)) >> L;
		uint32_t dx_mask, dx_shift, fb_size;
		switch (fb_fmt) {
		case FB_I4: fb_size = 0; dx_mask = 0; dx_shift = 0; break;
		case FB_I8: fb_size = 1; dx_mask = ~7u; dx_shift = 3; break;
		case FB_RGBA5551:
		case FB_IA88: fb_size = 2; dx_mask = ~3u; dx_shift = 2; break;
		default: fb_size = 4; dx_mask = ~1u; dx_shift = 1; break;
		}
		int s_offset = dxp - (dxp & (int)dx_mask);
		int lerp_dx = (dxp >> dx_shift) * (flip ? 1 : -1);
		int32_t ss, tt;
		bool ovf = false;
		stw_to_st(span.stzw[0] + (ps->dstzw_dx[0] & ~0x1f) * lerp_dx, span.stzw[1] + (ps->dstzw_dx[1] & ~0x1f) * lerp_dx,
		          span.stzw[3] + (ps->dstzw_dx[3] & ~0x1f) * lerp_dx, perspective, ss, tt, ovf);
		const TileDesc t0 = ldc(tiles + (ps->tile & 7));
		int s = shift_coord(ss, (int)t0.slo, t0.shift_s) >> 5;
		int t = shift_coord(tt, (int)t0.tlo, t0.shift_t) >> 5;
		int texel;
		if (fb_size == 0)
			texel = 0;
		else if (fb_size == 1)
			texel = (sample_copy_word(tx, t0, s, t, s_offset >> 1, tlut) >> (8 - 8 * (s_offset & 1))) & 0xff;
		else
			texel = sample_copy_word(tx, t0, s, t, s_offset, tlut);
		out.z_dith = texel;
		out.coverage = COVERAGE_COPY;
		if (alpha_test && fb_size == 2 && (texel & 1) == 0) return false;
		return true;
	} else if (V::ras(f, RAS_FILL)) {
		out.coverage = COVERAGE_FILL;
		return span.valid && x >= span.start_x && x <= span.end_x;
	}

	const int coverage = compute_coverage(span, x);
	if (!span.valid || coverage == 0) return false;
	int cvg_count = __popc(coverage);
	if (!aa && (coverage & 1) == 0) return false;

	const int dx = x - span.base_x;
	const int dir = flip ? 1 : -1;
	const int first = find_lsb((uint32_t)coverage);
	const int yoff = first >> 1, xoff = ((first & 1) << 1) + (yoff & 1);

	int32_t z;
	{
		// Whole native pixels step by the full gradient, samples by the
		// gradient >> L (interpolation.h:135).
		int32_t zz = span.stzw[2] + ps->dstzw_dx[2] * (dx >> L) + (ps->dstzw_dx[2] >> L) * (dx & ((1 << L) - 1));
		int32_t sn = (zz >> 10) << (2 + L);
		sn += xoff * (ps->dstzw_dx[2] >> 10) + yoff * (ps->dstzw_dy[2] >> 10);
		z = clamp_z(sn >> (5 + L));
	}
	// Early depth rejection: shading, texturing, LOD and the combiner only
	// feed a fragment the depth test is known to discard.
	if (V::db(h.db_flags, DB_DEPTH_TEST) && depth_rejects(depth_compare(px.depth, px.dz, z, h.dz), V::z_mode(h.z_mode)))
		return false;

	int32_t shade[4];
#pragma unroll
	for (int i = 0; i < 4; i++) {
		int32_t v = span.rgba[i] + ((ps->drgba_dx[i] & ~0x1f) >> L) * dx;
		int32_t sn = (v >> 14) << (2 + L);
		sn += xoff * (ps->drgba_dx[i] >> 14) + yoff * (ps->drgba_dy[i] >> 14);
		shade[i] = clamp_9bit(sn >> (4 + L));
	}

	const bool uses_lod = V::ras(f, RAS_USES_LOD);
	int32_t st0[2], stdx[2] = { 0, 0 }, stdy[2] = { 0, 0 };
	bool povf = false;
	{
		int32_t s = span.stzw[0] + ((ps->dstzw_dx[0] & ~0x1f) >> L) * dx;
		int32_t t = span.stzw[1] + ((ps->dstzw_dx[1] & ~0x1f) >> L) * dx;
		int32_t w = span.stzw[3] + ((ps->dstzw_dx[3] & ~0x1f) >> L) * dx;
		stw_to_st(s, t, w, perspective, st0[0], st0[1], povf);
		if (uses_lod) {
			// shading.h:192-198: with the native-LOD quirk the LOD neighbours
			// are a native pixel away; interpolation.h:99-105.
			const int ydir = (L && (ps->flags & TRI_NATIVE_LOD)) ? (1 << L) : 1;
			const int tdir = dir * ydir;
			stw_to_st(s + tdir * ((ps->dstzw_dx[0] & ~0x1f) >> L), t + tdir * ((ps->dstzw_dx[1] & ~0x1f) >> L),
			          w + tdir * ((ps->dstzw_dx[3] & ~0x1f) >> L), perspective, stdx[0], stdx[1], povf);
			stw_to_st(s + ydir * ((ps->dstzw_dy[0] & ~0x7fff) >> L), t + ydir * ((ps->dstzw_dy[1] & ~0x7fff) >> L),
			          w + ydir * ((ps->dstzw_dy[3] & ~0x7fff) >> L), perspective, stdy[0], stdy[1], povf);
		}
	}

	uint32_t tile0 = ps->tile & 7, tile1 = (tile0 + 1) & 7;
	int32_t lod_frac = 0;
	if (uses_lod)
		compute_lod(tile0, tile1, lod_frac, ps->tile >> 3, (int)h.min_lod, st0, stdx, stdy, povf, V::ras(f, RAS_TEX_LOD),
		            V::ras(f, RAS_SHARPEN_LOD), V::ras(f, RAS_DETAIL_LOD));

	int32_t texel0[4] = { 0, 0, 0, 0 }, texel1[4] = { 0, 0, 0, 0 };
	if (V::ras(f, RAS_USES_TEXEL0))
	{
		const TileDesc td = ldc(tiles + tile0);
		sample_texture<V::texfmt>(tx, td, st0[0], st0[1], tlut, tlut_type, sample_quad, mid_texel, false, bilerp0, factors,
		               texel1, texel0);
	}

	if (V::ras(f, RAS_USES_PIPELINED_TEXEL1)) {
		// The next span's record is read only for the last pixel of a long
		// span, the one case whose LOD texel comes from it.
		const bool long_span = span.lodlength >= 8;
		const bool end_span = x == (flip ? span.end_x : span.start_x);
		bool from_next = false;
		int32_t s = 0, t = 0, w = 0;
		if (end_span && long_span && next_span->valid) {
			s = next_span->stzw[0] >> 16;
			t = next_span->stzw[1] >> 16;
			w = next_span->stzw[3] >> 16;
			from_next = true;
		}
		if (from_next) {
			if (perspective) {
				bool o2 = false;
				perspective_divide(s, t, w, st0[0], st0[1], o2);
			} else {
				st0[0] = s;
				st0[1] = t;
			}
		} else {
			bool o2 = false;
			const int d2 = dx + dir * (1 << L);
			stw_to_st(span.stzw[0] + ((ps->dstzw_dx[0] & ~0x1f) >> L) * d2,
			          span.stzw[1] + ((ps->dstzw_dx[1] & ~0x1f) >> L) * d2,
			          span.stzw[3] + ((ps->dstzw_dx[3] & ~0x1f) >> L) * d2, perspective, st0[0], st0[1], o2);
		}
		tile1 = tile0;
		uses_texel1 = true;
	}

	if (uses_texel1) {
		if (convert_one && !bilerp1)
			convert_factors(texel0, factors, texel1);
		else {
			const TileDesc td = ldc(tiles + tile1);
			sample_texture<V::texfmt>(tx, td, st0[0], st0[1], tlut, tlut_type, sample_quad, mid_texel, convert_one, bilerp1,
			               factors, texel0, texel1);
		}
	}

	int rgb_dith, alpha_dith;
	dither_coefficients(x, y >> (interlace ? 1 : 0), (int)(h.dither >> 2), (int)(h.dither & 3), noise, rgb_dith,
	                    alpha_dith);

	CombIn ci;
#pragma unroll
	for (int i = 0; i < 4; i++) {
		ci.shade[i] = shade[i];
		ci.texel0[i] = texel0[i];
		ci.texel1[i] = texel1[i];
		ci.combined[i] = 0;
	}
	ci.lod_frac = lod_frac;
	ci.noise = (int32_t)(((noise & 7) << 6) | 0x20);

	int32_t combined[4];
	int32_t alpha_ref;
	if (multi) {
		int32_t c0[4];
		if constexpr (V::comb != COMB_ANY)
			combine(ci, KnownSel<V::comb, 0>{}, ps->constants[0], c0);
		else
			combine(ci, RuntimeSel{ ps->comb[0] }, ps->constants[0], c0);
		if (alpha_test) {
			int32_t cl = clamp_9bit(c0[3]);
			int32_t ex = cl + ((cl + 1) >> 8);
			if (alpha_cvg_select)
				ex = cvg_times_alpha ? ((ex * cvg_count + 4) >> 3) : (cvg_count << 5);
			else
				ex += alpha_dith;
			alpha_ref = iclamp(ex, 0, 0xff);
		} else
			alpha_ref = 0;
#pragma unroll
		for (int i = 0; i < 4; i++) {
			ci.combined[i] = c0[i];
			int32_t tmp = ci.texel0[i];
			ci.texel0[i] = ci.texel1[i];
			ci.texel1[i] = tmp;
		}
		if (V::ras(f, RAS_NEED_NOISE_DUAL)) {
			uint32_t n2 = reseed_noise((uint32_t)(x + 1023), (uint32_t)(y + 7), h.prim_index + 11);
			ci.noise = (int32_t)(((n2 & 7) << 6) | 0x20);
		}
		if constexpr (V::comb != COMB_ANY)
			combine(ci, KnownSel<V::comb, 1>{}, ps->constants[1], combined);
		else
			combine(ci, RuntimeSel{ ps->comb[1] }, ps->constants[1], combined);
	} else if constexpr (V::comb != COMB_ANY) {
		combine(ci, KnownSel<V::comb, 1>{}, ps->constants[1], combined);
	} else {
		combine(ci, RuntimeSel{ ps->comb[1] }, ps->constants[1], combined);
	}
	// combiner_cycle1 tail (combiner.h:256-281)
#pragma unroll
	for (int i = 0; i < 4; i++) combined[i] = clamp_9bit(combined[i]);
	{
		int32_t ex = combined[3] + ((combined[3] + 1) >> 8);
		int32_t mod;
		if (cvg_times_alpha) {
			mod = (ex * cvg_count + 4) >> 3;
			cvg_count = mod >> 5;
		} else
			mod = cvg_count << 5;
		ex = alpha_cvg_select ? mod : ex + alpha_dith;
		combined[3] = iclamp(ex, 0, 0xff);
	}
	if (!multi) alpha_ref = combined[3];

	if (aa && cvg_count == 0) return false;
	if (alpha_test) {
		int32_t thr = V::ras(f, RAS_ALPHA_TEST_DITHER) ? (int32_t)(noise & 0xff) : ps->blendc[3];
		if (alpha_ref < thr) return false;
	}
#pragma unroll
	for (int i = 0; i < 4; i++) out.combined[i] = combined[i];
	out.z_dith = (z << 9) | rgb_dith;
	out.coverage = cvg_count;
	out.shade_alpha = imin(shade[3] + alpha_dith, 0xff);
	return true;
}

// memory_interfacing.h:488-571
template <class V>
DEV void depth_blend(const AS_GLOBAL uint8_t *lut, const AS_CONST RasterPrim *ps, const RasterPrimHead &h, uint32_t fmt,
                     bool alias, Pixel &p,
                     const Shaded &sh)
{
	const int32_t z = sh.z_dith >> 9;
	const int dith = sh.z_dith & 0x1ff;
	int cvg = sh.coverage;
	const uint32_t f = h.db_flags;
	const bool force_blend = V::db(f, DB_FORCE_BLEND), z_compare = V::db(f, DB_DEPTH_TEST);
	const bool z_update = V::db(f, DB_DEPTH_UPDATE), image_read = V::db(f, DB_IMAGE_READ);
	const bool color_on_cvg = V::db(f, DB_COLOR_ON_CVG), multi = V::db(f, DB_MULTI_CYCLE);
	const bool aa = V::db(f, DB_AA), dither_en = V::db(f, DB_DITHER);

	int32_t memory[4];
	int32_t mem_cov = image_read ? (p.c[3] & 0xe0) : 0xe0;
	switch (fmt) {
	case FB_I4: memory[0] = memory[1] = memory[2] = 0; mem_cov = 0xe0; break;
	case FB_I8: memory[0] = memory[1] = memory[2] = p.c[0] & 0xff; mem_cov = 0xe0; break;
	case FB_RGBA5551:
		memory[0] = p.c[0] & 0xf8;
		memory[1] = p.c[1] & 0xf8;
		memory[2] = p.c[2] & 0xf8;
		break;
	case FB_IA88: memory[0] = memory[1] = memory[2] = p.c[0] & 0xff; break;
	default:
		memory[0] = p.c[0] & 0xff;
		memory[1] = p.c[1] & 0xff;
		memory[2] = p.c[2] & 0xff;
		break;
	}
	memory[3] = mem_cov;
	const int memory_coverage = mem_cov >> 5;

	bool blend_en, cvg_wrap, pass;
	int shift0, shift1;
	if (z_compare) {
		const DepthCompare d = depth_compare(p.depth, p.dz, z, h.dz);
		shift0 = iclamp(h.dz_compressed - (int32_t)p.dz, 0, 4);
		shift1 = iclamp((int32_t)p.dz - h.dz_compressed, 0, 4);
		const bool overflow = (cvg + memory_coverage) >= 8;
		blend_en = force_blend || (!overflow && aa && d.farther);
		cvg_wrap = overflow;
		switch (V::z_mode(h.z_mode)) {
		case 0: pass = d.max_z || (overflow ? d.front : d.nearer); break;
		case 1:
			if (!d.front || !d.farther || !overflow)
				pass = d.max_z || (overflow ? d.front : d.nearer);
			else {
				int32_t c = imax(find_msb(d.comb_ip & 0xffff), 0);
				int32_t coeff = ((d.memory_z >> c) - (z >> c)) & 0xf;
				cvg = imin((coeff * cvg) >> 3, 8);
				pass = true;
			}
			break;
		case 2: pass = d.front || d.max_z; break;
		default: pass = d.farther && d.nearer && !d.max_z; break;
		}
	} else {
		shift0 = 0;
		shift1 = imin(0xf - h.dz_compressed, 4);
		const bool overflow = (cvg + memory_coverage) >= 8;
		blend_en = force_blend || (!overflow && aa);
		cvg_wrap = overflow;
		pass = true;
	}

	if (pass && (!aa || cvg != 0)) {
		int32_t pixel[4] = { sh.combined[0], sh.combined[1], sh.combined[2], sh.combined[3] };
		int32_t rgb[3];
		if (multi) {
			if constexpr (V::dbp != DBP_ANY)
				blender(lut, pixel, memory, ps->fog, ps->blendc, sh.shade_alpha, KnownBlend<V::dbp, 0>{}, force_blend,
				        blend_en, color_on_cvg, cvg_wrap, shift0, shift1, false, rgb);
			else
				blender(lut, pixel, memory, ps->fog, ps->blendc, sh.shade_alpha, RuntimeBlend{ ps->blend[0] }, force_blend,
				        blend_en, color_on_cvg, cvg_wrap, shift0, shift1, false, rgb);
			pixel[0] = rgb[0];
			pixel[1] = rgb[1];
			pixel[2] = rgb[2];
		}
		// The final cycle uses cycle 1's selectors in two-cycle mode and
		// cycle 0's in one-cycle mode.
		if constexpr (V::dbp != DBP_ANY) {
			if (multi)
				blender(lut, pixel, memory, ps->fog, ps->blendc, sh.shade_alpha, KnownBlend<V::dbp, 1>{}, force_blend,
				        blend_en, color_on_cvg, cvg_wrap, shift0, shift1, true, rgb);
			else
				blender(lut, pixel, memory, ps->fog, ps->blendc, sh.shade_alpha, KnownBlend<V::dbp, 0>{}, force_blend,
				        blend_en, color_on_cvg, cvg_wrap, shift0, shift1, true, rgb);
		} else
			blender(lut, pixel, memory, ps->fog, ps->blendc, sh.shade_alpha, RuntimeBlend{ ps->blend[multi ? 1 : 0] },
			        force_blend, blend_en, color_on_cvg, cvg_wrap, shift0, shift1, true, rgb);
		if (dither_en) {
#pragma unroll
			for (int i = 0; i < 3; i++) {
				int d = (dith >> (3 * i)) & 7;
				int32_t orig = rgb[i];
				int32_t v = orig > 247 ? 255 : (orig & 0xf8) + 8;
				int32_t sign = (d - (orig & 7)) >> 31;
				rgb[i] = (orig + ((v - orig) & sign)) & 0xff;
			}
		}
		int new_cov;
		switch (V::coverage_mode(h.coverage_mode)) {
		case 0: new_cov = blend_en ? imin(7, memory_coverage + cvg) : ((cvg - 1) & 7); break;
		case 1: new_cov = (cvg + memory_coverage) & 7; break;
		case 2: new_cov = 7; break;
		default: new_cov = memory_coverage; break;
		}
		write_color(p, fmt, rgb[0], rgb[1], rgb[2], new_cov << 5);
		if (z_update) {
			p.depth = z_compress(z);
			p.dz = (uint32_t)h.dz_compressed;
			p.dirty |= PIXEL_DEPTH_DIRTY;
			if (alias) alias_depth_to_color(p, fmt);
		} else if (alias)
			alias_color_to_depth(p, fmt);
	}
}

DEV void fill_pixel(Pixel &p, uint32_t fmt, bool alias, uint32_t color_index, uint32_t col)
{
	switch (fmt) {
	case FB_RGBA8888:
		write_color(p, fmt, (col >> 24) & 0xff, (col >> 16) & 0xff, (col >> 8) & 0xff, col & 0xff);
		break;
	case FB_RGBA5551:
		col >>= ((color_index & 1) ^ 1) * 16;
		write_color(p, fmt, (col >> 8) & 0xf8, (col >> 3) & 0xf8, (col << 2) & 0xf8, (col & 1) * 0xe0);
		break;
	case FB_IA88: {
		col = (col >> (((color_index & 1) ^ 1) * 16)) & 0xffff;
		int32_t r = (int32_t)((col >> 8) & 0xff);
		write_color(p, fmt, r, r, r, (int32_t)(col & 0xff));
		break;
	}
	case FB_I8: {
		int32_t v = (int32_t)((col >> (((color_index & 3) ^ 3) * 8)) & 0xff);
		write_color(p, fmt, v, v, v, v);
		break;
	}
	default: break;
	}
	if (alias) alias_color_to_depth(p, fmt);
}

DEV void copy_pixel(Pixel &p, uint32_t fmt, bool alias, uint32_t word)
{
	switch (fmt) {
	case FB_I4:
		p.c[0] = p.c[1] = p.c[2] = p.c[3] = 0;
		p.dirty |= PIXEL_COLOR_DIRTY;
		break;
	case FB_I8: {
		int32_t v = (int32_t)(word & 0xff);
		write_color(p, fmt, v, v, v, v);
		break;
	}
	case FB_RGBA5551:
		write_color(p, fmt, (word >> 8) & 0xf8, (word >> 3) & 0xf8, (word << 2) & 0xf8, (word & 1) * 0xe0);
		break;
	default: break;
	}
	if (alias) alias_color_to_depth(p, fmt);
}

// One primitive at one pixel: span lookup, shading, then fill / copy /
// depth test + blend into the pixel's registers.
template <class V>
DEV void raster_prim(const AS_CONST RasterPrim *ps, const RasterPrimHead &h, const AS_CONST DrawState *states,
                     const AS_GLOBAL SpanSetup *spans,
                     const Tex &tex, const AS_GLOBAL uint8_t *lut, uint32_t fmt, bool alias, uint32_t ci, int x, int y,
                     Pixel &px, int L = 0)
{
	// shading.h:45-52: a native-resolution primitive shades every sample as
	// its pixel's first one.
	if (L && (ps->flags & TRI_DISABLE_UPSCALING)) {
		x &= ~((1 << L) - 1);
		y &= ~((1 << L) - 1);
	}
	const AS_GLOBAL SpanSetup *sp = spans + (h.span_offset + (uint32_t)(y - h.ylo));
	// The span's valid flag is tested with the first coverage test in
	// shade_pixel, so every field the test needs is fetched at once.
	const SpanSetup span = ldg16(sp);
	Shaded sh;
	if constexpr (!V::fill_or_copy) {
		// Fetch the shade and depth gradients (and the texture ones of a
		// textured variant) now, beside the span load.
		int32_t gr[10];
#pragma unroll
		for (int i = 0; i < 4; i++) {
			gr[i] = ps->drgba_dx[i];
			gr[4 + i] = ps->drgba_dy[i];
		}
		gr[8] = ps->dstzw_dx[2];
		gr[9] = ps->dstzw_dy[2];
#pragma unroll
		for (int i = 0; i < 10; i++) asm volatile("" ::"s"(gr[i]));
		if constexpr (V::textured) {
			int32_t gt[6] = { ps->dstzw_dx[0], ps->dstzw_dx[1], ps->dstzw_dx[3],
			                  ps->dstzw_dy[0], ps->dstzw_dy[1], ps->dstzw_dy[3] };
#pragma unroll
			for (int i = 0; i < 6; i++) asm volatile("" ::"s"(gt[i]));
		}
	}
	asm volatile("" : "+v"(px.dirty));
	if (shade_pixel<V>(ps, h, states[h.state_index].tiles, tex, fmt, span, sp + 1, x, y, px, sh, L)) {
		if (sh.coverage & COVERAGE_FILL)
			fill_pixel(px, fmt, alias, ci, h.fill_color);
		else if (sh.coverage & COVERAGE_COPY)
			copy_pixel(px, fmt, alias, (uint32_t)sh.z_dith);
		else
			depth_blend<V>(lut, ps, h, fmt, alias, px, sh);
	}
	asm volatile("" : "+v"(px.dirty));
}

// Waves per SIMD each occupancy class is compiled for (the register budget
// follows: 512 / waves, in steps of 8).
#ifndef RDP_RASTER_WAVES_CLASS0
#define RDP_RASTER_WAVES_CLASS0 7
#endif
#ifndef RDP_RASTER_WAVES_CLASS1
#define RDP_RASTER_WAVES_CLASS1 7
#endif
#ifndef RDP_RASTER_WAVES_CLASS2
#define RDP_RASTER_WAVES_CLASS2 6
#endif

// Workgroups are dispatched to the 8 XCDs round robin (block b runs on XCD
// b % 8) and every XCD has its own L2.  Neighbouring tiles read the same
// primitive and span records, so consecutive chunks of `chunk` work items are
// kept on one XCD: chunk c goes to XCD c % 8.  chunk 0 gives each XCD one
// contiguous eighth of the items; a negative chunk keeps block order.  The
// chunk is a launch parameter chosen by the host (autotuned per workload).
#define RDP_NUM_XCDS 8

DEV uint32_t xcd_local_item(uint32_t b, uint32_t n, int32_t chunk)
{
	const uint32_t xcd = b % RDP_NUM_XCDS, k = b / RDP_NUM_XCDS;
	if (chunk < 0) return b;
	if (chunk == 0) {
		const uint32_t q = n / RDP_NUM_XCDS, r = n % RDP_NUM_XCDS;
		// Item order: XCD x owns items [x*q + min(x, r), ...) of length q (+1 if x < r).
		const uint32_t len = q + (xcd < r ? 1u : 0u);
		return k < len ? xcd * q + (xcd < r ? xcd : r) + k : b;
	}
	// Block b is the k-th block of its XCD; that XCD's k-th item is in its
	// (k / chunk)-th chunk.  The last, partial round of chunks keeps the
	// identity mapping.
	const uint32_t c = (uint32_t)chunk;
	const uint32_t per_round = RDP_NUM_XCDS * c;
	if (b >= (n / per_round) * per_round) return b;
	const uint32_t ch = (k / c) * RDP_NUM_XCDS + xcd;
	return ch * c + k % c;
}

// The blender divider table (one byte per 15-bit index), uploaded once per
// device by upload_blender_lut.  As a symbol its address is a constant the
// compiler rematerializes instead of keeping it in registers.
__device__ uint8_t g_blender_lut[0x8000];

hipError_t upload_blender_lut(const uint8_t *lut)
{
	return hipMemcpyToSymbol(HIP_SYMBOL(g_blender_lut), lut, sizeof(g_blender_lut), 0, hipMemcpyHostToDevice);
}

// Orders a wave's LDS staging writes against its own later reads (and its
// earlier reads against the next staging): a workgroup barrier when the
// workgroup is the one wave, a wave-scope fence and barrier when each wave of
// the workgroup stages into its own LDS slice and runs its own primitive list.
template <bool QUAD>
DEV void staging_sync()
{
	if constexpr (QUAD) {
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	} else
		__syncthreads();
}

// CLS: occupancy class; the kernel holds the variants of classes <= CLS (and
// the generic code in the top class).  The host launches each pass in the
// class of its heaviest primitive.
// QUAD: a work item is a 16x16 tile and the workgroup its four 8x8 quadrants,
// one wave each (quadrant q = wave q: x + (q & 1), y + (q >> 1)), so the
// quadrants of a 16x16 tile share one CU and its L1 for the primitive and
// span records they have in common.  Each wave walks its own quadrant's bin
// masks and stages TMEM into its own LDS slice; waves never wait for each
// other.  Otherwise a work item is one 8x8 tile and the workgroup one wave.
// One work item: the whole tile (or, with QUAD, this wave's quadrant of it).
// s_tmem_all: the workgroup's LDS staging area, 2048 halfwords per wave.
// UP: the pass renders into the sample planes at 2^scale_log2 times the
// resolution (submit_render_pass_upscaled, rdp_renderer.cpp:2287-2342), and
// work items are pass << 20 | tile x << 10 | tile y.
// emit_scaled_write_masks (memory_interfacing.h:305-336): bit 0 (colour)
// and bit 1 (depth) of the native pixel `key` for a sample this lane wrote.
// At 4x and 8x the 16 or 64 lanes of one native pixel merge their bits first
// and one lane per pixel sets them; at 2x (four lanes per pixel) each lane
// sets its own.
DEV void emit_scaled_write_masks(uint32_t *wmask, uint32_t key, uint32_t bits, int L, int lane)
{
	uint64_t todo = __ballot(bits != 0);
	if (L == 1) {
		if (bits) atomicOr(wmask + key, bits);
		todo = 0;
	}
	while (todo) {
		const int leader = __ffsll((unsigned long long)todo) - 1;
		const uint32_t k = __builtin_amdgcn_readlane(key, leader);
		const uint64_t same = __ballot(key == k && bits != 0);
		const uint32_t merged =
		    (__ballot(key == k && (bits & 1u)) ? 1u : 0u) | (__ballot(key == k && (bits & 2u)) ? 2u : 0u);
		if (lane == leader) atomicOr(wmask + k, merged);
		todo &= ~same;
	}
}

template <uint32_t CLS, bool QUAD, bool UP = false>
DEV void raster_tile(const PassDesc *__restrict__ passes, uint32_t item, uint16_t *s_tmem_all)
{
	const PassDesc &P = passes[UP ? (item >> 20) : (item >> 14)];
	uint32_t tx = UP ? (item >> 10) & 0x3ff : (item >> 7) & 0x7f, ty = UP ? item & 0x3ff : item & 0x7f;
	const int lane = threadIdx.x & 63;
	uint16_t *s_tmem = s_tmem_all;
	if constexpr (QUAD) {
		const uint32_t q = uniform_u32(threadIdx.x >> 6);
		tx = tx * 2 + (q & 1);
		ty = ty * 2 + (q >> 1);
		// A quadrant past the framebuffer's last tile column or row: the wave
		// has nothing to do (no other wave waits for it).
		if (tx >= P.tiles_x || ty >= P.tiles_y) return;
		s_tmem += q * 2048;
	}
	const int x = (int)(tx * 8) + (lane & 7);
	const int y = (int)(ty * 8) + (lane >> 3);
	const uint32_t fmt = P.fb_fmt;
	const bool in_fb = x < (int)P.fb_width && y < (int)P.fb_height;
	const bool alias = P.fb_addr == P.depth_addr;
	// Upscaled: pixel (x, y) is sample (x mod S, y mod S) of native pixel
	// (x / S, y / S), stored in plane (y mod S) * S + x mod S at the native
	// pixel's address (memory_interfacing.h:286-362).
	const int L = UP ? (int)uniform_u32(P.scale_log2) : 0;
	const uint32_t smask = (1u << L) - 1;
	const uint32_t slice = UP ? (((uint32_t)y & smask) << L) + ((uint32_t)x & smask) : 0;
	const size_t plane8 = UP ? (size_t)slice * ((size_t)P.rdram_mask + 1) : 0;
	const Mem m{ gptr(P.rdram) + plane8, gptr(P.hidden) + (plane8 >> 1), P.rdram_mask };
	const AS_CONST RasterPrim *rprims = cptr(P.rprims);
	const AS_CONST DrawState *states = cptr(P.states);
	const AS_GLOBAL SpanSetup *spans = gptr(P.spans);
	const AS_GLOBAL uint8_t *lut = (const AS_GLOBAL uint8_t *)g_blender_lut;
	const Tex tex{ (const AS_LDS uint16_t *)(s_tmem) };

	uint32_t addr_index;
	switch (fmt) {
	case FB_I4:
	case FB_I8: addr_index = P.fb_addr; break;
	case FB_RGBA5551:
	case FB_IA88: addr_index = P.fb_addr >> 1; break;
	default: addr_index = P.fb_addr >> 2; break;
	}
	const uint32_t nrow = P.fb_width >> L, nx = (uint32_t)x >> L, ny = (uint32_t)y >> L;
	const uint32_t ci = addr_index + nrow * ny + nx;
	const uint32_t di = (P.depth_addr >> 1) + nrow * ny + nx;

	Pixel px;
	px.dirty = 0;
	load_color(m, fmt, ci, px);
	{
		uint32_t w = m.rd16i(di);
		px.depth = w >> 2;
		px.dz = (m.hid(di) | ((w & 3) << 2)) & 0xff;
	}

	const uint32_t G = P.num_groups;
	const uint32_t tile = ty * P.tiles_x + tx;
	uint32_t staged = 0xffffffffu;
	const uint32_t needs_tmem = RAS_USES_TEXEL0 | RAS_USES_TEXEL1 | RAS_USES_PIPELINED_TEXEL1 | RAS_COPY;
	// Lane l holds the mask of group gbase + l; up to two rounds of 64 groups.
	for (uint32_t gbase = 0; gbase < G; gbase += 64) {
		uint64_t my_mask = gbase + (uint32_t)lane < G ? gptr(P.bin_fine)[(size_t)tile * G + gbase + lane] : 0;
		uint64_t coarse = __ballot(my_mask != 0);
		while (coarse) {
			const int g = __ffsll((unsigned long long)coarse) - 1;
			coarse &= coarse - 1;
			const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)my_mask, g);
			const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(my_mask >> 32), g);
			uint64_t mask = ((uint64_t)hi << 32) | lo;
			while (mask) {
				const int bit = __ffsll((unsigned long long)mask) - 1;
				mask &= mask - 1;
				const AS_CONST RasterPrim *ps = rprims + ((gbase + (uint32_t)g) * 64 + (uint32_t)bit);
				const RasterPrimHead h = ldc(reinterpret_cast<const AS_CONST RasterPrimHead *>(&ps->ylo));

				// Stage the primitive's TMEM snapshot in LDS (wave-uniform).
				if ((h.ras_flags & needs_tmem) && h.tmem_instance != staged) {
					staging_sync<QUAD>();
					// The snapshot base is read from the descriptor here, not kept
					// live through the loop.
					const AS_CONST PassDesc *pt = cptr(&P);
					asm volatile("" : "+s"(pt));
					const AS_GLOBAL uint32_t *src =
					    (const AS_GLOBAL uint32_t *)gptr(pt->tmem_instances) + (size_t)h.tmem_instance * 1024;
					uint32_t *dst = reinterpret_cast<uint32_t *>(s_tmem);
#pragma unroll
					for (int k = 0; k < 16; k++) dst[lane * 4 + (k & 3) + 256 * (k >> 2)] = src[lane * 4 + (k & 3) + 256 * (k >> 2)];
					staging_sync<QUAD>();
					staged = h.tmem_instance;
				}

				if (!in_fb || y < h.ylo || y > h.yhi) continue;
				switch (h.variant) {
#define RDP_DISPATCH_VARIANT(id, known, value, comb, texfmt, dbp, cls)                                            \
		case id:                                                                                                    \
			if constexpr ((cls) <= CLS)                                                                             \
				raster_prim<RasterVariant<(known), (value), (int)(comb), (texfmt), (int)(dbp)>>(ps, h, states, spans, tex, lut, fmt, alias, ci, \
				                                                                    x, y, px, L);                     \
			break;
					RDP_RASTER_VARIANTS(RDP_DISPATCH_VARIANT)
#undef RDP_DISPATCH_VARIANT
				default:
					if constexpr (CLS == kRasterClasses - 1)
						raster_prim<GenericVariant>(ps, h, states, spans, tex, lut, fmt, alias, ci, x, y, px, L);
					break;
				}
			}
		}
	}

	// The stream memory pointers and the depth index are read again here
	// rather than kept live through the primitive loop.
	const AS_CONST PassDesc *pe = cptr(&P);
	asm volatile("" : "+s"(pe));
	const Mem me{ gptr(pe->rdram) + plane8, gptr(pe->hidden) + (plane8 >> 1), pe->rdram_mask };
	AS_GLOBAL uint8_t *wm = gptr(pe->wmask);
	if (in_fb && (px.dirty & PIXEL_COLOR_DIRTY)) {
		store_color(me, fmt, ci, px);
		if (wm) mark_color(wm, pe->rdram_mask, fmt, ci);
	}
	if constexpr (UP) {
		if (pe->domain_flags & DOMAIN_SSAA_RESOLVE)
			emit_scaled_write_masks(pe->up_wmask, ny * (pe->fb_width >> L) + nx,
			                        in_fb ? px.dirty : 0u, L, lane);
	}
	if (in_fb && !alias && (px.dirty & PIXEL_DEPTH_DIRTY)) {
		const uint32_t de = (pe->depth_addr >> 1) + (pe->fb_width >> L) * ny + nx;
		me.wr16i(de, (px.depth << 2) | (px.dz >> 2));
		me.set_hid(de, px.dz & 3);
		if (wm) mark_depth(wm, pe->rdram_mask, de);
	}
}

#define RDP_RASTER_WAVES(CLS) \
	(CLS == 0 ? RDP_RASTER_WAVES_CLASS0 : (CLS == 1 ? RDP_RASTER_WAVES_CLASS1 : RDP_RASTER_WAVES_CLASS2))

template <uint32_t CLS, bool QUAD, bool UP = false>
__global__ __launch_bounds__(QUAD ? 256 : 64) __attribute__((amdgpu_waves_per_eu(RDP_RASTER_WAVES(CLS))))
void k_raster(const PassDesc *__restrict__ passes, const uint32_t *__restrict__ items, int32_t xcd_chunk)
{
	__shared__ uint16_t s_tmem_all[(QUAD ? 4 : 1) * 2048];
	persp_table_init();
	raster_tile<CLS, QUAD, UP>(passes, items[xcd_local_item(blockIdx.x, gridDim.x, xcd_chunk)], s_tmem_all);
}

// Passes whose every primitive is a FILL-mode primitive (variant 1, the
// first row of the variant table).  A fill writes every channel of a pixel
// it covers and reads neither colour nor depth, so the tile's framebuffer is
// not loaded; a wave takes kFillTiles horizontally adjacent tiles, reads
// their bin masks up front and runs them back to back.
constexpr uint32_t kFillTiles = kFillTilesPerItem;
using FillVariant = RasterVariant<RAS_FILL | RAS_COPY, RAS_FILL, (int)COMB_ANY, TEXFMT_ANY, (int)DBP_ANY>;

// UP: an upscaled pass's item (pass << 20 | tile group << 10 | tile row),
// every lane a sample of its native pixel in the sample planes, as in
// raster_tile.
template <bool UP>
DEV void fill_item(const PassDesc *__restrict__ passes, uint32_t item, uint16_t *s_unused)
{
	const PassDesc &P = passes[UP ? (item >> 20) : (item >> 14)];
	const uint32_t tx0 = (UP ? (item >> 10) & 0x3ff : (item >> 7) & 0x7f) * kFillTiles, ty = UP ? item & 0x3ff : item & 0x7f;
	const int lane = threadIdx.x & 63;
	const uint32_t fmt = P.fb_fmt;
	const bool alias = P.fb_addr == P.depth_addr;
	// Tiles start at multiples of 8, so a lane's sample slot (x mod S,
	// y mod S) is the same in every tile of the item.
	const int L = UP ? (int)uniform_u32(P.scale_log2) : 0;
	const uint32_t smask = (1u << L) - 1;
	const uint32_t slice = UP ? ((((uint32_t)lane >> 3) & smask) << L) + ((uint32_t)lane & smask) : 0;
	const size_t plane8 = UP ? (size_t)slice * ((size_t)P.rdram_mask + 1) : 0;
	const Mem m{ gptr(P.rdram) + plane8, gptr(P.hidden) + (plane8 >> 1), P.rdram_mask };
	const AS_CONST RasterPrim *rprims = cptr(P.rprims);
	const AS_CONST DrawState *states = cptr(P.states);
	const AS_GLOBAL SpanSetup *spans = gptr(P.spans);
	const AS_GLOBAL uint8_t *lut = (const AS_GLOBAL uint8_t *)g_blender_lut;
	const AS_GLOBAL uint64_t *bins = gptr(P.bin_fine);
	const Tex tex{ (const AS_LDS uint16_t *)(s_unused) };
	const uint32_t G = P.num_groups, tiles_x = P.tiles_x;
	const uint32_t y = ty * 8 + ((uint32_t)lane >> 3);

	uint64_t first[kFillTiles];
#pragma unroll
	for (uint32_t t = 0; t < kFillTiles; t++)
		first[t] = tx0 + t < tiles_x && (uint32_t)lane < G ? bins[(size_t)(ty * tiles_x + tx0 + t) * G + lane] : 0;

#pragma unroll
	for (uint32_t t = 0; t < kFillTiles; t++) {
		const uint32_t tx = tx0 + t;
		if (tx >= tiles_x) break;
		const uint32_t x = tx * 8 + ((uint32_t)lane & 7);
		const bool in_fb = x < P.fb_width && y < P.fb_height;
		uint32_t addr_index;
		switch (fmt) {
		case FB_I4:
		case FB_I8: addr_index = P.fb_addr; break;
		case FB_RGBA5551:
		case FB_IA88: addr_index = P.fb_addr >> 1; break;
		default: addr_index = P.fb_addr >> 2; break;
		}
		const uint32_t nx = x >> L, ny = y >> L;
		const uint32_t ci = addr_index + (P.fb_width >> L) * ny + nx;
		Pixel px;
		px.c[0] = px.c[1] = px.c[2] = px.c[3] = 0;
		px.depth = px.dz = 0;
		px.dirty = 0;
		for (uint32_t gbase = 0; gbase < G; gbase += 64) {
			const uint64_t my_mask =
			    gbase == 0 ? first[t] : (gbase + (uint32_t)lane < G ? bins[(size_t)(ty * tiles_x + tx) * G + gbase + lane] : 0);
			uint64_t coarse = __ballot(my_mask != 0);
			while (coarse) {
				const int g = __ffsll((unsigned long long)coarse) - 1;
				coarse &= coarse - 1;
				const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)my_mask, g);
				const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(my_mask >> 32), g);
				uint64_t mask = ((uint64_t)hi << 32) | lo;
				while (mask) {
					const int bit = __ffsll((unsigned long long)mask) - 1;
					mask &= mask - 1;
					const AS_CONST RasterPrim *ps = rprims + ((gbase + (uint32_t)g) * 64 + (uint32_t)bit);
					const RasterPrimHead h = ldc(reinterpret_cast<const AS_CONST RasterPrimHead *>(&ps->ylo));
					if (!in_fb || (int)y < h.ylo || (int)y > h.yhi) continue;
					raster_prim<FillVariant>(ps, h, states, spans, tex, lut, fmt, alias, ci, (int)x, (int)y, px, L);
				}
			}
		}
		if (in_fb && (px.dirty & PIXEL_COLOR_DIRTY)) {
			store_color(m, fmt, ci, px);
			if (P.wmask) mark_color(gptr(P.wmask), P.rdram_mask, fmt, ci);
		}
		// A fill writes no depth field (with colour and depth aliased the
		// colour store is the depth store).
		if constexpr (UP) {
			if (P.domain_flags & DOMAIN_SSAA_RESOLVE)
				emit_scaled_write_masks(P.up_wmask, ny * (P.fb_width >> L) + nx, in_fb ? px.dirty & PIXEL_COLOR_DIRTY : 0u, L,
				                        lane);
		}
	}
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8)))
void k_raster_fill(const PassDesc *__restrict__ passes, const uint32_t *__restrict__ items, int32_t xcd_chunk)
{
	__shared__ uint16_t s_unused[2];
	// Fill-mode primitives never divide; the table is filled all the same so
	// every kernel that can reach perspective_divide has it.
	persp_table_init();
	fill_item<false>(passes, items[xcd_local_item(blockIdx.x, gridDim.x, xcd_chunk)], s_unused);
}

// The FILL-only passes of an upscaled round: k_raster_fill in the sample planes.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8)))
void k_raster_fill_up(const PassDesc *__restrict__ passes, const uint32_t *__restrict__ items, int32_t xcd_chunk)
{
	__shared__ uint16_t s_unused[2];
	persp_table_init();
	fill_item<true>(passes, items[xcd_local_item(blockIdx.x, gridDim.x, xcd_chunk)], s_unused);
}

// One launch for every raster work item of a round (persistent): the grid
// holds at most as many one-wave workgroups as the GPU keeps resident, and
// workgroup b takes items b, b + grid, b + 2 grid, ... of the round's list:
// first the heavier class's tiles (hi, the densest passes first), then the
// lighter class's (lo), then the fill items.  (Taking items from one shared
// counter instead serialised on the counter's atomics across the XCDs: three
// times slower.)
template <uint32_t CLS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RDP_RASTER_WAVES(CLS))))
void k_raster_round(const PassDesc *__restrict__ passes, const uint32_t *__restrict__ items_hi, uint32_t n_hi,
                    const uint32_t *__restrict__ items_lo, uint32_t n_lo, const uint32_t *__restrict__ fill_items,
                    uint32_t n_fill)
{
	__shared__ uint16_t s_tmem[2048];
	persp_table_init();
	const uint32_t total = n_hi + n_lo + n_fill;
	for (uint32_t k = blockIdx.x; k < total; k += gridDim.x) {
		if (k < n_hi)
			raster_tile<CLS, false>(passes, items_hi[k], s_tmem);
		else if (k < n_hi + n_lo)
			raster_tile<CLS, false>(passes, items_lo[k - n_hi], s_tmem);
		else
			fill_item<false>(passes, fill_items[k - n_hi - n_lo], s_tmem);
		// The next item's staging must not overwrite texels this item's
		// lanes still read.
		__syncthreads();
	}
}
// ==========================================================================
// Launchers (host side).
// ==========================================================================
hipError_t launch_setup(const PassDesc *passes, const uint32_t *items, uint32_t n, const uint32_t *span_items,
                        uint32_t n_span, hipStream_t s)
{
	if (n) hipLaunchKernelGGL(k_setup, dim3(n), dim3(64), 0, s, passes, items);
	if (n_span) hipLaunchKernelGGL(k_span, dim3(n_span), dim3(64), 0, s, passes, span_items);
	return hipGetLastError();
}

hipError_t launch_tmem(const PassDesc *passes, const uint32_t *gather_items, uint32_t n_gather,
                       const uint32_t *scan_items, uint32_t n_scan, hipStream_t s)
{
	if (n_gather) hipLaunchKernelGGL(k_tmem_gather, dim3(n_gather), dim3(256), 0, s, passes, gather_items);
	if (n_scan) hipLaunchKernelGGL(k_tmem_scan, dim3(n_scan), dim3(256), 0, s, passes, scan_items);
	return hipGetLastError();
}

hipError_t launch_bin(const PassDesc *passes, const uint32_t *items, uint32_t n, hipStream_t s)
{
	if (!n) return hipSuccess;
	hipLaunchKernelGGL(k_bin, dim3(n), dim3(64), 0, s, passes, items);
	return hipGetLastError();
}

hipError_t launch_raster_fill(const PassDesc *passes, const uint32_t *items, uint32_t n, int32_t xcd_chunk, hipStream_t s)
{
	if (!n) return hipSuccess;
	hipLaunchKernelGGL(k_raster_fill, dim3(n), dim3(64), 0, s, passes, items, xcd_chunk);
	return hipGetLastError();
}

hipError_t launch_raster_fill_up(const PassDesc *passes, const uint32_t *items, uint32_t n, int32_t xcd_chunk,
                                 hipStream_t s)
{
	if (!n) return hipSuccess;
	hipLaunchKernelGGL(k_raster_fill_up, dim3(n), dim3(64), 0, s, passes, items, xcd_chunk);
	return hipGetLastError();
}

// The persistent raster launch of a round (k_raster_round): cls is the
// heavier class present (0 or 1), grid the number of one-wave workgroups.
hipError_t launch_raster_round(uint32_t cls, const PassDesc *passes, const uint32_t *items_hi, uint32_t n_hi,
                               const uint32_t *items_lo, uint32_t n_lo, const uint32_t *fill_items, uint32_t n_fill,
                               uint32_t grid, hipStream_t s)
{
	if (!(n_hi + n_lo + n_fill)) return hipSuccess;
	if (cls == 0)
		hipLaunchKernelGGL(k_raster_round<0>, dim3(grid), dim3(64), 0, s, passes, items_hi, n_hi, items_lo, n_lo, fill_items,
		                   n_fill);
	else
		hipLaunchKernelGGL(k_raster_round<1>, dim3(grid), dim3(64), 0, s, passes, items_hi, n_hi, items_lo, n_lo, fill_items,
		                   n_fill);
	return hipGetLastError();
}

// quad: the items are 16x16 tiles, four waves per workgroup.
hipError_t launch_raster(uint32_t cls, const PassDesc *passes, const uint32_t *items, uint32_t n, int32_t xcd_chunk,
                         bool quad, hipStream_t s)
{
	if (!n) return hipSuccess;
	if (quad) {
		switch (cls) {
		case 0: hipLaunchKernelGGL((k_raster<0, true>), dim3(n), dim3(256), 0, s, passes, items, xcd_chunk); break;
		case 1: hipLaunchKernelGGL((k_raster<1, true>), dim3(n), dim3(256), 0, s, passes, items, xcd_chunk); break;
		default: hipLaunchKernelGGL((k_raster<2, true>), dim3(n), dim3(256), 0, s, passes, items, xcd_chunk); break;
		}
	} else {
		switch (cls) {
		case 0: hipLaunchKernelGGL((k_raster<0, false>), dim3(n), dim3(64), 0, s, passes, items, xcd_chunk); break;
		case 1: hipLaunchKernelGGL((k_raster<1, false>), dim3(n), dim3(64), 0, s, passes, items, xcd_chunk); break;
		default: hipLaunchKernelGGL((k_raster<2, false>), dim3(n), dim3(64), 0, s, passes, items, xcd_chunk); break;
		}
	}
	return hipGetLastError();
}

// Upscaled passes (k_raster<cls, false, true>): 8x8 tiles of the S-times
// framebuffer.
hipError_t launch_raster_up(uint32_t cls, const PassDesc *passes, const uint32_t *items, uint32_t n, int32_t xcd_chunk,
                            hipStream_t s)
{
	if (!n) return hipSuccess;
	switch (cls) {
	case 0: hipLaunchKernelGGL((k_raster<0, false, true>), dim3(n), dim3(64), 0, s, passes, items, xcd_chunk); break;
	case 1: hipLaunchKernelGGL((k_raster<1, false, true>), dim3(n), dim3(64), 0, s, passes, items, xcd_chunk); break;
	default: hipLaunchKernelGGL((k_raster<2, false, true>), dim3(n), dim3(64), 0, s, passes, items, xcd_chunk); break;
	}
	return hipGetLastError();
}

// ==========================================================================
// Upscaled domain updates: update_upscaled_domain_pre.comp (post = false)
// and update_upscaled_domain_post.comp (post = true), one lane per element.
// ==========================================================================
struct DomainMem {
	const AS_GLOBAL uint8_t *rdram, *hidden;
	AS_GLOBAL uint8_t *ref, *planes, *hplanes;
	uint32_t rdram_size, num_planes;
};

// index counts host-order elements of 1 << size_log2 bytes.  Pre: where the
// native element differs from the reference copy, write it and its hidden
// bits to every plane and update the reference.  Post: reference := native.
template <int SIZE_LOG2>
DEV void domain_element(const DomainMem &d, bool post, uint32_t index)
{
	using T = typename std::conditional<SIZE_LOG2 == 0, uint8_t, typename std::conditional<SIZE_LOG2 == 1, uint16_t,
	                                                                                       uint32_t>::type>::type;
	index &= (d.rdram_size >> SIZE_LOG2) - 1;
	const T real = ((const AS_GLOBAL T *)d.rdram)[index];
	AS_GLOBAL T *ref = (AS_GLOBAL T *)d.ref;
	if (post) {
		ref[index] = real;
		return;
	}
	if (ref[index] == real) return;
	const size_t plane_elems = d.rdram_size >> SIZE_LOG2, hplane = d.rdram_size >> 1;
	for (uint32_t i = 0; i < d.num_planes; i++) {
		((AS_GLOBAL T *)d.planes)[index + i * plane_elems] = real;
		AS_GLOBAL uint8_t *uh = d.hplanes + i * hplane;
		if (SIZE_LOG2 == 0) {
			const uint32_t m = index ^ 3u;
			if (m & 1u) uh[m >> 1] = d.hidden[m >> 1];
		} else if (SIZE_LOG2 == 1) {
			uh[index ^ 1u] = d.hidden[index ^ 1u];
		} else {
			uh[2 * index] = d.hidden[2 * index];
			uh[2 * index + 1] = d.hidden[2 * index + 1];
		}
	}
	ref[index] = real;
}

DEV void domain_update(const DomainMem &d, bool post, uint32_t addr, uint32_t depth_addr, uint32_t num_pixels,
                       uint32_t size_log2, uint32_t i)
{
	// submit_update_upscaled_domain (rdp_renderer.cpp:1943-1975): whole words.
	addr &= ~3u;
	depth_addr &= ~3u;
	const uint32_t align = 4u >> size_log2;
	num_pixels = (num_pixels + align - 1u) & ~(align - 1u);
	if (i >= num_pixels) return;
	const uint32_t ci = (addr >> size_log2) + i;
	switch (size_log2) {
	case 0: domain_element<0>(d, post, ci); break;
	case 1: domain_element<1>(d, post, ci); break;
	default: domain_element<2>(d, post, ci); break;
	}
	if (addr != depth_addr) domain_element<1>(d, post, (depth_addr >> 1) + i);
}

// Per pass: the framebuffer region fb_width x domain_height; work items are
// pass << 17 | block of 256 elements.
__global__ __launch_bounds__(256) void k_domain(const PassDesc *__restrict__ passes, const uint32_t *__restrict__ items,
                                                uint32_t post)
{
	const uint32_t item = items[blockIdx.x];
	const PassDesc &P = passes[item >> 17];
	const uint32_t i = (item & 0x1ffff) * 256 + threadIdx.x;
	const uint32_t fmt = P.fb_fmt;
	const uint32_t size_log2 = fmt == FB_RGBA8888 ? 2 : (fmt == FB_RGBA5551 || fmt == FB_IA88) ? 1 : 0;
	const DomainMem d{ gptr(P.rdram), gptr(P.hidden), gptr(P.up_ref), gptr(P.up_rdram), gptr(P.up_hidden),
		               P.rdram_mask + 1, 1u << (2 * P.domain_log2) };
	domain_update(d, post != 0, P.fb_addr, P.depth_addr, P.fb_width * P.domain_height, size_log2, i);
	// clear_super_sampled_write_mask.comp
	if (!post && (P.domain_flags & DOMAIN_SSAA_CLEAR) && i < P.fb_width * P.domain_height) gptr(P.up_wmask)[i] = 0;
}

// update_upscaled_domain_resolve.comp:165-276, one lane per native pixel of
// the region (fb_width >> scale_log2) x domain_height of an upscaled pass:
// a pixel any sample wrote gets the mean of its samples in native RDRAM and
// the reference copy (8-bit: plane 0's byte; depth: plane 0's value).
__constant__ uint8_t c_ssaa_dither[16] = { 0, 4, 1, 5, 4, 0, 5, 1, 3, 7, 2, 6, 7, 3, 6, 2 };

__global__ __launch_bounds__(256) void k_ssaa_resolve(const PassDesc *__restrict__ passes,
                                                      const uint32_t *__restrict__ items)
{
	const uint32_t item = items[blockIdx.x];
	const PassDesc &P = passes[item >> 17];
	const uint32_t width = P.fb_width >> P.scale_log2;
	const uint32_t index = (item & 0x1ffff) * 256 + threadIdx.x;
	if (index >= width * P.domain_height) return;
	const uint32_t x = index % width, y = index / width;
	const uint32_t m = gptr(P.up_wmask)[index];
	if (!m) return;
	const uint32_t fmt = P.fb_fmt;
	const uint32_t size_log2 = fmt == FB_RGBA8888 ? 2 : (fmt == FB_RGBA5551 || fmt == FB_IA88) ? 1 : 0;
	const uint32_t N = 1u << (2 * P.scale_log2);
	const size_t R = (size_t)P.rdram_mask + 1;
	// An upscaled pass's rdram points at the planes.
	const AS_GLOBAL uint8_t *planes = gptr(P.rdram);
	AS_GLOBAL uint8_t *native = gptr(P.resolve_rdram);
	AS_GLOBAL uint8_t *ref = gptr(P.up_ref);
	if (m & 1u) {
		const uint32_t fb_index = P.fb_addr >> size_log2;
		if (size_log2 == 0) {
			const uint32_t ci = ((fb_index + index) & (uint32_t)(R - 1)) ^ 3u;
			uint32_t r = 0;
			for (uint32_t i = 0; i < N; i++) r += planes[ci];
			r = (r + (N >> 1)) / N;
			native[ci] = ref[ci] = (uint8_t)r;
		} else if (size_log2 == 1) {
			const uint32_t ci = ((fb_index + index) & (uint32_t)(R / 2 - 1)) ^ 1u;
			uint32_t c[4] = { 0, 0, 0, 0 };
			for (uint32_t i = 0; i < N; i++) {
				const uint32_t w = ((const AS_GLOBAL uint16_t *)(planes + i * R))[ci];
				c[0] += (w >> 11) & 31;
				c[1] += (w >> 6) & 31;
				c[2] += (w >> 1) & 31;
				c[3] += w & 1;
			}
			const bool dither = (P.domain_flags & DOMAIN_SSAA_DITHER) != 0;
			const uint32_t d = c_ssaa_dither[(y & 3) * 4 + (x & 3)] * N;
#pragma unroll
			for (int k = 0; k < 4; k++) c[k] = dither ? (8 * c[k] + d) / (8 * N) : (c[k] + (N >> 1)) / N;
			const uint16_t e = (uint16_t)((c[0] << 11) | (c[1] << 6) | (c[2] << 1) | c[3]);
			((AS_GLOBAL uint16_t *)native)[ci] = e;
			((AS_GLOBAL uint16_t *)ref)[ci] = e;
		} else {
			const uint32_t ci = (fb_index + index) & (uint32_t)(R / 4 - 1);
			uint32_t c[4] = { 0, 0, 0, 0 };
			for (uint32_t i = 0; i < N; i++) {
				const uint32_t w = ((const AS_GLOBAL uint32_t *)(planes + i * R))[ci];
#pragma unroll
				for (int k = 0; k < 4; k++) c[k] += (w >> (24 - 8 * k)) & 0xff;
			}
			uint32_t e = 0;
#pragma unroll
			for (int k = 0; k < 4; k++) e |= ((c[k] + (N >> 1)) / N) << (24 - 8 * k);
			((AS_GLOBAL uint32_t *)native)[ci] = e;
			((AS_GLOBAL uint32_t *)ref)[ci] = e;
		}
	}
	if (P.fb_addr != P.depth_addr && (m & 2u)) {
		const uint32_t di = (((P.depth_addr >> 1) + index) & (uint32_t)(R / 2 - 1)) ^ 1u;
		const uint16_t v = ((const AS_GLOBAL uint16_t *)planes)[di];
		((AS_GLOBAL uint16_t *)native)[di] = v;
		((AS_GLOBAL uint16_t *)ref)[di] = v;
	}
}

hipError_t launch_ssaa_resolve(const PassDesc *passes, const uint32_t *items, uint32_t n, hipStream_t s)
{
	if (!n) return hipSuccess;
	hipLaunchKernelGGL(k_ssaa_resolve, dim3(n), dim3(256), 0, s, passes, items);
	return hipGetLastError();
}

// One range, colour only (submit_update_upscaled_domain_external,
// rdp_renderer.cpp:1936-1941: the VI's scanout range before its fetch).
__global__ __launch_bounds__(256) void k_domain_range(const uint8_t *rdram, const uint8_t *hidden, uint8_t *ref,
                                                      uint8_t *planes, uint8_t *hplanes, uint32_t rdram_size,
                                                      uint32_t num_planes, uint32_t addr, uint32_t num_pixels,
                                                      uint32_t size_log2)
{
	const DomainMem d{ gptr(rdram), gptr(hidden), gptr(ref), gptr(planes), gptr(hplanes), rdram_size, num_planes };
	domain_update(d, false, addr, addr, num_pixels, size_log2, blockIdx.x * 256 + threadIdx.x);
}

hipError_t launch_domain(const PassDesc *passes, const uint32_t *items, uint32_t n, bool post, hipStream_t s)
{
	if (!n) return hipSuccess;
	hipLaunchKernelGGL(k_domain, dim3(n), dim3(256), 0, s, passes, items, post ? 1u : 0u);
	return hipGetLastError();
}

hipError_t launch_domain_range(const uint8_t *rdram, const uint8_t *hidden, uint8_t *ref, uint8_t *planes,
                               uint8_t *hplanes, uint32_t rdram_size, uint32_t num_planes, uint32_t addr,
                               uint32_t num_pixels, uint32_t size_log2, hipStream_t s)
{
	const uint32_t align = 4u >> size_log2;
	const uint32_t n = (((num_pixels + align - 1u) & ~(align - 1u)) + 255u) / 256u;
	if (!n) return hipSuccess;
	hipLaunchKernelGGL(k_domain_range, dim3(n), dim3(256), 0, s, rdram, hidden, ref, planes, hplanes, rdram_size,
	                   num_planes, addr, num_pixels, size_log2);
	return hipGetLastError();
}

}  // namespace rdp_mi

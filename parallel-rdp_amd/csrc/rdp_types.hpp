// Data layouts shared by the host front end and the gfx950 kernels.
//
// A render pass is the unit of GPU work: a run of primitives that target one
// colour/depth framebuffer, plus the TMEM uploads interleaved with them.  The
// host front end records, per pass:
//   * the raw command words of every primitive (decoded on the GPU),
//   * a table of resolved draw states (appended whenever a state command
//     changed something since the previous primitive),
//   * one PrimRecord per primitive (word offset, state index, TMEM snapshot,
//     span-line range),
//   * the TMEM uploads (UploadInfo) in command order.
// Everything else (triangle decode, span setup, binning, TMEM snapshots,
// shading, depth/blend) is computed on the device.
#pragma once

#include <stdint.h>

namespace rdp_mi {

// Triangle setup flags (parallel-rdp/rdp_data_structures.hpp:33-43).
enum : uint32_t {
	TRI_FLIP = 1u << 0,
	TRI_DO_OFFSET = 1u << 1,
	TRI_SKIP_XFRAC = 1u << 2,
	TRI_INTERLACE_FIELD = 1u << 3,
	TRI_INTERLACE_KEEP_ODD = 1u << 4,
	TRI_DISABLE_UPSCALING = 1u << 5,
	TRI_NATIVE_LOD = 1u << 6,
	TRI_FILL_COPY_RASTER = 1u << 7,
};

// Rasterisation flags; bit positions follow rdp_data_structures.hpp:45-78.
enum : uint32_t {
	RAS_INTERLACE_FIELD = 1u << 0,
	RAS_INTERLACE_KEEP_ODD = 1u << 1,
	RAS_AA = 1u << 2,
	RAS_PERSPECTIVE = 1u << 3,
	RAS_TLUT = 1u << 4,
	RAS_TLUT_TYPE = 1u << 5,
	RAS_CVG_TIMES_ALPHA = 1u << 6,
	RAS_ALPHA_CVG_SELECT = 1u << 7,
	RAS_MULTI_CYCLE = 1u << 8,
	RAS_TEX_LOD = 1u << 9,
	RAS_SHARPEN_LOD = 1u << 10,
	RAS_DETAIL_LOD = 1u << 11,
	RAS_FILL = 1u << 12,
	RAS_COPY = 1u << 13,
	RAS_SAMPLE_QUAD = 1u << 14,
	RAS_ALPHA_TEST = 1u << 15,
	RAS_ALPHA_TEST_DITHER = 1u << 16,
	RAS_MID_TEXEL = 1u << 17,
	RAS_USES_TEXEL0 = 1u << 18,
	RAS_USES_TEXEL1 = 1u << 19,
	RAS_USES_LOD = 1u << 20,
	RAS_USES_PIPELINED_TEXEL1 = 1u << 21,
	RAS_CONVERT_ONE = 1u << 22,
	RAS_BILERP0 = 1u << 23,
	RAS_BILERP1 = 1u << 24,
	RAS_NEED_NOISE_DUAL = 1u << 25,
	RAS_NEED_NOISE = 1u << 28,
	RAS_USE_STATIC_TEXTURE_SIZE_FORMAT = 1u << 29,
};

// Depth/blend flags (rdp_data_structures.hpp:80-90).
enum : uint32_t {
	DB_DEPTH_TEST = 1u << 0,
	DB_DEPTH_UPDATE = 1u << 1,
	DB_FORCE_BLEND = 1u << 3,
	DB_IMAGE_READ = 1u << 4,
	DB_COLOR_ON_CVG = 1u << 5,
	DB_MULTI_CYCLE = 1u << 6,
	DB_AA = 1u << 7,
	DB_DITHER = 1u << 8,
};

enum : uint32_t { FB_I4 = 0, FB_I8 = 1, FB_RGBA5551 = 2, FB_IA88 = 3, FB_RGBA8888 = 4 };
enum : uint32_t { TF_RGBA = 0, TF_YUV = 1, TF_CI = 2, TF_IA = 3, TF_I = 4 };
enum : uint32_t { UP_TILE = 0, UP_TLUT = 1, UP_BLOCK = 2 };
enum : uint32_t { TILE_CLAMP_S = 1, TILE_MIRROR_S = 2, TILE_CLAMP_T = 4, TILE_MIRROR_T = 8 };

constexpr int kTileSize = 8;                 // 8x8 pixel screen tile = one wave64
constexpr int kMaxPrimsPerPass = 8192;       // 128 bin groups x 64 primitives
constexpr int kBinGroups = kMaxPrimsPerPass / 64;
constexpr int kMaxUploadsPerPass = 1023;     // TMEM snapshots per pass (4 KiB each)
constexpr int kMaxWidth = 1024;
constexpr int kMaxHeight = 1024;
constexpr int kMaxTilesX = kMaxWidth / kTileSize;
constexpr int kMaxTilesY = kMaxHeight / kTileSize;

// Texture tile descriptor (rdp_common.hpp:263-288), 32 bytes.
struct TileDesc {
	uint32_t slo, shi, tlo, thi;
	uint32_t offset, stride;
	uint8_t fmt, size, palette, mask_s;
	uint8_t shift_s, mask_t, shift_t, flags;
};
static_assert(sizeof(TileDesc) == 32, "TileDesc layout");

// Raster code variants.  The raster kernel carries one instantiation of its
// per-primitive shading and blending code per variant; in each, the flag bits
// in RAS_KNOWN hold the values in RAS_VALUE as compile-time constants (all
// other bits are read from the primitive at run time), a known combiner
// replaces the run-time selector decode, and a known texture format
// (fmt * 4 + size) replaces the per-texel format decode.  Every primitive is
// given the first variant whose constants equal its resolved state and
// tiles, so the result is the same as the generic code's by construction.
#if defined(__HIPCC__)
#define RDP_HD __host__ __device__
#else
#define RDP_HD
#endif

enum : uint32_t {
	COMB_ANY = 0,        // selectors read at run time
	COMB_SHADE = 1,      // 1 cycle: rgb = SHADE, alpha = SHADE
	COMB_TEX0_SHADE = 2, // 1 cycle: TEXEL0 * SHADE (rgb and alpha)
	COMB_TEX0 = 3,       // 1 cycle: TEXEL0 (decal)
	COMB_LOD_SHADE = 4,  // 2 cycles: lerp(TEXEL0, TEXEL1, LOD_FRAC), then COMBINED * SHADE
	COMB_TEX0_ENV = 5,   // 2 cycles: TEXEL0, then rgb COMBINED * ENV, alpha COMBINED
	COMB_COUNT = 6,
};

// Raw SetCombine selectors of the known combiners, per cycle (rgb muladd,
// mulsub, mul, add, alpha muladd, mulsub, mul, add).  One-cycle combiners
// use cycle 1 only; for two-cycle ones both cycles must match.
constexpr uint8_t kKnownCombiners[COMB_COUNT][2][8] = {
	{ { 0, 0, 0, 0, 0, 0, 0, 0 }, { 0, 0, 0, 0, 0, 0, 0, 0 } },
	{ { 0, 0, 0, 0, 0, 0, 0, 0 }, { 8, 8, 16, 4, 7, 7, 7, 4 } },
	{ { 0, 0, 0, 0, 0, 0, 0, 0 }, { 1, 8, 4, 7, 1, 7, 4, 7 } },
	{ { 0, 0, 0, 0, 0, 0, 0, 0 }, { 8, 8, 16, 1, 7, 7, 7, 1 } },
	{ { 2, 1, 13, 1, 2, 1, 0, 1 }, { 0, 8, 4, 7, 0, 7, 4, 7 } },
	{ { 8, 8, 16, 1, 7, 7, 7, 1 }, { 0, 8, 5, 7, 7, 7, 7, 0 } },
};
constexpr bool kKnownCombinerTwoCycle[COMB_COUNT] = { false, false, false, false, true, true };

// Depth/blend profiles a raster variant can fix at compile time: all
// depth/blend flags, both cycles' blender selectors (1a, 1b, 2a, 2b), the Z
// mode and the coverage mode.
struct DepthBlendProfile {
	uint32_t db_flags;
	uint8_t blend[2][4];
	uint32_t z_mode, coverage_mode;
};
enum : uint32_t {
	DBP_ANY = 0,            // read at run time
	DBP_Z_OPAQUE = 1,       // Z test + update, no blending, 1 cycle
	DBP_Z_FORCE_BLEND = 2,  // Z test + update, forced pixel/memory blend by pixel alpha, 1 cycle
	DBP_Z_AA_FOG = 3,       // Z test + update, coverage AA, 2 cycles: fog by shade alpha, then pixel/memory
	DBP_AA_SPRITE = 4,      // coverage AA without Z, 2 cycles: pixel, then pixel/memory
	DBP_COUNT = 5,
};
constexpr DepthBlendProfile kDepthBlendProfiles[DBP_COUNT] = {
	{ 0, { { 0, 0, 0, 0 }, { 0, 0, 0, 0 } }, 0, 0 },
	{ DB_DEPTH_TEST | DB_DEPTH_UPDATE, { { 0, 0, 1, 0 }, { 0, 0, 1, 0 } }, 0, 0 },
	{ DB_DEPTH_TEST | DB_DEPTH_UPDATE | DB_FORCE_BLEND, { { 0, 0, 1, 0 }, { 0, 0, 1, 0 } }, 0, 0 },
	{ DB_DEPTH_TEST | DB_DEPTH_UPDATE | DB_IMAGE_READ | DB_MULTI_CYCLE | DB_AA, { { 3, 2, 0, 0 }, { 0, 0, 1, 1 } }, 0, 0 },
	{ DB_IMAGE_READ | DB_MULTI_CYCLE | DB_AA, { { 0, 0, 0, 0 }, { 0, 0, 1, 1 } }, 0, 0 },
};

RDP_HD inline bool depth_blend_profile_matches(uint32_t dbp, uint32_t db_flags, const uint8_t blend[2][4], uint32_t z_mode,
                                               uint32_t coverage_mode)
{
	if (dbp == DBP_ANY) return true;
	const DepthBlendProfile &p = kDepthBlendProfiles[dbp];
	if (p.db_flags != db_flags || p.z_mode != z_mode || p.coverage_mode != coverage_mode) return false;
	for (int c = 0; c < 2; c++)
		for (int i = 0; i < 4; i++)
			if (p.blend[c][i] != blend[c][i]) return false;
	return true;
}

constexpr int TEXFMT_ANY = -1;
constexpr int TEXFMT_RGBA16 = 0 * 4 + 2;
constexpr int TEXFMT_CI4 = 2 * 4 + 0;
constexpr int TEXFMT_CI8 = 2 * 4 + 1;

constexpr uint32_t RAS_STRUCT_1CYC = RAS_MULTI_CYCLE | RAS_COPY | RAS_FILL | RAS_USES_TEXEL0 | RAS_USES_TEXEL1 |
                                     RAS_USES_PIPELINED_TEXEL1 | RAS_USES_LOD;
// Coverage-to-alpha, alpha test, AA and noise all off: the plain shaded
// state, with every one of those paths folded away at compile time.
constexpr uint32_t RAS_PLAIN_KNOWN = RAS_AA | RAS_ALPHA_TEST | RAS_CVG_TIMES_ALPHA | RAS_ALPHA_CVG_SELECT | RAS_NEED_NOISE;
// Coverage AA with alpha from coverage (alpha-coverage select), the other
// RAS_PLAIN_KNOWN paths off.
constexpr uint32_t RAS_AA_CVG_SELECT = RAS_AA | RAS_ALPHA_CVG_SELECT;
constexpr uint32_t RAS_SAMPLING = RAS_PERSPECTIVE | RAS_SAMPLE_QUAD | RAS_BILERP0 | RAS_BILERP1 | RAS_MID_TEXEL |
                                  RAS_CONVERT_ONE | RAS_TLUT | RAS_TLUT_TYPE | RAS_TEX_LOD | RAS_SHARPEN_LOD |
                                  RAS_DETAIL_LOD;
// 1 cycle, texel 0 only, perspective bilinear without TLUT.
constexpr uint32_t RAS_TEX0_BILINEAR = RAS_USES_TEXEL0 | RAS_PERSPECTIVE | RAS_SAMPLE_QUAD | RAS_BILERP0;
// 2 cycles, two texels, mip LOD, perspective bilinear with an RGBA16 TLUT.
constexpr uint32_t RAS_2CYC_LOD_TLUT = RAS_MULTI_CYCLE | RAS_USES_TEXEL0 | RAS_USES_TEXEL1 | RAS_USES_LOD |
                                       RAS_PERSPECTIVE | RAS_SAMPLE_QUAD | RAS_BILERP0 | RAS_BILERP1 | RAS_TLUT |
                                       RAS_TEX_LOD;
// 2 cycles, texel 0 only, bilinear with an RGBA16 TLUT, no perspective or LOD
// (sprites drawn as texture rectangles).
constexpr uint32_t RAS_2CYC_SPRITE_TLUT = RAS_MULTI_CYCLE | RAS_USES_TEXEL0 | RAS_SAMPLE_QUAD | RAS_BILERP0 |
                                          RAS_BILERP1 | RAS_TLUT;

// X(id, ras_known, ras_value, combiner, texture format, depth/blend profile,
// occupancy class).
// The raster kernel is built once per occupancy class with only the variants
// of that class and below (the top class adds the generic code), and a pass
// runs in the lowest class that holds all its primitives' variants: light
// passes then get the register budget, and so the waves per SIMD, of their
// own code rather than of the largest variant.
#define RDP_RASTER_VARIANTS(X)                                                                                        \
	X(1, RAS_FILL | RAS_COPY, RAS_FILL, COMB_ANY, TEXFMT_ANY, DBP_ANY, 0)                                                \
	X(2, RAS_FILL | RAS_COPY, RAS_COPY, COMB_ANY, TEXFMT_ANY, DBP_ANY, 0)                                                \
	X(15, RAS_STRUCT_1CYC | RAS_PLAIN_KNOWN, 0u, COMB_SHADE, TEXFMT_ANY, DBP_Z_OPAQUE, 0)                                \
	X(3, RAS_STRUCT_1CYC, 0u, COMB_SHADE, TEXFMT_ANY, DBP_Z_OPAQUE, 0)                                                   \
	X(4, RAS_STRUCT_1CYC, 0u, COMB_ANY, TEXFMT_ANY, DBP_ANY, 1)                                                          \
	X(16, RAS_STRUCT_1CYC | RAS_SAMPLING | RAS_PLAIN_KNOWN, RAS_TEX0_BILINEAR, COMB_TEX0_SHADE, TEXFMT_RGBA16, DBP_Z_FORCE_BLEND, 0) \
	X(5, RAS_STRUCT_1CYC | RAS_SAMPLING, RAS_TEX0_BILINEAR, COMB_TEX0_SHADE, TEXFMT_RGBA16, DBP_Z_FORCE_BLEND, 0)        \
	X(14, RAS_STRUCT_1CYC | RAS_SAMPLING, RAS_TEX0_BILINEAR, COMB_TEX0_SHADE, TEXFMT_RGBA16, DBP_ANY, 0)                 \
	X(6, RAS_STRUCT_1CYC, RAS_USES_TEXEL0, COMB_TEX0_SHADE, TEXFMT_ANY, DBP_ANY, 1)                                      \
	X(7, RAS_STRUCT_1CYC, RAS_USES_TEXEL0, COMB_TEX0, TEXFMT_ANY, DBP_ANY, 1)                                            \
	X(8, RAS_STRUCT_1CYC, RAS_USES_TEXEL0, COMB_ANY, TEXFMT_ANY, DBP_ANY, 1)                                             \
	X(17, RAS_STRUCT_1CYC | RAS_SAMPLING | RAS_PLAIN_KNOWN, RAS_2CYC_LOD_TLUT | RAS_AA_CVG_SELECT, COMB_LOD_SHADE, TEXFMT_CI8, DBP_Z_AA_FOG, 1) \
	X(18, RAS_STRUCT_1CYC | RAS_SAMPLING | RAS_PLAIN_KNOWN, RAS_2CYC_LOD_TLUT | RAS_AA_CVG_SELECT, COMB_LOD_SHADE, TEXFMT_CI4, DBP_Z_AA_FOG, 1) \
	X(9, RAS_STRUCT_1CYC | RAS_SAMPLING, RAS_2CYC_LOD_TLUT, COMB_LOD_SHADE, TEXFMT_CI8, DBP_Z_AA_FOG, 1)                 \
	X(10, RAS_STRUCT_1CYC | RAS_SAMPLING, RAS_2CYC_LOD_TLUT, COMB_LOD_SHADE, TEXFMT_CI4, DBP_Z_AA_FOG, 1)                \
	X(19, RAS_STRUCT_1CYC | RAS_SAMPLING | RAS_PLAIN_KNOWN, RAS_2CYC_SPRITE_TLUT | RAS_AA_CVG_SELECT, COMB_TEX0_ENV, TEXFMT_CI8, DBP_AA_SPRITE, 1) \
	X(20, RAS_STRUCT_1CYC | RAS_SAMPLING | RAS_PLAIN_KNOWN, RAS_2CYC_SPRITE_TLUT | RAS_AA_CVG_SELECT, COMB_TEX0_ENV, TEXFMT_CI4, DBP_AA_SPRITE, 1) \
	X(12, RAS_STRUCT_1CYC | RAS_SAMPLING, RAS_2CYC_SPRITE_TLUT, COMB_TEX0_ENV, TEXFMT_CI8, DBP_AA_SPRITE, 1)             \
	X(13, RAS_STRUCT_1CYC | RAS_SAMPLING, RAS_2CYC_SPRITE_TLUT, COMB_TEX0_ENV, TEXFMT_CI4, DBP_AA_SPRITE, 1)             \
	X(11, RAS_MULTI_CYCLE | RAS_COPY | RAS_FILL, RAS_MULTI_CYCLE, COMB_ANY, TEXFMT_ANY, DBP_ANY, 2)
constexpr uint32_t kRasterClasses = 3;  // the generic variant (id 0) is in the top class
// Variant 1 is the FILL-mode row.  A pass whose every primitive uses it reads
// nothing from the framebuffer and is rasterized by its own light kernel.
constexpr uint32_t kFillVariant = 1;
#ifndef RDP_FILL_TILES
#define RDP_FILL_TILES 4
#endif
constexpr uint32_t kFillTilesPerItem = RDP_FILL_TILES;  // horizontally adjacent tiles per k_raster_fill wave

// The format every tile a primitive can sample shares, or TEXFMT_ANY.
// Without texture LOD the sampled tiles are base (texel 0, pipelined texel 1)
// and base + 1 (texel 1 in two-cycle mode); with it, compute_lod
// (texture.h:925-1003) selects tiles base .. base + max_level, one further
// with detail LOD.
RDP_HD inline int static_texture_format(uint32_t ras_flags, const TileDesc *tiles, uint32_t setup_tile)
{
	const uint32_t base = setup_tile & 7, max_level = setup_tile >> 3;
	uint32_t range = (ras_flags & RAS_TEX_LOD) ? max_level + ((ras_flags & RAS_DETAIL_LOD) ? 1u : 0u)
	                                           : ((ras_flags & RAS_USES_TEXEL1) ? 1u : 0u);
	if (range > 7) range = 7;
	for (uint32_t i = 1; i <= range; i++) {
		const TileDesc &t = tiles[(base + i) & 7];
		if (t.fmt != tiles[base].fmt || t.size != tiles[base].size) return TEXFMT_ANY;
	}
	return (int)tiles[base].fmt * 4 + (int)tiles[base].size;
}

RDP_HD inline uint32_t select_raster_variant(uint32_t ras_flags, const uint8_t comb[2][8], const TileDesc *tiles,
                                             uint32_t setup_tile, uint32_t db_flags, const uint8_t blend[2][4],
                                             uint32_t z_mode, uint32_t coverage_mode)
{
	const int tf = static_texture_format(ras_flags, tiles, setup_tile);
#define RDP_MATCH_VARIANT(id, known, value, kcomb, texfmt, dbp, cls)                                       \
	if ((ras_flags & (known)) == (value) && ((texfmt) == TEXFMT_ANY || (texfmt) == tf) &&                \
	    depth_blend_profile_matches((dbp), db_flags, blend, z_mode, coverage_mode)) {                      \
		bool ok = true;                                                                                     \
		if ((kcomb) != COMB_ANY)                                                                            \
			for (int i = 0; i < 8; i++) {                                                                   \
				ok = ok && comb[1][i] == kKnownCombiners[(kcomb)][1][i];                                     \
				if (kKnownCombinerTwoCycle[(kcomb)]) ok = ok && comb[0][i] == kKnownCombiners[(kcomb)][0][i]; \
			}                                                                                               \
		if (ok) return (id);                                                                                \
	}
	RDP_RASTER_VARIANTS(RDP_MATCH_VARIANT)
#undef RDP_MATCH_VARIANT
	return 0;
}

RDP_HD inline uint32_t raster_variant_class(uint32_t variant)
{
#define RDP_VARIANT_CLASS(id, known, value, comb, texfmt, dbp, cls) \
	if (variant == (id)) return (cls);
	RDP_RASTER_VARIANTS(RDP_VARIANT_CLASS)
#undef RDP_VARIANT_CLASS
	return kRasterClasses - 1;
}

// Resolved draw state, snapshotted by the host whenever any state command
// ran since the previous primitive.  Combiner constants are already resolved
// per selector the way Renderer::build_combiner_constants does it.
struct DrawState {
	uint32_t scissor[4];              // xlo, ylo, xhi, yhi (10.2)
	uint32_t ras_flags;               // normalised static flags incl. USES_* / NEED_NOISE*
	uint32_t dither;                  // rgb mode << 2 | alpha mode
	uint32_t db_flags;
	uint32_t coverage_mode;
	uint32_t z_mode;
	uint8_t comb[2][8];               // per cycle: rgb muladd, mulsub, mul, add, alpha muladd, mulsub, mul, add
	uint8_t blend[2][4];              // per cycle: 1a, 1b, 2a, 2b
	uint8_t constants[2][4][4];       // per cycle: muladd, mulsub, mul, add constant RGBA
	uint8_t fog[4], blendc[4];
	uint32_t fill_color;
	int32_t factors[4];               // convert K0..K3 as 2*sext9(k)+1
	uint32_t min_lod;
	uint32_t use_prim_depth;
	int32_t prim_depth;               // already << 16
	uint32_t prim_dz;
	uint32_t quirks;
	uint32_t raw_ras_flags;           // static flags before normalisation (interlace bits)
	uint32_t pad[2];
	TileDesc tiles[8];
};

// One primitive of a pass.
struct PrimRecord {
	uint32_t word_offset;   // into the pass word buffer
	uint16_t state_index;   // into the pass draw-state table
	uint16_t tmem_instance; // TMEM snapshot index (0 = state at pass start)
	int32_t ylo, yhi;       // active scanline range (yhi < ylo: nothing to draw)
	uint32_t span_offset;   // first SpanSetup record (lines ylo .. yhi+1)
	uint32_t tilerow_offset;// first tile-row range record (rows ylo>>3 .. yhi>>3)
	uint32_t prim_index;    // stream-global primitive counter (noise seed)
	uint8_t opcode;
	uint8_t raster_class;   // occupancy class of the primitive's raster variant
	uint16_t variant;       // raster variant (select_raster_variant)
};
static_assert(sizeof(PrimRecord) == 32, "PrimRecord layout");

// Decoded per-primitive setup written by the setup kernel.
struct PrimSetup {
	int32_t xh, xm, xl;
	int32_t dxhdy, dxmdy, dxldy;
	int32_t yh, ym, yl;
	uint32_t flags;
	uint32_t tile;
	uint32_t state_index;
	int32_t rgba[4], drgba_dx[4], drgba_de[4], drgba_dy[4];
	int32_t stzw[4], dstzw_dx[4], dstzw_de[4], dstzw_dy[4];
	int32_t dz, dz_compressed;
	uint32_t tmem_instance;
	uint32_t prim_index;
	int32_t ylo, yhi;
	uint32_t span_offset;
	uint32_t tilerow_offset;
	uint32_t scissor[4];    // the draw state's scissor (xlo, ylo, xhi, yhi in 10.2)
};
static_assert(sizeof(PrimSetup) == 224, "PrimSetup layout");

// Everything the per-tile kernel needs about one primitive, 256 bytes, read
// with scalar loads: the per-pixel attribute gradients, the span range and
// the resolved draw state of DrawState minus the texture tiles.
struct RasterPrim {
	int32_t drgba_dx[4], drgba_dy[4];
	int32_t dstzw_dx[4], dstzw_dy[4];
	uint32_t flags;          // TRI_* flags
	uint32_t tile;           // tile index | max level << 3
	int32_t ylo, yhi;
	uint32_t span_offset;
	uint32_t tmem_instance;
	uint32_t prim_index;
	uint32_t state_index;
	int32_t dz, dz_compressed;
	uint32_t ras_flags;
	uint32_t dither;
	uint32_t db_flags;
	uint32_t coverage_mode;
	uint32_t z_mode;
	uint32_t fill_color;
	uint32_t min_lod;
	uint32_t variant;        // raster code variant (select_raster_variant)
	uint32_t pad0[2];
	uint8_t comb[2][8];
	uint8_t blend[2][4];
	uint8_t constants[2][4][4];
	uint8_t fog[4], blendc[4];
	int32_t factors[4];
	uint32_t pad1[8];
};
static_assert(sizeof(RasterPrim) == 256, "RasterPrim layout");

// One scanline of one primitive (span_setup.comp output), 64 bytes.
struct SpanSetup {
	int32_t rgba[4];
	int32_t stzw[4];
	uint16_t xleft[4];
	uint16_t xright[4];
	int32_t base_x;
	int32_t start_x;
	int32_t end_x;
	int16_t lodlength;
	uint16_t valid;
};
static_assert(sizeof(SpanSetup) == 64, "SpanSetup layout");

// TMEM upload descriptor (rdp_data_structures.hpp:182-202), with the float
// reciprocals replaced by the integer divisors they stand for.
struct UploadInfo {
	int32_t width, height;
	int32_t min_line_stride, max_line_stride;
	int32_t vram_addr, vram_width, vram_size, vram_effective_width;
	int32_t tmem_offset, tmem_stride_words, tmem_size, tmem_fmt;
	int32_t mode, dxt;
	int32_t pad[2];
};
static_assert(sizeof(UploadInfo) == 64, "UploadInfo layout");

// Per-pass device descriptor.  All pointers are device addresses.
struct PassDesc {
	// inputs
	const uint32_t *words;
	const DrawState *states;
	const PrimRecord *prims;
	const UploadInfo *uploads;
	// per-pass scratch / outputs
	RasterPrim *rprims;
	SpanSetup *spans;
	uint32_t *tilerows;        // per (prim, tile row): lo tile x | hi tile x << 16, 0xffff lo = empty
	uint64_t *bin_fine;        // [tile][group] primitive masks
	uint32_t *tmem_blocks;     // [block of 32 uploads][2048]: first write in block << 16 | value after it
	uint16_t *tmem_instances;  // (num_uploads + 1) x 2048 halfwords
	PrimSetup *psetups;        // decoded edge and attribute setup per primitive
	uint16_t *tilerow_prim;    // per tile-row entry: the primitive it belongs to
	// stream memory
	uint8_t *rdram;
	uint8_t *hidden;
	uint16_t *tmem;            // persistent TMEM of the stream (2048 halfwords)
	uint8_t *wmask;            // host-coherent streams: 0xff over every RDRAM byte the raster writes (else null)
	uint32_t rdram_mask;       // rdram_size - 1
	uint32_t num_prims;
	uint32_t num_uploads;
	uint32_t fb_addr, fb_width, fb_height, fb_fmt, depth_addr;
	uint32_t tiles_x, tiles_y;
	uint32_t num_spans;
	uint32_t num_tilerows;
	uint32_t num_groups;       // ceil(num_prims / 64) rounded up to a multiple of 8
	// Upscaled rendering.  scale_log2 > 0: the pass renders at 2^scale_log2
	// times the resolution (fb_width, fb_height, tiles_* and the primitives'
	// lines are of that size) into the sample planes at rdram / hidden.
	uint32_t scale_log2;
	// The domain of a native pass (k_domain): planes, the reference copy and
	// the region height; domain_log2 is the factor of the domain.
	uint8_t *up_rdram, *up_hidden, *up_ref;
	uint32_t domain_height, domain_log2;
	// Super-sampled readback: per native pixel of the region, bit 0 colour
	// and bit 1 depth written by any sample; DOMAIN_* flags.
	uint32_t *up_wmask;
	uint32_t domain_flags;
	uint8_t *resolve_rdram;  // an upscaled pass with DOMAIN_SSAA_RESOLVE: native RDRAM
};

enum : uint32_t {
	DOMAIN_SSAA_CLEAR = 1u << 0,    // the pre update also clears the write masks of the region
	DOMAIN_SSAA_RESOLVE = 1u << 1,  // the upscaled raster records write masks; the pass is resolved
	DOMAIN_SSAA_DITHER = 1u << 2,   // ordered dither in the 16-bit resolve
};

}  // namespace rdp_mi

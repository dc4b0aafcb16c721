// Host front end: RDP command decode, state tracking and render-pass building.
//
// Mirrors the sequential half of the reference's CommandProcessor and
// Renderer (parallel-rdp/rdp_device.cpp:268-959, rdp_renderer.cpp:679-1590,
// 3051-3486): state commands update register state, draw commands are
// recorded into the open render pass together with a snapshot of the
// resolved draw state, and TMEM loads are planned into UploadInfo records.
// Triangle decode and everything per pixel run on the GPU.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "rdp_types.hpp"

namespace rdp_mi {

// Validation errors that would hang a real RDP (rdp_renderer.hpp:75-95).
enum ValidationError : uint32_t {
	VALIDATION_FILL_4BPP = 0,
	VALIDATION_FILL_DEPTH_TEST = 1,
	VALIDATION_FILL_DEPTH_WRITE = 2,
	VALIDATION_FILL_IMAGE_READ = 3,
	VALIDATION_COPY_32BPP = 4,
	VALIDATION_LOAD_TILE_4BPP = 5,
	VALIDATION_MULTILINE_TLUT = 6,
};

// A point inside an open pass where the reference's renderer ends a render
// pass but this front end kept going (the pass boundary is not observable
// there).  Enough is recorded to cut the pass at that point later.
struct PassMark {
	uint32_t prims, uploads, words, states, spans, tilerows, fb_height;
	bool color_written, depth_written;
	uint32_t domain_height;  // upscaled domain rows of the reference passes up to here
};

struct HostPass {
	std::vector<uint32_t> words;
	std::vector<DrawState> states;
	std::vector<PrimRecord> prims;
	std::vector<UploadInfo> uploads;
	std::vector<PassMark> marks;
	uint32_t fb_addr = 0, fb_width = 0, fb_fmt = 0, depth_addr = 0, fb_height = 0;
	uint32_t num_spans = 0, num_tilerows = 0;
	uint32_t raster_class = 0;  // highest raster occupancy class of its primitives
	// Upscaled rendering.  A native pass that renders: domain_pre runs the
	// domain's pre update over fb_width x domain_height before its raster,
	// domain_post the post update after it.  scale_log2 > 0: the pass is the
	// upscaled copy of the native pass before it (fb_width, fb_height and
	// the primitives' lines already scaled), rendering into the sample planes
	// with that pass's native_uploads + 1 TMEM snapshots.
	bool domain_pre = false, domain_post = false;
	uint32_t domain_height = 0;
	uint32_t scale_log2 = 0, native_uploads = 0;
	// Super-sampled readback: the native pass renders nothing and clears the
	// per-pixel write masks with its pre update (ssaa_clear); its upscaled
	// copy records which pixels it writes and resolves them into native
	// RDRAM over fb_width / 2^scale_log2 x domain_height (ssaa_resolve).
	bool ssaa_clear = false, ssaa_resolve = false;
	bool empty() const { return prims.empty() && uploads.empty(); }
};

// What the reference's Renderer tracks about its open render pass to decide
// when to end it (Renderer::need_flush and tmem_upload_needs_flush,
// rdp_renderer.cpp:1464-1590, 3051-3097, 3484-3485; the state caches are
// StateCache<T, N>, rdp_data_structures.hpp:256-312).
struct RefPassState {
	struct StaticKey {
		uint8_t comb[2][8];
		uint32_t flags, dither, texture_size, texture_fmt;
	};
	struct DepthBlendKey {
		uint8_t blend[2][4];
		uint32_t flags, coverage_mode, z_mode;
	};
	std::vector<StaticKey> statics;
	std::vector<DepthBlendKey> depth_blends;
	std::vector<TileDesc> tiles;
	uint32_t prims = 0, span_jobs = 0, uploads = 0, max_shaded_tiles = 0, deduced_height = 0;
	bool color_write_pending = false, depth_write_pending = false;
	// The caches already hold the keys of a draw made at state generation
	// key_gen with this setup tile (all-ones: none since the reset).
	uint64_t key_gen = ~0ull;
	unsigned key_setup_tile = 0;
	void reset();
};

class Frontend {
public:
	explicit Frontend(uint32_t rdram_size);

	// Number of words of the command that starts with this word (0: unknown).
	static unsigned command_length(uint32_t first_word);

	// Execute one complete command.
	void command(const uint32_t *words);

	// Close the open pass (if it holds any work).
	void flush();

	// Completed passes, in order; the caller takes ownership.
	std::vector<HostPass> &completed() { return done; }

	bool has_pending_work() const { return !done.empty() || !cur.empty(); }

	void set_validation_callback(std::function<void(uint32_t, const char *)> cb) { validation = std::move(cb); }

	uint64_t primitive_counter() const { return prim_counter; }

	// Largest pass size the front end will build (tunable for tests).
	uint32_t max_prims_per_pass = kMaxPrimsPerPass;
	uint32_t max_spans_per_pass = 1u << 20;

	// When set, receives the primitive counter at the end of every render
	// pass of the reference's renderer that held a draw (for tests).
	std::vector<uint64_t> *pass_end_log = nullptr;

	// log2 of the upscaling factor (0: native only) and super-sampled
	// readback; set before the first command.
	uint32_t scale_log2 = 0;
	bool ssaa = false;

private:
	uint32_t rdram_size;

	// Register state (rdp_device.hpp:236-246, rdp_renderer.hpp:294-312).
	uint32_t scissor[4] = {};
	uint32_t ras_flags = 0, dither = 0;
	uint8_t comb[2][8] = {};
	uint32_t db_flags = 0;
	uint8_t blend[2][4] = {};
	uint32_t coverage_mode = 0, z_mode = 0;
	struct { uint32_t addr, width, fmt, size; } teximg = {};
	TileDesc tiles[8] = {};
	uint32_t blend_color = 0, fog_color = 0, env_color = 0, prim_color = 0, fill_color = 0;
	uint32_t min_level = 0, prim_lod_frac = 0;
	int32_t prim_depth = 0;
	uint32_t prim_dz = 0;
	bool use_prim_depth = false;
	int32_t convert[6] = {};
	uint32_t key_width[3] = {}, key_center[3] = {}, key_scale[3] = {};
	uint32_t fb_addr = 0, fb_width = 0, fb_fmt = 0, depth_addr = 0;
	uint32_t quirks = 1;

	uint64_t prim_counter = 0;
	bool state_dirty = true;
	uint64_t state_gen = 0;  // bumped by every command that is not a draw
	// Raster class of the last (draw state, setup tile) pair.
	uint64_t class_state_serial = 0, state_serial = 1;
	uint32_t class_setup_tile = 0, class_value = 0, variant_value = 0;
	bool color_write_pending = false, depth_write_pending = false;

	HostPass cur;
	// Upscaled domain rows of the reference passes that ended inside cur.
	uint32_t domain_height = 0;
	std::vector<HostPass> done;
	RefPassState ref;
	std::function<void(uint32_t, const char *)> validation;

	void draw(unsigned op, const uint32_t *words, unsigned len);
	void resolve_state(DrawState &d) const;
	void load(uint32_t mode, const uint32_t *words);
	void load_tile_iteration(uint32_t tile, uint32_t mode, uint32_t slo, uint32_t tlo, uint32_t shi, uint32_t thi,
	                         uint32_t tmem_offset);
	bool load_reads_marked_writes(uint32_t mode, uint32_t slo, uint32_t tlo, uint32_t shi, uint32_t thi) const;
	bool ref_upload_needs_flush(uint32_t addr) const;
	bool ref_account_draw(unsigned op, const uint32_t *words);
	void add_ref_keys(unsigned setup_tile);
	void ref_pass_end();
	void split_at_last_mark();
	void split_or_flush();
	HostPass upscaled_copy(const HostPass &native) const;
	uint32_t ref_domain_height() const;
	void complete_pass(HostPass &&p, uint32_t domain_height);
	void validate_draw() const;
	void report(uint32_t err, const char *msg) const
	{
		if (validation) validation(err, msg);
	}
};

}  // namespace rdp_mi
